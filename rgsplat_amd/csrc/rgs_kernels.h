// rgs_kernels.h -- launchers of the gfx950 kernels (internal to librgsplat_hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stddef.h>
#include <stdint.h>

#include "rgs_math.h"

namespace rgs {

// Per (view, Gaussian) projection results at index vi = v * cap + i.
struct SplatArrays {
  float4* xy_ab;    // sx, sy, conic a, conic b
  float4* c_op_rg;  // conic c, opacity, red, green
  float* col_b;     // blue
  uint32_t* cnt;    // bins touched (0 when culled)
  uint2* rect;      // tiles: (tx0 | ty0 << 16, tx1 | ty1 << 16)
  float* reach;     // R: alpha >= 1/255 needs conic form <= R (+inf: no bound)
  float* depth;     // camera-space depth (visible splats)
};

// Status bits written by kernels into a device word.
// ST_NEED_BACK: a step run without the depth cut's back lists flagged a bin
// (its results are void; the parameter update is skipped).
enum : uint32_t { ST_NAN_DEPTH = 1u, ST_PAIR_OVERFLOW = 2u, ST_NEED_BACK = 4u, ST_SCAN_STALL = 8u };

// ---- projection (k_project.hip)
// Projects n Gaussians (SoA params P[59][cap]) into n_views views. Writes the
// splat arrays and the depth-sort input (key = depth bits, 0xffffffff when
// culled; value = v * cap + i) at dense index v * n + i. tight != 0: each
// splat's tile rect is the reference's 3-sigma rect cut down to the tiles its
// alpha-reach ellipse (alpha >= 1/255) can reach; the dropped (tile, splat)
// pairs change no pixel's blend, only the list indices. The reference
// composites a tile (rgb = C + T bg) iff its own list is non-empty, which a
// tight or binned list no longer tells: with cover non-null, every tile of a
// splat's reference rect gets cover[tile] = 1. Rects are in tiles; counts
// and pair keys in bins of 2^bin_shift x 2^bin_shift tiles (CamDev), the
// bins the rect meets.
void launch_project(const float* P, int cap, int n, const CamDev* cams, int n_views, SplatArrays S,
                    uint32_t* sort_keys, uint32_t* sort_vals, uint32_t* status, int tight, uint8_t* cover,
                    hipStream_t st);

// Device expf over n values (diagnostic).
void launch_eval_expf(const float* x, float* out, size_t n, hipStream_t st);

// ---- sorting and binning (k_sort.hip)
struct RadixScratch {
  uint32_t* hist;       // 256 * nb_max, hist[d * nb_max + b]
  uint32_t* row_total;  // 256
  int nb_max;           // >= radix_nb(n_max), a multiple of 256
  // optional: single-pass digit offsets (radix_os_words(nb_max) words) and
  // the status word that flags a look-back that gave up (ST_SCAN_STALL);
  // without them each pass is a histogram, a row scan and a scatter
  uint32_t* os = nullptr;
  uint32_t* status = nullptr;
  bool os_zeroed = false;  // the caller zeroed os's first radix_os_used_words words
};
constexpr int RS_TILE = 4096;
inline int radix_nb(size_t n_max) { return (int)((n_max + RS_TILE - 1) / RS_TILE); }
inline int radix_nb_max(size_t n_max) { return (radix_nb(n_max) + 255) / 256 * 256; }
// Single-pass digit offsets: the digit totals of up to OS_PASSES passes,
// then per pass OS_HEAD words (the first a block ticket) and 256 words per
// block (bits 30-31 a flag, bits 0-29 a count: sorts of fewer than 2^30
// keys).
constexpr int OS_PASSES = 4;
constexpr int OS_HEAD = 64;
inline size_t radix_os_pass_words(int nb) { return (size_t)OS_HEAD + (size_t)256 * (size_t)nb; }
inline size_t radix_os_words(int nb_max) { return (size_t)256 * OS_PASSES + OS_PASSES * radix_os_pass_words(nb_max); }
// The words a sort of at most nb tiles in `passes` passes uses.
inline size_t radix_os_used_words(int nb, int passes) {
  return (size_t)256 * OS_PASSES + (size_t)passes * radix_os_pass_words(nb);
}

// Logical block ids for per-radix-tile kernels. The hardware deals blocks to
// the 8 XCDs round-robin; this mapping gives each XCD runs of XG consecutive
// logical blocks, so the XG histogram columns that share a cache line
// (hist[d * nb_max + b .. b + XG)) and the neighbouring digit runs a scatter
// writes are merged in one XCD's L2 instead of being written back in
// fragments by all eight. The grid must be xcd_grid(n) blocks (a multiple of
// 8 * XG); kernels exit for logical ids past their work.
constexpr uint32_t XG = 32;
__device__ __forceinline__ uint32_t xcd_chunked(uint32_t b) {
  const uint32_t x = b & 7u, k = b >> 3;
  return (((k / XG) << 3) + x) * XG + (k % XG);
}
inline unsigned xcd_grid(size_t nb) { return (unsigned)((nb + 8 * XG - 1) / (8 * XG) * (8 * XG)); }

// Tile ranges produced by the last pass of the tile sort: for every key k
// present, rstart[k] / rend[k] = first / one-past-last position of k in the
// sorted order. Keys absent from the input keep the init values written by
// the sort (rstart 0xffffffff, rend 0), i.e. an empty range end < start.
struct TileRanges {
  uint32_t* rstart;
  uint32_t* rend;
  uint32_t n_keys;
  uint32_t prefilled = 0;  // nonzero: rstart/rend already hold ~0u / 0 (the caller filled them)
};

// Digit width of the LSD radix sorts over `bits` key bits: the fewest passes
// of at most 8 bits, equally wide (a 13-bit key: 7 + 6). A narrower digit
// gives each radix tile fewer, longer runs to write.
inline int radix_digit_bits(int bits) {
  const int passes = (bits + 7) / 8;
  return passes ? (bits + passes - 1) / passes : 8;
}
// Whether the pair generation and tile sort keep the bin keys as 16 bits
// (at most 2^16 bins).
inline bool tile_keys_16(int bits) { return bits <= 16; }
// Stable LSD radix sort of (key, value) pairs over bits [0, bits); count read
// from device memory (*n_dev <= n_max). Ping-pongs between the two buffer
// pairs; returns the index (0/1) of the buffer pair holding the result.
int radix_sort_pairs(uint32_t* keys[2], uint32_t* vals[2], const uint32_t* n_dev, size_t n_max, int bits,
                     RadixScratch rs, hipStream_t st);

constexpr int SCAN_TILE = 2048;
// Total written by a 32-bit scan whose true total is 2^32 - 1 or more.
constexpr uint32_t SCAN_SATURATED = 0xffffffffu;
inline int scan_nb(size_t n) { return (int)((n + SCAN_TILE - 1) / SCAN_TILE); }
// Exclusive scan of cnt[vals[r]] over r < n (pair offsets in depth order).
// block_sums needs scan_nb(n) + 1 entries; total written to *total.
void scan_gathered_counts(const uint32_t* vals, const uint32_t* cnt, size_t n, uint32_t* out,
                          uint32_t* block_sums, uint32_t* total, hipStream_t st);
// Exclusive scan of a u32 array (in may alias out).
void scan_u32(const uint32_t* in, size_t n, uint32_t* out, uint32_t* block_sums, uint32_t* total, hipStream_t st);
// Exclusive scan of a u64 array.
void scan_u64(const uint64_t* in, size_t n, uint64_t* out, uint64_t* block_sums, uint64_t* total, hipStream_t st);

// Bin rect of a splat's tile rect (bins of 2^sh x 2^sh tiles).
__device__ __forceinline__ uint2 bin_rect(uint2 r, int sh) {
  const uint32_t x0 = r.x & 0xffffu, y0 = r.x >> 16, x1 = r.y & 0xffffu, y1 = r.y >> 16;
  if (!sh || x1 <= x0 || y1 <= y0) return r;
  return make_uint2((x0 >> sh) | ((y0 >> sh) << 16), (((x1 - 1) >> sh) + 1) | ((((y1 - 1) >> sh) + 1) << 16));
}

// Writes the (tile, splat) pairs in depth order, one radix tile (RS_TILE
// pairs) per block, and the first radix pass's per-tile digit histogram
// (hist[d * nb_max + b] for the low 8 key bits). K = min(total, k_cap) goes to
// *k_out; ST_PAIR_OVERFLOW is raised when total > k_cap. first_splat needs
// radix_nb(k_cap) entries. n_dev (optional): only the first *n_dev sorted
// entries are valid, and first_splat was written by scan_pair_offsets
// (without n_dev it is made here from the offsets of all n_sorted entries).
void launch_gen_pairs(const uint32_t* sorted_vals, size_t n_sorted, const uint32_t* offsets, const uint32_t* total,
                      const SplatArrays S, int cap, const CamDev* cams, uint32_t* pair_keys, uint32_t* pair_vals,
                      size_t k_cap, uint32_t* k_out, uint32_t* status, uint32_t* first_splat, RadixScratch rs,
                      int key_bits, hipStream_t st, const uint32_t* n_dev = nullptr);
// Tile sort of the pairs from launch_gen_pairs (whose first-pass histogram is
// already written; the keys are 16-bit when tile_keys_16(bits), in the
// same buffers). Only the values of the result are written (the keys are
// implied by the ranges); the ranges come out of the last pass.
int radix_sort_tiles(uint32_t* keys[2], uint32_t* vals[2], const uint32_t* n_dev, size_t n_max, int bits,
                     RadixScratch rs, TileRanges tr, hipStream_t st);

// Depth-cut binning of the training step. The pairs of the splats in front
// of a cut depth are generated, sorted and rasterised first; a tile whose
// front list ends before all its pixels saturate is flagged, gets the pairs
// of the splats behind the cut (in depth order, after its front list) and
// the raster resumes it. A tile's front and back lists concatenated are its
// whole list, so images, last-contributor indices and gradients are those of
// whole lists; the cut only decides how many pairs are sorted. The cut of a
// step is the depth at which the previous step's tiles saturated (the
// deepest such splat, plus 5%).
struct CutState {       // persistent device words
  uint32_t key;         // f2u(depth) of the next cut; 0xffffffff: no cut
  uint32_t zneed;       // max f2u(depth) of a splat at which a tile saturated
  uint32_t mode;        // 0 adaptive, 1 no cut, 2 fixed key
  uint32_t pad;
};
// The training step's partition at the cut and its pair offsets, each a
// single-pass (look-back) scan. state: lookback_words(n) u64 words, zero
// when the call starts; status gets ST_SCAN_STALL if a look-back gave up.
constexpr int LB_HEAD = 1;
constexpr int LB_TILE = 2048;  // entries per look-back tile (at most)
inline int lookback_tiles(size_t n) { return (int)((n + LB_TILE - 1) / LB_TILE); }
inline size_t lookback_words(size_t n) { return (size_t)LB_HEAD + (size_t)std::max(lookback_tiles(n), 1); }
// Stable partition of the projected (key, value) entries at the cut into
// front (kf, vf: *n_front entries, keys[r] < cs->key) and back (kb, vb:
// *n_back = n - *n_front entries; may be null: nothing written) arrays;
// *cut_on = 1 when a cut is set (a front list may then be partial); the
// back entries that are not culled are added to *n_back_visible.
void partition_at_cut(const uint32_t* keys, const uint32_t* vals, size_t n, const CutState* cs, uint32_t* kf,
                      uint32_t* vf, uint32_t* kb, uint32_t* vb, uint32_t* n_front, uint32_t* n_back, uint32_t* cut_on,
                      uint32_t* n_back_visible, uint64_t* state, uint32_t* status, hipStream_t st);
// Exclusive scan of cnt[vals[r]] over the first *n_dev of n_max entries
// (pair offsets in depth order); the total goes to *total. first_splat[b]
// gets the entry holding pair b * RS_TILE, for every such pair below the
// total and k_cap.
void scan_pair_offsets(const uint32_t* vals, const uint32_t* cnt, size_t n_max, const uint32_t* n_dev,
                       uint32_t* out, uint32_t* total, uint64_t* state, uint32_t* status, hipStream_t st,
                       uint32_t* first_splat, size_t k_cap);
// Exclusive count of the back entries (first *n_dev) whose rect meets a
// flagged bin (all 0 when no bin is flagged).
void scan_touch_flagged(const uint32_t* vals, size_t n, const uint32_t* n_dev, const uint32_t* n_flagged,
                        const SplatArrays S, int cap, const CamDev* cams, const uint8_t* flag, uint32_t* out,
                        uint32_t* block_sums, uint32_t* total, hipStream_t st);
// Keeps the entries a 0/1 scan (off, *total) selected, in order.
void launch_compact(const uint32_t* keys, const uint32_t* vals, size_t n, const uint32_t* off, const uint32_t* total,
                    uint32_t* kc, uint32_t* vc, hipStream_t st);
// Back-pair count of the first *n_dev sorted back entries: flagged bins in
// the splat's bin rect (0 past them and when nothing is flagged).
void launch_back_count(const uint32_t* sorted_vals, size_t n, const uint32_t* n_dev, const uint32_t* n_flagged,
                       const SplatArrays S, int cap, const CamDev* cams, const uint8_t* flag, uint32_t* count,
                       hipStream_t st);
// Writes the back pairs at the exclusive offsets of those counts (total in
// *total); K = min(total, k_cap) goes to *k_out, ST_PAIR_OVERFLOW when
// total > k_cap.
void launch_back_gen(const uint32_t* sorted_vals, size_t n, const uint32_t* n_dev, const uint32_t* offsets,
                     const uint32_t* total, const uint32_t* n_flagged, const SplatArrays S, int cap,
                     const CamDev* cams, const uint8_t* flag, uint32_t* pair_keys, uint32_t* pair_vals, size_t k_cap,
                     uint32_t* k_out, uint32_t* status, hipStream_t st);
// Raises ST_NEED_BACK when *n_flagged != 0.
void launch_need_back(const uint32_t* n_flagged, uint32_t* status, hipStream_t st);
// Next cut from this step's saturation depths (mode 0); clears zneed.
void launch_cut_end(CutState* cs, hipStream_t st);
// Tile sort whose first-pass histogram is not yet computed.
int radix_sort_tiles_full(uint32_t* keys[2], uint32_t* vals[2], const uint32_t* n_dev, size_t n_max, int bits,
                          RadixScratch rs, TileRanges tr, hipStream_t st);

// ---- rasterization (k_raster.hip)
struct RasterIn {
  const uint32_t* rstart;
  const uint32_t* rend;
  const uint32_t* pair_vals;
  const int32_t* tile_order;  // schedule slot -> global tile id
  SplatArrays S;
  const CamDev* cams;
  int n_views;
  int sparse_lanes;  // backward: at most this many contributing lanes -> LDS atomics
  int batch;         // binned training raster: splats staged per batch (32, 64, 128; 0 = 64)
  size_t acc_stride;  // backward accumulators: value k of entry vi at accum[k * acc_stride + vi]
  int total_bins;   // workgroups: one per bin (2^bin_shift x 2^bin_shift tiles)
  int bin_shift;
  int fwd_tile_blocks;  // forward with bins: one workgroup per tile instead of per bin
  float bg[3];
  const uint8_t* cover;  // null, or per tile: composited by the reference (launch_project)
  // depth-cut binning (null when every list is whole)
  const uint32_t* cut_on;     // device word: 1 when a front list may be partial
  uint8_t* flag;              // per bin: a front list ended before all its pixels saturated
  uint8_t* tile_flag;         // per tile: kept its raw state for the resumed forward
  uint32_t* n_flagged;        // number of flagged workgroups
  uint32_t* zneed;            // CutState::zneed
  const uint32_t* bstart;     // back lists: ranges and values
  const uint32_t* bend;
  const uint32_t* back_vals;
  unsigned long long* counts; // launch_raster_counts only: 6 work counters
  // longest-first schedule of the backward (null: the static tile_order):
  // the forward writes each tile's cost (how far down its list it
  // composited); backward block b takes bin lpt_bins[b]
  uint32_t* tile_cost;
  const int32_t* lpt_bins;
  // persistent training raster (null: one workgroup per schedule slot): 8
  // per-XCD slot counters at queue[0], queue[32], ..., queue[224], zero at
  // launch; the training forward (mode 3) and the binned backward take
  // their bins from them
  uint32_t* queue;
};
// mode 0: training forward (rgb = C + T bg; final T; last contributor); with
//   cut_on set, a tile whose (front) list ends before all its pixels
//   saturate is flagged and keeps its raw state (rgb = C, T, last);
// mode 1: rasterizer.rs RGBA (rgba = bg + C, A = 1 - T; empty tiles left alone);
// mode 2: resumes the flagged tiles of mode 3 on their back lists;
// mode 3: mode 0 of the training step, alpha from the hardware exp2 with the
//   exact expf near the alpha thresholds (modes 0 and 1 use the exact expf).
void launch_raster_fwd(const RasterIn& in, int mode, float* rgb, float* final_t, int32_t* last, hipStream_t st);
// Accumulates the 9 per-splat screen-space gradients (dL/dcolour[3],
// dL/dopacity, dL/dconic[3], dL/dsx, dL/dsy) into accum[k * acc_stride + vi]:
// one row per value, so a splat's nine atomics land on nine different cache
// lines (L2 channels) instead of queueing on one.
void launch_raster_bwd(const RasterIn& in, const float* final_t, const int32_t* last, const float* dl_dpix,
                       float* accum, hipStream_t st);

// The backward's longest-first bin table from the tile costs of a forward:
// lpt_bins (total_bins entries), a permutation, costliest bin first.
void launch_lpt_order(const uint32_t* tile_cost, const CamDev* cams, int n_views, int total_bins, int shift,
                      int32_t* lpt_bins, hipStream_t st);

// The training step's forward (mode 3) and backward raster again, counting
// their own work into in.counts: [0] forward lane evaluations of a (pixel,
// splat) pair, [1] forward wave visits of a (splat, 4-row pixel group), [2]
// forward composites, [3]-[5] the same for the backward (contributing pairs
// in [5]). The forward rewrites its outputs with the same values (in.zneed
// and in.cut_on must be null); the backward writes no accumulator.
void launch_raster_counts(const RasterIn& in, float* rgb, float* final_t, int32_t* last, const float* dl_dpix,
                          hipStream_t st);

// out[0] += forward pair evaluations, out[1] += backward pair evaluations
// (list entries, of the lists the render used); one block per tile.
void launch_work_counts(const RasterIn& in, const float* final_t, const int32_t* last, unsigned long long* out,
                        int total_tiles, hipStream_t st);

// ---- loss (k_loss.hip)
// The loss sweeps each view in units of LSW output columns x LSG output rows
// (one workgroup per unit and colour channel).
constexpr int LSW = 54;
constexpr int LSG = 155;
RGS_HD int loss_units(int w, int h) { return ((w + LSW - 1) / LSW) * ((h + LSG - 1) / LSG); }
struct LossIn {
  const float* rgb;     // total_pix * 3
  const float* target;  // target store, RGBA; view v starts at pixel cams[v].tgt_base
  const CamDev* cams;   // null: one view, described by cam0
  int n_views;
  int total_lunits;     // loss units over all views
  size_t total_pix;
  float lambda;
  CamDev cam0;          // the view when cams is null (w, h, pix_base, tgt_base, lunit_base)
  // null, or the target store's statistics (launch_target_stats), planar
  // [channel][mu_y | E[y^2]][tgt_total]
  const float* tstats;
  size_t tgt_total;     // pixels of the target store
};
// dl: 3 * total_pix floats; partial: 2 * 3 * total_lunits floats (scratch);
// loss_out: n_views floats. Allocates nothing.
void launch_loss(const LossIn& in, float* dl, float* partial, float* loss_out, hipStream_t st);
// The blurred target and squared target of in's views into tstats
// (6 * tgt_total floats), for later launch_loss calls on those targets.
void launch_target_stats(const LossIn& in, float* tstats, hipStream_t st);

// ---- optimizer and parameter kernels (k_optim.hip)
struct AdamConsts {
  float lr[5];  // position, rotation, scale, opacity, sh_dc
  float lr_sh_rest;
  float beta1, beta2, one_minus_beta1, one_minus_beta2, bc1, bc2, eps;
  float inv_bc1, inv_bc2;  // 1 / bc1, 1 / bc2 (the training step's Adam)
};
// AoS (Gaussian3D, 59 floats) <-> SoA (P[59][cap], Adam flat order).
void launch_aos_to_soa(const float* aos, size_t n, float* P, size_t cap, hipStream_t st);
void launch_soa_to_aos(const float* P, size_t cap, size_t n, float* aos, hipStream_t st);
// GaussianGrad AoS (60 floats, zero when culled) for one view.
void launch_geometry_aos(const float* P, int cap, int n, const CamDev* cams, int view, const uint32_t* cnt,
                         const float* accum, size_t acc_stride, float* grads, hipStream_t st);
// Adam on the reference layouts: AoS params/grads, row-major moments.
void launch_adam_aos(float* g, const float* grads, size_t n, float* m1, float* m2, float* accum, uint32_t* denom,
                     const AdamConsts& k, hipStream_t st);
// Per-Gaussian geometry backprop over all views + Adam on SoA buffers (one
// kernel, one thread per Gaussian). Skips everything when *status has
// ST_PAIR_OVERFLOW, ST_NAN_DEPTH or ST_NEED_BACK. Clears accum. grad_out
// (optional, 60 rows of cap floats): the step's gradient of each of the 59
// parameter rows, summed over the views, and the screen-gradient norm sum.
void launch_train_adam(float* P, float* M1, float* M2, int cap, int n, const CamDev* cams, int n_views,
                       const uint32_t* cnt, float* accum, size_t acc_stride, float* gaccum, uint32_t* gdenom,
                       const AdamConsts* k_dev, const uint32_t* status, hipStream_t st, float* grad_out = nullptr);
// Model files (k_io.hip). A PLY vertex property: byte offset in the record,
// byte width (1, 2, 4 or 8) and the SoA row it sets (-1: ignored).
struct PlyField {
  int32_t offset, width, row;
};
constexpr int PLY_FIELDS = 62;  // floats per record written by save_ply
void launch_ply_encode(const float* P, size_t cap, size_t n, float* rec, hipStream_t st);
// n records; those from n_body on are all zero bytes (not read from body).
void launch_ply_decode(const uint8_t* body, size_t n, size_t n_body, int rec_bytes, const PlyField* fields,
                       int n_fields, int big_endian, float* P, size_t cap, hipStream_t st);
// Tiled decoder for records of at most ply_decode_tiled_fits bytes;
// row_field[59]: per SoA row the last property that sets it (width 0: none).
int ply_decode_tiled_fits(int rec_bytes);
void launch_ply_decode_tiled(const uint8_t* body, size_t n, size_t n_body, int rec_bytes, const PlyField* row_field,
                             int big_endian, float* P, size_t cap, hipStream_t st);
void launch_splat_encode(const float* P, size_t cap, size_t n, void* out /* n x 32 bytes */, hipStream_t st);
// image 0.25's resize_exact with the Triangle filter, RGBA8 (k_io.hip):
// taps are (left, count) per output row (v) / column (h), weights K per row.
void launch_resize_rgba8(const uint8_t* src, uint32_t sw, uint32_t sh, uint8_t* dst, uint32_t dw, uint32_t dh,
                         const int2* vtaps, const float* vw, int vK, const int2* htaps, const float* hw, int hK,
                         float* tmp, hipStream_t st);
// Preview renderer (k_preview.hip): GpuRenderer's project and composite.
struct PreviewCam {
  float view[12];  // rows 0-2 of the world-to-camera matrix, row-major
  float fx, fy, cx, cy;
};
void launch_preview_project(const float* P, int cap, uint32_t n, const PreviewCam& cam, float* splats,
                            uint32_t* keys, uint32_t* vals, hipStream_t st);
void launch_preview_rects(const float* splats, const uint32_t* keys, const uint32_t* order, uint32_t n, int w, int h,
                          int4* rects, int4* trects, uint64_t* cnt, hipStream_t st);
// blk: (m + 255) / 256 + 1 entries of scratch
void launch_preview_pairs(const int4* trects, const uint64_t* off, uint32_t n, uint32_t m, int ntx, uint32_t* blk,
                          uint32_t* pk, uint32_t* pv, hipStream_t st);
void launch_preview_ranges(const uint32_t* pk, uint32_t m, uint32_t* rs, uint32_t* re, hipStream_t st);
void launch_preview_composite(const float* splats, const uint32_t* order, const int4* rects, const uint32_t* pv,
                              const uint32_t* rs, const uint32_t* re, int w, int h, int ntx, int nty, const float bg[3],
                              float* rgb, hipStream_t st);
// SfM initialisation (k_sfm.hip). status bit 1: a NaN coordinate or distance.
void launch_sfm_keys(const double* xyz, uint32_t n, uint32_t* klo, uint32_t* khi, uint32_t* idx, uint32_t* status,
                     hipStream_t st);
void launch_sfm_gather_hi(const uint32_t* khi, const uint32_t* order, uint32_t n, uint32_t* keys, hipStream_t st);
void launch_sfm_sorted_xyz(const double* xyz, const uint32_t* order, uint32_t n, double* sxyz, hipStream_t st);
void launch_sfm_scale(const double* sxyz, const uint32_t* order, uint32_t n, float* scale, uint32_t* status,
                      hipStream_t st);
void launch_sfm_gaussians(const double* xyz, const uint8_t* rgb, const float* scale, uint32_t n, const float* dc,
                          float opacity_logit, float* P, size_t cap, hipStream_t st);
// Colour conversions (k_io.hip). lut: 512 floats, [0, 256) sRGB -> linear of
// c / 255, [256, 512) c / 255.
void launch_rgba8_to_linear(const uint8_t* rgba8, size_t n_pixels, const float* lut, float* out, hipStream_t st);
void launch_to_u8_srgb(const float* rgba, size_t n_pixels, uint8_t* out, hipStream_t st);
// reset_opacities on SoA (row 10) or AoS (offset 3) params.
void launch_reset_opacity(float* p, size_t n, size_t stride, hipStream_t st);
// Densify/prune on SoA params: P -> P_out (compacted + children). Writes the
// new count to *n_out. Scratch sized for n Gaussians.
struct DensifyScratch {
  uint32_t* want;        // n
  uint32_t* cnt;         // n
  uint32_t* W;           // n
  uint64_t* packed;      // n
  uint64_t* prefix;      // n
  uint32_t* bsum32;      // scan_nb(n) + 1
  uint64_t* bsum64;      // scan_nb(n) + 1
  uint32_t* total32;     // 1
  uint64_t* total64;     // 1
};
// split_div: 0 for the reference's split (log_scale[0] of the children
// becomes ln(max(s0 / sqrt 2, 1e-6))); > 0 divides every scale by it.
void launch_densify(const float* P, float* P_out, int cap, int n, const float* gaccum, const uint32_t* gdenom,
                    float grad_thresh, float prune_thresh, uint64_t max_gaussians, float extent, DensifyScratch ds,
                    uint32_t* n_out, hipStream_t st, float split_div = 0.0f);
void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st);
// Up to 8 fills (p[s][0 .. n[s]) = v[s]) in one launch.
struct FillSegs {
  uint32_t* p[8];
  uint32_t n[8];
  uint32_t v[8];
  int count;
  void add(uint32_t* ptr, size_t len, uint32_t val) {
    p[count] = ptr;
    n[count] = (uint32_t)len;
    v[count] = val;
    ++count;
  }
};
void launch_fill_multi(const FillSegs& f, hipStream_t st);
// The step's end of the front raster without the back-list stage: flags the
// step as needing that stage when a bin was flagged, then sets the next
// step's cut (launch_need_back and launch_cut_end in one launch).
void launch_need_back_cut_end(const uint32_t* n_flagged, uint32_t* status, CutState* cs, hipStream_t st);
// b[i] = a[i] for n16 16-byte elements (the HBM copy probe).
void launch_copy16(const void* a, void* b, size_t n16, hipStream_t st);
void launch_fill_f32(float* p, size_t n, float v, hipStream_t st);
// Zero rows [r0, r1) of each of the 59 rows of a SoA moment buffer.
void launch_zero_rows(float* M, size_t cap, size_t r0, size_t r1, hipStream_t st);

// xcd-aware block remap: consecutive logical tiles share an XCD (and its L2).
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  const int nx = 8;
  int per = nblocks / nx;
  int rem = nblocks - per * nx;
  int x = b % nx;  // hardware dispatches block b to XCD b % 8
  int k = b / nx;
  // XCD x owns logical tiles [x*per + min(x, rem), ...) of length per + (x < rem)
  int base = x * per + (x < rem ? x : rem);
  int len = per + (x < rem ? 1 : 0);
  if (k < len) return base + k;
  return b;  // unreachable for a full grid
}

}  // namespace rgs

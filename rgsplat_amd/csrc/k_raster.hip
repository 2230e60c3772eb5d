// k_raster.hip -- front-to-back alpha blending and its analytic backward.
//
// One wavefront per 16x16 tile (the reference's binning tile,
// backward.rs:21): each lane owns four pixels of one column, one in each
// 4-row pixel group. A workgroup rasterises one bin of S x S tiles (S = 1:
// the reference's tiles, API renders; S = 2 in the training step), wave w
// the tile (w % S, w / S) of it. Bin lists are consumed in batches of 64
// splats staged once in LDS for all the bin's waves; a splat whose
// alpha-reach ellipse misses a pixel group of a wave's tile is skipped for
// that group with a wave-uniform branch. The workgroup stops as soon as every
// pixel of the bin has saturated (T < 1e-4). The exp2 table of the exact expf
// lives in LDS.
//
// Forward follows backward.rs:296-385 (mode 0) and rasterizer.rs:72-192
// (mode 1, RGBA). Backward follows backward.rs:391-512: each pixel walks its
// list from last_contrib down to 0, recovering the transmittance in front of
// each splat by division (no per-fragment storage); a lane first sums its
// four pixels' gradients for the splat, the wave then reduces the nine sums
// across lanes into LDS, and each bin adds the sum over its waves to the
// per-splat accumulators with one FP32 atomic per value.
#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#include "rgs_kernels.h"

namespace rgs {

constexpr int RT = 256;

__device__ __forceinline__ int view_of_tile(const CamDev* cams, int n_views, int tile) {
  int v = 0;
  while (v + 1 < n_views && cams[v + 1].tile_base <= tile) ++v;
  return v;
}
__device__ __forceinline__ int view_of_bin(const CamDev* cams, int n_views, int bin) {
  int v = 0;
  while (v + 1 < n_views && cams[v + 1].bin_base <= bin) ++v;
  return v;
}

// Whether tile (tx, ty) lies in a splat's tile rect. A bin's list holds the
// splats whose rect meets the bin; a wave takes only those whose rect holds
// its tile, so that it sees exactly the tile's own list (the reference
// composites a splat only into the tiles of its rect, even where its alpha
// reaches beyond).
__device__ __forceinline__ bool in_rect(uint2 r, int tx, int ty) {
  return (uint32_t)tx >= (r.x & 0xffffu) && (uint32_t)tx < (r.y & 0xffffu) && (uint32_t)ty >= (r.x >> 16) &&
         (uint32_t)ty < (r.y >> 16);
}

// The tile that wave w of a bin's workgroup rasterises.
struct BinTile {
  int v;       // view
  int tx, ty;  // tile coordinates in the view
  int tile;    // global tile id (meaningful when ok)
  bool ok;     // inside the view's tile grid (bins at the right and bottom
               // edges may be partial)
};
template <int S>
__device__ __forceinline__ BinTile bin_tile(const RasterIn& in, int bin, int w) {
  BinTile b;
  b.v = view_of_bin(in.cams, in.n_views, bin);
  const CamDev& c = in.cams[b.v];
  const int lb = bin - c.bin_base;
  b.tx = (lb % c.nbx) * S + (w % S);
  b.ty = (lb / c.nbx) * S + (w / S);
  b.ok = b.tx < c.ntx && b.ty < c.nty;
  b.tile = c.tile_base + b.ty * c.ntx + b.tx;
  return b;
}

// Bit q of the result: the ellipse conic(d) <= R of the splat reaches the
// pixel centres of group q (tile rows 4q..4q+3, all 16 columns). With the
// conic [[ca, cb], [cb, cc]], the ellipse's slice at row offset dy is centred
// at dx = -(cb / ca) dy with half-width sqrt(max(0, R - k dy^2) / ca),
// k = cc - cb^2 / ca; over the group's rows the union of slices lies within
// the centre range widened by the widest slice. The approximate rcp/sqrt and
// the extra 1e-4 relative + 0.01 px slack are far inside R's own margin.
// Non-finite inputs give "reaches".
__device__ __forceinline__ uint32_t wave_mask(float sx, float sy, float ca, float cb, float cc, float R, int tx,
                                              int ty) {
  if (!__builtin_isfinite(R)) return 0xfu;
  const float inv_a = __builtin_amdgcn_rcpf(ca);
  const float slope = -cb * inv_a;
  const float k = cc - cb * cb * inv_a;
  const float X0 = ((float)(tx * TILE) + 0.5f) - sx, X1 = ((float)(tx * TILE) + 15.5f) - sx;
  uint32_t m = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float Y0 = ((float)(ty * TILE + 4 * q) + 0.5f) - sy, Y1 = ((float)(ty * TILE + 4 * q) + 3.5f) - sy;
    float dmin = (Y0 <= 0.0f && Y1 >= 0.0f) ? 0.0f : fmin_(__builtin_fabsf(Y0), __builtin_fabsf(Y1));
    dmin = fmax_(dmin - 0.01f, 0.0f);
    const float rem = R - k * dmin * dmin;
    const float w = __builtin_amdgcn_sqrtf(fmax_(rem, 0.0f) * inv_a) * 1.0001f + 0.01f;
    const float c0 = slope * Y0, c1 = slope * Y1;
    const bool out = (rem < 0.0f) || (fmax_(c0, c1) + w < X0) || (fmin_(c0, c1) - w > X1);
    m |= out ? 0u : (1u << q);
  }
  return m;
}

__device__ __forceinline__ int wave_max_i(int v) {
  v = max(v, __shfl_xor(v, 1, 64));
  v = max(v, __shfl_xor(v, 2, 64));
  v = max(v, __shfl_xor(v, 4, 64));
  v = max(v, __shfl_xor(v, 8, 64));
  v = max(v, __shfl_xor(v, 16, 64));
  v = max(v, __shfl_xor(v, 32, 64));
  return v;
}

// The raster takes the quadratic form qf = a dx^2 + 2 b dx dy + c dy^2 of the
// reference's power = -0.5 qf (summed in the reference's order). The
// reference skips power > 0, which holds exactly when qf < -2^-149 (the
// product by -0.5 rounds -2^-149 to +0 and keeps every other negative qf
// positive); NaN proceeds either way. And exp2(qf * (-0.5 log2 e)) is
// bit-identical to exp2(power * log2 e): the product by -0.5 is exact for
// every qf whose power matters (for |qf| < 2^-125 both arguments give 1).
constexpr float POWER_POSITIVE_QF = -0x1p-149f;
constexpr float NEG_HALF_LOG2E = -0.5f * 1.44269504088896341f;

// alpha_raw = op exp(power) from the hardware exp2 (relative error < 1e-6
// where alpha can reach either threshold: power in [-5.6, 0]); within 1e-5
// of a threshold the exact expf decides, so the alpha tests and the clamp
// agree with the reference's. The training forward and the backward both
// use it, with the same operations, so they agree with each other. The
// caller clamps: alpha = min(alpha_raw, 0.99) (clamp_alpha).
__device__ __forceinline__ void alpha_fast(float qf, float op, const uint64_t* s_tab, float& g_exp,
                                           float& alpha_raw) {
  g_exp = __builtin_amdgcn_exp2f(qf * NEG_HALF_LOG2E);
  alpha_raw = op * g_exp;
  const bool near = __builtin_fabsf(alpha_raw - ALPHA_THRESHOLD) <= 1e-5f * ALPHA_THRESHOLD ||
                    __builtin_fabsf(alpha_raw - MAX_ALPHA) <= 1e-5f * MAX_ALPHA;
  // a wave-uniform test first: the common case (no lane near a threshold)
  // leaves the execution mask untouched
  if (__builtin_expect(__builtin_amdgcn_ballot_w64(near) != 0, 0)) {
    const float ge = expf_tab(-0.5f * qf, s_tab);
    g_exp = near ? ge : g_exp;
    alpha_raw = op * g_exp;  // unchanged where not near
  }
}
// f32::min(alpha_raw, 0.99) (NaN -> 0.99); *clamped is set when the clamp
// applied, which is also when the reference zeroes dL/d(alpha_raw).
// With the flag wanted (the backward), the clamp is that flag's select on
// the bit patterns, one instruction; a float select would be matched to a
// minimum and cost a canonicalising maximum besides.
__device__ __forceinline__ float clamp_alpha(float alpha_raw, bool* clamped = nullptr) {
  const bool c = !(alpha_raw <= MAX_ALPHA);
  if (!clamped) return c ? MAX_ALPHA : alpha_raw;
  *clamped = c;
  uint32_t bits = __float_as_uint(alpha_raw);
  asm volatile("" : "+v"(bits));
  return __uint_as_float(c ? __float_as_uint(MAX_ALPHA) : bits);
}

// Forward: four pixels per lane (pixel group q = tile rows 4q..4q+3, lane l
// at column l & 15, row 4q + (l >> 4)); batches of 64 splats staged in LDS,
// each of the S*S waves loading one of the four splat fields; the
// workgroup stops when all pixels of the bin have saturated. Modes 0 and 1
// evaluate alpha with the exact expf (images bit-identical to the
// reference's); the training step's modes 3 (as 0) and 2 use alpha_fast
// (images equal to float rounding, alpha decisions exact). MODE 2 resumes a
// bin flagged by MODE 3 on its back list: the state (C, T, last) comes from
// the output buffers and list indices continue after the front list.
constexpr int FT = 64;
// Adds a wave's work counts to counts[0..2]: lane evaluations (summed over
// the lanes), wave visits (lane 0's count, the same in every lane) and lane
// composites.
__device__ __forceinline__ void count_add(unsigned long long* counts, uint32_t n_eval, uint32_t n_visit,
                                          uint32_t n_comp, int lane) {
  unsigned long long e = n_eval, c = n_comp;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    e += __shfl_xor(e, o, 64);
    c += __shfl_xor(c, o, 64);
  }
  if (lane == 0) {
    atomicAdd(&counts[0], e);
    atomicAdd(&counts[1], (unsigned long long)n_visit);
    atomicAdd(&counts[2], c);
  }
}
// NW waves per workgroup: S * S (the bin's tiles share the staged splats) or
// 1 (one workgroup per tile of the bin, each staging the bin list itself and
// ending on its own).
// COUNT: the same pass, also adding this kernel's own work to in.counts
// (0: lane evaluations of a (pixel, splat) pair, 1: wave visits of a
// (splat, pixel group), 2: pairs that composite).
// FB: splats staged per batch (32, 64 or 128); a batch of 128 is staged as
// two halves of 64 (one per lane each) and walked half by half.
// lb: the schedule slot (NW == 1: bin slot * S * S + the tile in the bin).
template <int MODE, int S, int NW, bool COUNT, int FB>
__device__ __forceinline__ void fwd_body(const RasterIn& in, const int lb, float* __restrict__ out_rgb,
                                         float* __restrict__ out_t, int32_t* __restrict__ out_last) {
  static_assert(NW == 1 || NW == S * S, "waves per workgroup");
  static_assert(FB == 32 || FB == 64 || FB == 128, "batch");
  constexpr bool FAST = MODE >= 2;
  constexpr bool RGBA = MODE == 1;
  constexpr int NH = FB > FT ? FB / FT : 1;  // halves of 64 per batch
  constexpr int LW = FB < FT ? FB : FT;      // lanes staging each half
  __shared__ uint64_t s_tab[32];
  __shared__ float4 s_a[NH * FT];
  __shared__ float4 s_b[NH * FT];
  __shared__ float s_c[NH * FT];
  __shared__ float s_r[NH * FT];
  __shared__ uint2 s_rect[NH * FT];
  __shared__ int32_t s_fin[NW];
  const int lane = threadIdx.x & 63;
  // this wave's tile in the bin (wave-uniform), and its LDS slot
  const int w = NW == 1 ? lb % (S * S) : __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int ws = NW == 1 ? 0 : w;
  const int bin = in.tile_order[NW == 1 ? lb / (S * S) : lb];
  if (MODE == 2 && !in.flag[bin]) return;  // block-uniform
  const BinTile bt = bin_tile<S>(in, bin, w);
  // a single-tile workgroup resumes only a tile that kept its raw state
  if (MODE == 2 && NW == 1 && !(bt.ok && in.tile_flag[bt.tile])) return;
  if (threadIdx.x < 32) s_tab[threadIdx.x] = kExp2Tab[threadIdx.x];
  const CamDev& c = in.cams[bt.v];
  const int tx = bt.tx, ty = bt.ty;
  const int px = tx * TILE + (lane & 15);
  const float fpx = (float)px + 0.5f;
  uint32_t start = in.rstart[bin], end = in.rend[bin];
  if (end < start) start = end;  // no pair reached this bin
  // the list walked here: the bin's list, or for MODE 2 its back list, whose
  // entries follow the front list's (li0) in the last indices
  const uint32_t* vals = in.pair_vals;
  int32_t li0 = 0;
  uint32_t total_len = end - start;
  if (MODE == 2) {
    li0 = (int32_t)(end - start);
    start = in.bstart[bin];
    end = in.bend[bin];
    if (end < start) start = end;
    vals = in.back_vals;
    total_len += end - start;
  }
  const float bg0 = in.bg[0], bg1 = in.bg[1], bg2 = in.bg[2];
  // T[q] < TRANSMITTANCE_STOP marks a pixel that takes no more splats: one
  // that saturated, or (with T = 0) one outside the image
  float T[4], C0[4], C1[4], C2[4];
  int32_t last[4];
  // pixel-centre row of group 0; group q's is fpy0 + 4 q (exact)
  const float fpy0 = (float)(ty * TILE + (lane >> 4)) + 0.5f;
  // (T stays in [+0, 1], where the bit patterns order as the values do)
  auto pixels_left = [&]() {
    const uint32_t tm = max(max(f2u(T[0]), f2u(T[1])), max(f2u(T[2]), f2u(T[3])));
    return tm >= f2u(TRANSMITTANCE_STOP);
  };
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int py = ty * TILE + 4 * q + (lane >> 4);
    const bool inside = bt.ok && px < c.w && py < c.h;
    T[q] = 1.0f;
    C0[q] = RGBA ? bg0 : 0.0f;
    C1[q] = RGBA ? bg1 : 0.0f;
    C2[q] = RGBA ? bg2 : 0.0f;
    last[q] = -1;
    if (MODE == 2 && inside) {
      const size_t pix = (size_t)c.pix_base + (size_t)py * c.w + px;
      T[q] = out_t[pix];
      C0[q] = out_rgb[pix * 3 + 0];
      C1[q] = out_rgb[pix * 3 + 1];
      C2[q] = out_rgb[pix * 3 + 2];
      last[q] = out_last[pix];
    }
    if (!inside) T[q] = 0.0f;
  }
  int32_t fin = -1;  // list index at which this wave's last pixel saturated
  uint32_t n_eval = 0, n_visit = 0, n_comp = 0;
  for (uint32_t b0 = start; b0 < end; b0 += FB) {
    // stop when every pixel of the bin has saturated; the barrier also ends
    // the previous batch's reads of the staged splats
    if (!__syncthreads_or(pixels_left())) break;
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const uint32_t idx = b0 + h * FT + lane;
      if (lane < LW && idx < end) {
        const uint32_t vi = vals[idx];
        const int sl = h * FT + lane;
        if (NW == 1 || ws == 0) s_a[sl] = in.S.xy_ab[vi];
        if (NW == 1 || ws == 1 % NW) s_b[sl] = in.S.c_op_rg[vi];
        if (NW == 1 || ws == 2 % NW) s_c[sl] = in.S.col_b[vi];
        if (NW == 1 || ws == 3 % NW) {
          s_r[sl] = in.S.reach[vi];
          if (S > 1) s_rect[sl] = in.S.rect[vi];
        }
      }
    }
    __syncthreads();
    // all lanes take part in the ballot: the lanes past the list's end may be
    // the ones with pixels left
    const bool wave_live = __builtin_amdgcn_ballot_w64(pixels_left()) != 0;
    bool done = false;
#pragma unroll 1
    for (int h = 0; h < NH && !done; ++h) {
    const uint32_t idx = b0 + h * FT + lane;
    uint32_t my_m = 0;  // group mask of this lane's staged splat for this wave's tile
    if (lane < LW && idx < end && wave_live && (S == 1 || in_rect(s_rect[h * FT + lane], tx, ty))) {
      const float4 A = s_a[h * FT + lane];
      const float4 B = s_b[h * FT + lane];
      my_m = wave_mask(A.x, A.y, A.z, A.w, B.x, s_r[h * FT + lane], tx, ty);
    }
    // splats that reach some pixel group, visited front to back as set bits
    // of a scalar mask (their LDS reads are not prefetched: holding a second
    // register set costs more than the other waves need to cover the wait)
    uint64_t live = __builtin_amdgcn_ballot_w64(my_m != 0);
    while (live) {
      const int jl = __builtin_ctzll(live);
      live &= live - 1;
      const uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)my_m, jl);
      const int j = h * FT + jl;
      const float4 A = s_a[j];
      const float4 B = s_b[j];
      const float cb = s_c[j];
      const int32_t li = li0 + (int32_t)(b0 - start) + j;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (__builtin_expect(!(m >> q & 1u), 0)) continue;  // wave-uniform; most visits set it
        if (COUNT) n_visit += 1;
        if (COUNT && !(T[q] < TRANSMITTANCE_STOP)) n_eval += 1;
        float dx = fpx - A.x;
        float dy = (fpy0 + (float)(4 * q)) - A.y;
        const float qf = A.z * dx * dx + 2.0f * A.w * dx * dy + B.x * dy * dy;
        float alpha;
        if (FAST) {
          float g_exp, alpha_raw;
          alpha_fast(qf, B.y, s_tab, g_exp, alpha_raw);
          alpha = clamp_alpha(alpha_raw);
        } else {
          alpha = fmin_(B.y * expf_tab(-0.5f * qf, s_tab), MAX_ALPHA);
        }
        // a pixel that takes no more splats, power > 0 or alpha below
        // 1/255: no composite. One test after alpha rather than a branch
        // each: every branch is an execution-mask region, and those, not the
        // arithmetic, bound this loop (alpha of a dropped lane is wasted)
        if (T[q] < TRANSMITTANCE_STOP || qf < POWER_POSITIVE_QF || alpha < ALPHA_THRESHOLD) continue;
        if (COUNT) n_comp += 1;
        // the reference's operations (no fused multiply-adds): the training
        // forward's images then differ from the reference's only where the
        // hardware exp moved alpha, which keeps the L1 term's sign where a
        // pixel sits on its target
        float weight = alpha * T[q];
        C0[q] += B.z * weight;
        C1[q] += B.w * weight;
        C2[q] += cb * weight;
        T[q] *= 1.0f - alpha;
        last[q] = li;
      }
      if (__builtin_expect(__builtin_amdgcn_ballot_w64(pixels_left()) == 0, 0)) {
        fin = li;
        done = true;
        break;
      }
    }
    }
  }
  if (COUNT) count_add(in.counts, n_eval, n_visit, n_comp, lane);
  if (!RGBA && in.tile_cost && bt.ok) {
    // the tile's cost for the backward's longest-first schedule: how far
    // down the list its pixels composited (the backward walks that far back)
    const int lm = wave_max_i(max(max(last[0], last[1]), max(last[2], last[3])) + 1);
    if (lane == 0) in.tile_cost[bt.tile] = (uint32_t)lm;
  }
  if (lane == 0) s_fin[ws] = fin;
  // (with NW = 1: this tile's pixels)
  const bool all_done = !__syncthreads_or(pixels_left());
  // MODE 3 under a depth cut: a list that ended with pixels left is a front
  // list; the workgroup's tiles keep their raw state (tile_flag) and the bin
  // gets its back list (flag)
  const bool partial = (MODE == 0 || MODE == 3) && in.cut_on && *in.cut_on && !all_done;
  if (threadIdx.x == 0) {
    if (!RGBA && in.zneed && all_done) {
      // the depth of the splat that saturated the tiles, for the next step's cut
      int32_t f = -1;
      for (int k = 0; k < NW; ++k) f = max(f, s_fin[k]);
      if (f >= 0) atomicMax(in.zneed, f2u(in.S.depth[vals[start + (uint32_t)(f - li0)]]));
    }
    if (partial) {
      in.flag[bin] = 1;
      atomicAdd(in.n_flagged, 1u);
    }
  }
  if (partial && lane == 0 && bt.ok) in.tile_flag[bt.tile] = 1;
  // a tile the reference never binned keeps rgb 0 (it composites only the
  // tiles its lists reach; cover says which when the lists here are not the
  // reference's)
  const bool empty = in.cover ? !(bt.ok && in.cover[bt.tile]) : total_len == 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int py = ty * TILE + 4 * q + (lane >> 4);
    if (!(bt.ok && px < c.w && py < c.h)) continue;
    const size_t pix = (size_t)c.pix_base + (size_t)py * c.w + px;
    if (!RGBA) {
      if (partial) {
        out_rgb[pix * 3 + 0] = C0[q];
        out_rgb[pix * 3 + 1] = C1[q];
        out_rgb[pix * 3 + 2] = C2[q];
      } else if (empty) {
        out_rgb[pix * 3 + 0] = 0.0f;
        out_rgb[pix * 3 + 1] = 0.0f;
        out_rgb[pix * 3 + 2] = 0.0f;
      } else {
        out_rgb[pix * 3 + 0] = C0[q] + T[q] * bg0;
        out_rgb[pix * 3 + 1] = C1[q] + T[q] * bg1;
        out_rgb[pix * 3 + 2] = C2[q] + T[q] * bg2;
      }
      out_t[pix] = T[q];
      out_last[pix] = last[q];
    } else {
      if (start == end) continue;
      out_rgb[pix * 4 + 0] = C0[q];
      out_rgb[pix * 4 + 1] = C1[q];
      out_rgb[pix * 4 + 2] = C2[q];
      out_rgb[pix * 4 + 3] = 1.0f - T[q];
    }
  }
}

// Persistent raster launches (in.queue set): a grid of the workgroups the
// chip holds at once, each taking schedule slots from in.queue until none is
// left. The n slots are split into eight shares, one per XCD (the hardware
// deals workgroups to the XCDs in turn, so workgroup b runs on XCD b % 8):
// a contiguous run of the schedule (strided = false; the run's bins are
// neighbours that share splats in the XCD's L2) or every eighth slot
// (strided = true; a descending share of a longest-first table). A
// workgroup drains its own XCD's share first, then takes from the other
// shares in turn, so an XCD that runs ahead helps the others with their
// last slots. Counter x lives at queue[32 x] (its own 128-byte line) and
// starts at zero. Every slot is handed out exactly once, whatever the grid
// size and however many of its workgroups are resident.
__device__ __forceinline__ int queue_slot(int x, int n, bool strided, uint32_t k) {
  if (strided) {
    const uint32_t s = (uint32_t)x + 8u * k;
    return s < (uint32_t)n ? (int)s : -1;
  }
  const int per = n >> 3, rem = n & 7;
  const uint32_t len = (uint32_t)(per + (x < rem ? 1 : 0));
  return k < len ? x * per + min(x, rem) + (int)k : -1;
}
// The next slot for a workgroup of XCD x0 (from: the share it drains, x0 at
// first); -1 when every share is drained.
__device__ __forceinline__ int queue_pop(uint32_t* queue, int n, bool strided, int x0, int& from) {
  for (;;) {
    const int s = queue_slot(from, n, strided, atomicAdd(&queue[32 * from], 1u));
    if (s >= 0) return s;
    from = (from + 1) & 7;
    if (from == x0) return -1;
  }
}
// The slot of the workgroup's next turn, the same in every thread (NW waves).
template <int NW>
__device__ __forceinline__ int queue_next(uint32_t* queue, int n, bool strided, int x0, int& from) {
  if (NW == 1) {
    int v = 0;
    if ((threadIdx.x & 63) == 0) v = queue_pop(queue, n, strided, x0, from);
    return __builtin_amdgcn_readlane(v, 0);
  }
  __shared__ int s_slot;
  if (threadIdx.x == 0) s_slot = queue_pop(queue, n, strided, x0, from);
  __syncthreads();
  const int v = s_slot;
  __syncthreads();  // every thread has read it before thread 0 writes the next
  return v;
}

// Q: persistent (slots from in.queue) instead of one workgroup per slot.
template <int MODE, int S, int NW, bool COUNT = false, int FB = FT, bool Q = false>
__global__ __launch_bounds__(FT * NW) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_raster_fwd(
    RasterIn in, float* __restrict__ out_rgb, float* __restrict__ out_t, int32_t* __restrict__ out_last) {
  if (!Q) {
    fwd_body<MODE, S, NW, COUNT, FB>(in, xcd_remap(blockIdx.x, gridDim.x), out_rgb, out_t, out_last);
    return;
  }
  const int n = in.total_bins * (NW == 1 ? S * S : 1);
  const int x0 = (int)(blockIdx.x & 7);
  int from = x0;
  for (;;) {
    const int slot = queue_next<NW>(in.queue, n, false, x0, from);
    if (slot < 0) break;
    fwd_body<MODE, S, NW, COUNT, FB>(in, slot, out_rgb, out_t, out_last);
  }
}

// Workgroups of kernel fn (threads per workgroup) that the device holds at
// once, rounded down to a multiple of 8 (cached per kernel and device): the
// grid of a persistent launch.
static int resident_grid(const void* fn, int threads) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, int> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  const auto it = cache.find({fn, dev});
  if (it != cache.end()) return it->second;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 8;
  const int g = std::max(8, per_cu * cus / 8 * 8);
  cache[{fn, dev}] = g;
  return g;
}
// Grid of a persistent launch over n slots.
static dim3 queue_grid(const void* fn, int threads, int n) {
  return dim3((unsigned)std::min(resident_grid(fn, threads), (n + 7) / 8 * 8));
}

// The binned training forward (MODE 2, 3) in staging batch FB; the training
// forward (MODE 3) runs persistent when in.queue is set.
template <int MODE, int NW, int FB>
static void fwd_binned_fb(const RasterIn& in, dim3 grid, float* rgb, float* final_t, int32_t* last,
                          hipStream_t st) {
  if constexpr (MODE == 3) {
    if (in.queue) {
      const auto k = k_raster_fwd<MODE, 2, NW, false, FB, true>;
      hipLaunchKernelGGL(k, queue_grid((const void*)k, FT * NW, (int)grid.x), dim3(FT * NW), 0, st, in, rgb,
                         final_t, last);
      return;
    }
  }
  hipLaunchKernelGGL((k_raster_fwd<MODE, 2, NW, false, FB>), grid, dim3(FT * NW), 0, st, in, rgb, final_t, last);
}
template <int MODE, int NW>
static void fwd_binned(const RasterIn& in, dim3 grid, float* rgb, float* final_t, int32_t* last, hipStream_t st) {
  const int b = MODE >= 2 ? in.batch : FT;
  if (b == 32)
    fwd_binned_fb<MODE, NW, 32>(in, grid, rgb, final_t, last, st);
  else if (b == 128)
    fwd_binned_fb<MODE, NW, 128>(in, grid, rgb, final_t, last, st);
  else
    fwd_binned_fb<MODE, NW, FT>(in, grid, rgb, final_t, last, st);
}

template <int MODE>
static void fwd_launch(const RasterIn& in, float* rgb, float* final_t, int32_t* last, hipStream_t st) {
  if (in.bin_shift == 0)
    hipLaunchKernelGGL((k_raster_fwd<MODE, 1, 1>), dim3(in.total_bins), dim3(FT), 0, st, in, rgb, final_t, last);
  else if (in.fwd_tile_blocks)
    fwd_binned<MODE, 1>(in, dim3(in.total_bins * 4), rgb, final_t, last, st);
  else
    fwd_binned<MODE, 4>(in, dim3(in.total_bins), rgb, final_t, last, st);
}

void launch_raster_fwd(const RasterIn& in, int mode, float* rgb, float* final_t, int32_t* last, hipStream_t st) {
  if (in.total_bins <= 0) return;
  if (mode == 1)  // API renders only: the reference's tiles
    hipLaunchKernelGGL((k_raster_fwd<1, 1, 1>), dim3(in.total_bins), dim3(FT), 0, st, in, rgb, final_t, last);
  else if (mode == 0)
    fwd_launch<0>(in, rgb, final_t, last, st);
  else if (mode == 2)
    fwd_launch<2>(in, rgb, final_t, last, st);
  else
    fwd_launch<3>(in, rgb, final_t, last, st);
}

// Lane exchange inside rows of 16 (quad_perm, row_ror, row_mirror and
// row_half_mirror patterns: every lane has a source, so bound_ctrl changes
// nothing and lets the move fold into the consuming add).
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, true));
}
// ds_swizzle in bit-mask mode: lane i reads lane i ^ X inside groups of 32.
template <int X>
__device__ __forceinline__ float swz_xor(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), 0x1F | (X << 10)));
}
// Sums eight per-lane values over each row of 16 lanes as a butterfly that
// halves the number of live values at each exchange (4 + 2 + 1 exchanges,
// then one for lane ^ 8, instead of 8 x 4). On return lanes l and l + 8 of
// each row hold that row's sum of g[rev3(l & 7)], where rev3 reverses the
// three low bits (the four rows' sums are then added in LDS).
__device__ __forceinline__ float row_sum8_scatter(const float g[8], int lane) {
  const bool b0 = lane & 1, b1 = lane & 2, b2 = lane & 4;
  float r[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float snd = b0 ? g[i] : g[4 + i];
    float kp = b0 ? g[4 + i] : g[i];
    r[i] = kp + dpp_f<0xB1>(snd);  // partner lane ^ 1
  }
  float s[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    float snd = b1 ? r[i] : r[2 + i];
    float kp = b1 ? r[2 + i] : r[i];
    s[i] = kp + dpp_f<0x4E>(snd);  // partner lane ^ 2
  }
  float snd = b2 ? s[0] : s[1];
  float kp = b2 ? s[1] : s[0];
  float t = kp + swz_xor<4>(snd);
  return t + dpp_f<0x128>(t);  // row_ror:8 = lane ^ 8 inside a row of 16
}
// Sum over the four rows of 16 lanes, complete in row 3 (lanes 48-63). The
// first result of v_permlane16_swap(v, v) holds rows (0, 0, 2, 2) and that
// of v_permlane32_swap(v, v) rows (0, 1, 0, 1); adding each to v leaves row
// 1 + 0 and 3 + 2, then row 3 + 2 + 1 + 0. (Only the first result of each
// swap is used.)
__device__ __forceinline__ float cross_row_sum3(float v) {
  const uint32_t u = __builtin_bit_cast(uint32_t, v);
  v += __builtin_bit_cast(float, __builtin_amdgcn_permlane16_swap(u, u, false, false)[0]);
  const uint32_t u2 = __builtin_bit_cast(uint32_t, v);
  return v + __builtin_bit_cast(float, __builtin_amdgcn_permlane32_swap(u2, u2, false, false)[0]);
}
// Sum over each row of 16 lanes, in every lane of the row.
__device__ __forceinline__ float row_sum(float v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f<0x141>(v);  // row_half_mirror
  return v + dpp_f<0x140>(v);  // row_mirror
}
__device__ __forceinline__ int rev3(int l) { return ((l & 1) << 2) | (l & 2) | ((l >> 2) & 1); }

// Backward: four pixels per lane as in the forward. A lane sums its four
// pixels' gradients for a splat in registers, the wave reduces them across
// lanes into its LDS slot, and after each batch the bin adds its waves' sums
// to the accumulators (one atomic per value and splat per bin).
constexpr int BT = 64;
// COUNT: adds this kernel's own work to in.counts (3: lane evaluations, 4:
// wave visits, 5: contributing pairs) and writes no accumulator.
// FB: splats staged per batch (32, 64 or 128; 128 as two halves of 64,
// walked back to front).
template <int S, bool COUNT, int FB>
__device__ __forceinline__ void bwd_body(const RasterIn& in, const int bin, const float* __restrict__ final_t,
                                         const int32_t* __restrict__ last_contrib,
                                         const float* __restrict__ dl_dpix, float* __restrict__ accum) {
  static_assert(FB == 32 || FB == 64 || FB == 128, "batch");
  constexpr int NW = S * S;
  constexpr int NH = FB > BT ? FB / BT : 1;  // halves of 64 per batch
  constexpr int LW = FB < BT ? FB : BT;      // lanes staging each half
  constexpr int GS = NH * BT + 1;            // s_g row stride
  __shared__ uint64_t s_tab[32];
  __shared__ float4 s_a[NH * BT];
  __shared__ float4 s_b[NH * BT];
  __shared__ float s_c[NH * BT];
  __shared__ float s_r[NH * BT];
  __shared__ uint2 s_rect[NH * BT];
  __shared__ uint32_t s_vi[NH * BT];
  // rows padded to an odd number of floats: the eight lanes that store a
  // splat's reduced sums (values 0-7 of one column) land in eight different
  // banks
  __shared__ float s_g[NW][9][GS];
  __shared__ float s_sink[NW][64 + NH * BT];  // written, never read (see g_dst)
  __shared__ int s_top[NW];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // wave-uniform
  if (threadIdx.x < 32) s_tab[threadIdx.x] = kExp2Tab[threadIdx.x];
  const BinTile bt = bin_tile<S>(in, bin, w);
  const CamDev& c = in.cams[bt.v];
  const int tx = bt.tx, ty = bt.ty;
  uint32_t start = in.rstart[bin], end = in.rend[bin];
  if (end < start) start = end;  // no pair reached this bin
  // list index p < nf: pair_vals[start + p]; else back_vals[bs + p - nf]
  const uint32_t nf = end - start;
  uint32_t bs = 0, nbk = 0;
  if (in.bstart) {
    bs = in.bstart[bin];
    const uint32_t be = in.bend[bin];
    nbk = be > bs ? be - bs : 0u;
  }
  const int px = tx * TILE + (lane & 15);
  const float fpx = (float)px + 0.5f;
  const float bg0 = in.bg[0], bg1 = in.bg[1], bg2 = in.bg[2];
  int lc[4];
  // H: sum over channels of dL/dpixel_c times the colour that the splats
  // behind the current one (and the background) add, i.e. d . h with the
  // reference's per-channel accumulators h (backward.rs:470-480)
  float T[4], d0[4], d1[4], d2[4], H[4];
  // pixel-centre row of group 0; group q's is fpy0 + 4 q (exact), formed
  // where it is used rather than held in four registers
  const float fpy0 = (float)(ty * TILE + (lane >> 4)) + 0.5f;
  int lc_max = -1;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int py = ty * TILE + 4 * q + (lane >> 4);
    lc[q] = -1;
    T[q] = 1.0f;
    d0[q] = d1[q] = d2[q] = 0.0f;
    if (bt.ok && px < c.w && py < c.h && nf + nbk != 0) {
      const size_t pix = (size_t)c.pix_base + (size_t)py * c.w + px;
      lc[q] = last_contrib[pix];
      T[q] = final_t[pix];
      d0[q] = dl_dpix[pix * 3 + 0];
      d1[q] = dl_dpix[pix * 3 + 1];
      d2[q] = dl_dpix[pix * 3 + 2];
    }
    H[q] = T[q] * ((d0[q] * bg0 + d1[q] * bg1) + d2[q] * bg2);
    lc_max = max(lc_max, lc[q]);
  }
  // wave-uniform (scalar) so that the splat loops are scalar-controlled
  const int top_w = __builtin_amdgcn_readfirstlane(wave_max_i(lc_max));
  if (lane == 0) s_top[w] = top_w;
  __syncthreads();
  int top = s_top[0];
#pragma unroll
  for (int k = 1; k < NW; ++k) top = max(top, s_top[k]);
  if (top < 0) return;  // block-uniform
  // the per-pixel inputs are needed by the first contributing splat; wait for
  // them here rather than inside the loop, where the wait would also cover
  // the previous batch's accumulator atomics
#pragma unroll
  for (int q = 0; q < 4; ++q) asm volatile("" ::"v"(d0[q]), "v"(d1[q]), "v"(d2[q]));
  float g[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) g[k] = 0.0f;
  // where this lane stores after a wave reduction: lanes 48-56 (row 3,
  // which holds the wave's sums) to the splat's column of value rev3(l & 15)
  // (lane 56: value 8), the others to a sink of their own, so that every
  // lane stores and the execution mask stays whole
  const int lr = lane & 15;
  float* const g_dst =
      lane >= 48 && lr <= 8 ? &s_g[w][lr < 8 ? rev3(lr) : 8][0] : &s_sink[w][lane];
  uint32_t n_eval = 0, n_visit = 0, n_comp = 0;
  for (int hi = top; hi >= 0; hi -= FB) {
    const int lo = max(0, hi - (FB - 1));
    const int nb = hi - lo + 1;
    __syncthreads();  // previous batch fully read and flushed
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const int sl = h * BT + lane;
      if (lane < LW && sl < nb) {
        const uint32_t p = (uint32_t)(lo + sl);
        const uint32_t vi = p < nf ? in.pair_vals[start + p] : in.back_vals[bs + (p - nf)];
        if (NW == 1 || w == 0) {
          s_vi[sl] = vi;
          s_a[sl] = in.S.xy_ab[vi];
        }
        if (NW == 1 || w == 1 % NW) s_b[sl] = in.S.c_op_rg[vi];
        if (NW == 1 || w == 2 % NW) s_c[sl] = in.S.col_b[vi];
        if (NW == 1 || w == 3 % NW) {
          s_r[sl] = in.S.reach[vi];
          if (S > 1) s_rect[sl] = in.S.rect[vi];
        }
      }
#pragma unroll
      for (int k = 0; k < 9; ++k) s_g[w][k][sl] = 0.0f;
    }
    __syncthreads();
    // the halves back to front
#pragma unroll 1
    for (int h = NH - 1; h >= 0; --h) {
    const int sl = h * BT + lane;
    uint32_t my_m = 0;  // group mask of this lane's staged splat for this wave's tile
    if (lane < LW && sl < nb && lo + h * BT <= top_w && (S == 1 || in_rect(s_rect[sl], tx, ty))) {
      const float4 A = s_a[sl];
      const float4 B = s_b[sl];
      my_m = wave_mask(A.x, A.y, A.z, A.w, B.x, s_r[sl], tx, ty);
    }
    // splats that reach some pixel group, visited back to front as set bits
    // of a scalar mask (their LDS reads are not prefetched, as in the forward)
    // (jl: the splat's lane in the half, j: its slot in the batch)
    auto splat = [&](const int jl, const float4 A, const float4 B, const float cb) {
      const uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)my_m, jl);
      const int j = h * BT + jl;
      const int li = lo + j;
      const float dx = fpx - A.x;
      bool contrib = false;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (__builtin_expect(!(m >> q & 1u), 0)) continue;  // wave-uniform; most visits set it
        if (COUNT) n_visit += 1;
        if (COUNT && !(li > lc[q])) n_eval += 1;
        float dy = (fpy0 + (float)(4 * q)) - A.y;
        const float qf = A.z * dx * dx + 2.0f * A.w * dx * dy + B.x * dy * dy;
        float g_exp, alpha_raw;
        alpha_fast(qf, B.y, s_tab, g_exp, alpha_raw);
        bool clamped;
        const float alpha = clamp_alpha(alpha_raw, &clamped);
        // past the pixel's last contributor, power > 0 (NaN proceeds, as in
        // the reference) or alpha below 1/255: one test, as in the forward
        if (li > lc[q] || qf < POWER_POSITIVE_QF || !(alpha >= ALPHA_THRESHOLD)) continue;
        if (COUNT) n_comp += 1;
        contrib = true;
        {
        // Gradient arithmetic only (the decisions above are uncontracted,
        // as in the forward): fused multiply-adds are within the gradient
        // tolerance.
#pragma clang fp contract(fast)
        // transmittance in front of this splat (only gradients use it)
        float inv_1ma = __builtin_amdgcn_rcpf(1.0f - alpha);
        float Tq = T[q] * inv_1ma;
        T[q] = Tq;
        float weight = alpha * Tq;
        g[0] += d0[q] * weight;
        g[1] += d1[q] * weight;
        g[2] += d2[q] * weight;
        const float dcol = d0[q] * B.z + d1[q] * B.w + d2[q] * cb;
        float dl_dalpha = Tq * dcol - H[q] * inv_1ma;
        float dl_draw = clamped ? 0.0f : dl_dalpha;
        float e = dl_draw * g_exp;
        // dL/dpower = opacity * e; the lane's pixels share dx, so only the
        // dy moments of e are summed here (sum e, sum e dy, sum e dy^2), and
        // the dx factors follow per splat, the opacity at the flush
        g[3] += e;
        float ty_ = e * dy;
        g[5] += ty_;
        g[8] += ty_ * dy;
        H[q] += weight * dcol;
        }
      }
      const uint64_t bm = __builtin_amdgcn_ballot_w64(contrib);
      if (bm == 0) return;
      {
#pragma clang fp contract(fast)
        g[4] = dx * g[3];
        g[6] = dx * g[4];
        g[7] = dx * g[5];
      }
      if (__popcll(bm) <= in.sparse_lanes) {
        if (contrib) {
#pragma unroll
          for (int k = 0; k < 9; ++k) atomicAdd(&s_g[w][k][j], g[k]);
        }
      } else {
        // row sums (lanes 0-7 of each row: value rev3(l), lane 8: value 8),
        // then the four rows summed in registers; each slot of the splat is
        // this wave's alone, so a plain store (no LDS atomic) sets it
        const float t8 = row_sum8_scatter(g, lane);
        const float s8 = row_sum(g[8]);
        g_dst[j] = cross_row_sum3(lr < 8 ? t8 : s8);
      }
      // the accumulators live across splats and restart here, which spares
      // re-initialising them for every splat (4, 6 and 7 are set above)
      g[0] = g[1] = g[2] = g[3] = g[5] = g[8] = 0.0f;
    };
    uint64_t live = __builtin_amdgcn_ballot_w64(my_m != 0);
    while (live) {
      const int jl = 63 - __builtin_clzll(live);
      live &= ~(1ull << jl);
      splat(jl, s_a[h * BT + jl], s_b[h * BT + jl], s_c[h * BT + jl]);
    }
    }
    __syncthreads();
    // flush: raw sums (dL/dcolour[3], dL/dopacity, S1..S5 with S the
    // moments sum(e dx), sum(e dy), sum(e dx^2), sum(e dx dy), sum(e dy^2) of
    // e = dL/dpower / opacity) over the waves; dL/dconic and dL/dmean follow
    // from the conic and the opacity. Output o of splat lane goes to thread
    // (o % NW) * 64 + lane, for each half.
#pragma unroll
    for (int h = 0; h < NH; ++h) {
    const int sl = h * BT + lane;
    if (lane < LW && sl < nb) {
      const uint32_t vi = s_vi[sl];
      auto raw = [&](int k) {
        float t = s_g[0][k][sl];
#pragma unroll
        for (int u = 1; u < NW; ++u) t += s_g[u][k][sl];
        return t;
      };
      for (int o = w; o < 9; o += NW) {  // wave-uniform o
        float val;
        if (o < 4) {
          val = raw(o);
        } else {
          const float op = s_b[sl].y;
          if (o == 4) {
            val = -0.5f * (op * raw(6));
          } else if (o == 5) {
            val = -(op * raw(7));
          } else if (o == 6) {
            val = -0.5f * (op * raw(8));
          } else {
            const float4 A = s_a[sl];
            const float cc = s_b[sl].x;
            const float s1 = op * raw(4), s2 = op * raw(5);
            val = o == 7 ? A.z * s1 + A.w * s2 : A.w * s1 + cc * s2;
          }
        }
        if (!COUNT && val != 0.0f) atomicAdd(&accum[(size_t)o * in.acc_stride + vi], val);
      }
    }
    }
  }
  if (COUNT) count_add(in.counts + 3, n_eval, n_visit, n_comp, lane);
}

// Q: persistent (slots from in.queue; every eighth slot per XCD under the
// longest-first table, a contiguous run of the schedule otherwise) instead
// of one workgroup per slot.
template <int S, bool COUNT = false, int FB = BT, bool Q = false>
__global__ __launch_bounds__(BT * S * S) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_raster_bwd(
    RasterIn in, const float* __restrict__ final_t, const int32_t* __restrict__ last_contrib,
    const float* __restrict__ dl_dpix, float* __restrict__ accum) {
  if (!Q) {
    const int bin = in.lpt_bins ? in.lpt_bins[blockIdx.x] : in.tile_order[xcd_remap(blockIdx.x, gridDim.x)];
    bwd_body<S, COUNT, FB>(in, bin, final_t, last_contrib, dl_dpix, accum);
    return;
  }
  const bool strided = in.lpt_bins != nullptr;
  const int x0 = (int)(blockIdx.x & 7);
  int from = x0;
  for (;;) {
    const int slot = queue_next<S * S>(in.queue, in.total_bins, strided, x0, from);
    if (slot < 0) break;
    const int bin = strided ? in.lpt_bins[slot] : in.tile_order[slot];
    bwd_body<S, COUNT, FB>(in, bin, final_t, last_contrib, dl_dpix, accum);
  }
}

// ---- longest-first schedule of the backward
// Orders the bins by cost (the largest tile cost among their tiles: the
// bin's waves walk that far back), descending, by a counting sort over
// LPT_BUCKETS buckets scaled to the largest cost (order inside a bucket is
// arbitrary). Every bin appears exactly once, so the table is a permutation
// whatever the costs hold. Block b of the backward then takes the b-th
// costliest bin; the hardware deals blocks to the XCDs in turn, so each XCD
// gets a descending share and the launch ends on short bins.
constexpr int LPT_THREADS = 1024;
constexpr int LPT_BUCKETS = 4096;
__global__ __launch_bounds__(LPT_THREADS) void k_lpt_order(const uint32_t* __restrict__ tile_cost,
                                                           const CamDev* __restrict__ cams, int n_views,
                                                           int total_bins, int shift, int32_t* __restrict__ lpt_bins) {
  __shared__ uint32_t s_cnt[LPT_BUCKETS];
  __shared__ uint32_t s_max[LPT_THREADS / 64];
  __shared__ uint32_t s_part[LPT_THREADS / 64];
  const int n = total_bins;
  const int t = threadIdx.x;
  auto cost = [&](int id) -> uint32_t {
    const int v = view_of_bin(cams, n_views, id);
    const CamDev& c = cams[v];
    const int lb = id - c.bin_base, S = 1 << shift;
    const int bx = lb % c.nbx, by = lb / c.nbx;
    uint32_t m = 0;
    for (int j = 0; j < S; ++j)
      for (int i = 0; i < S; ++i) {
        const int tx = bx * S + i, ty = by * S + j;
        if (tx < c.ntx && ty < c.nty) m = max(m, tile_cost[c.tile_base + ty * c.ntx + tx]);
      }
    return m;
  };
  for (int k = t; k < LPT_BUCKETS; k += LPT_THREADS) s_cnt[k] = 0;
  uint32_t m = 0;
  for (int id = t; id < n; id += LPT_THREADS) m = max(m, cost(id));
  for (int o = 32; o >= 1; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
  if ((t & 63) == 0) s_max[t >> 6] = m;
  __syncthreads();
  uint32_t cmax = 0;
  for (int k = 0; k < LPT_THREADS / 64; ++k) cmax = max(cmax, s_max[k]);
  // bucket LPT_BUCKETS - 1 - floor(cost (LPT_BUCKETS - 1) / cmax): costliest first
  auto bucket = [&](uint32_t c) -> int {
    return cmax ? LPT_BUCKETS - 1 - (int)(((uint64_t)c * (LPT_BUCKETS - 1)) / cmax) : 0;
  };
  for (int id = t; id < n; id += LPT_THREADS) atomicAdd(&s_cnt[bucket(cost(id))], 1u);
  __syncthreads();
  // exclusive scan of the bucket counts: each thread owns PER consecutive buckets
  constexpr int PER = LPT_BUCKETS / LPT_THREADS;
  uint32_t v[PER], sum = 0;
  for (int k = 0; k < PER; ++k) {
    v[k] = s_cnt[t * PER + k];
    sum += v[k];
  }
  uint32_t x = sum;
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
    if ((t & 63) >= d) x += y;
  }
  if ((t & 63) == 63) s_part[t >> 6] = x;
  __syncthreads();
  uint32_t base = x - sum;
  for (int k = 0; k < (t >> 6); ++k) base += s_part[k];
  for (int k = 0; k < PER; ++k) {
    s_cnt[t * PER + k] = base;
    base += v[k];
  }
  __syncthreads();
  for (int id = t; id < n; id += LPT_THREADS) lpt_bins[atomicAdd(&s_cnt[bucket(cost(id))], 1u)] = id;
}

void launch_lpt_order(const uint32_t* tile_cost, const CamDev* cams, int n_views, int total_bins, int shift,
                      int32_t* lpt_bins, hipStream_t st) {
  if (total_bins <= 0) return;
  hipLaunchKernelGGL(k_lpt_order, dim3(1), dim3(LPT_THREADS), 0, st, tile_cost, cams, n_views, total_bins, shift,
                     lpt_bins);
}

void launch_raster_counts(const RasterIn& in, float* rgb, float* final_t, int32_t* last, const float* dl_dpix,
                          hipStream_t st) {
  if (in.total_bins <= 0 || !in.counts) return;
  if (in.bin_shift == 0) {
    hipLaunchKernelGGL((k_raster_fwd<3, 1, 1, true>), dim3(in.total_bins), dim3(FT), 0, st, in, rgb, final_t, last);
    hipLaunchKernelGGL((k_raster_bwd<1, true>), dim3(in.total_bins), dim3(BT), 0, st, in, final_t, last, dl_dpix,
                       nullptr);
  } else {
    if (in.fwd_tile_blocks)
      hipLaunchKernelGGL((k_raster_fwd<3, 2, 1, true>), dim3(in.total_bins * 4), dim3(FT), 0, st, in, rgb, final_t,
                         last);
    else
      hipLaunchKernelGGL((k_raster_fwd<3, 2, 4, true>), dim3(in.total_bins), dim3(FT * 4), 0, st, in, rgb, final_t,
                         last);
    hipLaunchKernelGGL((k_raster_bwd<2, true>), dim3(in.total_bins), dim3(BT * 4), 0, st, in, final_t, last, dl_dpix,
                       nullptr);
  }
}

// The binned training backward in staging batch FB, persistent when
// in.queue is set.
template <int FB>
static void bwd_binned(const RasterIn& in, const float* final_t, const int32_t* last, const float* dl_dpix,
                       float* accum, hipStream_t st) {
  if (in.queue) {
    const auto k = k_raster_bwd<2, false, FB, true>;
    hipLaunchKernelGGL(k, queue_grid((const void*)k, BT * 4, in.total_bins), dim3(BT * 4), 0, st, in, final_t,
                       last, dl_dpix, accum);
  } else {
    hipLaunchKernelGGL((k_raster_bwd<2, false, FB>), dim3(in.total_bins), dim3(BT * 4), 0, st, in, final_t, last,
                       dl_dpix, accum);
  }
}

void launch_raster_bwd(const RasterIn& in, const float* final_t, const int32_t* last, const float* dl_dpix,
                       float* accum, hipStream_t st) {
  if (in.total_bins <= 0) return;
  if (in.bin_shift == 0)
    hipLaunchKernelGGL((k_raster_bwd<1>), dim3(in.total_bins), dim3(BT), 0, st, in, final_t, last, dl_dpix, accum);
  else if (in.batch == 32)
    bwd_binned<32>(in, final_t, last, dl_dpix, accum, st);
  else if (in.batch == 128)
    bwd_binned<128>(in, final_t, last, dl_dpix, accum, st);
  else
    bwd_binned<BT>(in, final_t, last, dl_dpix, accum, st);
}

// Algorithmic work of the last render: pairs the forward evaluates per pixel
// (up to and including the saturating splat, else the whole bin list) and
// pairs the backward evaluates (last contributor + 1). One block per tile.
__global__ __launch_bounds__(RT) void k_work_counts(RasterIn in, const float* __restrict__ final_t,
                                                    const int32_t* __restrict__ last,
                                                    unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s_f[RT / 64], s_b[RT / 64];
  const int t = threadIdx.x;
  const int tile = blockIdx.x;
  const int v = view_of_tile(in.cams, in.n_views, tile);
  const CamDev& c = in.cams[v];
  const int lt = tile - c.tile_base;
  const int ttx = lt % c.ntx, tty = lt / c.ntx;
  const int bin = c.bin_base + (tty >> in.bin_shift) * c.nbx + (ttx >> in.bin_shift);
  const int px = ttx * TILE + (t & 15), py = tty * TILE + (t >> 4);
  unsigned long long f = 0, b = 0;
  if (px < c.w && py < c.h) {
    const uint32_t s0 = in.rstart[bin], e0 = in.rend[bin];
    uint32_t len = e0 > s0 ? e0 - s0 : 0u;
    if (in.bstart) {
      const uint32_t s1 = in.bstart[bin], e1 = in.bend[bin];
      len += e1 > s1 ? e1 - s1 : 0u;
    }
    const size_t pix = (size_t)c.pix_base + (size_t)py * c.w + px;
    const int32_t lc = last[pix];
    f = final_t[pix] < TRANSMITTANCE_STOP ? (unsigned long long)(lc + 1) : (unsigned long long)len;
    b = (unsigned long long)(lc + 1);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    f += __shfl_xor(f, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  if ((t & 63) == 0) {
    s_f[t >> 6] = f;
    s_b[t >> 6] = b;
  }
  __syncthreads();
  if (t == 0) {
    unsigned long long sf = 0, sb = 0;
    for (int w = 0; w < RT / 64; ++w) {
      sf += s_f[w];
      sb += s_b[w];
    }
    atomicAdd(&out[0], sf);
    atomicAdd(&out[1], sb);
  }
}

void launch_work_counts(const RasterIn& in, const float* final_t, const int32_t* last, unsigned long long* out,
                        int total_tiles, hipStream_t st) {
  if (total_tiles <= 0) return;
  hipLaunchKernelGGL(k_work_counts, dim3(total_tiles), dim3(RT), 0, st, in, final_t, last, out);
}

}  // namespace rgs

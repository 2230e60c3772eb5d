// rgs_api.hip -- the C ABI (include/rgsplat_hip.h): contexts, device buffers,
// the reference-shaped entry points and the device-resident batched training
// step with hipGraph replay.
//
// The step is one stream of kernels whose sizes come from device memory:
// project -> depth sort -> pair-count scan -> pair generation -> tile sort ->
// tile ranges -> forward raster -> L1+D-SSIM -> backward raster -> fused
// geometry backprop + Adam. The only host synchronisation is at the end of a
// step (loss and status readback); the first step of a configuration runs
// eagerly to size the pair buffers, later steps replay a captured graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/rgsplat_hip.h"
#include "rgs_kernels.h"

using namespace rgs;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Runs f, turning a host exception (an allocation that fails) into an error
// code: no exception crosses the C ABI.
template <class F>
int guarded(const char* what, F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(RGS_E_HOST, std::string(what) + ": out of host memory");
  } catch (const std::exception& e) {
    return fail(RGS_E_HOST, std::string(what) + ": " + e.what());
  }
}

#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) return fail(RGS_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define CHK(x)            \
  do {                    \
    int r_ = (x);         \
    if (r_) return r_;    \
  } while (0)

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace
struct rgs_ctx;
namespace {
uint64_t* ctx_gen(rgs_ctx* c);

int ensure(rgs_ctx* c, DBuf& b, size_t bytes) {
  if (bytes <= b.bytes && b.p) return 0;
  ++*ctx_gen(c);  // a captured graph holding the old pointer is now stale
  if (b.p) {
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes ? bytes : 256;
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(RGS_E_HIP, std::string("hipMalloc(") + std::to_string(want) + "): " + hipGetErrorString(e));
  }
  b.bytes = want;
  return 0;
}
template <class T>
T* P_(DBuf& b) {
  return reinterpret_cast<T*>(b.p);
}

// Largest (tile, splat) pair count a render takes. Pair offsets, sort
// positions and tile ranges are 32-bit; a 32-bit scan total saturates at
// SCAN_SATURATED, so a larger count reads as at least this and is refused.
constexpr uint64_t PAIR_LIMIT = 0x7fffffffull;
int check_pairs(uint64_t k) {
  if (k < PAIR_LIMIT) return 0;
  return fail(RGS_E_CAPACITY, "the render needs " +
                                  (k >= rgs::SCAN_SATURATED ? std::string("2^32 or more")
                                                            : std::to_string((unsigned long long)k)) +
                                  " (tile, splat) pairs; at most 2^31 - 2 are supported");
}

// Matrix4::try_inverse of the row-major c2w (math/mod.rs:155-160), evaluated
// as the cofactor expansion on column-major storage; false when det == 0.
bool view_from_c2w(const float* c2w, float view[16]) {
  float m[16];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) m[c * 4 + r] = c2w[r * 4 + c];
  float inv[16];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  float det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  if (det == 0.0f) return false;
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
           m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
           m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
           m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
           m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
           m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
            m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
            m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
            m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
            m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  float inv_det = 1.0f / det;
  for (int i = 0; i < 16; ++i) inv[i] = inv[i] * inv_det;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) view[r * 4 + c] = inv[c * 4 + r];
  return true;
}

int make_camdev(const rgs_camera& cam, CamDev& out) {
  float view[16];
  if (!cam.width || !cam.height) return fail(RGS_E_ARG, "camera with zero width or height");
  if (!view_from_c2w(cam.c2w, view)) return fail(RGS_E_CAMERA, "c2w must be invertible");
  for (int i = 0; i < 12; ++i) out.view[i] = view[i];
  out.fx = cam.fx;
  out.fy = cam.fy;
  out.cx = cam.cx;
  out.cy = cam.cy;
  out.cam_pos[0] = cam.c2w[3];
  out.cam_pos[1] = cam.c2w[7];
  out.cam_pos[2] = cam.c2w[11];
  out.w = (int32_t)cam.width;
  out.h = (int32_t)cam.height;
  out.ntx = (out.w + TILE - 1) / TILE;
  out.nty = (out.h + TILE - 1) / TILE;
  out.tile_base = 0;
  out.pix_base = 0;
  out.tgt_base = 0;
  out.lunit_base = 0;
  return 0;
}

// A set of views rendered by one launch sequence.
// shift: pairs are binned per (1 << shift)^2 tiles (0: the reference's tiles,
// API renders; the training step uses its own setting).
struct ViewSet {
  std::vector<CamDev> cams;
  int shift = 0;
  int total_tiles = 0;
  int total_bins = 0;
  int total_lunits = 0;
  size_t total_pix = 0;
  void layout() {
    int tb = 0, bb = 0, lb = 0;
    size_t pb = 0;
    for (auto& c : cams) {
      c.tile_base = tb;
      c.bin_shift = shift;
      c.nbx = (c.ntx + (1 << shift) - 1) >> shift;
      c.nby = (c.nty + (1 << shift) - 1) >> shift;
      c.bin_base = bb;
      c.lunit_base = lb;
      c.pix_base = (int32_t)pb;
      tb += c.ntx * c.nty;
      bb += c.nbx * c.nby;
      lb += loss_units(c.w, c.h);
      pb += (size_t)c.w * c.h;
    }
    total_tiles = tb;
    total_bins = bb;
    total_lunits = lb;
    total_pix = pb;
  }
  int bits() const {
    int b = 1;
    while ((1ll << b) < (long long)total_bins) ++b;
    return b;
  }
};

// Device scalars of a render or step (cleared at its start). Depth cut:
// front / back entry counts, cut active, back pairs, flagged workgroups and
// the back entries that meet a flagged bin.
enum {
  SC_DEPTH_N = 0, SC_TOTAL = 1, SC_K = 2, SC_STATUS = 3, SC_NOUT = 4,
  SC_NF = 5, SC_NB = 6, SC_CUTON = 7, SC_BTOTAL = 8, SC_KB = 9, SC_NFLAG = 10, SC_NC = 11, SC_NBV = 12,
  SC_COUNT = 16
};

}  // namespace

struct rgs_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t st = nullptr;

  // render workspace (shared by every entry point)
  DBuf cams_d, xy_ab, c_op_rg, col_b, cnt, rect, reach, depth;
  DBuf dk[2], dv[2], offsets, bsums, scalars;
  DBuf pk[2], pv[2], hist, rowtot, os;
  DBuf rstart, rend, rgb, final_t, last, dl, partial, loss;
  DBuf first_splat;
  DBuf accum;
  size_t k_cap = 0;
  size_t depth_cap = 0;  // capacity (entries) of the depth-sort arrays
  int hist_nb = 0;
  const uint32_t* sorted_vals = nullptr;
  const uint32_t* sorted_keys = nullptr;
  const uint32_t* pair_vals = nullptr;
  // depth-cut binning of the training step
  DBuf flag, tile_flag, rbstart, rbend, pv2, cntb, cutst;
  DBuf part, kB, vB, kC[2], vC[2];  // cut partition: front offsets, back entries, their flagged subset
  const uint32_t* back_vals = nullptr;
  int cut_mode = 0;              // CutState::mode
  float cut_depth = 0.0f;        // mode 2: fixed cut depth
  uint64_t last_front_pairs = 0, last_back_pairs = 0, last_flagged = 0;
  bool last_cut = false;         // the last render was a training step (two-part lists)
  bool last_back_valid = false;  // ... that ran the back-list stage
  uint32_t* h_scalars = nullptr;  // pinned
  float* h_loss = nullptr;        // pinned, 64 entries

  // reference-API staging
  DBuf api_aos, api_P, api_grads, api_target, api_m1, api_m2, api_acc, api_den, api_P2;
  DBuf dens_want, dens_cnt, dens_W, dens_packed, dens_prefix, dens_b32, dens_b64, dens_t32, dens_t64;

  // forward state kept for rgs_backward_dev
  ViewSet fwd_views;
  int fwd_n = 0;
  float fwd_bg[3] = {0, 0, 0};
  bool fwd_valid = false;

  // device-resident scene
  DBuf P, P2, M1, M2, gacc, gden;
  DBuf io_soa, io_rec, io_fields;  // model files: rows of host records, records, PLY property table
  DBuf srgb_lut;                   // RGBA8 -> linear table (512 floats), filled once
  DBuf sfm_xyz, sfm_rgb, sfm_k[2], sfm_v[2], sfm_khi, sfm_sx, sfm_scale, sfm_hist, sfm_rowtot, sfm_misc;
  DBuf pv_blk;                           // preview: first splat of every 256 pairs
  DBuf rs_taps, rs_tmp, rs_src, rs_dst;  // resize: taps and weights, f32 rows, source and result
  uint32_t* pv_pairs = nullptr;  // the tile-sorted pair values of the last preview render
  uint64_t pv_last_pairs = 0;
  DBuf pv_sp, pv_k[2], pv_v[2], pv_rect, pv_cnt, pv_off, pv_bs, pv_misc, pv_pk[2], pv_pv[2], pv_hist, pv_rowtot,
      pv_rs, pv_re, pv_rgb;
  size_t n = 0, cap = 0, m_rows = 0;
  uint64_t t = 0;
  std::vector<CamDev> reg_cams;
  DBuf targets;
  DBuf tstats;  // the targets' blurred statistics for the loss (launch_target_stats)
  size_t target_pix = 0;
  DBuf adam_d;
  AdamConsts* h_adam = nullptr;  // pinned
  CamDev* h_cams = nullptr;      // pinned, 64 entries
  uint64_t last_pairs = 0;
  size_t last_cap = 1;      // row capacity of the last render (vi = v * cap + i)
  int last_tiles = 0;       // total bins of the last render
  int last_shift = 0;       // its bin shift (0: bins are the reference's tiles)
  int last_ntiles = 0;      // its total tiles
  size_t last_pix = 0;      // its total pixels
  int last_nviews = 0;      // views of the last render
  float last_bg[3] = {0, 0, 0};

  // captured step
  // captured step, per variant: [1] with the depth cut's back lists, [0]
  // without (used while the previous step flagged no bin; a step that does
  // flag one is void and re-run with [1])
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::string gkey[2];
  uint64_t gen = 0;        // bumped on every device (re)allocation
  uint64_t graph_gen[2] = {0, 0};  // gen at capture time
  bool need_back = true;   // the next step runs the back-list stage
  // adaptive cut paused (steps left; the device runs without a cut), and
  // steps that run the back-list stage whatever the last one flagged (after
  // a pause, so the first step with the cut again is not voided)
  int cut_paused = 0;
  int force_back = 0;
  uint64_t back_reruns = 0;  // steps re-run because a bin needed its back list
  DBuf api_acc9;           // raster accumulators of the reference-API paths
  DBuf tile_order[2];      // raster schedule per bin shift: slot -> global bin id
  DBuf cover;              // per tile: composited although its tight list is empty
  std::vector<int32_t> h_order;
  std::string order_key[2];
  int bin_shift = 1;       // training step: bins of 2^shift x 2^shift tiles
  int fwd_tile_blocks = 1; // training forward with bins: a workgroup per tile (1) or per bin (0)
  int super_tile = 8;      // tiles per side of a schedule super-tile; 0: the backward longest first
  // longest-first schedule of the training step's backward: per-tile costs
  // of the step's forward and the bins ordered by them
  DBuf tile_cost, lpt_bins;
  int sparse_lanes = 2;    // backward: LDS atomics up to this many lanes
  int raster_batch = 64;   // binned training raster: splats staged per batch
  int raster_queue = 0;    // persistent training raster: bit 0 the forward, bit 1 the backward
  DBuf rqueue;             // their slot counters: forward at [0, 256), backward at [256, 512)
  float* scratch_dev = nullptr;  // this device's copy of the no-alloc scratch array
  float split_div = 0.0f;         // densify: 0 the reference's split, > 0 every scale / split_div
  bool grad_capture = false;  // the step writes its summed gradients into gcap
  DBuf gcap;                  // 60 rows x cap: gradient rows 0-58, screen-gradient norm
  DBuf work_counts;           // rgs_raster_work_counts: 6 u64

  // stage timing
  bool profiling = false;
  uint32_t prof_mask = 0;  // stages bracketed by events while profiling
  hipEvent_t ev_b[RGS_N_STAGES] = {}, ev_e[RGS_N_STAGES] = {};
  double stage_ms[RGS_N_STAGES] = {};
  uint64_t prof_steps = 0;
  uint64_t prof_graph_steps = 0;  // profiled steps that were graph launches
  bool prof_graph_failed = false; // capture with event nodes was refused
};

namespace {

uint64_t* ctx_gen(rgs_ctx* c) { return &c->gen; }

enum Stage { S_PROJECT, S_DSORT, S_SCAN, S_GEN, S_TSORT, S_FWD, S_LOSS, S_BWD, S_ADAM, S_BACK };
// Stage events. Under stream capture they become event-record nodes of the
// graph (hipEventRecordExternal), so replays are timed stage by stage too.
inline void record(rgs_ctx* c, hipEvent_t e) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(c->st, &cs);
  if (cs == hipStreamCaptureStatusActive) (void)hipEventRecordWithFlags(e, c->st, hipEventRecordExternal);
  else (void)hipEventRecord(e, c->st);
}
inline void stage_begin(rgs_ctx* c, int s) {
  if (c->profiling && (c->prof_mask >> s & 1u)) record(c, c->ev_b[s]);
}
inline void stage_end(rgs_ctx* c, int s) {
  if (c->profiling && (c->prof_mask >> s & 1u)) record(c, c->ev_e[s]);
}

void drop_graph(rgs_ctx* c) {
  for (int v = 0; v < 2; ++v) {
    if (c->gexec[v]) {
      (void)hipGraphExecDestroy(c->gexec[v]);
      c->gexec[v] = nullptr;
    }
    c->gkey[v].clear();
  }
}

// Raster schedule for the tiles of vs: views in order, each view walked in
// super-tiles of super_tile x super_tile tiles (row-major inside and across),
// so that the blocks one XCD holds in flight cover a compact patch of the
// image and share the splats (and accumulator lines) of that patch in L2.
int ensure_tile_order(rgs_ctx* c, const ViewSet& vs) {
  const int sh = vs.shift;
  std::string key = std::to_string(c->super_tile) + ":";
  for (const auto& cd : vs.cams) key += std::to_string(cd.nbx) + "x" + std::to_string(cd.nby) + ",";
  size_t bytes = (size_t)std::max(vs.total_bins, 1) * 4;
  if (key == c->order_key[sh] && c->tile_order[sh].bytes >= bytes) return 0;
  c->h_order.assign(std::max(vs.total_bins, 1), 0);
  const int s = std::max(c->super_tile, 1);
  size_t k = 0;
  for (const auto& cd : vs.cams)
    for (int sy = 0; sy < cd.nby; sy += s)
      for (int sx = 0; sx < cd.nbx; sx += s)
        for (int by = sy; by < std::min(sy + s, (int)cd.nby); ++by)
          for (int bx = sx; bx < std::min(sx + s, (int)cd.nbx); ++bx) c->h_order[k++] = cd.bin_base + by * cd.nbx + bx;
  CHK(ensure(c, c->tile_order[sh], bytes));
  HIPCHK(hipStreamSynchronize(c->st));  // a running raster may still read the old schedule
  HIPCHK(hipMemcpy(c->tile_order[sh].p, c->h_order.data(), bytes, hipMemcpyHostToDevice));
  ++c->gen;  // graphs captured with the old schedule are stale
  c->order_key[sh] = key;
  return 0;
}

// Buffers of the backward's longest-first schedule for the view set vs (the
// forward of the same step fills the costs and the order kernel the table
// before the backward reads them).
int ensure_lpt(rgs_ctx* c, const ViewSet& vs) {
  CHK(ensure(c, c->tile_cost, (size_t)std::max(vs.total_tiles, 1) * 4));
  CHK(ensure(c, c->lpt_bins, (size_t)std::max(vs.total_bins, 1) * 4));
  return 0;
}
// The forward records tile costs (when the schedule is longest first).
void lpt_forward(rgs_ctx* c, RasterIn& in) {
  if (c->super_tile == 0) in.tile_cost = P_<uint32_t>(c->tile_cost);
}
// Orders the bins for the backward and points its input at the table.
void lpt_backward(rgs_ctx* c, const ViewSet& vs, RasterIn& in) {
  if (c->super_tile != 0) return;
  launch_lpt_order(P_<uint32_t>(c->tile_cost), P_<CamDev>(c->cams_d), (int)vs.cams.size(), vs.total_bins, vs.shift,
                   P_<int32_t>(c->lpt_bins), c->st);
  in.tile_cost = nullptr;
  in.lpt_bins = P_<int32_t>(c->lpt_bins);
}

// The persistent training raster's slot counters (zeroed by the step before
// its forward): forward and backward.
constexpr size_t RQ_WORDS = 512;
uint32_t* fwd_queue(rgs_ctx* c) { return (c->raster_queue & 1) ? P_<uint32_t>(c->rqueue) : nullptr; }
uint32_t* bwd_queue(rgs_ctx* c) { return (c->raster_queue & 2) ? P_<uint32_t>(c->rqueue) + 256 : nullptr; }

// Sizes every buffer the render sequence touches for n Gaussians (row
// capacity cap) in the view set vs, pairs capacity k_cap.
int ensure_render(rgs_ctx* c, size_t cap, int n, const ViewSet& vs, size_t k_cap) {
  size_t nv = vs.cams.size();
  size_t vc = nv * cap;
  size_t dn = nv * (size_t)n;
  // sort values and the projection's thread index are 32-bit (v * cap + i)
  if (vc >= (1ull << 32))
    return fail(RGS_E_CAPACITY, "views x Gaussian capacity must stay below 2^32");
  CHK(ensure(c, c->cams_d, nv * sizeof(CamDev)));
  CHK(ensure(c, c->rqueue, RQ_WORDS * 4));
  CHK(ensure(c, c->xy_ab, vc * sizeof(float4)));
  CHK(ensure(c, c->c_op_rg, vc * sizeof(float4)));
  CHK(ensure(c, c->col_b, vc * sizeof(float)));
  CHK(ensure(c, c->cnt, vc * sizeof(uint32_t)));
  CHK(ensure(c, c->rect, vc * sizeof(uint2)));
  CHK(ensure(c, c->reach, vc * sizeof(float)));
  CHK(ensure(c, c->depth, vc * sizeof(float)));
  for (int i = 0; i < 2; ++i) {
    CHK(ensure(c, c->dk[i], dn * 4));
    CHK(ensure(c, c->dv[i], dn * 4));
    CHK(ensure(c, c->pk[i], k_cap * 4));
    CHK(ensure(c, c->pv[i], k_cap * 4));
  }
  c->depth_cap = dn;
  CHK(ensure(c, c->offsets, dn * 4));
  CHK(ensure(c, c->bsums, ((size_t)scan_nb(dn) + 1) * 8));
  // the scalars, then the training step's two look-back scan states
  CHK(ensure(c, c->scalars, SC_COUNT * 4 + 2 * lookback_words(dn) * 8));
  int nb = std::max(radix_nb_max(dn), radix_nb_max(k_cap));
  if (nb < 256) nb = 256;
  CHK(ensure(c, c->hist, (size_t)256 * nb * 4));
  c->hist_nb = nb;
  CHK(ensure(c, c->rowtot, 256 * 4));
  CHK(ensure(c, c->os, radix_os_words(nb) * 4));
  CHK(ensure(c, c->rstart, (size_t)vs.total_bins * 4));
  CHK(ensure(c, c->rend, (size_t)vs.total_bins * 4));
  CHK(ensure(c, c->rbstart, (size_t)vs.total_bins * 4));
  CHK(ensure(c, c->rbend, (size_t)vs.total_bins * 4));
  CHK(ensure(c, c->flag, ((size_t)vs.total_bins + 3) / 4 * 4));
  CHK(ensure(c, c->tile_flag, ((size_t)vs.total_tiles + 3) / 4 * 4));
  CHK(ensure(c, c->pv2, k_cap * 4));
  CHK(ensure(c, c->cntb, dn * 4));
  CHK(ensure(c, c->part, dn * 4));
  CHK(ensure(c, c->kB, dn * 4));
  CHK(ensure(c, c->vB, dn * 4));
  for (int i = 0; i < 2; ++i) {
    CHK(ensure(c, c->kC[i], dn * 4));
    CHK(ensure(c, c->vC[i], dn * 4));
  }
  CHK(ensure(c, c->first_splat, ((size_t)radix_nb(k_cap) + 1) * 4));
  CHK(ensure(c, c->rgb, vs.total_pix * 4 * 4));
  CHK(ensure(c, c->final_t, vs.total_pix * 4));
  CHK(ensure(c, c->last, vs.total_pix * 4));
  CHK(ensure_tile_order(c, vs));
  c->k_cap = k_cap;
  return 0;
}

SplatArrays splats(rgs_ctx* c) {
  return SplatArrays{P_<float4>(c->xy_ab), P_<float4>(c->c_op_rg), P_<float>(c->col_b), P_<uint32_t>(c->cnt),
                     P_<uint2>(c->rect), P_<float>(c->reach), P_<float>(c->depth)};
}

// Projection, depth sort and the pair-count scan (sizes fixed by n and vs).
// tight: training-step binning (only the tiles a splat's alpha can reach).
// cover: null, or the per-tile flags of launch_project (cleared here).
void enqueue_project(rgs_ctx* c, const float* P, size_t cap, int n, const ViewSet& vs, bool tight,
                     uint8_t* cover = nullptr) {
  hipStream_t st = c->st;
  uint32_t* sc = P_<uint32_t>(c->scalars);
  int nv = (int)vs.cams.size();
  size_t dn = (size_t)nv * n;
  launch_fill_u32(sc + SC_DEPTH_N, 1, (uint32_t)dn, st);
  if (cover) launch_fill_u32(reinterpret_cast<uint32_t*>(cover), ((size_t)vs.total_tiles + 3) / 4, 0u, st);
  stage_begin(c, S_PROJECT);
  launch_project(P, (int)cap, n, P_<CamDev>(c->cams_d), nv, splats(c), P_<uint32_t>(c->dk[0]),
                 P_<uint32_t>(c->dv[0]), sc + SC_STATUS, tight ? 1 : 0, cover, st);
  stage_end(c, S_PROJECT);
  uint32_t* keys[2] = {P_<uint32_t>(c->dk[0]), P_<uint32_t>(c->dk[1])};
  uint32_t* vals[2] = {P_<uint32_t>(c->dv[0]), P_<uint32_t>(c->dv[1])};
  RadixScratch rs{P_<uint32_t>(c->hist), P_<uint32_t>(c->rowtot), c->hist_nb};
  stage_begin(c, S_DSORT);
  int cur = radix_sort_pairs(keys, vals, sc + SC_DEPTH_N, dn, 32, rs, st);
  stage_end(c, S_DSORT);
  c->sorted_vals = vals[cur];
  c->sorted_keys = keys[cur];
  stage_begin(c, S_SCAN);
  scan_gathered_counts(vals[cur], P_<uint32_t>(c->cnt), dn, P_<uint32_t>(c->offsets), P_<uint32_t>(c->bsums),
                       sc + SC_TOTAL, st);
  stage_end(c, S_SCAN);
}

// The training step's projection: the entries are split at the depth cut,
// and only the front ones are depth-sorted and scanned here (the back ones
// that meet a bin left unsaturated are sorted after the front raster).
// The training step's look-back scan states (k = 0: the partition at the
// cut, 1: the pair offsets), after the scalars; the step's first fill clears
// them with the scalars.
size_t step_scalar_words(const rgs_ctx* c) { return SC_COUNT + 4 * lookback_words(c->depth_cap); }
uint64_t* lb_state(rgs_ctx* c, int k) {
  return reinterpret_cast<uint64_t*>(P_<uint32_t>(c->scalars) + SC_COUNT) + (size_t)k * lookback_words(c->depth_cap);
}

// The training step's radix sorts: single-pass digit offsets, a look-back
// that gives up flagging the step's status.
RadixScratch step_radix_scratch(rgs_ctx* c) {
  RadixScratch rs{P_<uint32_t>(c->hist), P_<uint32_t>(c->rowtot), c->hist_nb};
  rs.os = P_<uint32_t>(c->os);
  rs.status = P_<uint32_t>(c->scalars) + SC_STATUS;
  return rs;
}

// The training step's projection, partition at the cut (back entries only
// when the step has the back-list stage), depth sort of the front entries
// and pair offsets.
void enqueue_project_cut(rgs_ctx* c, const float* P, size_t cap, int n, const ViewSet& vs, uint8_t* cover,
                         bool with_back) {
  hipStream_t st = c->st;
  uint32_t* sc = P_<uint32_t>(c->scalars);
  int nv = (int)vs.cams.size();
  size_t dn = (size_t)nv * n;
  const CutState* cs = P_<CutState>(c->cutst);
  // (cover was cleared by the step's first fill)
  stage_begin(c, S_PROJECT);
  launch_project(P, (int)cap, n, P_<CamDev>(c->cams_d), nv, splats(c), P_<uint32_t>(c->dk[0]),
                 P_<uint32_t>(c->dv[0]), sc + SC_STATUS, 1, cover, st);
  stage_end(c, S_PROJECT);
  RadixScratch rs = step_radix_scratch(c);
  rs.os_zeroed = true;  // by the step's first fill
  stage_begin(c, S_DSORT);
  partition_at_cut(P_<uint32_t>(c->dk[0]), P_<uint32_t>(c->dv[0]), dn, cs, P_<uint32_t>(c->dk[1]),
                   P_<uint32_t>(c->dv[1]), with_back ? P_<uint32_t>(c->kB) : nullptr,
                   with_back ? P_<uint32_t>(c->vB) : nullptr, sc + SC_NF, sc + SC_NB, sc + SC_CUTON, sc + SC_NBV,
                   lb_state(c, 0), sc + SC_STATUS, st);
  uint32_t* keys[2] = {P_<uint32_t>(c->dk[1]), P_<uint32_t>(c->dk[0])};
  uint32_t* vals[2] = {P_<uint32_t>(c->dv[1]), P_<uint32_t>(c->dv[0])};
  int cur = radix_sort_pairs(keys, vals, sc + SC_NF, dn, 32, rs, st);
  stage_end(c, S_DSORT);
  c->sorted_vals = vals[cur];
  c->sorted_keys = keys[cur];
  stage_begin(c, S_SCAN);
  scan_pair_offsets(vals[cur], P_<uint32_t>(c->cnt), dn, sc + SC_NF, P_<uint32_t>(c->offsets), sc + SC_TOTAL,
                    lb_state(c, 1), sc + SC_STATUS, st, P_<uint32_t>(c->first_splat), c->k_cap);
  stage_end(c, S_SCAN);
}

RasterIn raster_in_n(rgs_ctx* c, int n_views, int total_bins, int shift, const float bg[3]) {
  RasterIn in{};
  in.rstart = P_<uint32_t>(c->rstart);
  in.rend = P_<uint32_t>(c->rend);
  in.pair_vals = c->pair_vals;
  in.tile_order = P_<int32_t>(c->tile_order[shift]);
  in.S = splats(c);
  in.cams = P_<CamDev>(c->cams_d);
  in.n_views = n_views;
  in.sparse_lanes = c->sparse_lanes;
  in.batch = c->raster_batch;
  in.acc_stride = 0;  // set by the backward's callers
  in.total_bins = total_bins;
  in.bin_shift = shift;
  in.fwd_tile_blocks = c->fwd_tile_blocks;
  for (int k = 0; k < 3; ++k) in.bg[k] = bg[k];
  return in;
}
RasterIn raster_in(rgs_ctx* c, const ViewSet& vs, const float bg[3]) {
  return raster_in_n(c, (int)vs.cams.size(), vs.total_bins, vs.shift, bg);
}

// Pair generation, tile sort, ranges and the forward raster.
void enqueue_bin_and_raster(rgs_ctx* c, size_t cap, int n, const ViewSet& vs, const float bg[3], int mode,
                            float* rgb_out, const uint8_t* cover = nullptr) {
  hipStream_t st = c->st;
  uint32_t* sc = P_<uint32_t>(c->scalars);
  size_t dn = vs.cams.size() * (size_t)n;
  RadixScratch rs{P_<uint32_t>(c->hist), P_<uint32_t>(c->rowtot), c->hist_nb};
  stage_begin(c, S_GEN);
  launch_gen_pairs(c->sorted_vals, dn, P_<uint32_t>(c->offsets), sc + SC_TOTAL, splats(c), (int)cap,
                   P_<CamDev>(c->cams_d), P_<uint32_t>(c->pk[0]), P_<uint32_t>(c->pv[0]), c->k_cap, sc + SC_K,
                   sc + SC_STATUS, P_<uint32_t>(c->first_splat), rs, vs.bits(), st);
  stage_end(c, S_GEN);
  uint32_t* keys[2] = {P_<uint32_t>(c->pk[0]), P_<uint32_t>(c->pk[1])};
  uint32_t* vals[2] = {P_<uint32_t>(c->pv[0]), P_<uint32_t>(c->pv[1])};
  stage_begin(c, S_TSORT);
  TileRanges tr{P_<uint32_t>(c->rstart), P_<uint32_t>(c->rend), (uint32_t)vs.total_bins};
  int cur = radix_sort_tiles(keys, vals, sc + SC_K, c->k_cap, vs.bits(), rs, tr, st);
  stage_end(c, S_TSORT);
  c->pair_vals = vals[cur];
  RasterIn in = raster_in(c, vs, bg);
  in.cover = cover;
  stage_begin(c, S_FWD);
  launch_raster_fwd(in, mode, rgb_out, P_<float>(c->final_t), P_<int32_t>(c->last), st);
  stage_end(c, S_FWD);
}


// (Re)writes the persistent cut state from the host settings.
int write_cut_state(rgs_ctx* c) {
  CutState cs{0xffffffffu, 0u, (uint32_t)(c->cut_paused > 0 ? 1 : c->cut_mode), 0u};
  if (c->cut_mode == 2) cs.key = f2u(c->cut_depth);
  HIPCHK(hipMemcpyAsync(c->cutst.p, &cs, sizeof(cs), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

// The training step's binning and forward raster with the depth cut: front
// pairs, their sort and forward (flagging the tiles the front lists leave
// unsaturated), then the back pairs of the flagged tiles, their sort and the
// resumed forward, and the next step's cut. Every launch runs whatever the
// counts turn out to be (no host sync): with nothing flagged the back stage
// reduces to a count pass, a scan and empty launches.
void enqueue_bin_and_raster_cut(rgs_ctx* c, size_t cap, int n, const ViewSet& vs, const float bg[3], float* rgb_out,
                                const uint8_t* cover, bool with_back) {
  hipStream_t st = c->st;
  uint32_t* sc = P_<uint32_t>(c->scalars);
  size_t dn = vs.cams.size() * (size_t)n;
  RadixScratch rs = step_radix_scratch(c);
  CutState* cs = P_<CutState>(c->cutst);
  stage_begin(c, S_GEN);
  launch_gen_pairs(c->sorted_vals, dn, P_<uint32_t>(c->offsets), sc + SC_TOTAL, splats(c), (int)cap,
                   P_<CamDev>(c->cams_d), P_<uint32_t>(c->pk[0]), P_<uint32_t>(c->pv[0]), c->k_cap, sc + SC_K,
                   sc + SC_STATUS, P_<uint32_t>(c->first_splat), rs, vs.bits(), st, sc + SC_NF);
  stage_end(c, S_GEN);
  uint32_t* keys[2] = {P_<uint32_t>(c->pk[0]), P_<uint32_t>(c->pk[1])};
  uint32_t* vals[2] = {P_<uint32_t>(c->pv[0]), P_<uint32_t>(c->pv[1])};
  stage_begin(c, S_TSORT);
  // (the ranges, the flags and the queue counters were set by the step's
  // first fill)
  TileRanges tr{P_<uint32_t>(c->rstart), P_<uint32_t>(c->rend), (uint32_t)vs.total_bins, 1u};
  int cur = radix_sort_tiles(keys, vals, sc + SC_K, c->k_cap, vs.bits(), rs, tr, st);
  stage_end(c, S_TSORT);
  c->pair_vals = vals[cur];
  RasterIn in = raster_in(c, vs, bg);
  in.cover = cover;
  in.cut_on = sc + SC_CUTON;
  in.flag = P_<uint8_t>(c->flag);
  in.tile_flag = P_<uint8_t>(c->tile_flag);
  in.n_flagged = sc + SC_NFLAG;
  in.zneed = &cs->zneed;
  lpt_forward(c, in);
  in.queue = fwd_queue(c);
  stage_begin(c, S_FWD);
  launch_raster_fwd(in, 3, rgb_out, P_<float>(c->final_t), P_<int32_t>(c->last), st);
  stage_end(c, S_FWD);
  if (!with_back) {
    // no back-list stage: a flagged bin voids the step (Adam is skipped and
    // the host re-runs it with the stage)
    stage_begin(c, S_BACK);
    launch_need_back_cut_end(sc + SC_NFLAG, sc + SC_STATUS, cs, st);
    stage_end(c, S_BACK);
    c->back_vals = nullptr;
    return;
  }
  // back lists of the flagged tiles: keys ping-pong in the front sort's key
  // buffers, values in its spare value buffer and pv2
  stage_begin(c, S_BACK);
  // the back entries that meet a flagged bin, in depth order
  uint32_t* kC[2] = {P_<uint32_t>(c->kC[0]), P_<uint32_t>(c->kC[1])};
  uint32_t* vC[2] = {P_<uint32_t>(c->vC[0]), P_<uint32_t>(c->vC[1])};
  scan_touch_flagged(P_<uint32_t>(c->vB), dn, sc + SC_NB, sc + SC_NFLAG, splats(c), (int)cap, P_<CamDev>(c->cams_d),
                     P_<uint8_t>(c->flag), P_<uint32_t>(c->part), P_<uint32_t>(c->bsums), sc + SC_NC, st);
  launch_compact(P_<uint32_t>(c->kB), P_<uint32_t>(c->vB), dn, P_<uint32_t>(c->part), sc + SC_NC, kC[0], vC[0], st);
  const int curc = radix_sort_pairs(kC, vC, sc + SC_NC, dn, 32, rs, st);
  uint32_t* cntb = P_<uint32_t>(c->cntb);
  launch_back_count(vC[curc], dn, sc + SC_NC, sc + SC_NFLAG, splats(c), (int)cap, P_<CamDev>(c->cams_d),
                    P_<uint8_t>(c->flag), cntb, st);
  scan_u32(cntb, dn, cntb, P_<uint32_t>(c->bsums), sc + SC_BTOTAL, st);
  uint32_t* bvals[2] = {vals[cur ^ 1], P_<uint32_t>(c->pv2)};
  launch_back_gen(vC[curc], dn, sc + SC_NC, cntb, sc + SC_BTOTAL, sc + SC_NFLAG, splats(c), (int)cap,
                  P_<CamDev>(c->cams_d), P_<uint8_t>(c->flag), keys[0], bvals[0], c->k_cap, sc + SC_KB,
                  sc + SC_STATUS, st);
  TileRanges trb{P_<uint32_t>(c->rbstart), P_<uint32_t>(c->rbend), (uint32_t)vs.total_bins};
  int curb = radix_sort_tiles_full(keys, bvals, sc + SC_KB, c->k_cap, vs.bits(), rs, trb, st);
  c->back_vals = bvals[curb];
  in.bstart = P_<uint32_t>(c->rbstart);
  in.bend = P_<uint32_t>(c->rbend);
  in.back_vals = c->back_vals;
  launch_raster_fwd(in, 2, rgb_out, P_<float>(c->final_t), P_<int32_t>(c->last), st);
  launch_cut_end(cs, st);
  stage_end(c, S_BACK);
}

// Eager render of n Gaussians (SoA P, row capacity cap) into vs: sizes the
// pair buffers from the exact count.
int render_eager(rgs_ctx* c, const float* P, size_t cap, int n, ViewSet& vs, const float bg[3], int mode,
                 float* rgb_out) {
  vs.layout();
  CHK(ensure_render(c, cap, n, vs, std::max<size_t>(c->k_cap, 1)));
  HIPCHK(hipMemcpyAsync(c->cams_d.p, vs.cams.data(), vs.cams.size() * sizeof(CamDev), hipMemcpyHostToDevice, c->st));
  uint32_t* sc = P_<uint32_t>(c->scalars);
  launch_fill_u32(sc, SC_COUNT, 0u, c->st);
  enqueue_project(c, P, cap, n, vs, false);
  HIPCHK(hipMemcpyAsync(c->h_scalars, sc, SC_COUNT * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (c->h_scalars[SC_STATUS] & ST_NAN_DEPTH) return fail(RGS_E_NAN, "NaN depth in projection");
  size_t K = c->h_scalars[SC_TOTAL];
  CHK(check_pairs(K));
  if (K > c->k_cap) {
    size_t want = K + K / 4 + 1024;
    CHK(ensure_render(c, cap, n, vs, want));
  }
  enqueue_bin_and_raster(c, cap, n, vs, bg, mode, rgb_out ? rgb_out : P_<float>(c->rgb));
  c->last_pairs = K;
  c->last_cut = false;
  c->last_cap = cap;
  c->last_tiles = vs.total_bins;
  c->last_shift = vs.shift;
  c->last_ntiles = vs.total_tiles;
  c->last_nviews = (int)vs.cams.size();
  return 0;
}

// Uploads n AoS Gaussians and converts them to the SoA scratch scene.
int upload_api_scene(rgs_ctx* c, const void* g, size_t n, bool device_src) {
  size_t cap = n ? n : 1;
  CHK(ensure(c, c->api_aos, cap * 59 * 4));
  CHK(ensure(c, c->api_P, cap * 59 * 4));
  if (n) {
    if (device_src) {
      HIPCHK(hipMemcpyAsync(c->api_aos.p, g, n * 59 * 4, hipMemcpyDeviceToDevice, c->st));
    } else {
      HIPCHK(hipMemcpyAsync(c->api_aos.p, g, n * 59 * 4, hipMemcpyHostToDevice, c->st));
    }
    launch_aos_to_soa(P_<float>(c->api_aos), n, P_<float>(c->api_P), cap, c->st);
  }
  return 0;
}

int ensure_loss(rgs_ctx* c, const ViewSet& vs) {
  CHK(ensure(c, c->dl, vs.total_pix * 3 * 4));
  CHK(ensure(c, c->partial, (size_t)vs.total_lunits * 6 * 4));
  CHK(ensure(c, c->loss, 64 * 4));
  return 0;
}

// tstats: null, or the statistics of the target store `target` is (the
// registered views' targets).
void enqueue_loss(rgs_ctx* c, const ViewSet& vs, const float* rgb, const float* target, float lambda,
                  const float* tstats = nullptr) {
  LossIn li{rgb, target, P_<CamDev>(c->cams_d), (int)vs.cams.size(), vs.total_lunits, vs.total_pix, lambda};
  li.tstats = tstats;
  li.tgt_total = c->target_pix;
  launch_loss(li, P_<float>(c->dl), P_<float>(c->partial), P_<float>(c->loss), c->st);
}

// After the target store changed: registers the views and makes the
// targets' loss statistics (one sweep per view, stream order).
int finish_views(rgs_ctx* c, const std::vector<CamDev>& v, size_t tot) {
  CHK(ensure(c, c->tstats, tot * 6 * 4));
  for (const CamDev& cam : v) {
    LossIn li{};
    li.target = P_<float>(c->targets);
    li.n_views = 1;
    li.total_lunits = loss_units(cam.w, cam.h);
    li.cam0 = cam;
    li.cam0.lunit_base = 0;
    li.tgt_total = tot;
    launch_target_stats(li, P_<float>(c->tstats), c->st);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));
  c->reg_cams = v;
  c->target_pix = tot;
  return 0;
}

AdamConsts adam_consts(const rgs_train_cfg& cfg, uint64_t t) {
  AdamConsts k;
  k.lr[0] = cfg.lr_position;
  k.lr[1] = cfg.lr_rotation;
  k.lr[2] = cfg.lr_scale;
  k.lr[3] = cfg.lr_opacity;
  k.lr[4] = cfg.lr_sh_dc;
  k.lr_sh_rest = cfg.lr_sh_rest;
  k.beta1 = cfg.beta1;
  k.beta2 = cfg.beta2;
  k.one_minus_beta1 = 1.0f - cfg.beta1;
  k.one_minus_beta2 = 1.0f - cfg.beta2;
  // f32::powi -> compiler-rt __powisf2 (repeated squaring)
  auto powi = [](float a, int b) {
    const int recip = b < 0;
    float r = 1;
    while (1) {
      if (b & 1) r *= a;
      b /= 2;
      if (b == 0) break;
      a *= a;
    }
    return recip ? 1 / r : r;
  };
  k.bc1 = 1.0f - powi(cfg.beta1, (int)t);
  k.bc2 = 1.0f - powi(cfg.beta2, (int)t);
  k.eps = cfg.eps;
  k.inv_bc1 = 1.0f / k.bc1;
  k.inv_bc2 = 1.0f / k.bc2;
  return k;
}

// Scratch of rgs_loss_grad_dev_noalloc: the per-unit partial sums of the
// loss kernel, for at most NOALLOC_MAX_UNITS loss units of LSW x LSG pixels
// (any image up to 16384 x 16384 pixels), 1.5 MiB per GPU for the life of
// the process.
// Host code never uses the array's name as an address: each device's copy is
// resolved once with hipGetSymbolAddress (scratch_on) and that pointer is
// passed to the kernels.
constexpr size_t NOALLOC_MAX_UNITS = 65536;
static __device__ __align__(256) float scratch[NOALLOC_MAX_UNITS * 6];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float* g_scratch[MAX_DEVICES] = {};

// The address of the scratch array on `device`, which must be current.
int scratch_on(int device, float** out) {
  if (device < 0 || device >= MAX_DEVICES) return fail(RGS_E_ARG, "device id beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(RGS_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<float*>(ptr);
  }
  *out = g_scratch[device];
  return 0;
}

}  // namespace

// ============================================================== C ABI
extern "C" {

const char* rgs_last_error(void) { return g_err.c_str(); }

int rgs_create(int device, rgs_ctx** out) {
  if (!out) return fail(RGS_E_ARG, "null out");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return fail(RGS_E_HIP, "no HIP device available");
  if (device >= ndev) return fail(RGS_E_ARG, "device id out of range");
  if (device < 0) {
    HIPCHK(hipGetDevice(&device));
  }
  HIPCHK(hipSetDevice(device));
  float* scratch_dev = nullptr;
  CHK(scratch_on(device, &scratch_dev));
  auto* c = new rgs_ctx();
  c->device = device;
  c->scratch_dev = scratch_dev;
  if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(RGS_E_HIP, "hipStreamCreate failed");
  }
  c->st = c->own;
  if (hipHostMalloc((void**)&c->h_scalars, SC_COUNT * 4, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_loss, 64 * 4, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_adam, sizeof(AdamConsts), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_cams, 64 * sizeof(CamDev), hipHostMallocDefault) != hipSuccess) {
    rgs_destroy(c);
    return fail(RGS_E_HIP, "hipHostMalloc failed");
  }
  *out = c;
  return 0;
}

void rgs_destroy(rgs_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  drop_graph(c);
  DBuf* all[] = {&c->cams_d, &c->reach, &c->xy_ab, &c->c_op_rg, &c->col_b, &c->cnt, &c->rect, &c->dk[0], &c->dk[1],
                 &c->dv[0], &c->dv[1], &c->offsets, &c->bsums, &c->scalars, &c->pk[0], &c->pk[1], &c->pv[0],
                 &c->pv[1], &c->hist, &c->rowtot, &c->os, &c->rstart, &c->rend, &c->rgb, &c->final_t, &c->last, &c->dl,
                 &c->partial, &c->loss, &c->accum, &c->api_aos, &c->api_P, &c->api_grads,
                 &c->api_target, &c->api_m1, &c->api_m2, &c->api_acc, &c->api_den, &c->api_P2, &c->dens_want,
                 &c->dens_cnt, &c->dens_W, &c->dens_packed, &c->dens_prefix, &c->dens_b32, &c->dens_b64,
                 &c->dens_t32, &c->dens_t64, &c->first_splat, &c->P, &c->P2, &c->M1, &c->M2, &c->gacc, &c->gden, &c->targets, &c->tstats,
                 &c->io_soa, &c->io_rec, &c->io_fields, &c->srgb_lut, &c->sfm_xyz, &c->sfm_rgb, &c->sfm_k[0],
                 &c->sfm_k[1], &c->sfm_v[0], &c->sfm_v[1], &c->sfm_khi, &c->sfm_sx, &c->sfm_scale, &c->sfm_hist,
                 &c->sfm_rowtot, &c->sfm_misc, &c->pv_sp, &c->pv_k[0], &c->pv_k[1], &c->pv_v[0], &c->pv_v[1],
                 &c->pv_rect, &c->pv_cnt, &c->pv_off, &c->pv_bs, &c->pv_misc, &c->pv_pk[0], &c->pv_pk[1],
                 &c->pv_pv[0], &c->pv_pv[1], &c->pv_hist, &c->pv_rowtot, &c->pv_rs, &c->pv_re, &c->pv_rgb,
                 &c->rs_taps, &c->rs_tmp, &c->rs_src, &c->rs_dst, &c->pv_blk,
                 &c->adam_d, &c->api_acc9, &c->tile_order[0], &c->tile_order[1], &c->tile_cost, &c->lpt_bins, &c->rqueue, &c->cover, &c->depth,
                 &c->flag, &c->tile_flag, &c->rbstart, &c->rbend, &c->pv2, &c->cntb, &c->cutst, &c->part,
                 &c->kB, &c->vB, &c->kC[0], &c->kC[1], &c->vC[0], &c->vC[1], &c->gcap, &c->work_counts};
  for (DBuf* b : all)
    if (b->p) (void)hipFree(b->p);
  if (c->h_scalars) (void)hipHostFree(c->h_scalars);
  if (c->h_loss) (void)hipHostFree(c->h_loss);
  if (c->h_adam) (void)hipHostFree(c->h_adam);
  if (c->h_cams) (void)hipHostFree(c->h_cams);
  for (int s = 0; s < RGS_N_STAGES; ++s) {
    if (c->ev_b[s]) (void)hipEventDestroy(c->ev_b[s]);
    if (c->ev_e[s]) (void)hipEventDestroy(c->ev_e[s]);
  }
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
}

int rgs_set_stream(rgs_ctx* c, void* s) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  drop_graph(c);
  c->st = (hipStream_t)s;
  return 0;
}

int rgs_use_own_stream(rgs_ctx* c) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  drop_graph(c);
  c->st = c->own;
  return 0;
}

int rgs_set_raster_params(rgs_ctx* c, int super_tile, int sparse_lanes) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  if (super_tile < 0 || super_tile > 1024)
    return fail(RGS_E_ARG, "super_tile must be 0 (longest first) or in [1, 1024]");
  if (sparse_lanes < 0 || sparse_lanes > 64) return fail(RGS_E_ARG, "sparse_lanes must be in [0, 64]");
  if (super_tile != c->super_tile || sparse_lanes != c->sparse_lanes) drop_graph(c);
  c->super_tile = super_tile;
  c->sparse_lanes = sparse_lanes;
  return 0;
}

int rgs_get_raster_params(rgs_ctx* c, int* super_tile, int* sparse_lanes) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  if (super_tile) *super_tile = c->super_tile;
  if (sparse_lanes) *sparse_lanes = c->sparse_lanes;
  return 0;
}

int rgs_set_raster_batch(rgs_ctx* c, int batch) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  if (batch != 32 && batch != 64 && batch != 128) return fail(RGS_E_ARG, "batch must be 32, 64 or 128");
  if (batch != c->raster_batch) drop_graph(c);
  c->raster_batch = batch;
  return 0;
}

int rgs_get_raster_batch(rgs_ctx* c, int* batch) {
  if (!c || !batch) return fail(RGS_E_ARG, "null argument");
  *batch = c->raster_batch;
  return 0;
}

int rgs_set_raster_queue(rgs_ctx* c, int mode) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  if (mode < 0 || mode > 3) return fail(RGS_E_ARG, "queue mode must be in [0, 3]");
  if (mode != c->raster_queue) drop_graph(c);
  c->raster_queue = mode;
  return 0;
}

int rgs_get_raster_queue(rgs_ctx* c, int* mode) {
  if (!c || !mode) return fail(RGS_E_ARG, "null argument");
  *mode = c->raster_queue;
  return 0;
}

static int check_common(rgs_ctx* c, const void* g, size_t n, const rgs_camera* cam) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  if (!cam) return fail(RGS_E_ARG, "null camera");
  if (n && !g) return fail(RGS_E_ARG, "null gaussians");
  if (n >= (1u << 31)) return fail(RGS_E_ARG, "too many gaussians");
  (void)hipSetDevice(c->device);
  c->fwd_valid = false;  // the shared render workspace is about to change
  return 0;
}

int rgs_render_color(rgs_ctx* c, const rgs_gaussian* g, size_t n, const rgs_camera* cam, const float bg[3],
                     float* rgb_out) {
  CHK(check_common(c, g, n, cam));
  if (!rgb_out || !bg) return fail(RGS_E_ARG, "null output or bg");
  ViewSet vs;
  vs.cams.resize(1);
  CHK(make_camdev(*cam, vs.cams[0]));
  CHK(upload_api_scene(c, g, n, false));
  CHK(render_eager(c, P_<float>(c->api_P), n ? n : 1, (int)n, vs, bg, 0, nullptr));
  HIPCHK(hipMemcpyAsync(rgb_out, c->rgb.p, vs.total_pix * 3 * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int rgs_rasterize_rgba(rgs_ctx* c, const rgs_gaussian* g, size_t n, const rgs_camera* cam, const float bg[3],
                       float* rgba_out) {
  CHK(check_common(c, g, n, cam));
  if (!rgba_out || !bg) return fail(RGS_E_ARG, "null output or bg");
  ViewSet vs;
  vs.cams.resize(1);
  CHK(make_camdev(*cam, vs.cams[0]));
  vs.layout();
  CHK(ensure_render(c, n ? n : 1, (int)n, vs, std::max<size_t>(c->k_cap, 1)));
  launch_fill_f32(P_<float>(c->rgb), vs.total_pix * 4, 0.0f, c->st);
  CHK(upload_api_scene(c, g, n, false));
  CHK(render_eager(c, P_<float>(c->api_P), n ? n : 1, (int)n, vs, bg, 1, nullptr));
  HIPCHK(hipMemcpyAsync(rgba_out, c->rgb.p, vs.total_pix * 4 * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int rgs_render_and_backward(rgs_ctx* c, const rgs_gaussian* g, size_t n, const rgs_camera* cam,
                            const float* target_rgba, const float bg[3], float lambda, float* loss_out,
                            rgs_grad* grads_out) {
  CHK(check_common(c, g, n, cam));
  if (!target_rgba || !bg || !grads_out) return fail(RGS_E_ARG, "null target, bg or grads");
  ViewSet vs;
  vs.cams.resize(1);
  CHK(make_camdev(*cam, vs.cams[0]));
  size_t npix = (size_t)cam->width * cam->height;
  CHK(ensure(c, c->api_target, npix * 16));
  HIPCHK(hipMemcpyAsync(c->api_target.p, target_rgba, npix * 16, hipMemcpyHostToDevice, c->st));
  CHK(upload_api_scene(c, g, n, false));
  size_t cap = n ? n : 1;
  CHK(render_eager(c, P_<float>(c->api_P), cap, (int)n, vs, bg, 0, nullptr));
  CHK(ensure_loss(c, vs));
  enqueue_loss(c, vs, P_<float>(c->rgb), P_<float>(c->api_target), lambda);
  CHK(ensure(c, c->api_acc9, 9 * cap * 4));
  launch_fill_f32(P_<float>(c->api_acc9), 9 * cap, 0.0f, c->st);
  {
    RasterIn in = raster_in(c, vs, bg);
    in.acc_stride = cap;
    launch_raster_bwd(in, P_<float>(c->final_t), P_<int32_t>(c->last), P_<float>(c->dl), P_<float>(c->api_acc9),
                      c->st);
  }
  CHK(ensure(c, c->api_grads, cap * 60 * 4));
  launch_geometry_aos(P_<float>(c->api_P), (int)cap, (int)n, P_<CamDev>(c->cams_d), 0, P_<uint32_t>(c->cnt),
                      P_<float>(c->api_acc9), cap, P_<float>(c->api_grads), c->st);
  if (n) HIPCHK(hipMemcpyAsync(grads_out, c->api_grads.p, n * 60 * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(c->h_loss, c->loss.p, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (loss_out) *loss_out = c->h_loss[0];
  return 0;
}

int rgs_loss_grad(rgs_ctx* c, const float* rgb, const float* target_rgba, uint32_t w, uint32_t h, float lambda,
                  float* loss_out, float* dl_out) {
  if (!c || !rgb || !target_rgba || !dl_out) return fail(RGS_E_ARG, "null argument");
  if (!w || !h) return fail(RGS_E_ARG, "empty image");
  (void)hipSetDevice(c->device);
  ViewSet vs;
  vs.cams.resize(1);
  CamDev& cd = vs.cams[0];
  std::memset(&cd, 0, sizeof(cd));
  cd.w = (int32_t)w;
  cd.h = (int32_t)h;
  cd.ntx = (cd.w + TILE - 1) / TILE;
  cd.nty = (cd.h + TILE - 1) / TILE;
  vs.layout();
  size_t npix = (size_t)w * h;
  CHK(ensure(c, c->cams_d, sizeof(CamDev)));
  CHK(ensure(c, c->rgb, npix * 16));
  CHK(ensure(c, c->api_target, npix * 16));
  CHK(ensure_loss(c, vs));
  HIPCHK(hipMemcpyAsync(c->cams_d.p, vs.cams.data(), sizeof(CamDev), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->rgb.p, rgb, npix * 12, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->api_target.p, target_rgba, npix * 16, hipMemcpyHostToDevice, c->st));
  enqueue_loss(c, vs, P_<float>(c->rgb), P_<float>(c->api_target), lambda);
  HIPCHK(hipMemcpyAsync(dl_out, c->dl.p, npix * 12, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(c->h_loss, c->loss.p, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (loss_out) *loss_out = c->h_loss[0];
  return 0;
}

int rgs_loss_grad_dev_noalloc(rgs_ctx* c, const float* rgb_dev, const float* target_rgba_dev, uint32_t w,
                              uint32_t h, float lambda, float* loss_dev, float* dl_drgb_dev) {
  if (!c || !rgb_dev || !target_rgba_dev || !loss_dev || !dl_drgb_dev) return fail(RGS_E_ARG, "null argument");
  if (!w || !h) return fail(RGS_E_ARG, "empty image");
  if (w > (uint32_t)INT32_MAX || h > (uint32_t)INT32_MAX)
    return fail(RGS_E_CAPACITY, "image beyond the no-alloc scratch");
  const size_t units = (size_t)((w + LSW - 1) / LSW) * (size_t)((h + LSG - 1) / LSG);
  if (units > NOALLOC_MAX_UNITS)
    return fail(RGS_E_CAPACITY,
                "image beyond the no-alloc scratch (at most 65536 loss units of 54 x 155 pixels, e.g. 16384 x 16384)");
  if (!c->scratch_dev) return fail(RGS_E_HIP, "no scratch on this context's device");
  (void)hipSetDevice(c->device);
  LossIn li{};
  li.rgb = rgb_dev;
  li.target = target_rgba_dev;
  li.cams = nullptr;
  li.n_views = 1;
  li.total_lunits = (int)units;
  li.total_pix = (size_t)w * h;
  li.lambda = lambda;
  li.cam0.w = (int32_t)w;
  li.cam0.h = (int32_t)h;
  launch_loss(li, dl_drgb_dev, c->scratch_dev, loss_dev, c->st);
  HIPCHK(hipGetLastError());
  return 0;
}

int rgs_adam_step(rgs_ctx* c, rgs_gaussian* g, const rgs_grad* grads, size_t n, rgs_adam* st_,
                  const rgs_train_cfg* cfg) {
  if (!c || !st_ || !cfg || (n && (!g || !grads))) return fail(RGS_E_ARG, "null argument");
  if (n > st_->rows) return fail(RGS_E_ARG, "AdamState has fewer rows than Gaussians");
  (void)hipSetDevice(c->device);
  st_->t += 1;
  if (!n) return 0;
  AdamConsts k = adam_consts(*cfg, st_->t);
  CHK(ensure(c, c->api_aos, n * 59 * 4));
  CHK(ensure(c, c->api_grads, n * 60 * 4));
  CHK(ensure(c, c->api_m1, n * 59 * 4));
  CHK(ensure(c, c->api_m2, n * 59 * 4));
  CHK(ensure(c, c->api_acc, n * 4));
  CHK(ensure(c, c->api_den, n * 4));
  HIPCHK(hipMemcpyAsync(c->api_aos.p, g, n * 59 * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->api_grads.p, grads, n * 60 * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->api_m1.p, st_->m1, n * 59 * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->api_m2.p, st_->m2, n * 59 * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->api_acc.p, st_->grad_accum, n * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->api_den.p, st_->grad_denom, n * 4, hipMemcpyHostToDevice, c->st));
  launch_adam_aos(P_<float>(c->api_aos), P_<float>(c->api_grads), n, P_<float>(c->api_m1), P_<float>(c->api_m2),
                  P_<float>(c->api_acc), P_<uint32_t>(c->api_den), k, c->st);
  HIPCHK(hipMemcpyAsync(g, c->api_aos.p, n * 59 * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(st_->m1, c->api_m1.p, n * 59 * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(st_->m2, c->api_m2.p, n * 59 * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(st_->grad_accum, c->api_acc.p, n * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(st_->grad_denom, c->api_den.p, n * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

static int ensure_densify_scratch(rgs_ctx* c, size_t n) {
  size_t nb = (size_t)scan_nb(n) + 1;
  CHK(ensure(c, c->dens_want, n * 4));
  CHK(ensure(c, c->dens_cnt, n * 4));
  CHK(ensure(c, c->dens_W, n * 4));
  CHK(ensure(c, c->dens_packed, n * 8));
  CHK(ensure(c, c->dens_prefix, n * 8));
  CHK(ensure(c, c->dens_b32, nb * 4));
  CHK(ensure(c, c->dens_b64, nb * 8));
  CHK(ensure(c, c->dens_t32, 8));
  CHK(ensure(c, c->dens_t64, 8));
  CHK(ensure(c, c->scalars, SC_COUNT * 4));
  return 0;
}
static DensifyScratch dscratch(rgs_ctx* c) {
  return DensifyScratch{P_<uint32_t>(c->dens_want), P_<uint32_t>(c->dens_cnt), P_<uint32_t>(c->dens_W),
                        P_<uint64_t>(c->dens_packed), P_<uint64_t>(c->dens_prefix), P_<uint32_t>(c->dens_b32),
                        P_<uint64_t>(c->dens_b64), P_<uint32_t>(c->dens_t32), P_<uint64_t>(c->dens_t64)};
}

int rgs_densify_and_prune(rgs_ctx* c, rgs_gaussian* g, size_t* n_io, size_t capacity, rgs_adam* st_,
                          const rgs_train_cfg* cfg, float extent) {
  if (!c || !n_io || !st_ || !cfg || (*n_io && !g)) return fail(RGS_E_ARG, "null argument");
  size_t n = *n_io;
  if (n > capacity || n > st_->rows) return fail(RGS_E_ARG, "n exceeds capacity or AdamState rows");
  (void)hipSetDevice(c->device);
  size_t cap = capacity ? capacity : 1;
  CHK(ensure(c, c->api_aos, cap * 59 * 4));
  CHK(ensure(c, c->api_P, cap * 59 * 4));
  CHK(ensure(c, c->api_P2, cap * 59 * 4));
  CHK(ensure(c, c->api_acc, (n ? n : 1) * 4));
  CHK(ensure(c, c->api_den, (n ? n : 1) * 4));
  CHK(ensure_densify_scratch(c, n ? n : 1));
  if (n) {
    HIPCHK(hipMemcpyAsync(c->api_aos.p, g, n * 59 * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->api_acc.p, st_->grad_accum, n * 4, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->api_den.p, st_->grad_denom, n * 4, hipMemcpyHostToDevice, c->st));
    launch_aos_to_soa(P_<float>(c->api_aos), n, P_<float>(c->api_P), cap, c->st);
  }
  uint32_t* sc = P_<uint32_t>(c->scalars);
  launch_densify(P_<float>(c->api_P), P_<float>(c->api_P2), (int)cap, (int)n, P_<float>(c->api_acc),
                 P_<uint32_t>(c->api_den), cfg->densify_grad_threshold, cfg->prune_opacity_thresh,
                 cfg->max_gaussians, extent, dscratch(c), sc + SC_NOUT, c->st, c->split_div);
  HIPCHK(hipMemcpyAsync(c->h_scalars, sc, SC_COUNT * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  size_t nn = c->h_scalars[SC_NOUT];
  if (nn > capacity) return fail(RGS_E_CAPACITY, "densify result exceeds capacity");
  if (nn > st_->capacity) return fail(RGS_E_CAPACITY, "densify result exceeds AdamState capacity");
  if (nn) {
    launch_soa_to_aos(P_<float>(c->api_P2), cap, nn, P_<float>(c->api_aos), c->st);
    HIPCHK(hipMemcpyAsync(g, c->api_aos.p, nn * 59 * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  }
  // AdamState::resize (grow only) then the statistics reset
  if (nn > st_->rows) {
    std::memset(st_->m1 + st_->rows * 59, 0, (nn - st_->rows) * 59 * 4);
    std::memset(st_->m2 + st_->rows * 59, 0, (nn - st_->rows) * 59 * 4);
    st_->rows = nn;
  }
  std::memset(st_->grad_accum, 0, st_->rows * 4);
  std::memset(st_->grad_denom, 0, st_->rows * 4);
  *n_io = nn;
  return 0;
}

int rgs_reset_opacities(rgs_ctx* c, rgs_gaussian* g, size_t n) {
  if (!c || (n && !g)) return fail(RGS_E_ARG, "null argument");
  if (!n) return 0;
  (void)hipSetDevice(c->device);
  CHK(ensure(c, c->api_aos, n * 59 * 4));
  HIPCHK(hipMemcpyAsync(c->api_aos.p, g, n * 59 * 4, hipMemcpyHostToDevice, c->st));
  launch_reset_opacity(P_<float>(c->api_aos) + 3, n, 59, c->st);
  HIPCHK(hipMemcpyAsync(g, c->api_aos.p, n * 59 * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

// ---- device-pointer forms
int rgs_forward_dev(rgs_ctx* c, const float* g_dev, size_t n, const rgs_camera* cam, const float bg[3],
                    float* rgb_dev) {
  CHK(check_common(c, g_dev, n, cam));
  if (!rgb_dev || !bg) return fail(RGS_E_ARG, "null output or bg");
  c->fwd_valid = false;
  ViewSet vs;
  vs.cams.resize(1);
  CHK(make_camdev(*cam, vs.cams[0]));
  CHK(upload_api_scene(c, g_dev, n, true));
  CHK(render_eager(c, P_<float>(c->api_P), n ? n : 1, (int)n, vs, bg, 0, rgb_dev));
  c->fwd_views = vs;
  c->fwd_n = (int)n;
  for (int i = 0; i < 3; ++i) c->fwd_bg[i] = bg[i];
  c->fwd_valid = true;
  return 0;
}

int rgs_backward_dev(rgs_ctx* c, const float* dl_dev, float* grads_dev) {
  if (!c || !dl_dev || !grads_dev) return fail(RGS_E_ARG, "null argument");
  if (!c->fwd_valid) return fail(RGS_E_ARG, "rgs_backward_dev without a preceding rgs_forward_dev");
  (void)hipSetDevice(c->device);
  size_t n = (size_t)c->fwd_n, cap = n ? n : 1;
  CHK(ensure(c, c->api_acc9, 9 * cap * 4));
  launch_fill_f32(P_<float>(c->api_acc9), 9 * cap, 0.0f, c->st);
  RasterIn in = raster_in(c, c->fwd_views, c->fwd_bg);
  in.acc_stride = cap;
  launch_raster_bwd(in, P_<float>(c->final_t), P_<int32_t>(c->last), dl_dev, P_<float>(c->api_acc9), c->st);
  launch_geometry_aos(P_<float>(c->api_P), (int)cap, (int)n, P_<CamDev>(c->cams_d), 0, P_<uint32_t>(c->cnt),
                      P_<float>(c->api_acc9), cap, grads_dev, c->st);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---- device-resident scene
namespace {
// A fresh device scene of n Gaussians (rows not written here) with room for
// capacity: moments and densification statistics cleared, Adam's t reset.
int scene_reset(rgs_ctx* c, size_t n, size_t capacity) {
  if (capacity < n) capacity = n;
  if (capacity == 0) capacity = 1;
  if (capacity >= (1u << 31)) return fail(RGS_E_ARG, "capacity too large");
  drop_graph(c);
  CHK(ensure(c, c->P, capacity * 59 * 4));
  CHK(ensure(c, c->P2, capacity * 59 * 4));
  CHK(ensure(c, c->M1, capacity * 59 * 4));
  CHK(ensure(c, c->M2, capacity * 59 * 4));
  CHK(ensure(c, c->gacc, capacity * 4));
  CHK(ensure(c, c->gden, capacity * 4));
  CHK(ensure(c, c->api_aos, (n ? n : 1) * 59 * 4));
  c->cap = capacity;
  c->n = n;
  c->m_rows = n;
  c->t = 0;
  launch_fill_f32(P_<float>(c->M1), capacity * 59, 0.0f, c->st);
  launch_fill_f32(P_<float>(c->M2), capacity * 59, 0.0f, c->st);
  launch_fill_f32(P_<float>(c->gacc), capacity, 0.0f, c->st);
  launch_fill_u32(P_<uint32_t>(c->gden), capacity, 0u, c->st);
  return 0;
}
}  // namespace

int rgs_scene_upload(rgs_ctx* c, const rgs_gaussian* g, size_t n, size_t capacity) {
  if (!c || (n && !g)) return fail(RGS_E_ARG, "null argument");
  (void)hipSetDevice(c->device);
  CHK(scene_reset(c, n, capacity));
  if (n) {
    HIPCHK(hipMemcpyAsync(c->api_aos.p, g, n * 59 * 4, hipMemcpyHostToDevice, c->st));
    launch_aos_to_soa(P_<float>(c->api_aos), n, P_<float>(c->P), c->cap, c->st);
  }
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

// ---- model files (io/ply.rs) ----------------------------------------------
}  // extern "C"
namespace {

// The reference's header text (io/ply.rs:43-64).
std::string ply_header(size_t n) {
  std::string h = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(n) + "\n";
  for (const char* p : {"x", "y", "z", "nx", "ny", "nz"}) h += std::string("property float ") + p + "\n";
  for (int i = 0; i < 3; ++i) h += "property float f_dc_" + std::to_string(i) + "\n";
  for (int i = 0; i < 45; ++i) h += "property float f_rest_" + std::to_string(i) + "\n";
  h += "property float opacity\n";
  for (int i = 0; i < 3; ++i) h += "property float scale_" + std::to_string(i) + "\n";
  for (int i = 0; i < 4; ++i) h += "property float rot_" + std::to_string(i) + "\n";
  h += "end_header\n";
  return h;
}

// Rows [59][cap] of n host Gaussian3D records in c->io_soa (cap = n).
int host_records_to_rows(rgs_ctx* c, const rgs_gaussian* g, size_t n) {
  CHK(ensure(c, c->api_aos, (n ? n : 1) * 59 * 4));
  CHK(ensure(c, c->io_soa, (n ? n : 1) * 59 * 4));
  if (n) {
    HIPCHK(hipMemcpyAsync(c->api_aos.p, g, n * 59 * 4, hipMemcpyHostToDevice, c->st));
    launch_aos_to_soa(P_<float>(c->api_aos), n, P_<float>(c->io_soa), n, c->st);
  }
  return 0;
}

int write_file(const char* path, const std::string& head, const void* body, size_t bytes) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(RGS_E_IO, std::string("cannot create ") + path);
  bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
  if (ok && bytes) ok = std::fwrite(body, 1, bytes, f) == bytes;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : fail(RGS_E_IO, std::string("cannot write ") + path);
}

// PLY records of n Gaussians held as rows [59][cap], written to path.
int save_ply_rows_impl(rgs_ctx* c, const float* rows, size_t cap, size_t n, const char* path);
int save_ply_rows(rgs_ctx* c, const float* rows, size_t cap, size_t n, const char* path) {
  return guarded("writing the model file", [&] { return save_ply_rows_impl(c, rows, cap, n, path); });
}
int save_ply_rows_impl(rgs_ctx* c, const float* rows, size_t cap, size_t n, const char* path) {
  std::vector<float> rec(n * PLY_FIELDS);
  if (n) {
    CHK(ensure(c, c->io_rec, n * PLY_FIELDS * 4));
    launch_ply_encode(rows, cap, n, P_<float>(c->io_rec), c->st);
    HIPCHK(hipMemcpyAsync(rec.data(), c->io_rec.p, n * PLY_FIELDS * 4, hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  return write_file(path, ply_header(n), rec.data(), n * PLY_FIELDS * 4);  // little-endian host
}

int save_splat_rows_impl(rgs_ctx* c, const float* rows, size_t cap, size_t n, const char* path);
int save_splat_rows(rgs_ctx* c, const float* rows, size_t cap, size_t n, const char* path) {
  return guarded("writing the model file", [&] { return save_splat_rows_impl(c, rows, cap, n, path); });
}
int save_splat_rows_impl(rgs_ctx* c, const float* rows, size_t cap, size_t n, const char* path) {
  std::vector<uint8_t> rec(n * 32);
  if (n) {
    CHK(ensure(c, c->io_rec, n * 32));
    launch_splat_encode(rows, cap, n, c->io_rec.p, c->st);
    HIPCHK(hipMemcpyAsync(rec.data(), c->io_rec.p, n * 32, hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  return write_file(path, std::string(), rec.data(), n * 32);
}

// gaussian_from_props (io/ply.rs:216-247): the SoA row a property sets, -1 if none.
int ply_row(const std::string& name) {
  static const std::pair<const char*, int> named[] = {
      {"x", P_POS + 0}, {"y", P_POS + 1}, {"z", P_POS + 2}, {"f_dc_0", P_SH + 0}, {"f_dc_1", P_SH + 16},
      {"f_dc_2", P_SH + 32}, {"opacity", P_OP}, {"scale_0", P_LS + 0}, {"scale_1", P_LS + 1},
      {"scale_2", P_LS + 2}, {"rot_0", P_ROT + 0}, {"rot_1", P_ROT + 1}, {"rot_2", P_ROT + 2},
      {"rot_3", P_ROT + 3}};
  for (const auto& kv : named)
    if (name == kv.first) return kv.second;
  if (name.compare(0, 7, "f_rest_") != 0) return -1;
  // usize::from_str: optional '+', then decimal digits only
  size_t i = 7;
  if (i < name.size() && name[i] == '+') ++i;
  if (i == name.size()) return -1;
  unsigned long long idx = 0;
  for (; i < name.size(); ++i) {
    if (name[i] < '0' || name[i] > '9') return -1;
    idx = idx * 10 + (unsigned)(name[i] - '0');
    if (idx > (1ull << 40)) return -1;  // far past the 45 coefficients
  }
  const unsigned long long ch = idx / 15, k = idx % 15;  // channel-major, 15 per channel
  return ch < 3 ? P_SH + (int)(ch * 16 + 1 + k) : -1;
}

struct PlyFile {
  int format = 2;  // 0 binary LE, 1 binary BE, 2 ASCII
  size_t count = 0;
  std::vector<std::string> names;
  std::vector<int> widths;
  std::vector<uint8_t> body;  // n_body records of the properties' widths
  size_t n_body = 0;          // records stored; those up to count are zeros
  int rec_bytes = 0;
};

bool read_text_line(FILE* f, std::string& line) {
  line.clear();
  int ch;
  while ((ch = std::fgetc(f)) != EOF) {
    line.push_back((char)ch);
    if (ch == '\n') break;
  }
  return !line.empty();
}
std::vector<std::string> split_ws(const std::string& l) {
  std::vector<std::string> t;
  size_t i = 0;
  while (i < l.size()) {
    while (i < l.size() && std::isspace((unsigned char)l[i])) ++i;
    size_t j = i;
    while (j < l.size() && !std::isspace((unsigned char)l[j])) ++j;
    if (j > i) t.push_back(l.substr(i, j - i));
    i = j;
  }
  return t;
}
// f32::from_str of one token (decimal with exponent, inf/infinity/nan in any
// case, optional sign); the reference maps a failure to 0.
float parse_f32(const std::string& t) {
  std::string d = t;
  bool neg = false;
  if (!d.empty() && (d[0] == '+' || d[0] == '-')) {
    neg = d[0] == '-';
    d = d.substr(1);
  }
  std::string low = d;
  for (char& ch : low) ch = (char)std::tolower((unsigned char)ch);
  if (low == "inf" || low == "infinity") return neg ? -INFINITY : INFINITY;
  if (low == "nan") return NAN;
  for (char ch : d)
    if (!((ch >= '0' && ch <= '9') || ch == '.' || ch == 'e' || ch == 'E' || ch == '+' || ch == '-')) return 0.0f;
  char* end = nullptr;
  const float v = std::strtof(t.c_str(), &end);
  return (end == t.c_str() || *end) ? 0.0f : v;
}

// Header and body of a PLY as load_ply reads them (io/ply.rs:121-189).
int read_ply_impl(const char* path, PlyFile& pf);
int read_ply(const char* path, PlyFile& pf) {
  return guarded("reading the PLY", [&] { return read_ply_impl(path, pf); });
}

int read_ply_impl(const char* path, PlyFile& pf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(RGS_E_IO, std::string("cannot open ") + path);
  struct Closer {
    FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  std::string line;
  auto trimmed = [](const std::string& l) {
    size_t a = 0, b = l.size();
    while (a < b && std::isspace((unsigned char)l[a])) ++a;
    while (b > a && std::isspace((unsigned char)l[b - 1])) --b;
    return l.substr(a, b - a);
  };
  if (!read_text_line(f, line) || trimmed(line) != "ply") return fail(RGS_E_FORMAT, "not a PLY file (missing magic)");
  while (true) {
    if (!read_text_line(f, line)) return fail(RGS_E_FORMAT, "unexpected EOF in PLY header");
    const auto t = split_ws(line);
    if (t.size() == 3 && t[0] == "format") {
      if (t[1] == "binary_little_endian") pf.format = 0;
      else if (t[1] == "binary_big_endian") pf.format = 1;
      else if (t[1] == "ascii") pf.format = 2;
      else return fail(RGS_E_FORMAT, "unsupported PLY format: " + t[1]);
    } else if (t.size() == 3 && t[0] == "element" && t[1] == "vertex") {
      const char* s0 = t[2].c_str();
      char* end = nullptr;
      if (*s0 == '-' || !*s0) return fail(RGS_E_FORMAT, "vertex count");
      const unsigned long long v = std::strtoull(s0, &end, 10);
      if (*end) return fail(RGS_E_FORMAT, "vertex count");
      pf.count = (size_t)v;
    } else if (t.size() == 3 && t[0] == "property") {
      const std::string& ty = t[1];
      int w;
      if (ty == "float" || ty == "float32") w = 4;
      else if (ty == "double" || ty == "float64") w = 8;
      else if (ty == "uchar" || ty == "uint8" || ty == "char" || ty == "int8") w = 1;
      else if (ty == "short" || ty == "ushort" || ty == "int16" || ty == "uint16") w = 2;
      else if (ty == "int" || ty == "uint" || ty == "int32" || ty == "uint32") w = 4;
      else return fail(RGS_E_FORMAT, "unsupported PLY property type: " + ty);
      pf.names.push_back(t[2]);
      pf.widths.push_back(w);
    } else if (t.size() == 1 && t[0] == "end_header") {
      break;
    }  // other elements, comments, obj_info: ignored
  }
  const size_t np = pf.names.size();
  if (pf.count > ((size_t)1 << 31)) return fail(RGS_E_FORMAT, "vertex count too large");
  if (pf.format == 2) {
    // one line per vertex, values in property order; a missing value (or
    // a missing line at EOF) reads as 0; stored as little-endian floats
    // stored as little-endian floats. Only the lines present are stored:
    // the records after the last line are zeros (n_body), so host memory is
    // bounded by the file, whatever the header claims.
    pf.rec_bytes = (int)(4 * np);
    size_t i = 0;
    for (; i < pf.count; ++i) {
      if (!read_text_line(f, line)) break;
      const auto t = split_ws(line);
      pf.body.resize((i + 1) * np * 4);
      for (size_t k = 0; k < np; ++k) {
        const float v = k < t.size() ? parse_f32(t[k]) : 0.0f;
        std::memcpy(&pf.body[(i * np + k) * 4], &v, 4);
      }
    }
    pf.n_body = i;
  } else {
    size_t rb = 0;
    for (int w : pf.widths) rb += (size_t)w;
    if (rb > (1u << 20)) return fail(RGS_E_FORMAT, "PLY record too large");
    pf.rec_bytes = (int)rb;
    // the body must hold count records before anything is allocated for it
    const long here = std::ftell(f);
    if (here < 0 || std::fseek(f, 0, SEEK_END) != 0) return fail(RGS_E_IO, "cannot seek in the PLY file");
    const long end = std::ftell(f);
    if (end < here || std::fseek(f, here, SEEK_SET) != 0) return fail(RGS_E_IO, "cannot seek in the PLY file");
    if (rb && (unsigned long long)(end - here) / rb < pf.count)
      return fail(RGS_E_FORMAT, "PLY body shorter than its vertex count");
    pf.body.resize(pf.count * rb);
    if (!pf.body.empty() && std::fread(pf.body.data(), 1, pf.body.size(), f) != pf.body.size())
      return fail(RGS_E_FORMAT, "PLY body shorter than its vertex count");
    pf.n_body = pf.count;
  }
  return 0;
}

// Decode a read PLY into rows [59][cap] (count <= cap).
int decode_ply(rgs_ctx* c, const PlyFile& pf, float* rows, size_t cap) {
  if (!pf.count) return 0;
  std::vector<PlyField> fields;
  int off = 0;
  for (size_t k = 0; k < pf.names.size(); ++k) {
    const int w = pf.format == 2 ? 4 : pf.widths[k];
    fields.push_back(PlyField{off, w, ply_row(pf.names[k])});
    off += w;
  }
  CHK(ensure(c, c->io_rec, pf.body.empty() ? 4 : pf.body.size()));
  CHK(ensure(c, c->io_fields, (std::max<size_t>(fields.size(), P_N) + 1) * sizeof(PlyField)));
  if (!pf.body.empty())
    HIPCHK(hipMemcpyAsync(c->io_rec.p, pf.body.data(), pf.body.size(), hipMemcpyHostToDevice, c->st));
  if (ply_decode_tiled_fits(pf.rec_bytes)) {
    // per row, the last property that sets it
    std::vector<PlyField> row_field(P_N, PlyField{0, 0, -1});
    for (const PlyField& f : fields)
      if (f.row >= 0) row_field[f.row] = f;
    HIPCHK(hipMemcpyAsync(c->io_fields.p, row_field.data(), P_N * sizeof(PlyField), hipMemcpyHostToDevice, c->st));
    launch_ply_decode_tiled(P_<uint8_t>(c->io_rec), pf.count, pf.n_body, pf.rec_bytes, P_<PlyField>(c->io_fields),
                            pf.format == 1, rows, cap, c->st);
  } else {
    if (!fields.empty())
      HIPCHK(hipMemcpyAsync(c->io_fields.p, fields.data(), fields.size() * sizeof(PlyField), hipMemcpyHostToDevice,
                            c->st));
    launch_ply_decode(P_<uint8_t>(c->io_rec), pf.count, pf.n_body, pf.rec_bytes, P_<PlyField>(c->io_fields),
                      (int)fields.size(), pf.format == 1, rows, cap, c->st);
  }
  // the host buffers above are read by the copies: wait before they go
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

}  // namespace
extern "C" {

int rgs_save_ply(rgs_ctx* c, const rgs_gaussian* g, size_t n, const char* path) {
  if (!c || !path || (n && !g)) return fail(RGS_E_ARG, "null argument");
  (void)hipSetDevice(c->device);
  CHK(host_records_to_rows(c, g, n));
  return save_ply_rows(c, P_<float>(c->io_soa), n, n, path);
}

int rgs_scene_save_ply(rgs_ctx* c, const char* path) {
  if (!c || !path) return fail(RGS_E_ARG, "null argument");
  (void)hipSetDevice(c->device);
  return save_ply_rows(c, P_<float>(c->P), c->cap, c->n, path);
}

int rgs_save_splat(rgs_ctx* c, const rgs_gaussian* g, size_t n, const char* path) {
  if (!c || !path || (n && !g)) return fail(RGS_E_ARG, "null argument");
  (void)hipSetDevice(c->device);
  CHK(host_records_to_rows(c, g, n));
  return save_splat_rows(c, P_<float>(c->io_soa), n, n, path);
}

int rgs_scene_save_splat(rgs_ctx* c, const char* path) {
  if (!c || !path) return fail(RGS_E_ARG, "null argument");
  (void)hipSetDevice(c->device);
  return save_splat_rows(c, P_<float>(c->P), c->cap, c->n, path);
}

int rgs_load_ply(rgs_ctx* c, const char* path, rgs_gaussian* out, size_t capacity, size_t* n_out) {
  if (!c || !path || !n_out) return fail(RGS_E_ARG, "null argument");
  (void)hipSetDevice(c->device);
  PlyFile pf;
  CHK(read_ply(path, pf));
  *n_out = pf.count;
  if (pf.count && (!out || capacity < pf.count)) return fail(RGS_E_CAPACITY, "output capacity smaller than the file");
  if (!pf.count) return 0;
  CHK(ensure(c, c->io_soa, pf.count * 59 * 4));
  CHK(ensure(c, c->api_aos, pf.count * 59 * 4));
  CHK(decode_ply(c, pf, P_<float>(c->io_soa), pf.count));
  launch_soa_to_aos(P_<float>(c->io_soa), pf.count, pf.count, P_<float>(c->api_aos), c->st);
  HIPCHK(hipMemcpyAsync(out, c->api_aos.p, pf.count * 59 * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int rgs_scene_load_ply(rgs_ctx* c, const char* path, size_t capacity) {
  if (!c || !path) return fail(RGS_E_ARG, "null argument");
  (void)hipSetDevice(c->device);
  PlyFile pf;
  CHK(read_ply(path, pf));
  CHK(scene_reset(c, pf.count, capacity));
  CHK(decode_ply(c, pf, P_<float>(c->P), c->cap));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int rgs_scene_size(rgs_ctx* c, size_t* n) {
  if (!c || !n) return fail(RGS_E_ARG, "null argument");
  *n = c->n;
  return 0;
}

int rgs_scene_download(rgs_ctx* c, rgs_gaussian* g, size_t capacity, size_t* n) {
  if (!c || !n || (c->n && !g)) return fail(RGS_E_ARG, "null argument");
  if (capacity < c->n) return fail(RGS_E_CAPACITY, "output capacity smaller than the scene");
  (void)hipSetDevice(c->device);
  if (c->n) {
    CHK(ensure(c, c->api_aos, c->n * 59 * 4));
    launch_soa_to_aos(P_<float>(c->P), c->cap, c->n, P_<float>(c->api_aos), c->st);
    HIPCHK(hipMemcpyAsync(g, c->api_aos.p, c->n * 59 * 4, hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  *n = c->n;
  return 0;
}

int rgs_adam_download(rgs_ctx* c, rgs_adam* st_) {
  if (!c || !st_) return fail(RGS_E_ARG, "null argument");
  if (st_->capacity < c->m_rows) return fail(RGS_E_CAPACITY, "AdamState capacity smaller than its rows");
  (void)hipSetDevice(c->device);
  size_t rows = c->m_rows;
  if (rows) {
    // SoA [59][cap] -> row-major rows x 59
    std::vector<float> tmp(59 * c->cap);
    HIPCHK(hipMemcpy(tmp.data(), c->M1.p, 59 * c->cap * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < rows; ++i)
      for (int j = 0; j < 59; ++j) st_->m1[i * 59 + j] = tmp[(size_t)j * c->cap + i];
    HIPCHK(hipMemcpy(tmp.data(), c->M2.p, 59 * c->cap * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < rows; ++i)
      for (int j = 0; j < 59; ++j) st_->m2[i * 59 + j] = tmp[(size_t)j * c->cap + i];
    HIPCHK(hipMemcpy(st_->grad_accum, c->gacc.p, rows * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(st_->grad_denom, c->gden.p, rows * 4, hipMemcpyDeviceToHost));
  }
  st_->rows = rows;
  st_->t = c->t;
  return 0;
}

// ---- preview renderer (gpu/compute.rs:138-297) ------------------------------
}  // extern "C"
namespace {
// GpuRenderer::project of the n Gaussians in rows P (stride cap) into
// c->pv_sp, then, when rgb is set, composite into rgb (host, h*w*3).
int preview_run(rgs_ctx* c, const float* P, size_t cap, size_t n, const rgs_camera* cam, const float* bg,
                float* splats_host, float* rgb, bool rgb_on_device = false) {
  if (!cam->width || !cam->height) return fail(RGS_E_ARG, "camera with zero width or height");
  float view[16];
  if (!view_from_c2w(cam->c2w, view)) return fail(RGS_E_CAMERA, "c2w must be invertible");
  PreviewCam pc;
  for (int k = 0; k < 12; ++k) pc.view[k] = view[k];
  pc.fx = cam->fx;
  pc.fy = cam->fy;
  pc.cx = cam->cx;
  pc.cy = cam->cy;
  const int w = (int)cam->width, h = (int)cam->height;
  const int ntx = (w + TILE - 1) / TILE, nty = (h + TILE - 1) / TILE;
  const size_t ntiles = (size_t)ntx * nty;
  if (ntiles >= (1u << 24)) return fail(RGS_E_ARG, "image too large for the preview renderer");
  const uint32_t n32 = (uint32_t)n;
  const size_t nn = n ? n : 1;
  CHK(ensure(c, c->pv_sp, nn * 40));
  for (int b = 0; b < 2; ++b) {
    CHK(ensure(c, c->pv_k[b], nn * 4));
    CHK(ensure(c, c->pv_v[b], nn * 4));
  }
  CHK(ensure(c, c->pv_misc, 16));
  launch_preview_project(P, (int)cap, n32, pc, P_<float>(c->pv_sp), P_<uint32_t>(c->pv_k[0]),
                         P_<uint32_t>(c->pv_v[0]), c->st);
  if (splats_host) {
    if (n) HIPCHK(hipMemcpyAsync(splats_host, c->pv_sp.p, n * 40, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return 0;
  }
  // depth order
  int nbm = radix_nb_max(nn);
  CHK(ensure(c, c->pv_hist, (size_t)256 * nbm * 4));
  CHK(ensure(c, c->pv_rowtot, 256 * 4));
  uint32_t* n_dev = P_<uint32_t>(c->pv_misc);
  uint32_t misc[4] = {n32, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(n_dev, misc, 16, hipMemcpyHostToDevice, c->st));
  uint32_t* ka[2] = {P_<uint32_t>(c->pv_k[0]), P_<uint32_t>(c->pv_k[1])};
  uint32_t* va[2] = {P_<uint32_t>(c->pv_v[0]), P_<uint32_t>(c->pv_v[1])};
  const int a = n ? radix_sort_pairs(ka, va, n_dev, nn, 32, RadixScratch{P_<uint32_t>(c->pv_hist),
                                                                           P_<uint32_t>(c->pv_rowtot), nbm},
                                     c->st)
                  : 0;
  // rects, pair offsets, pairs
  CHK(ensure(c, c->pv_rect, nn * 32));  // composite rects, then the tile ranges to list
  CHK(ensure(c, c->pv_cnt, nn * 8));
  CHK(ensure(c, c->pv_off, nn * 8));
  CHK(ensure(c, c->pv_bs, ((size_t)scan_nb(nn) + 2) * 8));
  int4* rects = P_<int4>(c->pv_rect);
  int4* trects = rects + nn;
  launch_preview_rects(P_<float>(c->pv_sp), ka[a], va[a], n32, w, h, rects, trects, P_<uint64_t>(c->pv_cnt), c->st);
  uint64_t* total_d = P_<uint64_t>(c->pv_bs) + scan_nb(nn) + 1;
  uint64_t total = 0;
  if (n) {
    scan_u64(P_<uint64_t>(c->pv_cnt), n, P_<uint64_t>(c->pv_off), P_<uint64_t>(c->pv_bs), total_d, c->st);
    HIPCHK(hipMemcpyAsync(&total, total_d, 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  }
  if (total >= (1ull << 31)) return fail(RGS_E_CAPACITY, "more than 2^31 (tile, splat) pairs");
  const size_t m = total ? total : 1;
  for (int b = 0; b < 2; ++b) {
    CHK(ensure(c, c->pv_pk[b], m * 4));
    CHK(ensure(c, c->pv_pv[b], m * 4));
  }
  CHK(ensure(c, c->pv_rs, ntiles * 4));
  CHK(ensure(c, c->pv_re, ntiles * 4));
  launch_fill_u32(P_<uint32_t>(c->pv_rs), ntiles, 0u, c->st);
  launch_fill_u32(P_<uint32_t>(c->pv_re), ntiles, 0u, c->st);
  if (total) {
    uint32_t* pk[2] = {P_<uint32_t>(c->pv_pk[0]), P_<uint32_t>(c->pv_pk[1])};
    uint32_t* pv[2] = {P_<uint32_t>(c->pv_pv[0]), P_<uint32_t>(c->pv_pv[1])};
    CHK(ensure(c, c->pv_blk, ((size_t)(total + 255) / 256 + 1) * 4));
    launch_preview_pairs(trects, P_<uint64_t>(c->pv_off), n32, (uint32_t)total, ntx, P_<uint32_t>(c->pv_blk), pk[0],
                         pv[0], c->st);
    misc[1] = (uint32_t)total;
    HIPCHK(hipMemcpyAsync(n_dev, misc, 16, hipMemcpyHostToDevice, c->st));
    int bits = 8;
    while (bits < 32 && (ntiles - 1) >> bits) bits += 8;
    nbm = radix_nb_max(m);
    CHK(ensure(c, c->pv_hist, (size_t)256 * nbm * 4));
    const int q = radix_sort_pairs(pk, pv, n_dev + 1, m, bits,
                                   RadixScratch{P_<uint32_t>(c->pv_hist), P_<uint32_t>(c->pv_rowtot), nbm}, c->st);
    launch_preview_ranges(pk[q], (uint32_t)total, P_<uint32_t>(c->pv_rs), P_<uint32_t>(c->pv_re), c->st);
    c->pv_pairs = pv[q];
  }
  if (!rgb_on_device) CHK(ensure(c, c->pv_rgb, (size_t)w * h * 12));
  launch_preview_composite(P_<float>(c->pv_sp), va[a], rects, total ? c->pv_pairs : P_<uint32_t>(c->pv_pv[0]),
                           P_<uint32_t>(c->pv_rs), P_<uint32_t>(c->pv_re), w, h, ntx, nty, bg,
                           rgb_on_device ? rgb : P_<float>(c->pv_rgb), c->st);
  if (!rgb_on_device)
    HIPCHK(hipMemcpyAsync(rgb, c->pv_rgb.p, (size_t)w * h * 12, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->pv_last_pairs = total;
  return 0;
}
}  // namespace
extern "C" {

int rgs_preview_project(rgs_ctx* c, const rgs_gaussian* g, size_t n, const rgs_camera* cam, float* splats_out) {
  CHK(check_common(c, g, n, cam));
  if (!n) return 0;
  if (!splats_out) return fail(RGS_E_ARG, "null output");
  CHK(upload_api_scene(c, g, n, false));
  return preview_run(c, P_<float>(c->api_P), n, n, cam, nullptr, splats_out, nullptr);
}

int rgs_preview_render(rgs_ctx* c, const rgs_gaussian* g, size_t n, const rgs_camera* cam, const float bg[3],
                       float* rgb_out) {
  CHK(check_common(c, g, n, cam));
  if (!rgb_out || !bg) return fail(RGS_E_ARG, "null output or bg");
  CHK(upload_api_scene(c, g, n, false));
  return preview_run(c, P_<float>(c->api_P), n ? n : 1, n, cam, bg, nullptr, rgb_out);
}

int rgs_scene_preview_render(rgs_ctx* c, const rgs_camera* cam, const float bg[3], float* rgb_out) {
  if (!c || !cam) return fail(RGS_E_ARG, "null argument");
  if (!rgb_out || !bg) return fail(RGS_E_ARG, "null output or bg");
  (void)hipSetDevice(c->device);
  return preview_run(c, P_<float>(c->P), c->cap, c->n, cam, bg, nullptr, rgb_out);
}

int rgs_scene_preview_render_device(rgs_ctx* c, const rgs_camera* cam, const float bg[3], float* rgb_dev) {
  if (!c || !cam) return fail(RGS_E_ARG, "null argument");
  if (!rgb_dev || !bg) return fail(RGS_E_ARG, "null output or bg");
  (void)hipSetDevice(c->device);
  return preview_run(c, P_<float>(c->P), c->cap, c->n, cam, bg, nullptr, rgb_dev, true);
}

// ---- SfM initialisation (pipeline/sfm.rs:365-455) ---------------------------
}  // extern "C"
namespace {
// Rows [59][cap] of the Gaussians gaussians_from_sfm makes from n points.
int sfm_rows(rgs_ctx* c, const double* xyz, const uint8_t* rgb, size_t n, float* rows, size_t cap) {
  if (!n) return 0;
  if (n >= (1u << 31)) return fail(RGS_E_ARG, "too many points");
  const uint32_t n32 = (uint32_t)n;
  const int nbm = radix_nb_max(n);
  CHK(ensure(c, c->sfm_xyz, n * 24));
  CHK(ensure(c, c->sfm_rgb, (n * 3 + 3) / 4 * 4));
  for (int b = 0; b < 2; ++b) {
    CHK(ensure(c, c->sfm_k[b], n * 4));
    CHK(ensure(c, c->sfm_v[b], n * 4));
  }
  CHK(ensure(c, c->sfm_khi, n * 4));
  CHK(ensure(c, c->sfm_sx, n * 24));
  CHK(ensure(c, c->sfm_scale, n * 4));
  CHK(ensure(c, c->sfm_hist, (size_t)256 * nbm * 4));
  CHK(ensure(c, c->sfm_rowtot, 256 * 4));
  CHK(ensure(c, c->sfm_misc, (2 + 256) * 4));  // n, status, dc table
  // dc[c] = (powf(c / 255, 2.2) - 0.5) * SH_C0_INV with the reference's f32
  // constant 1 / 0.28209479177387814 and libm's powf; logit(0.1) likewise
  const float sh_c0_inv = 1.0f / 0.28209479177387814f;
  uint32_t misc[2 + 256];
  misc[0] = n32;
  misc[1] = 0;
  for (int v = 0; v < 256; ++v) {
    const float dcv = (powf((float)v / 255.0f, 2.2f) - 0.5f) * sh_c0_inv;
    std::memcpy(&misc[2 + v], &dcv, 4);
  }
  const float opacity_logit = logf(0.1f / (1.0f - 0.1f));
  HIPCHK(hipMemcpyAsync(c->sfm_xyz.p, xyz, n * 24, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->sfm_rgb.p, rgb, n * 3, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->sfm_misc.p, misc, sizeof(misc), hipMemcpyHostToDevice, c->st));
  uint32_t* n_dev = P_<uint32_t>(c->sfm_misc);
  uint32_t* status = n_dev + 1;
  const float* dc = reinterpret_cast<const float*>(n_dev + 2);
  const double* X = P_<double>(c->sfm_xyz);
  RadixScratch rs{P_<uint32_t>(c->sfm_hist), P_<uint32_t>(c->sfm_rowtot), nbm};
  // stable sort by x: low key word, then high key word
  launch_sfm_keys(X, n32, P_<uint32_t>(c->sfm_k[0]), P_<uint32_t>(c->sfm_khi), P_<uint32_t>(c->sfm_v[0]), status,
                  c->st);
  uint32_t* ka[2] = {P_<uint32_t>(c->sfm_k[0]), P_<uint32_t>(c->sfm_k[1])};
  uint32_t* va[2] = {P_<uint32_t>(c->sfm_v[0]), P_<uint32_t>(c->sfm_v[1])};
  const int a = radix_sort_pairs(ka, va, n_dev, n, 32, rs, c->st);
  launch_sfm_gather_hi(P_<uint32_t>(c->sfm_khi), va[a], n32, ka[1 - a], c->st);
  uint32_t* kb[2] = {ka[1 - a], ka[a]};
  uint32_t* vb[2] = {va[a], va[1 - a]};
  const int b = radix_sort_pairs(kb, vb, n_dev, n, 32, rs, c->st);
  const uint32_t* order = vb[b];
  launch_sfm_sorted_xyz(X, order, n32, P_<double>(c->sfm_sx), c->st);
  launch_sfm_scale(P_<double>(c->sfm_sx), order, n32, P_<float>(c->sfm_scale), status, c->st);
  uint32_t st_h = 0;
  HIPCHK(hipMemcpyAsync(&st_h, status, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (st_h) return fail(RGS_E_NAN, "NaN coordinate or point distance (the reference panics in its sorts)");
  launch_sfm_gaussians(X, P_<uint8_t>(c->sfm_rgb), P_<float>(c->sfm_scale), n32, dc, opacity_logit, rows, cap, c->st);
  return 0;
}
}  // namespace
extern "C" {

int rgs_gaussians_from_sfm(rgs_ctx* c, const double* xyz, const uint8_t* rgb, size_t n, rgs_gaussian* out) {
  if (!c || (n && (!xyz || !rgb || !out))) return fail(RGS_E_ARG, "null argument");
  if (!n) return 0;
  (void)hipSetDevice(c->device);
  CHK(ensure(c, c->io_soa, n * 59 * 4));
  CHK(ensure(c, c->api_aos, n * 59 * 4));
  CHK(sfm_rows(c, xyz, rgb, n, P_<float>(c->io_soa), n));
  launch_soa_to_aos(P_<float>(c->io_soa), n, n, P_<float>(c->api_aos), c->st);
  HIPCHK(hipMemcpyAsync(out, c->api_aos.p, n * 59 * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int rgs_scene_from_sfm(rgs_ctx* c, const double* xyz, const uint8_t* rgb, size_t n, size_t capacity) {
  if (!c || (n && (!xyz || !rgb))) return fail(RGS_E_ARG, "null argument");
  (void)hipSetDevice(c->device);
  CHK(scene_reset(c, n, capacity));
  CHK(sfm_rows(c, xyz, rgb, n, P_<float>(c->P), c->cap));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

// ---- colour conversions (io/images.rs:63-82, cpu/rasterizer.rs:38-49) -------
}  // extern "C"
namespace {
// srgb_to_linear (io/images.rs:76-82) of c / 255 for every byte c, with the
// libm powf the reference calls; then c / 255 (alpha).
int srgb_table(rgs_ctx* c) {
  if (c->srgb_lut.p) return 0;
  float t[512];
  for (int v = 0; v < 256; ++v) {
    const float x = (float)v / 255.0f;
    t[v] = x <= 0.04045f ? x / 12.92f : powf((x + 0.055f) / 1.055f, 2.4f);
    t[256 + v] = x;
  }
  CHK(ensure(c, c->srgb_lut, sizeof(t)));
  HIPCHK(hipMemcpy(c->srgb_lut.p, t, sizeof(t), hipMemcpyHostToDevice));
  return 0;
}

// The taps of one pass of image 0.25's resize (imageops/sample.rs,
// vertical_sample / horizontal_sample) with the Triangle filter (support 1):
// for output index o, the source window [left, right) around the centre
// (o + 0.5) * ratio and the normalised weights triangle((i - centre + 0.5) /
// sratio), all in f32 as the crate computes them.
struct Taps {
  std::vector<int32_t> lc;  // (left, count) per output index
  std::vector<float> w;     // K per output index
  int K = 1;
};
Taps resize_taps(uint32_t src, uint32_t dst) {
  Taps t;
  const float ratio = (float)src / (float)dst;
  const float sratio = ratio < 1.0f ? 1.0f : ratio;
  const float support = 1.0f * sratio;
  auto clampi = [](int64_t a, int64_t lo, int64_t hi) { return a < lo ? lo : (a > hi ? hi : a); };
  auto as_i64 = [](float v) -> int64_t {  // Rust `as i64` of a floored / ceiled value
    if (v != v) return 0;
    if (v >= 9.2e18f) return INT64_MAX;
    if (v <= -9.2e18f) return INT64_MIN;
    return (int64_t)v;
  };
  std::vector<std::vector<float>> ws(dst);
  t.lc.resize(2 * (size_t)dst);
  for (uint32_t o = 0; o < dst; ++o) {
    float centre = ((float)o + 0.5f) * ratio;
    const int64_t left = clampi(as_i64(floorf(centre - support)), 0, (int64_t)src - 1);
    const int64_t right = clampi(as_i64(ceilf(centre + support)), left + 1, (int64_t)src);
    centre = centre - 0.5f;
    float sum = 0.0f;
    for (int64_t i = left; i < right; ++i) {
      const float x = ((float)i - centre) / sratio;
      const float ax = fabsf(x);
      const float wv = ax < 1.0f ? 1.0f - ax : 0.0f;
      ws[o].push_back(wv);
      sum += wv;
    }
    for (float& wv : ws[o]) wv /= sum;
    t.lc[2 * o] = (int32_t)left;
    t.lc[2 * o + 1] = (int32_t)(right - left);
    t.K = std::max(t.K, (int)(right - left));
  }
  t.w.assign((size_t)dst * t.K, 0.0f);
  for (uint32_t o = 0; o < dst; ++o)
    for (size_t i = 0; i < ws[o].size(); ++i) t.w[(size_t)o * t.K + i] = ws[o][i];
  return t;
}

// Resizes the RGBA8 image at src (device, sw x sh) into dst (device, dw x
// dh) with image 0.25's resize_exact(.., Triangle).
int resize_device(rgs_ctx* c, const uint8_t* src, uint32_t sw, uint32_t sh, uint8_t* dst, uint32_t dw,
                  uint32_t dh) {
  if (dh > 65535) return fail(RGS_E_ARG, "resize: output taller than 65535 rows");  // grid rows
  const Taps v = resize_taps(sh, dh), h = resize_taps(sw, dw);
  const size_t vb = v.lc.size() * 4, vwb = v.w.size() * 4, hb = h.lc.size() * 4, hwb = h.w.size() * 4;
  CHK(ensure(c, c->rs_taps, vb + vwb + hb + hwb));
  CHK(ensure(c, c->rs_tmp, (size_t)sw * dh * 16));
  std::vector<uint8_t> blob(vb + vwb + hb + hwb);
  std::memcpy(blob.data(), v.lc.data(), vb);
  std::memcpy(blob.data() + vb, v.w.data(), vwb);
  std::memcpy(blob.data() + vb + vwb, h.lc.data(), hb);
  std::memcpy(blob.data() + vb + vwb + hb, h.w.data(), hwb);
  HIPCHK(hipMemcpyAsync(c->rs_taps.p, blob.data(), blob.size(), hipMemcpyHostToDevice, c->st));
  uint8_t* base = P_<uint8_t>(c->rs_taps);
  launch_resize_rgba8(src, sw, sh, dst, dw, dh, reinterpret_cast<const int2*>(base),
                      reinterpret_cast<const float*>(base + vb), v.K, reinterpret_cast<const int2*>(base + vb + vwb),
                      reinterpret_cast<const float*>(base + vb + vwb + hb), h.K, P_<float>(c->rs_tmp), c->st);
  HIPCHK(hipStreamSynchronize(c->st));  // the tap blob is pageable host memory
  return 0;
}
}  // namespace
extern "C" {

int rgs_resize_rgba8(rgs_ctx* c, const uint8_t* src, uint32_t sw, uint32_t sh, uint8_t* dst, uint32_t dw,
                     uint32_t dh) {
  if (!c || !src || !dst) return fail(RGS_E_ARG, "null argument");
  if (!sw || !sh || !dw || !dh) return fail(RGS_E_ARG, "zero image size");
  (void)hipSetDevice(c->device);
  const size_t sb = (size_t)sw * sh * 4, db = (size_t)dw * dh * 4;
  CHK(ensure(c, c->rs_src, sb));
  CHK(ensure(c, c->rs_dst, db));
  HIPCHK(hipMemcpyAsync(c->rs_src.p, src, sb, hipMemcpyHostToDevice, c->st));
  if (sw == dw && sh == dh) {  // resize copies when the size is unchanged
    HIPCHK(hipMemcpyAsync(dst, c->rs_src.p, sb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return 0;
  }
  CHK(resize_device(c, P_<uint8_t>(c->rs_src), sw, sh, P_<uint8_t>(c->rs_dst), dw, dh));
  HIPCHK(hipMemcpy(dst, c->rs_dst.p, db, hipMemcpyDeviceToHost));
  return 0;
}

int rgs_set_views_rgba8_sized(rgs_ctx* c, const rgs_camera* cams, const uint8_t* const* images,
                              const uint32_t* sizes, int nv) {
  if (!c || nv <= 0 || !cams || !images || !sizes) return fail(RGS_E_ARG, "null argument or no views");
  (void)hipSetDevice(c->device);
  drop_graph(c);
  std::vector<CamDev> v(nv);
  size_t tot = 0;
  for (int i = 0; i < nv; ++i) {
    if (!images[i]) return fail(RGS_E_ARG, "null image");
    if (!sizes[2 * i] || !sizes[2 * i + 1]) return fail(RGS_E_ARG, "zero image size");
    CHK(make_camdev(cams[i], v[i]));
    v[i].tgt_base = (uint32_t)tot;
    tot += (size_t)cams[i].width * cams[i].height;
  }
  if (tot >= (1ull << 32)) return fail(RGS_E_ARG, "targets too large");
  CHK(srgb_table(c));
  CHK(ensure(c, c->targets, tot * 16));
  CHK(ensure(c, c->io_soa, tot * 4));
  for (int i = 0; i < nv; ++i) {
    const uint32_t sw = sizes[2 * i], sh = sizes[2 * i + 1];
    uint8_t* dst = P_<uint8_t>(c->io_soa) + (size_t)v[i].tgt_base * 4;
    if (sw == cams[i].width && sh == cams[i].height) {
      HIPCHK(hipMemcpyAsync(dst, images[i], (size_t)sw * sh * 4, hipMemcpyHostToDevice, c->st));
    } else {
      CHK(ensure(c, c->rs_src, (size_t)sw * sh * 4));
      HIPCHK(hipMemcpyAsync(c->rs_src.p, images[i], (size_t)sw * sh * 4, hipMemcpyHostToDevice, c->st));
      CHK(resize_device(c, P_<uint8_t>(c->rs_src), sw, sh, dst, cams[i].width, cams[i].height));
    }
  }
  launch_rgba8_to_linear(P_<uint8_t>(c->io_soa), tot, P_<float>(c->srgb_lut), P_<float>(c->targets), c->st);
  return finish_views(c, v, tot);
}

int rgs_rgba8_to_linear(rgs_ctx* c, const uint8_t* rgba8, size_t n_pixels, float* out) {
  if (!c || (n_pixels && (!rgba8 || !out))) return fail(RGS_E_ARG, "null argument");
  if (!n_pixels) return 0;
  (void)hipSetDevice(c->device);
  CHK(srgb_table(c));
  CHK(ensure(c, c->io_soa, n_pixels * 4));
  CHK(ensure(c, c->io_rec, n_pixels * 16));
  HIPCHK(hipMemcpyAsync(c->io_soa.p, rgba8, n_pixels * 4, hipMemcpyHostToDevice, c->st));
  launch_rgba8_to_linear(P_<uint8_t>(c->io_soa), n_pixels, P_<float>(c->srgb_lut), P_<float>(c->io_rec), c->st);
  HIPCHK(hipMemcpyAsync(out, c->io_rec.p, n_pixels * 16, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int rgs_to_u8_srgb(rgs_ctx* c, const float* rgba, size_t n_pixels, uint8_t* out) {
  if (!c || (n_pixels && (!rgba || !out))) return fail(RGS_E_ARG, "null argument");
  if (!n_pixels) return 0;
  (void)hipSetDevice(c->device);
  CHK(ensure(c, c->io_rec, n_pixels * 16));
  CHK(ensure(c, c->io_soa, n_pixels * 4));
  HIPCHK(hipMemcpyAsync(c->io_rec.p, rgba, n_pixels * 16, hipMemcpyHostToDevice, c->st));
  launch_to_u8_srgb(P_<float>(c->io_rec), n_pixels, P_<uint8_t>(c->io_soa), c->st);
  HIPCHK(hipMemcpyAsync(out, c->io_soa.p, n_pixels * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int rgs_set_views(rgs_ctx* c, const rgs_camera* cams, const float* const* targets, int nv) {
  if (!c || nv <= 0 || !cams || !targets) return fail(RGS_E_ARG, "null argument or no views");
  (void)hipSetDevice(c->device);
  drop_graph(c);
  std::vector<CamDev> v(nv);
  size_t tot = 0;
  for (int i = 0; i < nv; ++i) {
    if (!targets[i]) return fail(RGS_E_ARG, "null target");
    CHK(make_camdev(cams[i], v[i]));
    v[i].tgt_base = (uint32_t)tot;
    tot += (size_t)cams[i].width * cams[i].height;
  }
  if (tot >= (1ull << 32)) return fail(RGS_E_ARG, "targets too large");
  CHK(ensure(c, c->targets, tot * 16));
  for (int i = 0; i < nv; ++i)
    HIPCHK(hipMemcpyAsync(P_<float>(c->targets) + (size_t)v[i].tgt_base * 4, targets[i],
                          (size_t)cams[i].width * cams[i].height * 16, hipMemcpyHostToDevice, c->st));
  return finish_views(c, v, tot);
}

int rgs_set_views_rgba8(rgs_ctx* c, const rgs_camera* cams, const uint8_t* const* images, int nv) {
  if (!c || nv <= 0 || !cams || !images) return fail(RGS_E_ARG, "null argument or no views");
  (void)hipSetDevice(c->device);
  drop_graph(c);
  std::vector<CamDev> v(nv);
  size_t tot = 0;
  for (int i = 0; i < nv; ++i) {
    if (!images[i]) return fail(RGS_E_ARG, "null image");
    CHK(make_camdev(cams[i], v[i]));
    v[i].tgt_base = (uint32_t)tot;
    tot += (size_t)cams[i].width * cams[i].height;
  }
  if (tot >= (1ull << 32)) return fail(RGS_E_ARG, "targets too large");
  CHK(srgb_table(c));
  CHK(ensure(c, c->targets, tot * 16));
  CHK(ensure(c, c->io_soa, tot * 4));
  for (int i = 0; i < nv; ++i)
    HIPCHK(hipMemcpyAsync(P_<uint8_t>(c->io_soa) + (size_t)v[i].tgt_base * 4, images[i],
                          (size_t)cams[i].width * cams[i].height * 4, hipMemcpyHostToDevice, c->st));
  launch_rgba8_to_linear(P_<uint8_t>(c->io_soa), tot, P_<float>(c->srgb_lut), P_<float>(c->targets), c->st);
  return finish_views(c, v, tot);
}

// Enqueues the step after the projection/depth-sort/scan prefix.
// Tiles the training step's tight binning may leave empty although the
// reference composites them: only a non-zero background tells them apart.
static uint8_t* step_cover(rgs_ctx* c, const float bg[3]) {
  return (bg[0] != 0.0f || bg[1] != 0.0f || bg[2] != 0.0f) ? P_<uint8_t>(c->cover) : nullptr;
}

static void enqueue_step_rest(rgs_ctx* c, const ViewSet& vs, const float bg[3], float lambda, bool with_back) {
  int n = (int)c->n;
  size_t cap = c->cap;
  uint32_t* sc = P_<uint32_t>(c->scalars);
  enqueue_bin_and_raster_cut(c, cap, n, vs, bg, P_<float>(c->rgb), step_cover(c, bg), with_back);
  stage_begin(c, S_LOSS);
  enqueue_loss(c, vs, P_<float>(c->rgb), P_<float>(c->targets), lambda, P_<float>(c->tstats));
  stage_end(c, S_LOSS);
  stage_begin(c, S_BWD);
  RasterIn in = raster_in(c, vs, bg);
  lpt_backward(c, vs, in);
  in.queue = bwd_queue(c);
  in.acc_stride = vs.cams.size() * cap;
  if (with_back) {
    in.bstart = P_<uint32_t>(c->rbstart);
    in.bend = P_<uint32_t>(c->rbend);
    in.back_vals = c->back_vals;
  }
  launch_raster_bwd(in, P_<float>(c->final_t), P_<int32_t>(c->last), P_<float>(c->dl), P_<float>(c->accum), c->st);
  stage_end(c, S_BWD);
  stage_begin(c, S_ADAM);
  launch_train_adam(P_<float>(c->P), P_<float>(c->M1), P_<float>(c->M2), (int)cap, n, P_<CamDev>(c->cams_d),
                    (int)vs.cams.size(), P_<uint32_t>(c->cnt), P_<float>(c->accum), vs.cams.size() * cap,
                    P_<float>(c->gacc), P_<uint32_t>(c->gden), P_<AdamConsts>(c->adam_d), sc + SC_STATUS, c->st,
                    c->grad_capture ? P_<float>(c->gcap) : nullptr);
  stage_end(c, S_ADAM);
}

// The adaptive cut pays for itself by sorting and binning only the entries
// in front of it; when some tiles never saturate, their bins need the
// back-list stage every step (about 30 launches, ~0.13 ms). After a step
// that needed it while the cut left fewer than an eighth as many visible
// entries behind it as in front (little saved), the cut pauses for CUT_PAUSE steps
// (whole lists, no back-list stage), then resumes with the back-list stage
// on its first two steps (no step voided for want of it).
constexpr int CUT_PAUSE = 256;
static int adapt_cut(rgs_ctx* c) {
  if (c->cut_mode != 0) return 0;
  if (c->cut_paused > 0) {
    if (--c->cut_paused == 0) {
      CHK(write_cut_state(c));
      c->force_back = 2;
    }
    return 0;
  }
  const uint64_t nf = c->h_scalars[SC_NF], nbv = c->h_scalars[SC_NBV];
  if (c->h_scalars[SC_NFLAG] != 0 && nbv * 8 <= nf) {
    c->cut_paused = CUT_PAUSE;
    c->need_back = false;
    CHK(write_cut_state(c));
  }
  return 0;
}

// Enqueues one whole training step on the context's stream (no host sync).
static void enqueue_step(rgs_ctx* c, const ViewSet& vs, const float bg[3], float lambda, bool with_back) {
  // one launch clears what the step accumulates into or flags: the scalars
  // and look-back states, the bin and tile flags of the depth cut, the tile
  // ranges (the tile sort's atomic minima and maxima), the persistent-raster
  // queue counters and the coverage of a background step
  FillSegs f{};
  f.add(P_<uint32_t>(c->scalars), step_scalar_words(c), 0u);
  f.add(P_<uint32_t>(c->flag), ((size_t)vs.total_bins + 3) / 4, 0u);
  f.add(P_<uint32_t>(c->tile_flag), ((size_t)vs.total_tiles + 3) / 4, 0u);
  f.add(P_<uint32_t>(c->rstart), (size_t)vs.total_bins, 0xffffffffu);
  f.add(P_<uint32_t>(c->rend), (size_t)vs.total_bins, 0u);
  if (c->raster_queue) f.add(P_<uint32_t>(c->rqueue), RQ_WORDS, 0u);
  if (uint8_t* cover = step_cover(c, bg)) f.add(reinterpret_cast<uint32_t*>(cover), ((size_t)vs.total_tiles + 3) / 4, 0u);
  // the state of the front depth sort's single-pass digit offsets
  f.add(P_<uint32_t>(c->os), radix_os_used_words(radix_nb(vs.cams.size() * c->n), 4), 0u);
  launch_fill_multi(f, c->st);
  enqueue_project_cut(c, P_<float>(c->P), c->cap, (int)c->n, vs, step_cover(c, bg), with_back);
  enqueue_step_rest(c, vs, bg, lambda, with_back);
}

static int train_step_impl(rgs_ctx* c, const int* ids, int n_ids, const rgs_train_cfg* cfg, const float bg[3],
                           float lambda, int use_graph, float* loss_out);

// A failed step leaves the parameters untouched (the device skips the update
// on every status the host turns into an error), so the Adam step count goes
// back too: the next step's bias corrections are those it would have had.
int rgs_train_step(rgs_ctx* c, const int* ids, int n_ids, const rgs_train_cfg* cfg, const float bg[3],
                   float lambda, int use_graph, float* loss_out) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  const uint64_t t0 = c->t;
  int rc;
  try {
    rc = train_step_impl(c, ids, n_ids, cfg, bg, lambda, use_graph, loss_out);
  } catch (const std::exception& e) {
    rc = fail(RGS_E_HOST, std::string("host failure in the training step: ") + e.what());
  }
  if (rc) c->t = t0;
  return rc;
}

static int train_step_impl(rgs_ctx* c, const int* ids, int n_ids, const rgs_train_cfg* cfg, const float bg[3],
                           float lambda, int use_graph, float* loss_out) {
  if (!ids || n_ids <= 0 || !cfg || !bg) return fail(RGS_E_ARG, "null argument");
  c->fwd_valid = false;
  if (n_ids > 64) return fail(RGS_E_ARG, "at most 64 views per step");
  if (!c->P.p) return fail(RGS_E_ARG, "no scene uploaded");
  (void)hipSetDevice(c->device);
  ViewSet vs;
  std::string key;
  for (int i = 0; i < n_ids; ++i) {
    if (ids[i] < 0 || ids[i] >= (int)c->reg_cams.size()) return fail(RGS_E_ARG, "view id out of range");
    vs.cams.push_back(c->reg_cams[ids[i]]);
    key += std::to_string(c->reg_cams[ids[i]].w) + "x" + std::to_string(c->reg_cams[ids[i]].h) + ",";
  }
  vs.shift = c->bin_shift;
  vs.layout();
  size_t cap = c->cap;
  int n = (int)c->n;
  char buf[160];
  std::snprintf(buf, sizeof(buf), "|n%d|bg%a,%a,%a|l%a|s%d", n, bg[0], bg[1], bg[2], lambda, vs.shift);
  key += buf;

  c->t += 1;
  *c->h_adam = adam_consts(*cfg, c->t);
  CHK(ensure(c, c->adam_d, sizeof(AdamConsts)));
  size_t acc_n = 9 * vs.cams.size() * cap;
  bool fresh_accum = c->accum.bytes < acc_n * 4;
  CHK(ensure(c, c->accum, acc_n * 4));
  CHK(ensure(c, c->cover, ((size_t)vs.total_tiles + 3) / 4 * 4));
  if (c->super_tile == 0) CHK(ensure_lpt(c, vs));
  if (!c->cutst.p) {
    CHK(ensure(c, c->cutst, sizeof(CutState)));
    CHK(write_cut_state(c));
  }
  if (fresh_accum) launch_fill_f32(P_<float>(c->accum), acc_n, 0.0f, c->st);
  CHK(ensure_loss(c, vs));
  std::memcpy(c->h_cams, vs.cams.data(), vs.cams.size() * sizeof(CamDev));

  if (c->profiling && c->prof_graph_failed) use_graph = 0;
  if (!c->st) use_graph = 0;  // no capture on the null stream
  if (c->profiling) key += "|prof" + std::to_string(c->prof_mask);
  if (c->grad_capture) {
    CHK(ensure(c, c->gcap, (size_t)60 * cap * 4));
    key += "|gcap";
  }
  for (int attempt = 0; attempt < 4; ++attempt) {
    const int var = (c->need_back || c->force_back > 0) ? 1 : 0;
    bool replay = use_graph && c->gexec[var] && c->gkey[var] == key && c->graph_gen[var] == c->gen;
    bool launched_graph = replay;
    if (!replay) {
      if (c->gexec[var]) {
        (void)hipGraphExecDestroy(c->gexec[var]);
        c->gexec[var] = nullptr;
      }
      // eager: exact pair count sizes the buffers, then the rest of the step
      CHK(ensure_render(c, cap, n, vs, std::max<size_t>(c->k_cap, 1)));
      HIPCHK(hipMemcpyAsync(c->cams_d.p, c->h_cams, vs.cams.size() * sizeof(CamDev), hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->adam_d.p, c->h_adam, sizeof(AdamConsts), hipMemcpyHostToDevice, c->st));
      uint32_t* sc = P_<uint32_t>(c->scalars);
      launch_fill_u32(sc, SC_COUNT, 0u, c->st);
      enqueue_project(c, P_<float>(c->P), cap, n, vs, true, step_cover(c, bg));
      HIPCHK(hipMemcpyAsync(c->h_scalars, sc, SC_COUNT * 4, hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      if (c->h_scalars[SC_STATUS] & ST_NAN_DEPTH) return fail(RGS_E_NAN, "NaN depth in projection");
      size_t K = c->h_scalars[SC_TOTAL];
      CHK(check_pairs(K));
      size_t want = K + K / 4 + 4096;  // headroom for the following steps
      if (K > c->k_cap) CHK(ensure_render(c, cap, n, vs, std::max(want, c->k_cap)));
      if (use_graph) {
        // capture the full step for replay (the projection is re-run inside)
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        enqueue_step(c, vs, bg, lambda, var == 1);
        hipError_t e = hipStreamEndCapture(c->st, &g);
        if (e == hipSuccess) {
          e = hipGraphInstantiate(&c->gexec[var], g, nullptr, nullptr, 0);
          (void)hipGraphDestroy(g);
          if (e != hipSuccess) c->gexec[var] = nullptr;
        }
        if (e != hipSuccess && c->profiling) {
          // the runtime refused event nodes: profile this context eagerly
          (void)hipGetLastError();
          c->prof_graph_failed = true;
          use_graph = 0;
          enqueue_step(c, vs, bg, lambda, var == 1);
        } else if (e != hipSuccess) {
          return fail(RGS_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        } else {
          c->gkey[var] = key;
          c->graph_gen[var] = c->gen;
          HIPCHK(hipGraphLaunch(c->gexec[var], c->st));
          launched_graph = true;
        }
      } else {
        enqueue_step(c, vs, bg, lambda, var == 1);
      }
    } else {
      HIPCHK(hipMemcpyAsync(c->cams_d.p, c->h_cams, vs.cams.size() * sizeof(CamDev), hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->adam_d.p, c->h_adam, sizeof(AdamConsts), hipMemcpyHostToDevice, c->st));
      HIPCHK(hipGraphLaunch(c->gexec[var], c->st));
    }
    HIPCHK(hipMemcpyAsync(c->h_scalars, c->scalars.p, SC_COUNT * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(c->h_loss, c->loss.p, vs.cams.size() * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    uint32_t status = c->h_scalars[SC_STATUS];
    if (status & (ST_NAN_DEPTH | ST_SCAN_STALL)) {
      // the update was skipped on the device; the raster's sums are void
      launch_fill_f32(P_<float>(c->accum), acc_n, 0.0f, c->st);
      (void)hipStreamSynchronize(c->st);
      if (status & ST_NAN_DEPTH) return fail(RGS_E_NAN, "NaN depth in projection");
      return fail(RGS_E_HIP, "a single-pass scan's look-back gave up waiting");
    }
    if (status & ST_PAIR_OVERFLOW) {
      // the step was skipped on the device (parameters untouched): grow and redo
      size_t K = std::max<size_t>(c->h_scalars[SC_TOTAL], c->h_scalars[SC_BTOTAL]);
      CHK(check_pairs(K));
      drop_graph(c);
      CHK(ensure_render(c, cap, n, vs, K + K / 2 + 4096));
      launch_fill_f32(P_<float>(c->accum), acc_n, 0.0f, c->st);
      continue;
    }
    if (status & ST_NEED_BACK) {
      // a bin was left unsaturated by its front list in a step run without
      // the back-list stage: the update was skipped on the device; redo the
      // step with the stage
      c->need_back = true;
      c->back_reruns += 1;
      launch_fill_f32(P_<float>(c->accum), acc_n, 0.0f, c->st);
      continue;
    }
    c->need_back = c->h_scalars[SC_NFLAG] != 0;
    if (c->force_back > 0) --c->force_back;
    CHK(adapt_cut(c));
    c->last_back_valid = var == 1;
    c->last_front_pairs = c->h_scalars[SC_K];
    c->last_back_pairs = c->h_scalars[SC_KB];
    c->last_flagged = c->h_scalars[SC_NFLAG];
    c->last_pairs = c->last_front_pairs + c->last_back_pairs;
    c->last_cut = true;
    c->last_cap = cap;
    c->last_tiles = vs.total_bins;
    c->last_shift = vs.shift;
    c->last_ntiles = vs.total_tiles;
    c->last_pix = vs.total_pix;
    c->last_nviews = (int)vs.cams.size();
    for (int k = 0; k < 3; ++k) c->last_bg[k] = bg[k];
    if (c->profiling) {
      for (int s = 0; s < RGS_N_STAGES; ++s) {
        if (!(c->prof_mask >> s & 1u)) continue;
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev_b[s], c->ev_e[s]));
        c->stage_ms[s] += ms;
      }
      c->prof_steps += 1;
      if (launched_graph) c->prof_graph_steps += 1;
    }
    if (loss_out)
      for (int i = 0; i < n_ids; ++i) loss_out[i] = c->h_loss[i];
    return 0;
  }
  return fail(RGS_E_HIP, "the step kept overflowing its pair buffers");
}

int rgs_autotune_raster(rgs_ctx* c, const int* ids, int n_ids, const float bg[3], float lambda, int* super_tile_out,
                        int* sparse_lanes_out, float* ms_out) {
  if (!c || !ids || n_ids <= 0 || !bg) return fail(RGS_E_ARG, "null argument or no views");
  if (n_ids > 64) return fail(RGS_E_ARG, "at most 64 views per step");
  if (!c->P.p) return fail(RGS_E_ARG, "no scene uploaded");
  if (!c->st) return fail(RGS_E_ARG, "the autotuner needs a non-null stream");
  if (!c->targets.p) return fail(RGS_E_ARG, "no views registered");
  (void)hipSetDevice(c->device);
  ViewSet vs;
  for (int i = 0; i < n_ids; ++i) {
    if (ids[i] < 0 || ids[i] >= (int)c->reg_cams.size()) return fail(RGS_E_ARG, "view id out of range");
    vs.cams.push_back(c->reg_cams[ids[i]]);
  }
  vs.shift = c->bin_shift;
  vs.layout();
  drop_graph(c);
  const size_t cap = c->cap;
  const int n = (int)c->n;
  const size_t acc_n = 9 * vs.cams.size() * cap;
  CHK(ensure(c, c->accum, acc_n * 4));
  CHK(ensure_loss(c, vs));
  CHK(ensure_render(c, cap, n, vs, std::max<size_t>(c->k_cap, 1)));
  std::memcpy(c->h_cams, vs.cams.data(), vs.cams.size() * sizeof(CamDev));
  HIPCHK(hipMemcpyAsync(c->cams_d.p, c->h_cams, vs.cams.size() * sizeof(CamDev), hipMemcpyHostToDevice, c->st));
  uint32_t* sc = P_<uint32_t>(c->scalars);
  launch_fill_u32(sc, SC_COUNT, 0u, c->st);
  const bool prof = c->profiling;
  c->profiling = false;
  enqueue_project(c, P_<float>(c->P), cap, n, vs, true);
  HIPCHK(hipMemcpyAsync(c->h_scalars, sc, SC_COUNT * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  size_t K = c->h_scalars[SC_TOTAL];
  CHK(check_pairs(K));
  if (K > c->k_cap) CHK(ensure_render(c, cap, n, vs, K + K / 4 + 4096));
  enqueue_bin_and_raster(c, cap, n, vs, bg, 0, P_<float>(c->rgb));
  enqueue_loss(c, vs, P_<float>(c->rgb), P_<float>(c->targets), lambda, P_<float>(c->tstats));
  hipEvent_t ev[4];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  auto time_once = [&](float* ms) -> int {
    CHK(ensure_tile_order(c, vs));
    if (c->super_tile == 0) CHK(ensure_lpt(c, vs));
    RasterIn in = raster_in(c, vs, bg);
    in.acc_stride = vs.cams.size() * cap;
    lpt_forward(c, in);
    launch_fill_u32(P_<uint32_t>(c->rqueue), RQ_WORDS, 0u, c->st);
    in.queue = fwd_queue(c);
    HIPCHK(hipEventRecord(ev[0], c->st));
    launch_raster_fwd(in, 3, P_<float>(c->rgb), P_<float>(c->final_t), P_<int32_t>(c->last), c->st);
    HIPCHK(hipEventRecord(ev[1], c->st));
    HIPCHK(hipEventRecord(ev[2], c->st));
    lpt_backward(c, vs, in);
    in.queue = bwd_queue(c);
    launch_raster_bwd(in, P_<float>(c->final_t), P_<int32_t>(c->last), P_<float>(c->dl), P_<float>(c->accum), c->st);
    HIPCHK(hipEventRecord(ev[3], c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    float a = 0.0f, b = 0.0f;
    HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
    HIPCHK(hipEventElapsedTime(&b, ev[2], ev[3]));
    *ms = a + b;
    return 0;
  };
  auto score = [&](int st_, int sl, int fb, int q, float* best) -> int {
    c->super_tile = st_;
    c->sparse_lanes = sl;
    c->raster_batch = fb;
    c->raster_queue = q;
    float t = 0.0f, m = 3.4e38f;
    for (int r = 0; r < 2; ++r) {
      CHK(time_once(&t));
      m = std::min(m, t);
    }
    *best = m;
    return 0;
  };
  int best_st = c->super_tile, best_sl = c->sparse_lanes, best_fb = c->raster_batch, best_q = c->raster_queue;
  float best_ms = 0.0f;
  int rc = score(best_st, best_sl, best_fb, best_q, &best_ms);
  // coordinate search: schedule patch size, the sparse/dense switch, then
  // (the binned training raster only) the staging batch and the persistent
  // queue, then the schedule once more under the chosen queue (a queue
  // changes which bins run together)
  const int st_cand[] = {0, 2, 4, 8, 16, 1024};
  const int sl_cand[] = {0, 2, 4, 8, 16, 32};
  const int fb_cand[] = {32, 64, 128};
  const int q_cand[] = {0, 1, 2, 3};
  for (int st_ : st_cand) {
    if (rc) break;
    float m = 0.0f;
    rc = score(st_, best_sl, best_fb, best_q, &m);
    if (!rc && m < best_ms) best_ms = m, best_st = st_;
  }
  for (int sl : sl_cand) {
    if (rc) break;
    float m = 0.0f;
    rc = score(best_st, sl, best_fb, best_q, &m);
    if (!rc && m < best_ms) best_ms = m, best_sl = sl;
  }
  for (int fb : fb_cand) {
    if (rc || c->bin_shift == 0) break;
    float m = 0.0f;
    rc = score(best_st, best_sl, fb, best_q, &m);
    if (!rc && m < best_ms) best_ms = m, best_fb = fb;
  }
  const int q0 = best_q;
  for (int q : q_cand) {
    if (rc || c->bin_shift == 0) break;
    float m = 0.0f;
    rc = score(best_st, best_sl, best_fb, q, &m);
    if (!rc && m < best_ms) best_ms = m, best_q = q;
  }
  for (int st_ : st_cand) {
    if (rc || best_q == q0) break;
    float m = 0.0f;
    rc = score(st_, best_sl, best_fb, best_q, &m);
    if (!rc && m < best_ms) best_ms = m, best_st = st_;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  c->profiling = prof;
  c->super_tile = best_st;
  c->sparse_lanes = best_sl;
  c->raster_batch = best_fb;
  c->raster_queue = best_q;
  if (rc) return rc;
  CHK(ensure_tile_order(c, vs));
  launch_fill_f32(P_<float>(c->accum), acc_n, 0.0f, c->st);
  HIPCHK(hipStreamSynchronize(c->st));
  drop_graph(c);
  if (super_tile_out) *super_tile_out = best_st;
  if (sparse_lanes_out) *sparse_lanes_out = best_sl;
  if (ms_out) *ms_out = best_ms;
  return 0;
}

int rgs_scene_densify(rgs_ctx* c, const rgs_train_cfg* cfg, float extent) {
  if (!c || !cfg) return fail(RGS_E_ARG, "null argument");
  (void)hipSetDevice(c->device);
  size_t n = c->n;
  CHK(ensure_densify_scratch(c, n ? n : 1));
  uint32_t* sc = P_<uint32_t>(c->scalars);
  launch_densify(P_<float>(c->P), P_<float>(c->P2), (int)c->cap, (int)n, P_<float>(c->gacc), P_<uint32_t>(c->gden),
                 cfg->densify_grad_threshold, cfg->prune_opacity_thresh, cfg->max_gaussians, extent, dscratch(c),
                 sc + SC_NOUT, c->st, c->split_div);
  HIPCHK(hipMemcpyAsync(c->h_scalars, sc, SC_COUNT * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  size_t nn = c->h_scalars[SC_NOUT];
  if (nn > c->cap) return fail(RGS_E_CAPACITY, "densify result exceeds the scene capacity");
  std::swap(c->P, c->P2);
  if (nn > c->m_rows) {
    launch_zero_rows(P_<float>(c->M1), c->cap, c->m_rows, nn, c->st);
    launch_zero_rows(P_<float>(c->M2), c->cap, c->m_rows, nn, c->st);
    c->m_rows = nn;
  }
  launch_fill_f32(P_<float>(c->gacc), c->m_rows, 0.0f, c->st);
  launch_fill_u32(P_<uint32_t>(c->gden), c->m_rows, 0u, c->st);
  HIPCHK(hipStreamSynchronize(c->st));
  c->n = nn;
  drop_graph(c);
  return 0;
}

int rgs_scene_reset_opacities(rgs_ctx* c) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  (void)hipSetDevice(c->device);
  if (c->n) launch_reset_opacity(P_<float>(c->P) + (size_t)P_OP * c->cap, c->n, 1, c->st);
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int rgs_set_profiling(rgs_ctx* c, int enable) {
  return rgs_set_profiling_stages(c, enable ? (1u << RGS_N_STAGES) - 1u : 0u);
}

int rgs_set_profiling_stages(rgs_ctx* c, uint32_t stage_mask) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  (void)hipSetDevice(c->device);
  stage_mask &= (1u << RGS_N_STAGES) - 1u;
  if (stage_mask && !c->ev_b[0]) {
    for (int s = 0; s < RGS_N_STAGES; ++s) {
      HIPCHK(hipEventCreate(&c->ev_b[s]));
      HIPCHK(hipEventCreate(&c->ev_e[s]));
    }
  }
  c->profiling = stage_mask != 0;
  c->prof_mask = stage_mask;
  for (int s = 0; s < RGS_N_STAGES; ++s) c->stage_ms[s] = 0.0;
  c->prof_steps = 0;
  c->prof_graph_steps = 0;
  return 0;
}

int rgs_stage_times(rgs_ctx* c, float* mean_ms, int n_stages, uint64_t* steps, uint64_t* graph_steps) {
  if (!c || !mean_ms) return fail(RGS_E_ARG, "null argument");
  for (int s = 0; s < n_stages && s < RGS_N_STAGES; ++s)
    mean_ms[s] = c->prof_steps ? (float)(c->stage_ms[s] / (double)c->prof_steps) : 0.0f;
  if (steps) *steps = c->prof_steps;
  if (graph_steps) *graph_steps = c->prof_graph_steps;
  return 0;
}

int rgs_step_stats(rgs_ctx* c, uint64_t* n_pairs, uint64_t* fwd_evals, uint64_t* bwd_evals) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  if (n_pairs) *n_pairs = c->last_pairs;
  if (fwd_evals || bwd_evals) {
    (void)hipSetDevice(c->device);
    CHK(ensure(c, c->api_den, 16));
    unsigned long long* d = P_<unsigned long long>(c->api_den);
    launch_fill_u32(P_<uint32_t>(c->api_den), 4, 0u, c->st);
    const float zero3[3] = {0.0f, 0.0f, 0.0f};
    RasterIn in = raster_in_n(c, c->last_nviews, c->last_tiles, c->last_shift, zero3);
    if (c->last_cut && c->last_back_valid) {
      in.bstart = P_<uint32_t>(c->rbstart);
      in.bend = P_<uint32_t>(c->rbend);
    }
    launch_work_counts(in, P_<float>(c->final_t), P_<int32_t>(c->last), d, c->last_ntiles, c->st);
    unsigned long long h[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(h, d, 16, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (fwd_evals) *fwd_evals = h[0];
    if (bwd_evals) *bwd_evals = h[1];
  }
  return 0;
}

int rgs_step_image(rgs_ctx* c, float* rgb, size_t n_floats) {
  if (!c || !rgb) return fail(RGS_E_ARG, "null argument");
  if (!c->last_cut) return fail(RGS_E_ARG, "no training step rendered yet");
  const size_t need = c->last_pix * 3;
  if (n_floats < need) return fail(RGS_E_ARG, "rgb too small");
  (void)hipSetDevice(c->device);
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(rgb, c->rgb.p, need * 4, hipMemcpyDeviceToHost));
  return 0;
}

int rgs_set_train_binning(rgs_ctx* c, int shift, int fwd_tile_blocks) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  if (shift < 0 || shift > 1) return fail(RGS_E_ARG, "bin shift must be 0 or 1");
  if (fwd_tile_blocks < 0 || fwd_tile_blocks > 1) return fail(RGS_E_ARG, "fwd_tile_blocks must be 0 or 1");
  c->bin_shift = shift;
  c->fwd_tile_blocks = fwd_tile_blocks;
  ++c->gen;  // a captured step holds the old launch shapes
  return 0;
}

int rgs_set_depth_cut(rgs_ctx* c, int mode, float depth) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  if (mode < 0 || mode > 2) return fail(RGS_E_ARG, "depth cut mode must be 0, 1 or 2");
  if (mode == 2 && !(depth > 0.0f)) return fail(RGS_E_ARG, "fixed depth cut needs a positive depth");
  (void)hipSetDevice(c->device);
  c->cut_mode = mode;
  c->cut_depth = depth;
  c->cut_paused = 0;
  c->force_back = 0;
  if (c->cutst.p) CHK(write_cut_state(c));
  return 0;
}

int rgs_cut_stats(rgs_ctx* c, uint64_t* front_pairs, uint64_t* back_pairs, uint64_t* flagged_tiles,
                  float* next_cut, uint64_t* reruns) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  if (reruns) *reruns = c->back_reruns;
  if (front_pairs) *front_pairs = c->last_front_pairs;
  if (back_pairs) *back_pairs = c->last_back_pairs;
  if (flagged_tiles) *flagged_tiles = c->last_flagged;
  if (next_cut) {
    *next_cut = __builtin_inff();
    if (c->cutst.p) {
      (void)hipSetDevice(c->device);
      CutState cs{};
      HIPCHK(hipMemcpyAsync(&cs, c->cutst.p, sizeof(cs), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      if (cs.key != 0xffffffffu) {
        uint32_t k = cs.key;
        std::memcpy(next_cut, &k, 4);
      }
    }
  }
  return 0;
}

int rgs_eval_expf(rgs_ctx* c, const float* x, float* out, size_t n) {
  if (!c || (n && (!x || !out))) return fail(RGS_E_ARG, "null argument");
  if (!n) return 0;
  (void)hipSetDevice(c->device);
  CHK(ensure(c, c->api_aos, n * 8));
  HIPCHK(hipMemcpyAsync(c->api_aos.p, x, n * 4, hipMemcpyHostToDevice, c->st));
  launch_eval_expf(P_<float>(c->api_aos), P_<float>(c->api_aos) + n, n, c->st);
  HIPCHK(hipMemcpyAsync(out, P_<float>(c->api_aos) + n, n * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int rgs_set_split_scale_divisor(rgs_ctx* c, float divisor) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  if (!(divisor == 0.0f || (divisor > 1.0f && divisor < 1e6f)))
    return fail(RGS_E_ARG, "split scale divisor must be 0 (the reference's split) or in (1, 1e6)");
  c->split_div = divisor;
  return 0;
}

int rgs_hbm_copy_probe(rgs_ctx* c, size_t bytes, int reps, float* gbs) {
  if (!c || !gbs || reps <= 0) return fail(RGS_E_ARG, "null argument or no repetitions");
  const size_t n16 = bytes / 16;
  if (n16 == 0) return fail(RGS_E_ARG, "probe size below 16 bytes");
  if (n16 > ((size_t)1 << 36)) return fail(RGS_E_ARG, "probe size above 1 TiB");
  if (!c->st) return fail(RGS_E_ARG, "the probe needs a non-null stream");
  (void)hipSetDevice(c->device);
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n16 * 16) != hipSuccess) return fail(RGS_E_HIP, "hipMalloc failed");
  if (hipMalloc(&b, n16 * 16) != hipSuccess) {
    (void)hipFree(a);
    return fail(RGS_E_HIP, "hipMalloc failed");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  float ms = 0.0f;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = fail(RGS_E_HIP, "hipEventCreate failed");
  } else {
    launch_fill_u32(static_cast<uint32_t*>(a), n16 * 4, 0u, c->st);
    for (int r = 0; r < 2; ++r) launch_copy16(a, b, n16, c->st);
    (void)hipEventRecord(e0, c->st);
    for (int r = 0; r < reps; ++r) launch_copy16(a, b, n16, c->st);
    (void)hipEventRecord(e1, c->st);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.0f)
      rc = fail(RGS_E_HIP, "probe timing failed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(c->st);
  (void)hipFree(a);
  (void)hipFree(b);
  if (rc) return rc;
  *gbs = (float)(2.0 * (double)(n16 * 16) / ((double)ms / reps * 1e-3) / 1e9);
  return 0;
}

int rgs_raster_work_counts(rgs_ctx* c, uint64_t* counts) {
  if (!c || !counts) return fail(RGS_E_ARG, "null argument");
  if (!c->last_cut) return fail(RGS_E_ARG, "no training step rendered yet");
  if (c->last_back_valid && c->last_back_pairs)
    return fail(RGS_E_ARG, "the last step resumed tiles on back lists; the counts cover front lists only");
  (void)hipSetDevice(c->device);
  CHK(ensure(c, c->work_counts, 6 * sizeof(unsigned long long)));
  launch_fill_u32(P_<uint32_t>(c->work_counts), 12, 0u, c->st);
  RasterIn in = raster_in_n(c, c->last_nviews, c->last_tiles, c->last_shift, c->last_bg);
  in.acc_stride = (size_t)c->last_nviews * c->last_cap;
  in.cover = step_cover(c, c->last_bg);
  if (c->last_back_valid) {
    in.bstart = P_<uint32_t>(c->rbstart);
    in.bend = P_<uint32_t>(c->rbend);
    in.back_vals = c->back_vals;
  }
  in.counts = P_<unsigned long long>(c->work_counts);
  launch_raster_counts(in, P_<float>(c->rgb), P_<float>(c->final_t), P_<int32_t>(c->last), P_<float>(c->dl), c->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(counts, c->work_counts.p, 6 * sizeof(uint64_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int rgs_set_grad_capture(rgs_ctx* c, int enable) {
  if (!c) return fail(RGS_E_ARG, "null ctx");
  c->grad_capture = enable != 0;
  return 0;
}

static int step_grads_impl(rgs_ctx* c, rgs_grad* out, size_t n);
int rgs_step_grads(rgs_ctx* c, rgs_grad* out, size_t n) {
  return guarded("rgs_step_grads", [&] { return step_grads_impl(c, out, n); });
}
static int step_grads_impl(rgs_ctx* c, rgs_grad* out, size_t n) {
  if (!c || !out) return fail(RGS_E_ARG, "null argument");
  if (!c->grad_capture || !c->gcap.p) return fail(RGS_E_ARG, "gradient capture is off or no step has run");
  if (n != c->n) return fail(RGS_E_ARG, "n must be the scene size");
  if (c->gcap.bytes < (size_t)60 * c->cap * 4) return fail(RGS_E_ARG, "no step ran since the scene grew");
  (void)hipSetDevice(c->device);
  HIPCHK(hipStreamSynchronize(c->st));
  std::vector<float> rows((size_t)60 * c->cap);
  HIPCHK(hipMemcpy(rows.data(), c->gcap.p, rows.size() * 4, hipMemcpyDeviceToHost));
  float* o = reinterpret_cast<float*>(out);
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < 60; ++j) o[i * 60 + j] = rows[j * c->cap + i];
  return 0;
}

static int debug_tiles_impl(rgs_ctx* c, uint64_t* offsets, size_t n_offsets, uint32_t* gis, size_t k_max,
                            uint64_t* k_out);
int rgs_debug_tiles(rgs_ctx* c, uint64_t* offsets, size_t n_offsets, uint32_t* gis, size_t k_max, uint64_t* k_out) {
  return guarded("rgs_debug_tiles", [&] { return debug_tiles_impl(c, offsets, n_offsets, gis, k_max, k_out); });
}
static int debug_tiles_impl(rgs_ctx* c, uint64_t* offsets, size_t n_offsets, uint32_t* gis, size_t k_max,
                            uint64_t* k_out) {
  if (!c || !offsets || !k_out) return fail(RGS_E_ARG, "null argument");
  if (c->last_shift != 0) return fail(RGS_E_ARG, "the last render binned per super-tile (a training step)");
  if (n_offsets < (size_t)c->last_tiles + 1) return fail(RGS_E_ARG, "offsets too small");
  (void)hipSetDevice(c->device);
  HIPCHK(hipStreamSynchronize(c->st));
  size_t nt = (size_t)c->last_tiles;
  std::vector<uint32_t> rs(nt), re(nt), pv(c->last_pairs);
  if (nt) {
    HIPCHK(hipMemcpy(rs.data(), c->rstart.p, nt * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(re.data(), c->rend.p, nt * 4, hipMemcpyDeviceToHost));
  }
  if (c->last_pairs) HIPCHK(hipMemcpy(pv.data(), c->pair_vals, c->last_pairs * 4, hipMemcpyDeviceToHost));
  size_t k = 0;
  for (size_t t = 0; t < nt; ++t) {
    offsets[t] = k;
    for (uint32_t p = rs[t]; p < re[t]; ++p) {
      if (gis && k < k_max) gis[k] = (uint32_t)(pv[p] % c->last_cap);
      ++k;
    }
  }
  offsets[nt] = k;
  *k_out = k;
  if (k > k_max && gis) return fail(RGS_E_ARG, "gis too small");
  return 0;
}

}  // extern "C"

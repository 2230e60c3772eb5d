// k_loss.hip -- L = (1 - lambda) L1 + lambda (1 - SSIM) and dL/dpixel.
//
// Follows backward.rs:646-807. SSIM uses the 11-tap separable Gaussian window
// (sigma 1.5) with zero padding. Its gradient with respect to x is the
// adjoint blur of five per-pixel coefficient maps (backward.rs:744-754),
//   g = B(P) + y B(Q) - B(Q mu_y) + 2 x B(R) - 2 B(R mu_x),
// which by linearity of the blur B is computed from three maps:
//   g = B(P - Q mu_y - 2 R mu_x) + y B(Q) + 2 x B(R).
//
// Two kernels compute the same dL/dpixel, bit for bit:
//   k_ssim_fwd + k_ssim_bwd (the training step and the reference-API entry
//     points): two passes over 32x32 output tiles with a 5-pixel halo staged
//     in LDS, one colour channel at a time. The forward blurs x, y, x^2,
//     y^2, xy (horizontal into LDS, then vertical), writes the three
//     coefficient maps, the L1 gradient and per-tile partial sums of |x - y|
//     and SSIM; the backward blurs the three maps and adds
//     -lambda dSSIM/dx / M. 36 bytes of maps per pixel go through HBM.
//   k_ssim (the entry point that allocates nothing): one pass per tile with
//     a 10-pixel halo; the products are blurred on the tile plus 5 pixels and
//     the maps never leave LDS. It needs no map buffer, at the price of
//     recomputing the products on the halo (0.15 ms per 1920x1080 view
//     against 0.12 ms for the two passes).
// Each thread makes several consecutive outputs of a row (horizontal blurs)
// or a column (vertical blurs) from one window held in registers, so every
// staged value is read once per window and the products are formed once per
// input. Taps are summed in order k = 0..10 (horizontal, then vertical) with
// fused multiply-adds; the per-pixel values agree with the reference's to
// float rounding. The loss scalar sums the per-tile partials in double.
#include "rgs_kernels.h"

namespace rgs {

constexpr int LT = 256;

constexpr int LH = LTILE + 10;  // 42: tile plus halo
constexpr int LN = LH * LH;     // staged values per map
constexpr int RUN = 4;          // outputs per thread per window
constexpr int WIN = RUN + 10;   // window length

struct Kern11 {
  float k[11];
};

struct LossTile {
  int v, tx, ty;  // view and loss-tile coordinates
};

// View v's record: from device memory, or the by-value copy of a one-view
// call (cams == nullptr).
__device__ __forceinline__ const CamDev& loss_cam(const LossIn& in, int v) {
  return in.cams ? in.cams[v] : in.cam0;
}

__device__ __forceinline__ LossTile loss_tile(const LossIn& in, int t) {
  int v = 0;
  while (v + 1 < in.n_views && loss_cam(in, v + 1).ltile_base <= t) ++v;
  const CamDev& c = loss_cam(in, v);
  const int nlx = (c.w + LTILE - 1) / LTILE;
  const int lt = t - c.ltile_base;
  return LossTile{v, lt % nlx, lt / nlx};
}

// Per-view constants, read once per block.
struct LossView {
  int w, h;
  size_t pb, tb;  // first pixel of the view in the image arrays / target store
  float m_total, lam;
};

__device__ __forceinline__ LossView loss_view(const LossIn& in, const LossTile& T) {
  const CamDev& c = loss_cam(in, T.v);
  LossView V;
  V.w = c.w;
  V.h = c.h;
  V.pb = (size_t)c.pix_base;
  V.tb = (size_t)c.tgt_base;
  V.m_total = (float)((size_t)c.w * c.h * 3);
  V.lam = in.lambda;
  return V;
}

constexpr int LPT = (LN + LT - 1) / LT;  // staged values per thread

// Pixel offset within the view of each staged value this thread loads (the
// tile plus a 5-pixel halo), -1 outside the image (zero padding).
__device__ __forceinline__ void stage_offsets(const LossView& V, const LossTile& T, int off[LPT]) {
  const int x0 = T.tx * LTILE - 5, y0 = T.ty * LTILE - 5;
#pragma unroll
  for (int i = 0; i < LPT; ++i) {
    const int idx = (int)threadIdx.x + i * LT;
    const int gy = y0 + idx / LH, gx = x0 + idx % LH;
    off[i] = (idx < LN && gx >= 0 && gy >= 0 && gx < V.w && gy < V.h) ? gy * V.w + gx : -1;
  }
}

__device__ __forceinline__ float block_sum_256(float x, float* s_tmp) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
  if ((threadIdx.x & 63) == 0) s_tmp[threadIdx.x >> 6] = x;
  __syncthreads();
  float r = (s_tmp[0] + s_tmp[1]) + (s_tmp[2] + s_tmp[3]);
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(LT) void k_ssim_fwd(LossIn in, Kern11 K, float* __restrict__ maps, float* __restrict__ partial,
                                                 size_t total_pix) {
#pragma clang fp contract(fast)
  __shared__ float s_x[LN];
  __shared__ float s_y[LN];
  __shared__ float s_h[5][LH][LTILE];
  __shared__ float s_tmp[4];
  const int t = threadIdx.x;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);  // halo-sharing neighbours on one XCD
  const LossTile T = loss_tile(in, tile);
  const LossView V = loss_view(in, T);
  const float C1 = 0.01f * 0.01f;
  const float C2 = 0.03f * 0.03f;
  // vertical pass / output mapping: column oc, rows 4 * oq .. 4 * oq + 3
  const int oc = t & (LTILE - 1), oq = t >> 5;
  int off[LPT];
  stage_offsets(V, T, off);
  // channel ch + 1 is fetched into registers while channel ch is blurred
  float vx[LPT], vy[LPT];
  auto fetch = [&](int ch) {
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      vx[i] = off[i] >= 0 ? in.rgb[(V.pb + (size_t)off[i]) * 3 + ch] : 0.0f;
      vy[i] = off[i] >= 0 ? in.target[(V.tb + (size_t)off[i]) * 4 + ch] : 0.0f;
    }
  };
  fetch(0);
  for (int ch = 0; ch < 3; ++ch) {
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      const int idx = t + i * LT;
      if (idx < LN) {
        s_x[idx] = vx[i];
        s_y[idx] = vy[i];
      }
    }
    __syncthreads();
    if (ch < 2) fetch(ch + 1);
    // horizontal: LH rows x (LTILE / RUN) runs
    for (int task = t; task < LH * (LTILE / RUN); task += LT) {
      const int r = task / (LTILE / RUN), c0 = (task % (LTILE / RUN)) * RUN;
      float a[5][RUN];
#pragma unroll
      for (int m = 0; m < 5; ++m)
#pragma unroll
        for (int o = 0; o < RUN; ++o) a[m][o] = 0.0f;
#pragma unroll
      for (int w = 0; w < WIN; ++w) {
        const float xv = s_x[r * LH + c0 + w], yv = s_y[r * LH + c0 + w];
        const float xx = xv * xv, yy = yv * yv, xy = xv * yv;
#pragma unroll
        for (int o = 0; o < RUN; ++o) {
          const int ki = w - o;
          if (ki < 0 || ki > 10) continue;
          const float kk = K.k[ki];
          a[0][o] += kk * xv;
          a[1][o] += kk * yv;
          a[2][o] += kk * xx;
          a[3][o] += kk * yy;
          a[4][o] += kk * xy;
        }
      }
#pragma unroll
      for (int m = 0; m < 5; ++m)
#pragma unroll
        for (int o = 0; o < RUN; ++o) s_h[m][r][c0 + o] = a[m][o];
    }
    __syncthreads();
    float mu[5][RUN];
#pragma unroll
    for (int m = 0; m < 5; ++m)
#pragma unroll
      for (int o = 0; o < RUN; ++o) mu[m][o] = 0.0f;
#pragma unroll
    for (int w = 0; w < WIN; ++w) {
      float hv[5];
#pragma unroll
      for (int m = 0; m < 5; ++m) hv[m] = s_h[m][oq * RUN + w][oc];
#pragma unroll
      for (int o = 0; o < RUN; ++o) {
        const int ki = w - o;
        if (ki < 0 || ki > 10) continue;
        const float kk = K.k[ki];
#pragma unroll
        for (int m = 0; m < 5; ++m) mu[m][o] += kk * hv[m];
      }
    }
    float l1 = 0.0f, ss = 0.0f;
    const int px = T.tx * LTILE + oc;
#pragma unroll
    for (int o = 0; o < RUN; ++o) {
      const int py = T.ty * LTILE + oq * RUN + o;
      if (px >= V.w || py >= V.h) continue;
      const size_t pix = V.pb + (size_t)py * V.w + px;
      float mx = mu[0][o], my = mu[1][o];
      float sx2 = mu[2][o] - mx * mx;
      float sy2 = mu[3][o] - my * my;
      float sxy = mu[4][o] - mx * my;
      float a1 = 2.0f * mx * my + C1;
      float a2 = 2.0f * sxy + C2;
      float b1 = mx * mx + my * my + C1;
      float b2 = sx2 + sy2 + C2;
      float num = a1 * a2;
      float den = b1 * b2;
      ss += num / den;
      float inv_den = 1.0f / den;
      float pc = inv_den * (2.0f * my * a2) - (num / (den * den)) * (2.0f * mx * b2);
      float qc = inv_den * (2.0f * a1);
      float rc = -(num / (den * den)) * b1;
      maps[(size_t)(0 * 3 + ch) * total_pix + pix] = pc - qc * my - 2.0f * (rc * mx);
      maps[(size_t)(1 * 3 + ch) * total_pix + pix] = qc;
      maps[(size_t)(2 * 3 + ch) * total_pix + pix] = rc;
      const float xv = s_x[(oq * RUN + o + 5) * LH + oc + 5];
      const float yv = s_y[(oq * RUN + o + 5) * LH + oc + 5];
      l1 += __builtin_fabsf(xv - yv);
    }
    float bl1 = block_sum_256(l1, s_tmp);
    float bss = block_sum_256(ss, s_tmp);
    if (t == 0) {
      partial[((size_t)tile * 3 + ch) * 2 + 0] = bl1;
      partial[((size_t)tile * 3 + ch) * 2 + 1] = bss;
    }
  }
}

__global__ __launch_bounds__(LT) void k_ssim_bwd(LossIn in, Kern11 K, float* __restrict__ dl,
                                                 const float* __restrict__ maps, size_t total_pix) {
#pragma clang fp contract(fast)
  __shared__ float s_m[SSIM_MAPS][LN];
  __shared__ float s_h[SSIM_MAPS][LH][LTILE];
  const int t = threadIdx.x;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);  // halo-sharing neighbours on one XCD
  const LossTile T = loss_tile(in, tile);
  const LossView V = loss_view(in, T);
  const int oc = t & (LTILE - 1), oq = t >> 5;
  int off[LPT];
  stage_offsets(V, T, off);
  // pixels this thread writes: column px, rows py0 .. py0 + RUN - 1
  const int px = T.tx * LTILE + oc, py0 = T.ty * LTILE + oq * RUN;
  float vm[SSIM_MAPS][LPT];
  auto fetch = [&](int ch) {
#pragma unroll
    for (int m = 0; m < SSIM_MAPS; ++m)
#pragma unroll
      for (int i = 0; i < LPT; ++i)
        vm[m][i] = off[i] >= 0 ? maps[(size_t)(m * 3 + ch) * total_pix + V.pb + (size_t)off[i]] : 0.0f;
  };
  fetch(0);
  for (int ch = 0; ch < 3; ++ch) {
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      const int idx = t + i * LT;
      if (idx < LN)
#pragma unroll
        for (int m = 0; m < SSIM_MAPS; ++m) s_m[m][idx] = vm[m][i];
    }
    __syncthreads();
    if (ch < 2) fetch(ch + 1);
    float ox[RUN], oy[RUN];
#pragma unroll
    for (int o = 0; o < RUN; ++o) {
      const bool in_img = px < V.w && py0 + o < V.h;
      const size_t q = (size_t)(py0 + o) * V.w + px;
      ox[o] = in_img ? in.rgb[(V.pb + q) * 3 + ch] : 0.0f;
      oy[o] = in_img ? in.target[(V.tb + q) * 4 + ch] : 0.0f;
    }
    for (int task = t; task < LH * (LTILE / RUN); task += LT) {
      const int r = task / (LTILE / RUN), c0 = (task % (LTILE / RUN)) * RUN;
      float a[SSIM_MAPS][RUN];
#pragma unroll
      for (int m = 0; m < SSIM_MAPS; ++m)
#pragma unroll
        for (int o = 0; o < RUN; ++o) a[m][o] = 0.0f;
#pragma unroll
      for (int w = 0; w < WIN; ++w) {
        float v[SSIM_MAPS];
#pragma unroll
        for (int m = 0; m < SSIM_MAPS; ++m) v[m] = s_m[m][r * LH + c0 + w];
#pragma unroll
        for (int o = 0; o < RUN; ++o) {
          const int ki = w - o;
          if (ki < 0 || ki > 10) continue;
          const float kk = K.k[ki];
#pragma unroll
          for (int m = 0; m < SSIM_MAPS; ++m) a[m][o] += kk * v[m];
        }
      }
#pragma unroll
      for (int m = 0; m < SSIM_MAPS; ++m)
#pragma unroll
        for (int o = 0; o < RUN; ++o) s_h[m][r][c0 + o] = a[m][o];
    }
    __syncthreads();
    float b[SSIM_MAPS][RUN];
#pragma unroll
    for (int m = 0; m < SSIM_MAPS; ++m)
#pragma unroll
      for (int o = 0; o < RUN; ++o) b[m][o] = 0.0f;
#pragma unroll
    for (int w = 0; w < WIN; ++w) {
      float hv[SSIM_MAPS];
#pragma unroll
      for (int m = 0; m < SSIM_MAPS; ++m) hv[m] = s_h[m][oq * RUN + w][oc];
#pragma unroll
      for (int o = 0; o < RUN; ++o) {
        const int ki = w - o;
        if (ki < 0 || ki > 10) continue;
        const float kk = K.k[ki];
#pragma unroll
        for (int m = 0; m < SSIM_MAPS; ++m) b[m][o] += kk * hv[m];
      }
    }
#pragma unroll
    for (int o = 0; o < RUN; ++o) {
      if (px >= V.w || py0 + o >= V.h) continue;
      const size_t pix = V.pb + (size_t)(py0 + o) * V.w + px;
      const float xv = ox[o], yv = oy[o];
      const float d = xv - yv;
      // Rust f32::signum: +1 for +0.0
      const float sg = d != d ? d : __builtin_copysignf(1.0f, d);
      const float grad = b[0][o] + yv * b[1][o] + 2.0f * xv * b[2][o];
      dl[pix * 3 + ch] = (1.0f - V.lam) * sg / V.m_total + -V.lam * grad / V.m_total;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- fused pass
// One workgroup per 32x32 output tile and one pass over HBM: the maps' 5-pixel
// halo needs the blurred products on the tile plus 5 pixels, which need the
// inputs on the tile plus 10. Per colour channel:
//   1. x, y of the 52x52 input window into LDS (zero outside the image);
//   2. the five products blurred horizontally (52 rows x 42 columns, in two
//      groups of at most three maps to bound LDS) and vertically (42 x 42,
//      seven outputs per thread held in registers);
//   3. the three coefficient maps on the 42x42 window (zero outside the
//      image, where the reference has no map) into LDS, and the tile's L1 and
//      SSIM sums;
//   4. the maps blurred horizontally (42 x 32) and vertically (32 x 32) and
//      dL/dpixel written.
// The taps, their order and the fused multiply-adds are those of the two
// separate passes, so the maps and dL/dpixel are the same values.
constexpr int FH = LTILE + 20;          // 52: input window side
constexpr int MH = LTILE + 10;          // 42: map window side
constexpr int FN = FH * FH;             // staged inputs per channel and image
constexpr int MN = MH * MH;             // map values per map
constexpr int FPT = (FN + LT - 1) / LT; // staged inputs per thread
constexpr int FR = 7;                   // outputs per thread of the product blurs (42 = 6 x 7)
constexpr int FW = FR + 10;             // their window
constexpr int FRUNS = MH / FR;          // 6 runs per row / column
constexpr int MR = 4;                   // outputs per thread of the map blurs
constexpr int MW = MR + 10;
static_assert(MH % FR == 0 && LTILE % MR == 0, "run lengths");
static_assert(FRUNS * MH <= LT, "one vertical product run per thread");
static_assert(2 * FN >= 3 * MN && 3 * FH * MH >= 3 * MH * LTILE, "LDS aliasing");

// Horizontal blur of NM products of the staged inputs over the 52 rows of
// the window into s_b[m][row][0..42): GRP 0 blurs x, y, x^2; GRP 1 y^2, xy.
template <int GRP, int NM>
__device__ __forceinline__ void product_rows(const float* __restrict__ s_a, float* __restrict__ s_b, const Kern11& K) {
#pragma clang fp contract(fast)
  for (int task = threadIdx.x; task < FH * FRUNS; task += LT) {
    const int r = task / FRUNS, c0 = (task % FRUNS) * FR;
    float a[NM][FR];
#pragma unroll
    for (int m = 0; m < NM; ++m)
#pragma unroll
      for (int o = 0; o < FR; ++o) a[m][o] = 0.0f;
#pragma unroll
    for (int w = 0; w < FW; ++w) {
      const float xv = s_a[r * FH + c0 + w], yv = s_a[FN + r * FH + c0 + w];
      float pr[NM];
      if (GRP == 0) {
        pr[0] = xv;
        pr[1] = yv;
        pr[NM - 1] = xv * xv;
      } else {
        pr[0] = yv * yv;
        pr[NM - 1] = xv * yv;
      }
#pragma unroll
      for (int o = 0; o < FR; ++o) {
        const int ki = w - o;
        if (ki < 0 || ki > 10) continue;
        const float kk = K.k[ki];
#pragma unroll
        for (int m = 0; m < NM; ++m) a[m][o] += kk * pr[m];
      }
    }
#pragma unroll
    for (int m = 0; m < NM; ++m)
#pragma unroll
      for (int o = 0; o < FR; ++o) s_b[(m * FH + r) * MH + c0 + o] = a[m][o];
  }
}

// Vertical blur of NM row sums: rows mr0 .. mr0 + FR - 1 of column mc.
template <int NM>
__device__ __forceinline__ void product_cols(const float* __restrict__ s_b, int mr0, int mc, const Kern11& K,
                                             float (&mu)[NM][FR]) {
#pragma clang fp contract(fast)
#pragma unroll
  for (int m = 0; m < NM; ++m)
#pragma unroll
    for (int o = 0; o < FR; ++o) mu[m][o] = 0.0f;
#pragma unroll
  for (int w = 0; w < FW; ++w) {
    float hv[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) hv[m] = s_b[(m * FH + mr0 + w) * MH + mc];
#pragma unroll
    for (int o = 0; o < FR; ++o) {
      const int ki = w - o;
      if (ki < 0 || ki > 10) continue;
      const float kk = K.k[ki];
#pragma unroll
      for (int m = 0; m < NM; ++m) mu[m][o] += kk * hv[m];
    }
  }
}

__global__ __launch_bounds__(LT) void k_ssim(LossIn in, Kern11 K, float* __restrict__ dl,
                                             float* __restrict__ partial) {
#pragma clang fp contract(fast)
  __shared__ float s_a[2 * FN];       // x | y windows; then the three maps (MN each)
  __shared__ float s_b[3 * FH * MH];  // product row sums [m][52][42]; then map row sums [m][42][32]
  __shared__ float s_tmp[4];
  const int t = threadIdx.x;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);  // halo-sharing neighbours on one XCD
  const LossTile T = loss_tile(in, tile);
  const LossView V = loss_view(in, T);
  const float C1 = 0.01f * 0.01f;
  const float C2 = 0.03f * 0.03f;
  // vertical product pass and maps: column mc, rows mr0 .. mr0 + FR - 1 of the
  // map window (threads past FRUNS * MH have none)
  const bool has_maps = t < FRUNS * MH;
  const int mc = t % MH, mr0 = (t / MH) * FR;
  // map blurs and output: column oc, rows oq * MR .. + MR - 1 of the tile
  const int oc = t & (LTILE - 1), oq = t >> 5;
  const int px = T.tx * LTILE + oc, py0 = T.ty * LTILE + oq * MR;
  // the 52x52 input window of a channel (zero outside the image), fetched
  // into registers one channel ahead
  float vx[FPT], vy[FPT];
  auto fetch = [&](int ch) {
    const int x0 = T.tx * LTILE - 10, y0 = T.ty * LTILE - 10;
#pragma unroll
    for (int i = 0; i < FPT; ++i) {
      const int idx = t + i * LT;
      const int gy = y0 + idx / FH, gx = x0 + idx % FH;
      const bool ok = idx < FN && gx >= 0 && gy >= 0 && gx < V.w && gy < V.h;
      const size_t q = (size_t)gy * V.w + gx;
      vx[i] = ok ? in.rgb[(V.pb + q) * 3 + ch] : 0.0f;
      vy[i] = ok ? in.target[(V.tb + q) * 4 + ch] : 0.0f;
    }
  };
  fetch(0);
  for (int ch = 0; ch < 3; ++ch) {
#pragma unroll
    for (int i = 0; i < FPT; ++i) {
      const int idx = t + i * LT;
      if (idx < FN) {
        s_a[idx] = vx[i];
        s_a[FN + idx] = vy[i];
      }
    }
    __syncthreads();
    float mu0[3][FR], mu1[2][FR];
    product_rows<0, 3>(s_a, s_b, K);
    __syncthreads();
    if (has_maps) product_cols<3>(s_b, mr0, mc, K, mu0);
    __syncthreads();
    product_rows<1, 2>(s_a, s_b, K);
    __syncthreads();
    if (has_maps) product_cols<2>(s_b, mr0, mc, K, mu1);
    // coefficient maps on the map window; L1 and SSIM of the tile's pixels
    float mp[3][FR];
    float l1 = 0.0f, ss = 0.0f;
    if (has_maps) {
      const int gx = T.tx * LTILE - 5 + mc;
#pragma unroll
      for (int o = 0; o < FR; ++o) {
        const int mr = mr0 + o;
        const int gy = T.ty * LTILE - 5 + mr;
        const bool in_img = gx >= 0 && gy >= 0 && gx < V.w && gy < V.h;
        const float mx = mu0[0][o], my = mu0[1][o];
        const float sx2 = mu0[2][o] - mx * mx;
        const float sy2 = mu1[0][o] - my * my;
        const float sxy = mu1[1][o] - mx * my;
        const float a1 = 2.0f * mx * my + C1;
        const float a2 = 2.0f * sxy + C2;
        const float b1 = mx * mx + my * my + C1;
        const float b2 = sx2 + sy2 + C2;
        const float num = a1 * a2;
        const float den = b1 * b2;
        const float inv_den = 1.0f / den;
        const float pc = inv_den * (2.0f * my * a2) - (num / (den * den)) * (2.0f * mx * b2);
        const float qc = inv_den * (2.0f * a1);
        const float rc = -(num / (den * den)) * b1;
        mp[0][o] = in_img ? pc - qc * my - 2.0f * (rc * mx) : 0.0f;
        mp[1][o] = in_img ? qc : 0.0f;
        mp[2][o] = in_img ? rc : 0.0f;
        if (in_img && mr >= 5 && mr < 5 + LTILE && mc >= 5 && mc < 5 + LTILE) {
          ss += num / den;
          const int q = (mr + 5) * FH + mc + 5;
          l1 += __builtin_fabsf(s_a[q] - s_a[FN + q]);
        }
      }
    }
    // this thread's output pixels, read before the maps replace the inputs
    float ox[MR], oy[MR];
#pragma unroll
    for (int o = 0; o < MR; ++o) {
      const int q = (oq * MR + o + 10) * FH + oc + 10;
      ox[o] = s_a[q];
      oy[o] = s_a[FN + q];
    }
    __syncthreads();
    if (has_maps) {
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int o = 0; o < FR; ++o) s_a[m * MN + (mr0 + o) * MH + mc] = mp[m][o];
    }
    const float bl1 = block_sum_256(l1, s_tmp);
    const float bss = block_sum_256(ss, s_tmp);
    if (t == 0) {
      partial[((size_t)tile * 3 + ch) * 2 + 0] = bl1;
      partial[((size_t)tile * 3 + ch) * 2 + 1] = bss;
    }
    if (ch < 2) fetch(ch + 1);
    // maps, horizontal: 42 rows x 32 columns
    for (int task = t; task < MH * (LTILE / MR); task += LT) {
      const int r = task / (LTILE / MR), c0 = (task % (LTILE / MR)) * MR;
      float a[3][MR];
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int o = 0; o < MR; ++o) a[m][o] = 0.0f;
#pragma unroll
      for (int w = 0; w < MW; ++w) {
        float v[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) v[m] = s_a[m * MN + r * MH + c0 + w];
#pragma unroll
        for (int o = 0; o < MR; ++o) {
          const int ki = w - o;
          if (ki < 0 || ki > 10) continue;
          const float kk = K.k[ki];
#pragma unroll
          for (int m = 0; m < 3; ++m) a[m][o] += kk * v[m];
        }
      }
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int o = 0; o < MR; ++o) s_b[(m * MH + r) * LTILE + c0 + o] = a[m][o];
    }
    __syncthreads();
    // maps, vertical, and dL/dpixel
    float b[3][MR];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int o = 0; o < MR; ++o) b[m][o] = 0.0f;
#pragma unroll
    for (int w = 0; w < MW; ++w) {
      float hv[3];
#pragma unroll
      for (int m = 0; m < 3; ++m) hv[m] = s_b[(m * MH + oq * MR + w) * LTILE + oc];
#pragma unroll
      for (int o = 0; o < MR; ++o) {
        const int ki = w - o;
        if (ki < 0 || ki > 10) continue;
        const float kk = K.k[ki];
#pragma unroll
        for (int m = 0; m < 3; ++m) b[m][o] += kk * hv[m];
      }
    }
#pragma unroll
    for (int o = 0; o < MR; ++o) {
      if (px >= V.w || py0 + o >= V.h) continue;
      const size_t pix = V.pb + (size_t)(py0 + o) * V.w + px;
      const float xv = ox[o], yv = oy[o];
      const float d = xv - yv;
      // Rust f32::signum: +1 for +0.0
      const float sg = d != d ? d : __builtin_copysignf(1.0f, d);
      const float grad = b[0][o] + yv * b[1][o] + 2.0f * xv * b[2][o];
      dl[pix * 3 + ch] = (1.0f - V.lam) * sg / V.m_total + -V.lam * grad / V.m_total;
    }
    __syncthreads();  // the next channel overwrites both LDS areas
  }
}

// One block per view: loss from the per-tile partial sums.
__global__ __launch_bounds__(256) void k_loss_reduce(LossIn in, const float* __restrict__ partial,
                                                     float* __restrict__ loss_out) {
  __shared__ double s_l1[256];
  __shared__ double s_ss[256];
  const int v = blockIdx.x;
  const CamDev& c = loss_cam(in, v);
  const int nt = ((c.w + LTILE - 1) / LTILE) * ((c.h + LTILE - 1) / LTILE);
  double l1 = 0.0, ss = 0.0;
  for (int i = threadIdx.x; i < nt * 3; i += 256) {
    size_t tile = (size_t)c.ltile_base + i / 3;
    int ch = i % 3;
    l1 += (double)partial[(tile * 3 + ch) * 2 + 0];
    ss += (double)partial[(tile * 3 + ch) * 2 + 1];
  }
  s_l1[threadIdx.x] = l1;
  s_ss[threadIdx.x] = ss;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s_l1[threadIdx.x] += s_l1[threadIdx.x + s];
      s_ss[threadIdx.x] += s_ss[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double m = (double)c.w * (double)c.h * 3.0;
    double lam = (double)in.lambda;
    loss_out[v] = (float)((1.0 - lam) * (s_l1[0] / m) + lam * (1.0 - s_ss[0] / m));
  }
}

static Kern11 make_kernel() {
  // gaussian_kernel_1d(5, 1.5) (backward.rs:759-772), evaluated once on the
  // host with the same f32 operations and libm expf.
  Kern11 K;
  float sum = 0.0f;
  const float sigma = 1.5f;
  for (int i = 0; i < 11; ++i) {
    float d = (float)i - 5.0f;
    float v = expf(-0.5f * d * d / (sigma * sigma));
    K.k[i] = v;
    sum += v;
  }
  for (int i = 0; i < 11; ++i) K.k[i] /= sum;
  return K;
}

void launch_loss(const LossIn& in, float* dl, float* ssim_maps, float* partial, float* loss_out, hipStream_t st) {
  if (in.total_ltiles <= 0) return;
  static const Kern11 K = make_kernel();
  hipLaunchKernelGGL(k_ssim_fwd, dim3(in.total_ltiles), dim3(LT), 0, st, in, K, ssim_maps, partial,
                     in.total_pix);
  hipLaunchKernelGGL(k_ssim_bwd, dim3(in.total_ltiles), dim3(LT), 0, st, in, K, dl, ssim_maps, in.total_pix);
  hipLaunchKernelGGL(k_loss_reduce, dim3(in.n_views), dim3(256), 0, st, in, partial, loss_out);
}

void launch_loss_fused(const LossIn& in, float* dl, float* partial, float* loss_out, hipStream_t st) {
  if (in.total_ltiles <= 0) return;
  static const Kern11 K = make_kernel();
  hipLaunchKernelGGL(k_ssim, dim3(in.total_ltiles), dim3(LT), 0, st, in, K, dl, partial);
  hipLaunchKernelGGL(k_loss_reduce, dim3(in.n_views), dim3(256), 0, st, in, partial, loss_out);
}

}  // namespace rgs

// k_loss.hip -- L = (1 - lambda) L1 + lambda (1 - SSIM) and dL/dpixel.
//
// Follows backward.rs:646-807. SSIM uses the 11-tap separable Gaussian window
// (sigma 1.5) with zero padding. Its gradient with respect to x is the
// adjoint blur of five per-pixel coefficient maps (backward.rs:744-754),
//   g = B(P) + y B(Q) - B(Q mu_y) + 2 x B(R) - 2 B(R mu_x),
// which by linearity of the blur B is computed from three maps:
//   g = B(P - Q mu_y - 2 R mu_x) + y B(Q) + 2 x B(R).
//
// One kernel, k_ssim_sweep, makes dL/dpixel in one pass over HBM: a workgroup
// owns one colour channel of a unit of LSW (54) output columns x LSG (155)
// output rows and sweeps down it eight rows per step. The maps' 5-pixel
// halo needs the products blurred on 64 columns, which need the inputs on
// 74. Per step, with two 18-row LDS rings carrying the vertical windows:
//   B: the new 8 input rows' products x, y, x^2, y^2, xy blurred along rows
//      (64 columns) into the product ring;
//   C: the product ring blurred down columns for the 8 map rows five rows
//      behind, and the three coefficient maps of those rows (zero outside the
//      image, where the reference has no map);
//   D: the maps blurred along rows (54 columns) into the map ring;
//   E: the map ring blurred down columns for the 8 output rows ten rows
//      behind the inputs, and dL/dpixel written.
// Nothing but the inputs, dL/dpixel and two partial sums per workgroup goes
// through HBM. Each thread makes two consecutive outputs of a row (B, D) or a
// column (C, E) from one window held in registers. Taps are summed in order
// k = 0..10 (horizontal, then vertical) with fused multiply-adds, the map
// quotients use the hardware reciprocal; the per-pixel values agree with the
// reference's to float rounding. The loss scalar sums the per-workgroup
// partials in double.
#include "rgs_kernels.h"

namespace rgs {

constexpr int LT = 256;

struct Kern11 {
  float k[11];
};

// View v's record: from device memory, or the by-value copy of a one-view
// call (cams == nullptr).
__device__ __forceinline__ const CamDev& loss_cam(const LossIn& in, int v) {
  return in.cams ? in.cams[v] : in.cam0;
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

constexpr int LMC = LSW + 10;           // 64 map columns
constexpr int LIC = LSW + 20;           // 74 input columns
constexpr int LRB = 8;                  // rows per step
constexpr int LRING = LRB + 10;         // ring rows: a step's 8 plus the window's 10
constexpr int LIN = LRB * LIC;          // staged inputs per image per step
constexpr int LIPT = (LIN + LT - 1) / LT;
static_assert(LMC == 64 && LT == 4 * 64 && LRB == 2 * (LT / 64), "a wave per row pair, a lane per map column");

// Sweep modes:
//   SW_FULL: the five products blurred (x, y staged; products formed in B);
//   SW_CACHED: the target's two blurs (mu_y, E[y^2]) read from a statistics
//     buffer made once per set of targets; x, x^2, xy staged and blurred;
//   SW_STATS: that buffer: y, y^2 staged and blurred, written for the unit's
//     own pixels (no maps, no dL/dpixel).
// The blurs of y and y^2 are the same fused multiply-add chains in SW_FULL
// and SW_STATS, so SW_CACHED computes the values SW_FULL does.
enum { SW_FULL = 0, SW_CACHED = 1, SW_STATS = 2 };
template <int MODE>
struct Sweep {
  static constexpr int NP = MODE == SW_FULL ? 5 : MODE == SW_CACHED ? 3 : 2;  // blurred products
  static constexpr int NS = MODE == SW_FULL ? 2 : NP;                         // staged values per input
  static constexpr bool MAPS = MODE != SW_STATS;
  static constexpr int ST = (NS * LIN > 3 * LRB * LMC || !MAPS) ? NS * LIN : 3 * LRB * LMC;
};

template <int MODE>
__global__ __launch_bounds__(LT) __attribute__((amdgpu_waves_per_eu(MODE == SW_CACHED ? 5 : 1))) void k_ssim_sweep(LossIn in, Kern11 K, float* __restrict__ out,
                                                   float* __restrict__ partial) {
#pragma clang fp contract(fast)
  using SP = Sweep<MODE>;
  constexpr int NP = SP::NP, NS = SP::NS;
  __shared__ float s_hp[NP][LRING][LMC];             // product ring: the products blurred along rows
  __shared__ float s_mh[SP::MAPS ? 3 : 1][LRING][LSW];  // map ring: the maps blurred along rows
  __shared__ float s_st[SP::ST];                     // the step's staged inputs [NS][8][74], then its maps [3][8][64]
  __shared__ float s_tmp[4];
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);  // row pair of C and E
  // the three channels of a unit are consecutive and on one XCD: they read the
  // same input lines
  const int id = xcd_remap(blockIdx.x, gridDim.x);
  const int ch = id % 3, unit = id / 3;
  int v = 0;
  while (v + 1 < in.n_views && loss_cam(in, v + 1).lunit_base <= unit) ++v;
  const CamDev& cam = loss_cam(in, v);
  const int W = cam.w, H = cam.h;
  const size_t pb = (size_t)cam.pix_base, tb = (size_t)cam.tgt_base;
  const int nsx = (W + LSW - 1) / LSW, local = unit - cam.lunit_base;
  const int x0 = (local % nsx) * LSW, y0 = (local / nsx) * LSG;
  const int yend = min(y0 + LSG, H);
  const int Y = y0 - 10;  // image row of input row 0; map row m and output row o are rows Y + m, Y + o
  // step k: input rows [8k, 8k + 8), map rows [8k - 5, 8k + 3), output rows
  // [8k - 10, 8k - 2); the last step is the first whose output rows reach
  // yend - 1 - Y
  const int nk = (yend - 1 - Y + 3 + LRB - 1) / LRB + 1;
  const float C1 = 0.01f * 0.01f;
  const float C2 = 0.03f * 0.03f;
  const float m_total = (float)((size_t)W * H * 3);
  const float lam = in.lambda;
  // (1 - lambda) sign(d) / M for the two signs: the L1 term of dL/dpixel;
  // -lambda / M scales the SSIM term
  const float l1_pos = (1.0f - lam) * 1.0f / m_total, l1_neg = (1.0f - lam) * -1.0f / m_total;
  const float s_scale = -lam / m_total;
  // the target statistics of channel ch: mu_y, then E[y^2]
  const float* ts_mu = in.tstats + ((size_t)ch * 2 + 0) * in.tgt_total + tb;
  const float* ts_yy = in.tstats + ((size_t)ch * 2 + 1) * in.tgt_total + tb;

  for (int i = t; i < NP * LRING * LMC; i += LT) (&s_hp[0][0][0])[i] = 0.0f;
  if (SP::MAPS)
    for (int i = t; i < 3 * LRING * LSW; i += LT) (&s_mh[0][0][0])[i] = 0.0f;

  // the staged inputs this thread loads: row ir[i] of the step, image column ix[i]
  int ir[LIPT], ix[LIPT];
  bool cok[LIPT];
#pragma unroll
  for (int i = 0; i < LIPT; ++i) {
    const int idx = t + i * LT;
    ir[i] = idx / LIC;
    ix[i] = x0 - 10 + idx % LIC;
    cok[i] = idx < LIN && ix[i] >= 0 && ix[i] < W;
  }
  float px[LIPT], py[LIPT];
  auto fetch = [&](int k) {
#pragma unroll
    for (int i = 0; i < LIPT; ++i) {
      const int gy = Y + LRB * k + ir[i];
      const bool ok = cok[i] && gy >= 0 && gy < H;
      const size_t q = (size_t)gy * W + ix[i];
      px[i] = (MODE != SW_STATS && ok) ? in.rgb[(pb + q) * 3 + ch] : 0.0f;
      py[i] = ok ? in.target[(tb + q) * 4 + ch] : 0.0f;
    }
  };
  fetch(0);
  float l1 = 0.0f, ss = 0.0f;
  for (int k = 0; k < nk; ++k) {
#pragma unroll
    for (int i = 0; i < LIPT; ++i) {
      const int idx = t + i * LT;
      if (idx < LIN) {
        const float xv = px[i], yv = py[i];
        if (MODE == SW_FULL) {
          s_st[idx] = xv;
          s_st[LIN + idx] = yv;
        } else if (MODE == SW_CACHED) {
          s_st[idx] = xv;
          s_st[LIN + idx] = xv * xv;
          s_st[2 * LIN + idx] = xv * yv;
        } else {
          s_st[idx] = yv;
          s_st[LIN + idx] = yv * yv;
        }
      }
    }
    __syncthreads();
    if (k + 1 < nk) fetch(k + 1);
    // the target statistics of this step's map rows (stage C), loaded ahead
    float t_mu[2] = {0.0f, 0.0f}, t_yy[2] = {0.0f, 0.0f};
    if (MODE == SW_CACHED && k > 0) {
      const int gx = x0 - 5 + lane;
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        const int gy = Y + LRB * k - 5 + 2 * wv + o;
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
          const size_t q = (size_t)gy * W + gx;
          t_mu[o] = ts_mu[q];
          t_yy[o] = ts_yy[q];
        }
      }
    }
    // the image and target values of this step's output pixels (stage E),
    // loaded ahead
    const int or0 = LRB * k - 10 + 2 * wv;
    const bool e_rows = SP::MAPS && Y + or0 + 1 >= y0 && lane < LSW;
    float e_x[2] = {0.0f, 0.0f}, e_y[2] = {0.0f, 0.0f};
    if (e_rows) {
      const int gx = x0 + lane;
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        const int gy = Y + or0 + o;
        if (gx < W && gy >= y0 && gy < yend) {
          const size_t q = (size_t)gy * W + gx;
          e_x[o] = in.rgb[(pb + q) * 3 + ch];
          e_y[o] = in.target[(tb + q) * 4 + ch];
        }
      }
    }
    // B: products of input row r blurred along the row, map columns c0, c0 + 1
    {
      const int r = t >> 5, c0 = (t & 31) * 2;
      float a[NP][2];
#pragma unroll
      for (int m = 0; m < NP; ++m) a[m][0] = a[m][1] = 0.0f;
#pragma unroll
      for (int w = 0; w < 12; ++w) {
        float pr[NP];
        if (MODE == SW_FULL) {
          const float xv = s_st[r * LIC + c0 + w], yv = s_st[LIN + r * LIC + c0 + w];
          const float p5[5] = {xv, yv, xv * xv, yv * yv, xv * yv};
#pragma unroll
          for (int m = 0; m < NP; ++m) pr[m] = p5[m];
        } else {
#pragma unroll
          for (int m = 0; m < NP; ++m) pr[m] = s_st[m * LIN + r * LIC + c0 + w];
        }
#pragma unroll
        for (int o = 0; o < 2; ++o) {
          const int ki = w - o;
          if (ki < 0 || ki > 10) continue;
          const float kk = K.k[ki];
#pragma unroll
          for (int m = 0; m < NP; ++m) a[m][o] += kk * pr[m];
        }
      }
      const int slot = (LRB * k + r) % LRING;
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        s_hp[m][slot][c0] = a[m][0];
        s_hp[m][slot][c0 + 1] = a[m][1];
      }
    }
    __syncthreads();
    // C: map rows mr0, mr0 + 1 of map column `lane`, from product rows
    // mr0 - 5 .. mr0 + 6
    if (k > 0) {
      const int mr0 = LRB * k - 5 + 2 * wv;
      float mu[NP][2];
#pragma unroll
      for (int m = 0; m < NP; ++m) mu[m][0] = mu[m][1] = 0.0f;
#pragma unroll
      for (int w = 0; w < 12; ++w) {
        const int slot = (mr0 - 5 + w + 2 * LRING) % LRING;
        float hv[NP];
#pragma unroll
        for (int m = 0; m < NP; ++m) hv[m] = s_hp[m][slot][lane];
#pragma unroll
        for (int o = 0; o < 2; ++o) {
          const int ki = w - o;
          if (ki < 0 || ki > 10) continue;
          const float kk = K.k[ki];
#pragma unroll
          for (int m = 0; m < NP; ++m) mu[m][o] += kk * hv[m];
        }
      }
      const int gx = x0 - 5 + lane;
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        const int gy = Y + mr0 + o;
        const bool in_img = gx >= 0 && gx < W && gy >= 0 && gy < H;
        const bool own = in_img && gy >= y0 && gy < yend && lane >= 5 && lane < 5 + LSW;
        if (MODE == SW_STATS) {
          if (own) {
            const size_t q = (size_t)gy * W + gx;
            out[((size_t)ch * 2 + 0) * in.tgt_total + tb + q] = mu[0][o];
            out[((size_t)ch * 2 + 1) * in.tgt_total + tb + q] = mu[1][o];
          }
          continue;
        }
        float mx, my, exx, eyy, exy;
        if (MODE == SW_FULL) {
          mx = mu[0][o], my = mu[1][o], exx = mu[2][o], eyy = mu[3][o], exy = mu[4][o];
        } else {
          mx = mu[0][o], exx = mu[1][o], exy = mu[2][o];
          my = t_mu[o];
          eyy = t_yy[o];
        }
        const float sx2 = exx - mx * mx;
        const float sy2 = eyy - my * my;
        const float sxy = exy - mx * my;
        const float a1 = 2.0f * mx * my + C1;
        const float a2 = 2.0f * sxy + C2;
        const float b1 = mx * mx + my * my + C1;
        const float b2 = sx2 + sy2 + C2;
        const float num = a1 * a2;
        const float den = b1 * b2;
        // den >= C1 C2 > 0: the hardware reciprocal (1 ulp) for the three
        // quotients, which feed only gradients and the SSIM sum
        const float inv_den = __builtin_amdgcn_rcpf(den);
        const float ssim = num * inv_den;
        const float nd2 = ssim * inv_den;  // num / den^2
        const float pc = inv_den * (2.0f * my * a2) - nd2 * (2.0f * mx * b2);
        const float qc = inv_den * (2.0f * a1);
        const float rc = -nd2 * b1;
        const int row = 2 * wv + o;
        s_st[(0 * LRB + row) * LMC + lane] = in_img ? pc - qc * my - 2.0f * (rc * mx) : 0.0f;
        s_st[(1 * LRB + row) * LMC + lane] = in_img ? qc : 0.0f;
        s_st[(2 * LRB + row) * LMC + lane] = in_img ? rc : 0.0f;
        if (own) ss += ssim;
      }
    }
    if (!SP::MAPS) {
      __syncthreads();  // the product ring's oldest rows may be rewritten
      continue;
    }
    __syncthreads();
    // D: map row r blurred along the row, output columns c0, c0 + 1
    if (k > 0) {
      const int r = t >> 5, c0 = (t & 31) * 2;
      if (c0 < LSW) {
        float b[3][2];
#pragma unroll
        for (int m = 0; m < 3; ++m) b[m][0] = b[m][1] = 0.0f;
#pragma unroll
        for (int w = 0; w < 12; ++w) {
          float vv[3];
#pragma unroll
          for (int m = 0; m < 3; ++m) vv[m] = s_st[(m * LRB + r) * LMC + c0 + w];
#pragma unroll
          for (int o = 0; o < 2; ++o) {
            const int ki = w - o;
            if (ki < 0 || ki > 10) continue;
            const float kk = K.k[ki];
#pragma unroll
            for (int m = 0; m < 3; ++m) b[m][o] += kk * vv[m];
          }
        }
        const int slot = (LRB * k - 5 + r + 2 * LRING) % LRING;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          s_mh[m][slot][c0] = b[m][0];
          s_mh[m][slot][c0 + 1] = b[m][1];
        }
      }
    }
    __syncthreads();
    // E: output rows or0, or0 + 1 of output column `lane`, from map rows
    // or0 - 5 .. or0 + 6; the first rows of the sweep are ahead of the unit
    if (e_rows) {
      float b[3][2];
#pragma unroll
      for (int m = 0; m < 3; ++m) b[m][0] = b[m][1] = 0.0f;
#pragma unroll
      for (int w = 0; w < 12; ++w) {
        const int slot = (or0 - 5 + w + 2 * LRING) % LRING;
        float hv[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) hv[m] = s_mh[m][slot][lane];
#pragma unroll
        for (int o = 0; o < 2; ++o) {
          const int ki = w - o;
          if (ki < 0 || ki > 10) continue;
          const float kk = K.k[ki];
#pragma unroll
          for (int m = 0; m < 3; ++m) b[m][o] += kk * hv[m];
        }
      }
      const int gx = x0 + lane;
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        const int gy = Y + or0 + o;
        if (gx >= W || gy < y0 || gy >= yend) continue;
        const size_t q = (size_t)gy * W + gx;
        const float xv = e_x[o], yv = e_y[o];
        const float d = xv - yv;
        l1 += __builtin_fabsf(d);
        // Rust f32::signum: +1 for +0.0, NaN for NaN
        const float l1_term = d != d ? d : (__builtin_signbitf(d) ? l1_neg : l1_pos);
        const float grad = b[0][o] + yv * b[1][o] + 2.0f * xv * b[2][o];
        out[(pb + q) * 3 + ch] = l1_term + s_scale * grad;
      }
    }
  }
  if (!SP::MAPS) return;
  const float bl1 = block_sum_256(l1, s_tmp);
  const float bss = block_sum_256(ss, s_tmp);
  if (t == 0) {
    partial[((size_t)unit * 3 + ch) * 2 + 0] = bl1;
    partial[((size_t)unit * 3 + ch) * 2 + 1] = bss;
  }
}

// One block per view: loss from the per-unit partial sums.
__global__ __launch_bounds__(256) void k_loss_reduce(LossIn in, const float* __restrict__ partial,
                                                     float* __restrict__ loss_out) {
  __shared__ double s_l1[256];
  __shared__ double s_ss[256];
  const int v = blockIdx.x;
  const CamDev& c = loss_cam(in, v);
  const int nt = loss_units(c.w, c.h);
  double l1 = 0.0, ss = 0.0;
  for (int i = threadIdx.x; i < nt * 3; i += 256) {
    const size_t j = (size_t)c.lunit_base * 3 + i;
    l1 += (double)partial[j * 2 + 0];
    ss += (double)partial[j * 2 + 1];
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

void launch_loss(const LossIn& in, float* dl, float* partial, float* loss_out, hipStream_t st) {
  if (in.total_lunits <= 0) return;
  static const Kern11 K = make_kernel();
  if (in.tstats)
    hipLaunchKernelGGL(k_ssim_sweep<SW_CACHED>, dim3(in.total_lunits * 3), dim3(LT), 0, st, in, K, dl, partial);
  else
    hipLaunchKernelGGL(k_ssim_sweep<SW_FULL>, dim3(in.total_lunits * 3), dim3(LT), 0, st, in, K, dl, partial);
  hipLaunchKernelGGL(k_loss_reduce, dim3(in.n_views), dim3(256), 0, st, in, partial, loss_out);
}

void launch_target_stats(const LossIn& in, float* tstats, hipStream_t st) {
  if (in.total_lunits <= 0) return;
  static const Kern11 K = make_kernel();
  hipLaunchKernelGGL(k_ssim_sweep<SW_STATS>, dim3(in.total_lunits * 3), dim3(LT), 0, st, in, K, tstats, nullptr);
}

}  // namespace rgs

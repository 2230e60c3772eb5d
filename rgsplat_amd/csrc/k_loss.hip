// k_loss.hip -- L = (1 - lambda) L1 + lambda (1 - SSIM) and dL/dpixel.
//
// Follows backward.rs:646-807. SSIM uses the 11-tap separable Gaussian window
// (sigma 1.5) with zero padding; its gradient is the adjoint blur of five
// coefficient maps (backward.rs:744-754). Two passes over 16x16 output tiles
// with a 5-pixel halo staged in LDS:
//   k_ssim_fwd: blurs x, y, x^2, y^2, xy (horizontal into LDS, then vertical),
//               writes the five coefficient maps, the L1 gradient and per-tile
//               partial sums of |x - y| and SSIM;
//   k_ssim_bwd: blurs the five maps and adds -lambda dSSIM/dx / M.
// Taps are summed in the reference's order (horizontal then vertical,
// k = 0..10) and out-of-image taps contribute +0, so the per-pixel values
// equal the reference's. The loss scalar sums the partials in double.
#include "rgs_kernels.h"

namespace rgs {

constexpr int LT = 256;
constexpr int HP = TILE + 10;      // 26: tile plus halo
constexpr int HN = HP * HP;        // 676 staged values per map
constexpr int HH = HP * TILE;      // 416 horizontal outputs per map

struct Kern11 {
  float k[11];
};

__device__ __forceinline__ float block_sum_256(float x, float* s_tmp) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
  if ((threadIdx.x & 63) == 0) s_tmp[threadIdx.x >> 6] = x;
  __syncthreads();
  float r = (s_tmp[0] + s_tmp[1]) + (s_tmp[2] + s_tmp[3]);
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(LT) void k_ssim_fwd(LossIn in, Kern11 K, float* __restrict__ dl,
                                                 float* __restrict__ maps, float* __restrict__ partial,
                                                 size_t total_pix) {
  __shared__ float s_xa[3][HN];
  __shared__ float s_ya[3][HN];
  __shared__ float s_h[5][HH];
  __shared__ float s_tmp[4];
  const int t = threadIdx.x;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);  // halo-sharing neighbours on one XCD
  int v = 0;
  while (v + 1 < in.n_views && in.cams[v + 1].tile_base <= tile) ++v;
  const CamDev& c = in.cams[v];
  const int lt = tile - c.tile_base;
  const int tx = lt % c.ntx, ty = lt / c.ntx;
  const int x0 = tx * TILE - 5, y0 = ty * TILE - 5;
  const int ox = t & 15, oy = t >> 4;
  const int px = tx * TILE + ox, py = ty * TILE + oy;
  const bool inside = px < c.w && py < c.h;
  const size_t pix = (size_t)c.pix_base + (size_t)py * c.w + px;
  const float m_total = (float)((size_t)c.w * c.h * 3);
  const float lam = in.lambda;
  const float C1 = 0.01f * 0.01f;
  const float C2 = 0.03f * 0.03f;
  // all three channels of the haloed tile, one pass over the images
  for (int idx = t; idx < HN; idx += LT) {
    int gy = y0 + idx / HP, gx = x0 + idx % HP;
    bool ok = gx >= 0 && gy >= 0 && gx < c.w && gy < c.h;
    size_t gp = (size_t)c.pix_base + (size_t)gy * c.w + gx;
    size_t tp = (size_t)c.tgt_base + (size_t)gy * c.w + gx;
    float4 yv = ok ? *reinterpret_cast<const float4*>(in.target + tp * 4) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) s_xa[ch][idx] = ok ? in.rgb[gp * 3 + ch] : 0.0f;
    s_ya[0][idx] = yv.x;
    s_ya[1][idx] = yv.y;
    s_ya[2][idx] = yv.z;
  }
  __syncthreads();
  for (int ch = 0; ch < 3; ++ch) {
    const float* s_x = s_xa[ch];
    const float* s_y = s_ya[ch];
    for (int idx = t; idx < HH; idx += LT) {
      int r = idx / TILE, cc = idx % TILE;
      float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f, a4 = 0.0f;
#pragma unroll
      for (int ki = 0; ki < 11; ++ki) {
        float xv = s_x[r * HP + cc + ki];
        float yv = s_y[r * HP + cc + ki];
        float kk = K.k[ki];
        a0 += kk * xv;
        a1 += kk * yv;
        a2 += kk * (xv * xv);
        a3 += kk * (yv * yv);
        a4 += kk * (xv * yv);
      }
      s_h[0][idx] = a0;
      s_h[1][idx] = a1;
      s_h[2][idx] = a2;
      s_h[3][idx] = a3;
      s_h[4][idx] = a4;
    }
    __syncthreads();
    float mu[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ki = 0; ki < 11; ++ki) {
      float kk = K.k[ki];
#pragma unroll
      for (int m = 0; m < 5; ++m) mu[m] += kk * s_h[m][(oy + ki) * TILE + ox];
    }
    float l1 = 0.0f, ss = 0.0f;
    if (inside) {
      float mx = mu[0], my = mu[1];
      float sx2 = mu[2] - mx * mx;
      float sy2 = mu[3] - my * my;
      float sxy = mu[4] - mx * my;
      float a1 = 2.0f * mx * my + C1;
      float a2 = 2.0f * sxy + C2;
      float b1 = mx * mx + my * my + C1;
      float b2 = sx2 + sy2 + C2;
      float num = a1 * a2;
      float den = b1 * b2;
      ss = num / den;
      float inv_den = 1.0f / den;
      float pc = inv_den * (2.0f * my * a2) - (num / (den * den)) * (2.0f * mx * b2);
      float qc = inv_den * (2.0f * a1);
      float rc = -(num / (den * den)) * b1;
      maps[(size_t)(0 * 3 + ch) * total_pix + pix] = pc;
      maps[(size_t)(1 * 3 + ch) * total_pix + pix] = qc;
      maps[(size_t)(2 * 3 + ch) * total_pix + pix] = rc;
      maps[(size_t)(3 * 3 + ch) * total_pix + pix] = qc * my;
      maps[(size_t)(4 * 3 + ch) * total_pix + pix] = rc * mx;
      float xv = s_x[(oy + 5) * HP + ox + 5];
      float yv = s_y[(oy + 5) * HP + ox + 5];
      float d = xv - yv;
      l1 = __builtin_fabsf(d);
      // Rust f32::signum: +1 for +0.0
      float sg = d != d ? d : __builtin_copysignf(1.0f, d);
      dl[pix * 3 + ch] = (1.0f - lam) * sg / m_total;
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
  __shared__ float s_m[5][HN];
  __shared__ float s_h[5][HH];
  const int t = threadIdx.x;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);  // halo-sharing neighbours on one XCD
  int v = 0;
  while (v + 1 < in.n_views && in.cams[v + 1].tile_base <= tile) ++v;
  const CamDev& c = in.cams[v];
  const int lt = tile - c.tile_base;
  const int tx = lt % c.ntx, ty = lt / c.ntx;
  const int x0 = tx * TILE - 5, y0 = ty * TILE - 5;
  const int ox = t & 15, oy = t >> 4;
  const int px = tx * TILE + ox, py = ty * TILE + oy;
  const bool inside = px < c.w && py < c.h;
  const size_t pix = (size_t)c.pix_base + (size_t)py * c.w + px;
  const float m_total = (float)((size_t)c.w * c.h * 3);
  const float lam = in.lambda;
  for (int ch = 0; ch < 3; ++ch) {
    for (int idx = t; idx < HN; idx += LT) {
      int gy = y0 + idx / HP, gx = x0 + idx % HP;
      bool ok = gx >= 0 && gy >= 0 && gx < c.w && gy < c.h;
      size_t gp = (size_t)c.pix_base + (size_t)gy * c.w + gx;
#pragma unroll
      for (int m = 0; m < 5; ++m) s_m[m][idx] = ok ? maps[(size_t)(m * 3 + ch) * total_pix + gp] : 0.0f;
    }
    __syncthreads();
    for (int idx = t; idx < HH; idx += LT) {
      int r = idx / TILE, cc = idx % TILE;
      float a[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ki = 0; ki < 11; ++ki) {
        float kk = K.k[ki];
#pragma unroll
        for (int m = 0; m < 5; ++m) a[m] += kk * s_m[m][r * HP + cc + ki];
      }
#pragma unroll
      for (int m = 0; m < 5; ++m) s_h[m][idx] = a[m];
    }
    __syncthreads();
    float b[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ki = 0; ki < 11; ++ki) {
      float kk = K.k[ki];
#pragma unroll
      for (int m = 0; m < 5; ++m) b[m] += kk * s_h[m][(oy + ki) * TILE + ox];
    }
    if (inside) {
      float xv = in.rgb[pix * 3 + ch];
      float yv = in.target[((size_t)c.tgt_base + (size_t)py * c.w + px) * 4 + ch];
      float grad = b[0] + yv * b[1] - b[3] + 2.0f * xv * b[2] - 2.0f * b[4];
      dl[pix * 3 + ch] += -lam * grad / m_total;
    }
    __syncthreads();
  }
}

// One block per view: loss from the per-tile partial sums.
__global__ __launch_bounds__(256) void k_loss_reduce(LossIn in, const float* __restrict__ partial,
                                                     float* __restrict__ loss_out) {
  __shared__ double s_l1[256];
  __shared__ double s_ss[256];
  const int v = blockIdx.x;
  const CamDev& c = in.cams[v];
  const int nt = c.ntx * c.nty;
  double l1 = 0.0, ss = 0.0;
  for (int i = threadIdx.x; i < nt * 3; i += 256) {
    size_t tile = (size_t)c.tile_base + i / 3;
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
  if (in.total_tiles <= 0) return;
  static const Kern11 K = make_kernel();
  hipLaunchKernelGGL(k_ssim_fwd, dim3(in.total_tiles), dim3(LT), 0, st, in, K, dl, ssim_maps, partial,
                     in.total_pix);
  hipLaunchKernelGGL(k_ssim_bwd, dim3(in.total_tiles), dim3(LT), 0, st, in, K, dl, ssim_maps, in.total_pix);
  hipLaunchKernelGGL(k_loss_reduce, dim3(in.n_views), dim3(256), 0, st, in, partial, loss_out);
}

}  // namespace rgs

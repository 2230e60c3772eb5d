// k_raster.hip -- front-to-back alpha blending and its analytic backward.
//
// One 256-thread workgroup per 16x16 tile (the reference's binning tile,
// backward.rs:21), one thread per pixel. Tile lists are consumed in batches of
// 256 splats staged through LDS; the workgroup stops as soon as every pixel
// has saturated (T < 1e-4). The exp2 table of the exact expf lives in LDS.
//
// Forward follows backward.rs:296-385 (mode 0) and rasterizer.rs:72-192
// (mode 1, RGBA). Backward follows backward.rs:391-512: each pixel walks its
// list from last_contrib down to 0, recovering the transmittance in front of
// each splat by division (no per-fragment storage); the per-splat gradient of
// the 64 pixels of a wave is reduced across lanes, summed over the four waves
// in LDS, and added to the per-splat accumulators with one FP32 atomic per
// value per tile.
#include "rgs_kernels.h"

namespace rgs {

constexpr int RT = 256;

__device__ __forceinline__ int view_of_tile(const CamDev* cams, int n_views, int tile) {
  int v = 0;
  while (v + 1 < n_views && cams[v + 1].tile_base <= tile) ++v;
  return v;
}

// Bit w of the result: the splat's alpha-reach box meets the pixel centres
// of wave w (rows 4w..4w+3 of the tile). NaN inputs give "meets".
__device__ __forceinline__ uint32_t wave_mask(float sx, float sy, float2 reach, int tx, int ty) {
  const float x_lo = (float)(tx * TILE) + 0.5f, x_hi = (float)(tx * TILE) + 15.5f;
  const bool out_x = (sx + reach.x < x_lo) || (sx - reach.x > x_hi);
  uint32_t m = 0;
#pragma unroll
  for (int w = 0; w < RT / 64; ++w) {
    const float y_lo = (float)(ty * TILE + 4 * w) + 0.5f, y_hi = (float)(ty * TILE + 4 * w) + 3.5f;
    const bool out = out_x || (sy + reach.y < y_lo) || (sy - reach.y > y_hi);
    m |= out ? 0u : (1u << w);
  }
  return m;
}

template <int MODE>
__global__ __launch_bounds__(RT) void k_raster_fwd(RasterIn in, float* __restrict__ out_rgb,
                                                   float* __restrict__ out_t, int32_t* __restrict__ out_last) {
  __shared__ uint64_t s_tab[32];
  __shared__ float4 s_a[RT];
  __shared__ float4 s_b[RT];
  __shared__ float s_c[RT];
  __shared__ uint32_t s_m[RT];
  const int t = threadIdx.x;
  const uint32_t wbit = 1u << (t >> 6);
  if (t < 32) s_tab[t] = kExp2Tab[t];
  const int tile = in.tile_order[xcd_remap(blockIdx.x, gridDim.x)];
  const int v = view_of_tile(in.cams, in.n_views, tile);
  const CamDev& c = in.cams[v];
  const int lt = tile - c.tile_base;
  const int tx = lt % c.ntx, ty = lt / c.ntx;
  const int px = tx * TILE + (t & 15), py = ty * TILE + (t >> 4);
  const bool inside = px < c.w && py < c.h;
  uint32_t start = in.rstart[tile], end = in.rend[tile];
  if (end < start) start = end;  // no pair reached this tile
  const float bg0 = in.bg[0], bg1 = in.bg[1], bg2 = in.bg[2];
  float T = 1.0f;
  float C0 = MODE == 1 ? bg0 : 0.0f, C1 = MODE == 1 ? bg1 : 0.0f, C2 = MODE == 1 ? bg2 : 0.0f;
  int32_t last = -1;
  bool done = !inside;
  const float fpx = (float)px + 0.5f, fpy = (float)py + 0.5f;
  for (uint32_t b0 = start; b0 < end; b0 += RT) {
    if (__syncthreads_count(done) == RT) break;
    uint32_t idx = b0 + t;
    if (idx < end) {
      uint32_t vi = in.pair_vals[idx];
      float4 A = in.S.xy_ab[vi];
      s_a[t] = A;
      s_b[t] = in.S.c_op_rg[vi];
      s_c[t] = in.S.col_b[vi];
      s_m[t] = wave_mask(A.x, A.y, in.S.reach[vi], tx, ty);
    }
    __syncthreads();
    const int nb = (int)min((uint32_t)RT, end - b0);
    if (!done) {
      for (int j = 0; j < nb; ++j) {
        if (!(s_m[j] & wbit)) continue;  // wave-uniform
        const float4 A = s_a[j];
        const float4 B = s_b[j];
        float dx = fpx - A.x;
        float dy = fpy - A.y;
        float power = -0.5f * (A.z * dx * dx + 2.0f * A.w * dx * dy + B.x * dy * dy);
        if (power > 0.0f) continue;
        float alpha = fmin_(B.y * expf_tab(power, s_tab), MAX_ALPHA);
        if (alpha < ALPHA_THRESHOLD) continue;
        float weight = alpha * T;
        C0 += B.z * weight;
        C1 += B.w * weight;
        C2 += s_c[j] * weight;
        T *= 1.0f - alpha;
        last = (int32_t)(b0 - start) + j;
        if (T < TRANSMITTANCE_STOP) {
          done = true;
          break;
        }
      }
    }
  }
  if (!inside) return;
  const size_t pix = (size_t)c.pix_base + (size_t)py * c.w + px;
  if (MODE == 0) {
    if (start == end) {
      out_rgb[pix * 3 + 0] = 0.0f;
      out_rgb[pix * 3 + 1] = 0.0f;
      out_rgb[pix * 3 + 2] = 0.0f;
    } else {
      out_rgb[pix * 3 + 0] = C0 + T * bg0;
      out_rgb[pix * 3 + 1] = C1 + T * bg1;
      out_rgb[pix * 3 + 2] = C2 + T * bg2;
    }
    out_t[pix] = T;
    out_last[pix] = last;
  } else {
    if (start == end) return;
    out_rgb[pix * 4 + 0] = C0;
    out_rgb[pix * 4 + 1] = C1;
    out_rgb[pix * 4 + 2] = C2;
    out_rgb[pix * 4 + 3] = 1.0f - T;
  }
}

void launch_raster_fwd(const RasterIn& in, int mode, float* rgb, float* final_t, int32_t* last, hipStream_t st) {
  if (in.total_tiles <= 0) return;
  if (mode == 0)
    hipLaunchKernelGGL(k_raster_fwd<0>, dim3(in.total_tiles), dim3(RT), 0, st, in, rgb, final_t, last);
  else
    hipLaunchKernelGGL(k_raster_fwd<1>, dim3(in.total_tiles), dim3(RT), 0, st, in, rgb, final_t, last);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ float read_lane_f(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
// Sum over the 64 lanes (all lanes active): butterfly inside each row of 16
// with DPP, then the four row sums through readlane. Wave-uniform result.
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f<0x141>(v);  // row_half_mirror
  v += dpp_f<0x140>(v);  // row_mirror
  return (read_lane_f(v, 0) + read_lane_f(v, 16)) + (read_lane_f(v, 32) + read_lane_f(v, 48));
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

__global__ __launch_bounds__(RT) void k_raster_bwd(RasterIn in, const float* __restrict__ final_t,
                                                   const int32_t* __restrict__ last_contrib,
                                                   const float* __restrict__ dl_dpix, float* __restrict__ accum) {
  __shared__ uint64_t s_tab[32];
  __shared__ float4 s_a[RT];
  __shared__ float4 s_b[RT];
  __shared__ float s_c[RT];
  __shared__ uint32_t s_vi[RT];
  __shared__ uint32_t s_m[RT];
  __shared__ float s_g[9][RT];
  __shared__ int s_max;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const uint32_t wbit = 1u << (t >> 6);
  if (t < 32) s_tab[t] = kExp2Tab[t];
  if (t == 0) s_max = -1;
  const int tile = in.tile_order[xcd_remap(blockIdx.x, gridDim.x)];
  const int v = view_of_tile(in.cams, in.n_views, tile);
  const CamDev& c = in.cams[v];
  const int lt = tile - c.tile_base;
  const int tx = lt % c.ntx, ty = lt / c.ntx;
  const int px = tx * TILE + (t & 15), py = ty * TILE + (t >> 4);
  const bool inside = px < c.w && py < c.h;
  uint32_t start = in.rstart[tile], end = in.rend[tile];
  if (end < start) start = end;  // no pair reached this tile
  int lc = -1;
  float T = 1.0f, dl0 = 0.0f, dl1 = 0.0f, dl2 = 0.0f;
  if (inside && start != end) {
    const size_t pix = (size_t)c.pix_base + (size_t)py * c.w + px;
    lc = last_contrib[pix];
    T = final_t[pix];
    dl0 = dl_dpix[pix * 3 + 0];
    dl1 = dl_dpix[pix * 3 + 1];
    dl2 = dl_dpix[pix * 3 + 2];
  }
  const float bg0 = in.bg[0], bg1 = in.bg[1], bg2 = in.bg[2];
  float bh0 = T * bg0, bh1 = T * bg1, bh2 = T * bg2;
  const int wave_lc = wave_max_i(lc);
  __syncthreads();
  if (lane == 0 && wave_lc >= 0) atomicMax(&s_max, wave_lc);
  __syncthreads();
  const int top = s_max;
  if (top < 0) return;
  const float fpx = (float)px + 0.5f, fpy = (float)py + 0.5f;
  for (int hi = top; hi >= 0; hi -= RT) {
    const int lo = max(0, hi - (RT - 1));
    const int nb = hi - lo + 1;
    __syncthreads();
    if (t < nb) {
      uint32_t vi = in.pair_vals[start + lo + t];
      s_vi[t] = vi;
      float4 A = in.S.xy_ab[vi];
      s_a[t] = A;
      s_b[t] = in.S.c_op_rg[vi];
      s_c[t] = in.S.col_b[vi];
      s_m[t] = wave_mask(A.x, A.y, in.S.reach[vi], tx, ty);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) s_g[k][t] = 0.0f;
    __syncthreads();
    // list entries above this wave's last contributor cannot contribute
    const int jtop = min(nb - 1, wave_lc - lo);
    for (int j = jtop; j >= 0; --j) {
      if (!(s_m[j] & wbit)) continue;  // wave-uniform
      const int li = lo + j;
      float g[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) g[k] = 0.0f;
      bool contrib = false;
      if (li <= lc) {
        const float4 A = s_a[j];
        const float4 B = s_b[j];
        const float cb = s_c[j];
        float dx = fpx - A.x;
        float dy = fpy - A.y;
        float power = -0.5f * (A.z * dx * dx + 2.0f * A.w * dx * dy + B.x * dy * dy);
        if (!(power > 0.0f)) {  // NaN proceeds, as in the reference
          float g_exp = expf_tab(power, s_tab);
          float alpha_raw = B.y * g_exp;
          float alpha = fmin_(alpha_raw, MAX_ALPHA);
          if (alpha >= ALPHA_THRESHOLD) {
            contrib = true;
            // transmittance in front of this splat (only gradients use it)
            float inv_1ma = __builtin_amdgcn_rcpf(1.0f - alpha);
            T = T * inv_1ma;
            float weight = alpha * T;
            g[0] = dl0 * weight;
            g[1] = dl1 * weight;
            g[2] = dl2 * weight;
            float dl_dalpha = dl0 * (T * B.z - bh0 * inv_1ma);
            dl_dalpha += dl1 * (T * B.w - bh1 * inv_1ma);
            dl_dalpha += dl2 * (T * cb - bh2 * inv_1ma);
            float dl_draw = alpha_raw > MAX_ALPHA ? 0.0f : dl_dalpha;
            g[3] = dl_draw * g_exp;
            float dl_dpower = dl_draw * B.y * g_exp;
            g[4] = dl_dpower * (-0.5f * dx * dx);
            g[5] = dl_dpower * (-dx * dy);
            g[6] = dl_dpower * (-0.5f * dy * dy);
            g[7] = dl_dpower * (A.z * dx + A.w * dy);
            g[8] = dl_dpower * (A.w * dx + B.x * dy);
            bh0 += weight * B.z;
            bh1 += weight * B.w;
            bh2 += weight * cb;
          }
        }
      }
      const uint64_t m = __ballot(contrib);
      if (m == 0) continue;
      if (__popcll(m) <= 2) {
        if (contrib) {
#pragma unroll
          for (int k = 0; k < 9; ++k) atomicAdd(&s_g[k][j], g[k]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          float s = wave_sum(g[k]);
          if (lane == 0) atomicAdd(&s_g[k][j], s);
        }
      }
    }
    __syncthreads();
    if (t < nb) {
      const uint32_t vi = s_vi[t];
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        float s = s_g[k][t];
        if (s != 0.0f) atomicAdd(&accum[(size_t)vi * 9 + k], s);
      }
    }
  }
}

void launch_raster_bwd(const RasterIn& in, const float* final_t, const int32_t* last, const float* dl_dpix,
                       float* accum, hipStream_t st) {
  if (in.total_tiles <= 0) return;
  hipLaunchKernelGGL(k_raster_bwd, dim3(in.total_tiles), dim3(RT), 0, st, in, final_t, last, dl_dpix, accum);
}

// Algorithmic work of the last render: pairs the forward evaluates per pixel
// (up to and including the saturating splat, else the whole tile list) and
// pairs the backward evaluates (last contributor + 1).
__global__ __launch_bounds__(RT) void k_work_counts(RasterIn in, const float* __restrict__ final_t,
                                                    const int32_t* __restrict__ last,
                                                    unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s_f[RT / 64], s_b[RT / 64];
  const int t = threadIdx.x;
  const int tile = blockIdx.x;
  const int v = view_of_tile(in.cams, in.n_views, tile);
  const CamDev& c = in.cams[v];
  const int lt = tile - c.tile_base;
  const int px = (lt % c.ntx) * TILE + (t & 15), py = (lt / c.ntx) * TILE + (t >> 4);
  unsigned long long f = 0, b = 0;
  if (px < c.w && py < c.h) {
    const size_t pix = (size_t)c.pix_base + (size_t)py * c.w + px;
    const uint32_t s0 = in.rstart[tile], e0 = in.rend[tile];
    const uint32_t len = e0 > s0 ? e0 - s0 : 0u;
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
                        hipStream_t st) {
  if (in.total_tiles <= 0) return;
  hipLaunchKernelGGL(k_work_counts, dim3(in.total_tiles), dim3(RT), 0, st, in, final_t, last, out);
}

}  // namespace rgs

// k_project.hip -- projection with EWA 2-D covariance and degree-3 SH colour.
//
// One thread per (Gaussian, view): a one-wave workgroup per 64 consecutive
// Gaussians and one view, the V views of a chunk in consecutive workgroups
// of one XCD (they read the same parameter lines, from its L2), so every
// load and store is 256 B contiguous per wave and no thread walks the views
// serially. Follows backward.rs:210-280 (project)
// with gaussian_aabb_2d (projection.rs:138-165) and the tile rect of
// backward.rs:275-278, keeping the reference's cull order and rounding so that
// depths and tile rects are bit-identical to its CPU path.
//
// Reads 59 floats per Gaussian from the SoA parameter rows and writes 64 B
// per (Gaussian, view); HBM-bound.
#include "rgs_kernels.h"

namespace rgs {

__global__ __launch_bounds__(64) void k_project(const float* __restrict__ P, int cap, int n,
                                                 const CamDev* __restrict__ cams, int n_views, SplatArrays S,
                                                 uint32_t* __restrict__ sort_keys, uint32_t* __restrict__ sort_vals,
                                                 uint32_t* __restrict__ status, int tight,
                                                 uint8_t* __restrict__ cover) {
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int v = lb % n_views;
  const int i = (lb / n_views) * 64 + (int)threadIdx.x;
  if (i >= n) return;
  const CamDev& c = cams[v];
  const size_t vi = (size_t)v * cap + i;
  const size_t e = (size_t)v * n + i;
  // every parameter but the SH rows, loaded up front
  const float op_logit = P[(size_t)P_OP * cap + i];
  const float px = P[(size_t)0 * cap + i], py = P[(size_t)1 * cap + i], pz = P[(size_t)2 * cap + i];
  const float qw = P[(size_t)(P_ROT + 0) * cap + i], qx = P[(size_t)(P_ROT + 1) * cap + i];
  const float qy = P[(size_t)(P_ROT + 2) * cap + i], qz = P[(size_t)(P_ROT + 3) * cap + i];
  const float l0 = P[(size_t)(P_LS + 0) * cap + i], l1s = P[(size_t)(P_LS + 1) * cap + i];
  const float l2s = P[(size_t)(P_LS + 2) * cap + i];
  sort_vals[e] = (uint32_t)vi;
  auto cull = [&]() {
    sort_keys[e] = 0xffffffffu;
    S.cnt[vi] = 0;
  };
  const float opacity = sigmoid(op_logit);
  if (opacity < ALPHA_THRESHOLD) {
    cull();
    return;
  }
  float pc[3];
  to_camera(c, px, py, pz, pc);
  float tz = pc[2];
  if (tz <= 0.001f) {
    cull();
    return;
  }
  float sx = c.fx * pc[0] / tz + c.cx;
  float sy = c.fy * pc[1] / tz + c.cy;
  const float margin = 128.0f;
  if (sx < -margin || sx > (float)c.w + margin || sy < -margin || sy > (float)c.h + margin) {
    cull();
    return;
  }
  const M3 sigma = covariance_3d(qw, qx, qy, qz, rgs_expf(l0), rgs_expf(l1s), rgs_expf(l2s));
  // project_covariance (projection.rs:24-69): tz clamped at 0.001 (no-op here)
  M3 T = ewa_T(c, pc[0], pc[1], fmax_(tz, 0.001f));
  M3 s2d = m3_mul(m3_mul(T, sigma), m3_t(T));
  float a = s2d.a[0][0] + 0.3f;
  float b = s2d.a[0][1];
  float cc = s2d.a[1][1] + 0.3f;
  float det = a * cc - b * b;
  if (__builtin_fabsf(det) < 1e-8f) {
    cull();
    return;
  }
  float inv_det = 1.0f / det;

  // view_dir = normalize(cam - pos) (sh.rs:122-139)
  float u0 = c.cam_pos[0] - px, u1 = c.cam_pos[1] - py, u2 = c.cam_pos[2] - pz;
  float len = __builtin_sqrtf(u0 * u0 + u1 * u1 + u2 * u2);
  float d0, d1, d2;
  if (len < 1e-8f) {
    d0 = 0.0f; d1 = 0.0f; d2 = 1.0f;
  } else {
    d0 = u0 / len; d1 = u1 / len; d2 = u2 / len;
  }
  // gaussian_aabb_2d(n_sigma = 3)
  float mid = (a + cc) * 0.5f;
  float disc = __builtin_sqrtf((a - cc) * (a - cc) * 0.25f + b * b);
  float l1 = fmax_(mid + disc, 0.0f);
  float l2 = fmax_(mid - disc, 0.0f);
  float r = 3.0f * fmax_(__builtin_sqrtf(l1), __builtin_sqrtf(l2));
  int32_t x0 = max(as_i32(sx - r), 0);
  int32_t y0 = max(as_i32(sy - r), 0);
  int32_t x1 = min((int32_t)((uint32_t)as_i32(sx + r) + 1u), c.w);
  int32_t y1 = min((int32_t)((uint32_t)as_i32(sy + r) + 1u), c.h);
  if (x0 >= x1 || y0 >= y1) {
    cull();
    return;
  }
  int32_t tx0 = x0 / TILE, ty0 = y0 / TILE;
  int32_t tx1 = (x1 + TILE - 1) / TILE, ty1 = (y1 + TILE - 1) / TILE;
  // binning clip (backward.rs:196-199)
  tx0 = max(tx0, 0);
  ty0 = max(ty0, 0);
  tx1 = min(tx1, c.ntx);
  ty1 = min(ty1, c.nty);
  uint32_t ntiles = (tx1 > tx0 && ty1 > ty0) ? (uint32_t)((tx1 - tx0) * (ty1 - ty0)) : 0u;
  if (cover && ntiles) {
    // the tiles whose reference lists hold this splat
    for (int32_t ty = ty0; ty < ty1; ++ty)
      for (int32_t tx = tx0; tx < tx1; ++tx) cover[c.tile_base + ty * c.ntx + tx] = 1;
  }

  // Reach of the alpha test: op * exp(-q/2) >= 1/255 needs the quadratic form
  // q <= R = 2 ln(255 op). The f32 conic and power carry relative errors that
  // grow with the condition number kappa = l1/l2 of the 2-D covariance
  // (cancellation in det and in the quadratic form, ~1e-6 kappa); R is widened
  // by 2% + 20x that estimate + 1e-3, and above kappa = 2e4 (or for any
  // non-finite input) R is +inf, which disables culling. Raster waves skip a
  // splat for pixel rows that the ellipse q <= R does not reach: those pixels
  // fail the alpha test anyway, so the results are unchanged.
  const float kappa = l1 / fmax_(l2, 1e-30f);
  const float q_err = 1e-6f * kappa;
  float reach_q = __builtin_inff();
  if (q_err <= 0.02f && a > 0.0f && cc > 0.0f && det > 0.0f) {
    reach_q = 2.0f * __logf(255.0f * opacity) * (1.02f + 20.0f * q_err) + 1e-3f;
    reach_q = fmax_(reach_q, 1e-3f);
    if (!(__builtin_isfinite(reach_q) && __builtin_isfinite(sx) && __builtin_isfinite(sy) &&
          __builtin_isfinite(cc * inv_det) && __builtin_isfinite(a * inv_det)))
      reach_q = __builtin_inff();
  }
  S.reach[vi] = reach_q;

  if (tight && __builtin_isfinite(reach_q)) {
    // tiles whose pixel centres [16 t + 0.5, 16 t + 15.5] meet the box of
    // the alpha-reach ellipse, |dx| <= sqrt(R a) + 0.5, |dy| <= sqrt(R c) + 0.5
    const float hx = __builtin_sqrtf(reach_q * a) + 0.5f, hy = __builtin_sqrtf(reach_q * cc) + 0.5f;
    const float fx0 = __builtin_ceilf((sx - hx - 15.5f) * (1.0f / TILE));
    const float fx1 = __builtin_floorf((sx + hx - 0.5f) * (1.0f / TILE)) + 1.0f;
    const float fy0 = __builtin_ceilf((sy - hy - 15.5f) * (1.0f / TILE));
    const float fy1 = __builtin_floorf((sy + hy - 0.5f) * (1.0f / TILE)) + 1.0f;
    // clamp in float before converting (the reference rect bounds the result)
    tx0 = max(tx0, (int32_t)fmax_(fx0, (float)tx0));
    ty0 = max(ty0, (int32_t)fmax_(fy0, (float)ty0));
    tx1 = min(tx1, (int32_t)fmin_(fx1, (float)tx1));
    ty1 = min(ty1, (int32_t)fmin_(fy1, (float)ty1));
    ntiles = (tx1 > tx0 && ty1 > ty0) ? (uint32_t)((tx1 - tx0) * (ty1 - ty0)) : 0u;
    if (ntiles == 0) tx1 = tx0, ty1 = ty0;
  }
  if (c.bin_shift && ntiles) {
    // pairs per bin of 2^shift x 2^shift tiles; the rect stays in tiles
    const int sh = c.bin_shift;
    ntiles = (uint32_t)(((((tx1 - 1) >> sh) + 1) - (tx0 >> sh)) * ((((ty1 - 1) >> sh) + 1) - (ty0 >> sh)));
  }
  if (tz != tz) atomicOr(status, ST_NAN_DEPTH);
  sort_keys[e] = f2u(tz);
  S.depth[vi] = tz;
  S.xy_ab[vi] = make_float4(sx, sy, cc * inv_det, -b * inv_det);
  S.cnt[vi] = ntiles;
  S.rect[vi] = make_uint2((uint32_t)tx0 | ((uint32_t)ty0 << 16), (uint32_t)tx1 | ((uint32_t)ty1 << 16));
  // colour, channel by channel (sh.rs's k = 0..15 order): 16 coefficients
  // live at a time
  float basis[16];
  sh_basis(d0, d1, d2, basis);
  float col[3];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    float cf[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) cf[k] = P[(size_t)(P_SH + ch * 16 + k) * cap + i];
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 16; ++k) acc += basis[k] * cf[k];
    float raw = 0.5f + acc;
    // f32::clamp(0, 1)
    if (raw < 0.0f) raw = 0.0f;
    if (raw > 1.0f) raw = 1.0f;
    col[ch] = raw;
  }
  S.c_op_rg[vi] = make_float4(a * inv_det, opacity, col[0], col[1]);
  S.col_b[vi] = col[2];
}

void launch_project(const float* P, int cap, int n, const CamDev* cams, int n_views, SplatArrays S,
                    uint32_t* sort_keys, uint32_t* sort_vals, uint32_t* status, int tight, uint8_t* cover,
                    hipStream_t st) {
  if (n <= 0 || n_views <= 0) return;
  const size_t blocks = (size_t)((n + 63) / 64) * (size_t)n_views;
  hipLaunchKernelGGL(k_project, dim3((unsigned)blocks), dim3(64), 0, st, P, cap, n, cams, n_views, S, sort_keys,
                     sort_vals, status, tight, cover);
}

__global__ void k_eval_expf(const float* __restrict__ x, float* __restrict__ out, size_t n) {
  __shared__ uint64_t s_tab[32];
  if (threadIdx.x < 32) s_tab[threadIdx.x] = kExp2Tab[threadIdx.x];
  __syncthreads();
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = expf_tab(x[i], s_tab);
}

void launch_eval_expf(const float* x, float* out, size_t n, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_eval_expf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, out, n);
}

}  // namespace rgs

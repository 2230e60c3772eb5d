// k_preview.hip -- the preview renderer: GpuRenderer::project and ::render
// (gpu/compute.rs:138-297, gpu/splat.wgsl:120-190, gpu/sort.rs:10-17).
//
// project: one thread per Gaussian evaluates splat.wgsl's `project` with
// each WGSL operation as one IEEE f32 operation in source order (WGSL's
// matrix constructors take columns, so the shader's Jacobian is the
// transposed perspective Jacobian and the 2-D covariance keeps no
// perspective terms; the quaternion is used unnormalised). It writes the
// SplatInfo2D record (zero where the shader culls, as wgpu's zeroed output
// buffer holds) and a depth key for depth_order: the depth's bits for
// splats with alpha >= 1/255 and depth > 0, ~0 otherwise.
//
// render: composite's loop over splats (nearest first, each over the pixels
// of its 3-sigma rect) becomes a loop per 16x16 tile over the splats whose
// rect meets the tile, in depth order, with the rect test per pixel: every
// pixel sees the same splats in the same order, so the f32 results equal
// composite's. A tile whose pixels a splat's alpha-reach ellipse cannot
// reach (alpha < 1/255 there, which composite skips) does not list it.
// Depth order: a stable radix sort of the depth keys (ties in index order;
// the reference's sort_unstable_by leaves them unspecified). (tile, sorted
// position) pairs are emitted in splat-major order and stably sorted by
// tile, so each tile's list stays in depth order.
#include "rgs_kernels.h"

namespace rgs {

__global__ __launch_bounds__(256) void k_preview_project(const float* __restrict__ P, int cap, uint32_t n,
                                                         PreviewCam cam, float* __restrict__ splats,
                                                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  auto row = [&](int j) { return P[(size_t)j * cap + i]; };
  float out[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) out[k] = 0.0f;
  uint32_t key = 0xffffffffu;
  const float px = row(P_POS), py = row(P_POS + 1), pz = row(P_POS + 2);
  float pc[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {  // camera.view * vec4(position, 1)
    float acc = cam.view[r * 4 + 0] * px;
    acc = acc + cam.view[r * 4 + 1] * py;
    acc = acc + cam.view[r * 4 + 2] * pz;
    acc = acc + cam.view[r * 4 + 3] * 1.0f;
    pc[r] = acc;
  }
  const float tx = pc[0], ty = pc[1], tz = pc[2];
  const float alpha = 1.0f / (1.0f + rgs_expf(-row(P_OP)));
  if (!(tz <= 0.001f) && !(alpha < ALPHA_THRESHOLD)) {
    const float sx = cam.fx * tx / tz + cam.cx;
    const float sy = cam.fy * ty / tz + cam.cy;
    const float w = row(P_ROT), x = row(P_ROT + 1), y = row(P_ROT + 2), z = row(P_ROT + 3);
    float R[3][3];
    R[0][0] = 1.0f - 2.0f * (y * y + z * z);
    R[1][0] = 2.0f * (x * y + w * z);
    R[2][0] = 2.0f * (x * z - w * y);
    R[0][1] = 2.0f * (x * y - w * z);
    R[1][1] = 1.0f - 2.0f * (x * x + z * z);
    R[2][1] = 2.0f * (y * z + w * x);
    R[0][2] = 2.0f * (x * z + w * y);
    R[1][2] = 2.0f * (y * z - w * x);
    R[2][2] = 1.0f - 2.0f * (x * x + y * y);
    const float S[3][3] = {{rgs_expf(row(P_LS)), 0.0f, 0.0f},
                           {0.0f, rgs_expf(row(P_LS + 1)), 0.0f},
                           {0.0f, 0.0f, rgs_expf(row(P_LS + 2))}};
    // C = A * B, summed over k in order (WGSL mat3x3 products)
    auto mul = [](const float A[3][3], const float B[3][3], float C[3][3]) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) C[r][c] = A[r][0] * B[0][c] + A[r][1] * B[1][c] + A[r][2] * B[2][c];
    };
    auto transpose = [](const float A[3][3], float B[3][3]) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) B[r][c] = A[c][r];
    };
    float RS[3][3], RSt[3][3], sig3[3][3];
    mul(R, S, RS);
    transpose(RS, RSt);
    mul(RS, RSt, sig3);
    const float J[3][3] = {{cam.fx / tz, 0.0f, 0.0f},
                           {0.0f, cam.fy / tz, 0.0f},
                           {-cam.fx * tx / (tz * tz), -cam.fy * ty / (tz * tz), 0.0f}};
    float W[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) W[r][c] = cam.view[r * 4 + c];
    float T[3][3], Tt[3][3], TS[3][3], sig2[3][3];
    mul(J, W, T);
    transpose(T, Tt);
    mul(T, sig3, TS);
    mul(TS, Tt, sig2);
    sig2[0][0] += 0.3f;
    sig2[1][1] += 0.3f;
    const float a = sig2[0][0], b = sig2[1][0], c = sig2[1][1];
    const float det = a * c - b * b;
    if (!(fabsf(det) < 1e-8f)) {
      const float inv_det = 1.0f / det;
      const float SH_C0 = 0.28209479177387814f;
      auto clamp01 = [](float v) { return fmin_(fmax_(v, 0.0f), 1.0f); };
      out[0] = sx;
      out[1] = sy;
      out[2] = c * inv_det;
      out[3] = -b * inv_det;
      out[4] = a * inv_det;
      out[5] = clamp01(0.5f + row(P_SH + 0) * SH_C0);
      out[6] = clamp01(0.5f + row(P_SH + 16) * SH_C0);
      out[7] = clamp01(0.5f + row(P_SH + 32) * SH_C0);
      out[8] = alpha;
      out[9] = tz;
      // depth_order's filter; a positive depth's bits order as the depth
      if (alpha >= ALPHA_THRESHOLD && tz > 0.0f) key = f2u(tz);
    }
  }
  float2* o = reinterpret_cast<float2*>(splats + (size_t)i * 10);
#pragma unroll
  for (int k = 0; k < 5; ++k) o[k] = make_float2(out[2 * k], out[2 * k + 1]);
  keys[i] = key;
  vals[i] = i;
}

void launch_preview_project(const float* P, int cap, uint32_t n, const PreviewCam& cam, float* splats,
                            uint32_t* keys, uint32_t* vals, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_preview_project, dim3((n + 255) / 256), dim3(256), 0, st, P, cap, n, cam, splats, keys, vals);
}

// Rust `v as usize` of a floored value >= 0 (NaN -> 0), capped at w (the
// pixel loops past w are empty either way).
__device__ __forceinline__ int index_capped(float v, int w) {
  if (!(v > 0.0f)) return 0;
  if (v >= (float)w) return w;
  return (int)v;
}

// Pixel range [lo, hi) of the centres px + 0.5 within [s - hw, s + hw],
// cut to [r0, r1) (hw finite and >= 0; huge values saturate).
__device__ __forceinline__ void reach_range(float s, float hw, int r0, int r1, int& lo, int& hi) {
  const float flo = ceilf(s - hw - 0.5f), fhi = floorf(s + hw - 0.5f) + 1.0f;
  lo = !(flo > (float)r0) ? r0 : (flo >= (float)r1 ? r1 : (int)flo);
  hi = !(fhi < (float)r1) ? r1 : (fhi <= (float)r0 ? r0 : (int)fhi);
}

// composite's rect (compute.rs:249-260) for the splat at each sorted
// position, and the tiles to list it in: those of the rect that the splat's
// alpha-reach ellipse meets. Outside the ellipse q <= R (q the conic form,
// R = 2 ln(255 alpha) with a margin for the f32 evaluation of q, which
// grows with the conic's condition number; no cut when that is too large)
// alpha stays below 1/255, so composite skips those pixels anyway: the cut
// drops no pixel's splat. The composite still tests the full rect per pixel.
__global__ __launch_bounds__(256) void k_preview_rects(const float* __restrict__ splats,
                                                       const uint32_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ order, uint32_t n, int w, int h,
                                                       int4* __restrict__ rects, int4* __restrict__ trects,
                                                       uint64_t* __restrict__ cnt) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int4 r = make_int4(0, 0, 0, 0), tr = make_int4(0, 0, 0, 0);
  uint64_t tiles = 0;
  if (keys[p] != 0xffffffffu) {
    const float* s = splats + (size_t)order[p] * 10;
    const float a = s[2], b = s[3], c = s[4];
    const float mid = 0.5f * (a + c);
    const float disc = __builtin_sqrtf(0.25f * (a - c) * (a - c) + b * b);
    const float lambda_min = fmax_(mid - disc, 1e-6f);
    const float rad = 3.0f * __builtin_sqrtf(1.0f / lambda_min);
    r.x = index_capped(fmax_(floorf(s[0] - rad), 0.0f), w);
    r.y = index_capped(fmax_(floorf(s[1] - rad), 0.0f), h);
    const float ex = ceilf(s[0] + rad), ey = ceilf(s[1] + rad);  // `as i64`, then clamp(0, w)
    r.z = ex != ex ? 0 : (int)fmin_(fmax_(ex, 0.0f), (float)w);
    r.w = ey != ey ? 0 : (int)fmin_(fmax_(ey, 0.0f), (float)h);
    // the alpha-reach box, in pixels, within the rect
    int x0 = r.x, x1 = r.z, y0 = r.y, y1 = r.w;
    const float l1 = mid + disc, l2 = mid - disc;  // conic eigenvalues
    const float q_err = 1e-6f * (fmax_(l1, 0.0f) / fmax_(l2, 1e-30f));
    const float det = a * c - b * b;
    if (q_err <= 0.02f && l2 > 0.0f && det > 0.0f) {
      const float R = fmax_(2.0f * logf(255.0f * s[8]) * (1.02f + 20.0f * q_err) + 1e-3f, 1e-3f);
      const float hx = __builtin_sqrtf(R * (c / det)) + 0.5f, hy = __builtin_sqrtf(R * (a / det)) + 0.5f;
      if (__builtin_isfinite(hx) && __builtin_isfinite(hy) && __builtin_isfinite(s[0]) && __builtin_isfinite(s[1])) {
        reach_range(s[0], hx, r.x, r.z, x0, x1);
        reach_range(s[1], hy, r.y, r.w, y0, y1);
      }
    }
    if (x1 > x0 && y1 > y0) {
      tr = make_int4(x0 / TILE, y0 / TILE, (x1 - 1) / TILE + 1, (y1 - 1) / TILE + 1);
      tiles = (uint64_t)(tr.z - tr.x) * (uint64_t)(tr.w - tr.y);
    }
  }
  rects[p] = r;
  trects[p] = tr;
  cnt[p] = tiles;
}

void launch_preview_rects(const float* splats, const uint32_t* keys, const uint32_t* order, uint32_t n, int w, int h,
                          int4* rects, int4* trects, uint64_t* cnt, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_preview_rects, dim3((n + 255) / 256), dim3(256), 0, st, splats, keys, order, n, w, h, rects,
                     trects, cnt);
}

// (tile, sorted position) pairs, one thread per pair: pair q belongs to the
// last splat p with off[p] <= q (off is the exclusive scan of the tile
// counts, so splats without tiles are passed over), and is that splat's
// (q - off[p])-th tile of [tx0, tx1) x [ty0, ty1) in row-major order. Two
// searches over off bound the splats of the workgroup's 256 pairs; their
// offsets are staged in LDS and each thread searches there (in global
// memory when more than 256 splats share the 256 pairs).
__device__ __forceinline__ uint32_t last_at_or_below(const uint64_t* off, uint32_t lo, uint32_t hi, uint64_t q) {
  // off[lo] <= q < off[hi] (off[hi] may be the virtual end)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (off[mid] <= q) lo = mid;
    else hi = mid;
  }
  return lo;
}

// blk[b] = the splat of pair min(256 b, m - 1), for b = 0 .. nb.
__global__ __launch_bounds__(256) void k_preview_pair_blocks(const uint64_t* __restrict__ off, uint32_t n, uint32_t m,
                                                             uint32_t nb, uint32_t* __restrict__ blk) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nb) return;
  blk[b] = last_at_or_below(off, 0, n, (uint64_t)min(b * 256u, m - 1));
}

__global__ __launch_bounds__(256) void k_preview_pairs(const int4* __restrict__ trects,
                                                       const uint64_t* __restrict__ off, uint32_t n, uint32_t m,
                                                       const uint32_t* __restrict__ blk, int ntx,
                                                       uint32_t* __restrict__ pk, uint32_t* __restrict__ pv) {
  __shared__ uint64_t s_off[256];
  const uint32_t q = blockIdx.x * 256u + threadIdx.x;
  const uint32_t lo = blk[blockIdx.x], span = blk[blockIdx.x + 1] - lo + 1;
  if (span <= 256 && threadIdx.x < span) s_off[threadIdx.x] = off[lo + threadIdx.x];
  __syncthreads();
  if (q >= m) return;
  uint32_t p;
  uint64_t first;
  if (span <= 256) {
    uint32_t a = 0, b = span;  // s_off[a] <= q < s_off[b] (s_off[span] virtual)
    while (b - a > 1) {
      const uint32_t mid = (a + b) / 2;
      if (s_off[mid] <= q) a = mid;
      else b = mid;
    }
    p = lo + a;
    first = s_off[a];
  } else {
    p = last_at_or_below(off, lo, lo + span, q);
    first = off[p];
  }
  const int4 r = trects[p];
  const uint32_t k = q - (uint32_t)first, wt = (uint32_t)(r.z - r.x);
  pk[q] = (uint32_t)((r.y + (int)(k / wt)) * ntx + r.x + (int)(k % wt));
  pv[q] = p;
}

void launch_preview_pairs(const int4* trects, const uint64_t* off, uint32_t n, uint32_t m, int ntx, uint32_t* blk,
                          uint32_t* pk, uint32_t* pv, hipStream_t st) {
  if (!n || !m) return;
  const uint32_t nb = (m + 255) / 256;
  hipLaunchKernelGGL(k_preview_pair_blocks, dim3(nb / 256 + 1), dim3(256), 0, st, off, n, m, nb, blk);
  hipLaunchKernelGGL(k_preview_pairs, dim3(nb), dim3(256), 0, st, trects, off, n, m, blk, ntx, pk, pv);
}

// [start, end) of every tile's run in the tile-sorted pairs (tiles without
// pairs keep the zero fill).
__global__ __launch_bounds__(256) void k_preview_ranges(const uint32_t* __restrict__ pk, uint32_t m,
                                                        uint32_t* __restrict__ rs, uint32_t* __restrict__ re) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const uint32_t k = pk[q];
  if (q == 0 || pk[q - 1] != k) rs[k] = q;
  if (q + 1 == m || pk[q + 1] != k) re[k] = q + 1;
}

void launch_preview_ranges(const uint32_t* pk, uint32_t m, uint32_t* rs, uint32_t* re, hipStream_t st) {
  if (!m) return;
  hipLaunchKernelGGL(k_preview_ranges, dim3((m + 255) / 256), dim3(256), 0, st, pk, m, rs, re);
}

// composite (compute.rs:243-297) for one 16x16 tile per workgroup, a pixel
// per thread. The tile's splats are staged in LDS 256 at a time.
__global__ __launch_bounds__(256) void k_preview_composite(const float* __restrict__ splats,
                                                           const uint32_t* __restrict__ order,
                                                           const int4* __restrict__ rects,
                                                           const uint32_t* __restrict__ pv,
                                                           const uint32_t* __restrict__ rs,
                                                           const uint32_t* __restrict__ re, int w, int h, int ntx,
                                                           float bg0, float bg1, float bg2, float* __restrict__ rgb) {
  __shared__ float s_sp[256][9];
  __shared__ int4 s_rect[256];
  const int tile = blockIdx.x;
  const int t = threadIdx.x;
  const int px = (tile % ntx) * TILE + (t % TILE), py = (tile / ntx) * TILE + (t / TILE);
  const uint32_t b = rs[tile], e = re[tile];
  float cr = 0.0f, cg = 0.0f, cb = 0.0f, tr = 1.0f;
  const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
  const bool inside = px < w && py < h;
  for (uint32_t base = b; base < e; base += 256) {
    // composite skips a pixel once T < 1e-4 (T never grows): stop when every
    // pixel of the tile is there
    if (__syncthreads_count(inside && !(tr < TRANSMITTANCE_STOP)) == 0) break;
    if (base + t < e) {
      const uint32_t p = pv[base + t];
      const float* s = splats + (size_t)order[p] * 10;
#pragma unroll
      for (int k = 0; k < 9; ++k) s_sp[t][k] = s[k];
      s_rect[t] = rects[p];
    }
    __syncthreads();
    const int m = (int)min(256u, e - base);
    for (int j = 0; j < m; ++j) {
      const int4 r = s_rect[j];
      if (px < r.x || px >= r.z || py < r.y || py >= r.w) continue;
      const float T = tr;
      if (T < TRANSMITTANCE_STOP) continue;
      const float* s = s_sp[j];
      const float dx = fx - s[0];
      const float dy = fy - s[1];
      const float power = -0.5f * (s[2] * dx * dx + 2.0f * s[3] * dx * dy + s[4] * dy * dy);
      if (power > 0.0f) continue;
      const float alpha = fmin_(s[8] * rgs_expf(power), MAX_ALPHA);
      if (alpha < ALPHA_THRESHOLD) continue;
      const float weight = alpha * T;
      cr += s[5] * weight;
      cg += s[6] * weight;
      cb += s[7] * weight;
      tr = T * (1.0f - alpha);
    }
  }
  if (inside) {
    float* o = rgb + ((size_t)py * w + px) * 3;
    o[0] = cr + tr * bg0;
    o[1] = cg + tr * bg1;
    o[2] = cb + tr * bg2;
  }
}

void launch_preview_composite(const float* splats, const uint32_t* order, const int4* rects, const uint32_t* pv,
                              const uint32_t* rs, const uint32_t* re, int w, int h, int ntx, int nty, const float bg[3],
                              float* rgb, hipStream_t st) {
  hipLaunchKernelGGL(k_preview_composite, dim3(ntx * nty), dim3(256), 0, st, splats, order, rects, pv, rs, re, w, h,
                     ntx, bg[0], bg[1], bg[2], rgb);
}

}  // namespace rgs

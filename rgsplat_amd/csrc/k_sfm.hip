// k_sfm.hip -- initial Gaussians from an SfM point cloud (gaussians_from_sfm
// and estimate_initial_scales, pipeline/sfm.rs:365-455).
//
// The points are sorted by x with the radix sort (a stable sort on the
// 64-bit order-preserving key of x, as two 32-bit passes: low word, then
// high word), so equal x keep their input order as in the reference's
// stable sort_by. Then one thread per point repeats the reference's window
// search: its own x binary-searched in the sorted x (Rust's slice
// binary_search_by of edition-2024 toolchains: no early exit, ending on the
// last of equal keys), the window of 50 sorted points from 25 before it, the
// f64 distances to the others, and the mean of the 3 smallest. A last kernel
// writes the parameter rows.
#include "rgs_kernels.h"

namespace rgs {

__global__ __launch_bounds__(256) void k_sfm_keys(const double* __restrict__ xyz, uint32_t n,
                                                  uint32_t* __restrict__ klo, uint32_t* __restrict__ khi,
                                                  uint32_t* __restrict__ idx, uint32_t* __restrict__ status) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = xyz[3 * (size_t)i];
  if (x != x) atomicOr(status, 1u);  // partial_cmp().unwrap() panics on NaN
  if (x == 0.0) x = 0.0;             // -0 and +0 compare equal
  uint64_t u = __builtin_bit_cast(uint64_t, x);
  u = (u >> 63) ? ~u : (u | (1ull << 63));
  klo[i] = (uint32_t)u;
  khi[i] = (uint32_t)(u >> 32);
  idx[i] = i;
}

void launch_sfm_keys(const double* xyz, uint32_t n, uint32_t* klo, uint32_t* khi, uint32_t* idx, uint32_t* status,
                     hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_sfm_keys, dim3((n + 255) / 256), dim3(256), 0, st, xyz, n, klo, khi, idx, status);
}

__global__ __launch_bounds__(256) void k_sfm_gather_hi(const uint32_t* __restrict__ khi,
                                                       const uint32_t* __restrict__ order, uint32_t n,
                                                       uint32_t* __restrict__ keys) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) keys[p] = khi[order[p]];
}

void launch_sfm_gather_hi(const uint32_t* khi, const uint32_t* order, uint32_t n, uint32_t* keys, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_sfm_gather_hi, dim3((n + 255) / 256), dim3(256), 0, st, khi, order, n, keys);
}

// The points in x order, as three f64 rows (x, y, z).
__global__ __launch_bounds__(256) void k_sfm_sorted_xyz(const double* __restrict__ xyz,
                                                        const uint32_t* __restrict__ order, uint32_t n,
                                                        double* __restrict__ sxyz) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const size_t i = order[p];
  sxyz[p] = xyz[3 * i];
  sxyz[(size_t)n + p] = xyz[3 * i + 1];
  sxyz[2 * (size_t)n + p] = xyz[3 * i + 2];
}

void launch_sfm_sorted_xyz(const double* xyz, const uint32_t* order, uint32_t n, double* sxyz, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_sfm_sorted_xyz, dim3((n + 255) / 256), dim3(256), 0, st, xyz, order, n, sxyz);
}

// estimate_initial_scales (:410-455). Thread p takes the point at sorted
// position p (point order[p]), so that the threads of a wave search and read
// overlapping windows. A window entry is the point itself exactly when its
// sorted position is p (order is a permutation). The distances are kept
// squared: a correctly rounded sqrt is monotone, so the 3 smallest distances
// are the square roots of the 3 smallest squares, and a distance is NaN
// exactly when its square is.
__global__ __launch_bounds__(256) void k_sfm_scale(const double* __restrict__ sxyz,
                                                   const uint32_t* __restrict__ order, uint32_t n,
                                                   float* __restrict__ scale, uint32_t* __restrict__ status) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double* __restrict__ sx = sxyz;
  const double* __restrict__ sy = sxyz + n;
  const double* __restrict__ sz = sxyz + 2 * (size_t)n;
  const double px = sx[p], py = sy[p], pz = sz[p];
  // binary_search_by(|s| s.x.partial_cmp(&px)), then unwrap_or(0)
  uint32_t size = n, base = 0;
  while (size > 1) {
    const uint32_t half = size / 2, mid = base + half;
    base = sx[mid] > px ? base : mid;
    size -= half;
  }
  const uint32_t idx = sx[base] == px ? base : 0u;
  const uint32_t window = min(50u, n);
  const uint32_t start = idx >= window / 2 ? idx - window / 2 : 0u;
  const uint32_t end = min(start + window, n);
  // the three smallest squared distances, ascending
  double q0 = __builtin_inf(), q1 = __builtin_inf(), q2 = __builtin_inf(), first = 0.0;
  uint32_t cnt = 0;
  bool nan = false;
  for (uint32_t k = start; k < end; ++k) {
    if (k == p) continue;
    const double dx = sx[k] - px, dy = sy[k] - py, dz = sz[k] - pz;
    const double q = dx * dx + dy * dy + dz * dz;
    nan |= q != q;
    if (cnt++ == 0) first = q;
    if (q < q2) {
      if (q < q1) {
        q2 = q1;
        if (q < q0) {
          q1 = q0;
          q0 = q;
        } else {
          q1 = q;
        }
      } else {
        q2 = q;
      }
    }
  }
  // the distance sort's partial_cmp().unwrap() panics on a NaN once it
  // compares anything, i.e. from two distances on; a single (NaN) distance
  // is the scale itself
  if (nan && cnt >= 2) atomicOr(status, 1u);
  float s;
  if (cnt == 0) {
    s = 0.01f;
  } else if (cnt == 1) {
    s = (float)(-0.0 + __builtin_sqrt(first));
  } else {
    const uint32_t k = min(3u, cnt);
    double sum = -0.0;  // f64 Sum over the k smallest, ascending
    sum += __builtin_sqrt(q0);
    if (k > 1) sum += __builtin_sqrt(q1);
    if (k > 2) sum += __builtin_sqrt(q2);
    s = (float)(sum / (double)k);
  }
  scale[order[p]] = s;
}

void launch_sfm_scale(const double* sxyz, const uint32_t* order, uint32_t n, float* scale, uint32_t* status,
                      hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_sfm_scale, dim3((n + 255) / 256), dim3(256), 0, st, sxyz, order, n, scale, status);
}

// gaussians_from_sfm (:365-407). dc[c]: (powf(c / 255, 2.2) - 0.5) * SH_C0_INV
// for every byte c, from the host's libm; opacity: logit(0.1). ln of the
// scale is the f64 log rounded to f32.
__global__ __launch_bounds__(256) void k_sfm_gaussians(const double* __restrict__ xyz, const uint8_t* __restrict__ rgb,
                                                       const float* __restrict__ scale, uint32_t n,
                                                       const float* __restrict__ dc, float opacity_logit,
                                                       float* __restrict__ P, size_t cap) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  auto row = [&](int j) -> float& { return P[(size_t)j * cap + i]; };
  for (int j = 0; j < P_N; ++j) row(j) = 0.0f;
  for (int d = 0; d < 3; ++d) row(P_POS + d) = (float)xyz[3 * (size_t)i + d];
  row(P_ROT) = 1.0f;
  for (int ch = 0; ch < 3; ++ch) row(P_SH + 16 * ch) = dc[rgb[3 * (size_t)i + ch]];
  const float s = (float)::log((double)fmaxf(scale[i], 1e-4f));
  row(P_LS + 0) = s;
  row(P_LS + 1) = s;
  row(P_LS + 2) = s;
  row(P_OP) = opacity_logit;
}

void launch_sfm_gaussians(const double* xyz, const uint8_t* rgb, const float* scale, uint32_t n, const float* dc,
                          float opacity_logit, float* P, size_t cap, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_sfm_gaussians, dim3((n + 255) / 256), dim3(256), 0, st, xyz, rgb, scale, n, dc, opacity_logit,
                     P, cap);
}

}  // namespace rgs

// rgs_math.h -- scalar math shared by the gfx950 kernels (and compiled for the
// host by the exhaustive expf check).
//
// Every cull and threshold decision of the training step (sigmoid < 1/255,
// tz <= 0.001, |det| < 1e-8, the 3-sigma tile rect, power > 0, alpha < 1/255,
// T < 1e-4) depends on f32 results that must round exactly as the reference's
// CPU path does. So this header keeps the reference's operation order and
// relies on the build's -ffp-contract=off, IEEE division and square root, and
// an expf that is bit-identical to glibc's:
//
//   rgs_expf: the exp2f-table algorithm glibc 2.35 ships for expf
//   (sysdeps/ieee754/flt-32/e_expf.c, from Arm's optimized-routines),
//   in the form glibc's FMA build evaluates it (fused range reduction and
//   polynomial). Table entries are 2^(i/32) rounded to double minus i<<47, as
//   that algorithm defines them; the polynomial is its published degree-3
//   fit. Checked against the host's expf on all 2^32 inputs.
//
// Reference citations are to the rgsplat sources.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define RGS_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#include <string.h>
#define RGS_HD static inline
#endif

namespace rgs {

// backward.rs:21-24
constexpr int TILE = 16;
constexpr float ALPHA_THRESHOLD = 1.0f / 255.0f;
constexpr float TRANSMITTANCE_STOP = 1e-4f;
constexpr float MAX_ALPHA = 0.99f;

// Adam flat parameter order (optimizer.rs:165-173): position 0-2,
// rotation (w,x,y,z) 3-6, log_scale 7-9, opacity_logit 10, sh 11-58.
constexpr int P_POS = 0, P_ROT = 3, P_LS = 7, P_OP = 10, P_SH = 11, P_N = 59;

// 2^(i/32) as double bits minus (i << 47).
#define RGS_EXP2F_TABLE                                                                     \
  {0x3ff0000000000000ULL, 0x3fefd9b0d3158574ULL, 0x3fefb5586cf9890fULL, 0x3fef9301d0125b51ULL, \
   0x3fef72b83c7d517bULL, 0x3fef54873168b9aaULL, 0x3fef387a6e756238ULL, 0x3fef1e9df51fdee1ULL, \
   0x3fef06fe0a31b715ULL, 0x3feef1a7373aa9cbULL, 0x3feedea64c123422ULL, 0x3feece086061892dULL, \
   0x3feebfdad5362a27ULL, 0x3feeb42b569d4f82ULL, 0x3feeab07dd485429ULL, 0x3feea47eb03a5585ULL, \
   0x3feea09e667f3bcdULL, 0x3fee9f75e8ec5f74ULL, 0x3feea11473eb0187ULL, 0x3feea589994cce13ULL, \
   0x3feeace5422aa0dbULL, 0x3feeb737b0cdc5e5ULL, 0x3feec49182a3f090ULL, 0x3feed503b23e255dULL, \
   0x3feee89f995ad3adULL, 0x3feeff76f2fb5e47ULL, 0x3fef199bdd85529cULL, 0x3fef3720dcef9069ULL, \
   0x3fef5818dcfba487ULL, 0x3fef7c97337b9b5fULL, 0x3fefa4afa2a490daULL, 0x3fefd0765b6e4540ULL}

RGS_HD uint32_t f2u(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_uint(f);
#else
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
#endif
}
RGS_HD uint64_t d2u(double f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint64_t)__double_as_longlong(f);
#else
  uint64_t u;
  memcpy(&u, &f, 8);
  return u;
#endif
}
RGS_HD double u2d(uint64_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __longlong_as_double((long long)u);
#else
  double f;
  memcpy(&f, &u, 8);
  return f;
#endif
}

// expf with an explicit exp2 table (so a kernel may pass a copy in LDS).
RGS_HD float expf_tab(float x, const uint64_t* tab) {
  const double kInvLn2N = 0x1.71547652b82fep+0 * 32.0;
  const double kShift = 0x1.8p+52;
  const double c0 = 0x1.c6af84b912394p-5 / 32.0 / 32.0 / 32.0;
  const double c1 = 0x1.ebfce50fac4f3p-3 / 32.0 / 32.0;
  const double c2 = 0x1.62e42ff0c52d6p-1 / 32.0;
  uint32_t abstop = (f2u(x) >> 20) & 0x7ff;
  if (abstop >= 0x42b) {  // |x| >= 88 or x is inf/nan
    if (f2u(x) == 0xff800000u) return 0.0f;
    if (abstop >= 0x7f8) return x + x;
    if (x > 0x1.62e42ep6f) return __builtin_inff();
    if (x < -0x1.9fe368p6f) return 0.0f;
  }
  double xd = (double)x;
  double kd = __builtin_fma(kInvLn2N, xd, kShift);
  uint64_t ki = d2u(kd);
  kd -= kShift;
  double r = __builtin_fma(kInvLn2N, xd, -kd);
  uint64_t t = tab[ki & 31];
  t += ki << 47;
  double s = u2d(t);
  double z = __builtin_fma(c0, r, c1);
  double r2 = r * r;
  double y = __builtin_fma(c2, r, 1.0);
  y = __builtin_fma(z, r2, y);
  y = y * s;
  return (float)y;
}

#if defined(__HIP_DEVICE_COMPILE__)
__device__ __constant__ static const uint64_t kExp2Tab[32] = RGS_EXP2F_TABLE;
#else
static const uint64_t kExp2Tab[32] = RGS_EXP2F_TABLE;
#endif

RGS_HD float rgs_expf(float x) { return expf_tab(x, kExp2Tab); }

RGS_HD float rsqrt_ieee(float x) { return __builtin_sqrtf(x); }  // sqrtf, correctly rounded
RGS_HD float fmax_(float a, float b) { return __builtin_fmaxf(a, b); }
RGS_HD float fmin_(float a, float b) { return __builtin_fminf(a, b); }

// math/mod.rs:183-185
RGS_HD float sigmoid(float x) { return 1.0f / (1.0f + rgs_expf(-x)); }

// Rust `f as i32`: truncate toward zero, saturate, NaN -> 0.
RGS_HD int32_t as_i32(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}

struct M3 {
  float a[3][3];
};
RGS_HD M3 m3_zero() {
  M3 m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.a[i][j] = 0.0f;
  return m;
}
// C = A B with products summed in k order (the column-wise gemv the
// reference's 3x3 products evaluate).
RGS_HD M3 m3_mul(const M3& A, const M3& B) {
  M3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      float acc = A.a[i][0] * B.a[0][j];
      acc = acc + A.a[i][1] * B.a[1][j];
      acc = acc + A.a[i][2] * B.a[2][j];
      C.a[i][j] = acc;
    }
  return C;
}
RGS_HD M3 m3_t(const M3& A) {
  M3 T;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T.a[i][j] = A.a[j][i];
  return T;
}

// Gaussian3D::covariance_3d (math/mod.rs:89-96): unit quaternion by the
// library normalisation ((x^2 + z^2) + (y^2 + w^2)), rotation matrix in the
// ww+ii-jj-kk form, Sigma = (R S)(R S)^T.
RGS_HD M3 covariance_3d(float qw, float qx, float qy, float qz, float s0, float s1, float s2) {
  float a = qx * qx, b = qy * qy, c = qz * qz, d = qw * qw;
  a = a + c;
  b = b + d;
  float n = __builtin_sqrtf(a + b);
  float i = qx / n, j = qy / n, k = qz / n, w = qw / n;
  float ww = w * w, ii = i * i, jj = j * j, kk = k * k;
  float ij = i * j * 2.0f, wk = w * k * 2.0f, wj = w * j * 2.0f;
  float ik = i * k * 2.0f, jk = j * k * 2.0f, wi = w * i * 2.0f;
  M3 r;
  r.a[0][0] = ww + ii - jj - kk; r.a[0][1] = ij - wk; r.a[0][2] = wj + ik;
  r.a[1][0] = ij + wk; r.a[1][1] = ww - ii + jj - kk; r.a[1][2] = jk - wi;
  r.a[2][0] = ik - wj; r.a[2][1] = wi + jk; r.a[2][2] = ww - ii - jj + kk;
  M3 s = m3_zero();
  s.a[0][0] = s0; s.a[1][1] = s1; s.a[2][2] = s2;
  M3 rs = m3_mul(r, s);
  return m3_mul(rs, m3_t(rs));
}

// quaternion.rs:10-30: normalise (sequential sum) then the 1-2(..) formula.
RGS_HD void quat_normalize(const float q[4], float out[4]) {
  float n = __builtin_sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (n < 1e-12f) {
    out[0] = 1.0f; out[1] = 0.0f; out[2] = 0.0f; out[3] = 0.0f;
  } else {
    out[0] = q[0] / n; out[1] = q[1] / n; out[2] = q[2] / n; out[3] = q[3] / n;
  }
}
RGS_HD M3 mat_from_quat(const float q[4]) {
  float w = q[0], x = q[1], y = q[2], z = q[3];
  M3 m;
  m.a[0][0] = 1.0f - 2.0f * (y * y + z * z);
  m.a[0][1] = 2.0f * (x * y - w * z);
  m.a[0][2] = 2.0f * (x * z + w * y);
  m.a[1][0] = 2.0f * (x * y + w * z);
  m.a[1][1] = 1.0f - 2.0f * (x * x + z * z);
  m.a[1][2] = 2.0f * (y * z - w * x);
  m.a[2][0] = 2.0f * (x * z - w * y);
  m.a[2][1] = 2.0f * (y * z + w * x);
  m.a[2][2] = 1.0f - 2.0f * (x * x + y * y);
  return m;
}

// sh.rs:9-32
constexpr float SH_C0 = 0.28209479177387814f;
constexpr float SH_C1 = 0.4886025119029199f;
constexpr float SH_C2_0 = 1.0925484305920792f, SH_C2_1 = -1.0925484305920792f, SH_C2_2 = 0.31539156525252005f,
                SH_C2_3 = -1.0925484305920792f, SH_C2_4 = 0.5462742152960396f;
constexpr float SH_C3_0 = -0.5900435899266435f, SH_C3_1 = 2.890611442640554f, SH_C3_2 = -0.4570457994644658f,
                SH_C3_3 = 0.3731763325901154f, SH_C3_4 = -0.4570457994644658f, SH_C3_5 = 1.445305721320277f,
                SH_C3_6 = -0.5900435899266435f;

// sh.rs:39-61
RGS_HD void sh_basis(float x, float y, float z, float o[16]) {
  float xx = x * x, yy = y * y, zz = z * z;
  float xy = x * y, xz = x * z, yz = y * z;
  o[0] = SH_C0;
  o[1] = SH_C1 * -y;
  o[2] = SH_C1 * z;
  o[3] = SH_C1 * -x;
  o[4] = SH_C2_0 * xy;
  o[5] = SH_C2_1 * yz;
  o[6] = SH_C2_2 * (2.0f * zz - xx - yy);
  o[7] = SH_C2_3 * xz;
  o[8] = SH_C2_4 * (xx - yy);
  o[9] = SH_C3_0 * y * (3.0f * xx - yy);
  o[10] = SH_C3_1 * xy * z;
  o[11] = SH_C3_2 * y * (4.0f * zz - xx - yy);
  o[12] = SH_C3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
  o[13] = SH_C3_4 * x * (4.0f * zz - xx - yy);
  o[14] = SH_C3_5 * z * (xx - yy);
  o[15] = SH_C3_6 * x * (xx - 3.0f * yy);
}

// sh.rs:68-89, contracted with dL/dbasis: returns sum_k dk[k] * d basis_k / d dir.
// Gradient-only (fused multiply-adds); the Jacobian's structural zeros
// (NZ) are skipped at compile time rather than multiplied.
RGS_HD void sh_basis_grad_contract(float x, float y, float z, const float dk[16], float dd[3]) {
#pragma clang fp contract(fast)
  constexpr bool NZ[16][3] = {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 0}, {1, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1},
                              {1, 1, 0}, {1, 1, 0}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 0}};
  float xx = x * x, yy = y * y, zz = z * z;
  const float g[16][3] = {
      {0.0f, 0.0f, 0.0f},
      {0.0f, -SH_C1, 0.0f},
      {0.0f, 0.0f, SH_C1},
      {-SH_C1, 0.0f, 0.0f},
      {SH_C2_0 * y, SH_C2_0 * x, 0.0f},
      {0.0f, SH_C2_1 * z, SH_C2_1 * y},
      {SH_C2_2 * -2.0f * x, SH_C2_2 * -2.0f * y, SH_C2_2 * 4.0f * z},
      {SH_C2_3 * z, 0.0f, SH_C2_3 * x},
      {SH_C2_4 * 2.0f * x, SH_C2_4 * -2.0f * y, 0.0f},
      {SH_C3_0 * 6.0f * x * y, SH_C3_0 * 3.0f * (xx - yy), 0.0f},
      {SH_C3_1 * y * z, SH_C3_1 * x * z, SH_C3_1 * x * y},
      {SH_C3_2 * -2.0f * x * y, SH_C3_2 * (4.0f * zz - xx - 3.0f * yy), SH_C3_2 * 8.0f * y * z},
      {SH_C3_3 * -6.0f * x * z, SH_C3_3 * -6.0f * y * z, SH_C3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy)},
      {SH_C3_4 * (4.0f * zz - 3.0f * xx - yy), SH_C3_4 * -2.0f * x * y, SH_C3_4 * 8.0f * x * z},
      {SH_C3_5 * 2.0f * x * z, SH_C3_5 * -2.0f * y * z, SH_C3_5 * (xx - yy)},
      {SH_C3_6 * 3.0f * (xx - yy), SH_C3_6 * -6.0f * x * y, 0.0f}};
  dd[0] = 0.0f; dd[1] = 0.0f; dd[2] = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (NZ[k][c]) dd[c] += dk[k] * g[k][c];
}

// Per-view camera constants for the kernels. view is the world-to-camera
// matrix (rows 0-2 of the 4x4 inverse of c2w), row-major.
struct CamDev {
  float view[12];
  float fx, fy, cx, cy;
  float cam_pos[3];
  int32_t w, h, ntx, nty;
  int32_t tile_base;  // first global tile id of this view
  int32_t pix_base;   // first pixel of this view in the batched image arrays
  uint32_t tgt_base;  // first pixel of this view's target in the target store
  int32_t lunit_base; // first loss unit of this view
  // binning grid: bins of (1 << bin_shift)^2 tiles, nbx x nby per view
  int32_t bin_shift, nbx, nby;
  int32_t bin_base;   // first global bin id of this view
};

// p_cam = view * [p, 1] (gemv order, math/mod.rs:225-226)
RGS_HD void to_camera(const CamDev& c, float x, float y, float z, float p[3]) {
  for (int i = 0; i < 3; ++i) {
    float acc = c.view[i * 4 + 0] * x;
    acc = acc + c.view[i * 4 + 1] * y;
    acc = acc + c.view[i * 4 + 2] * z;
    acc = acc + c.view[i * 4 + 3] * 1.0f;
    p[i] = acc;
  }
}

// T = J W for the EWA Jacobian at p_cam (projection.rs:47-61).
RGS_HD M3 ewa_T(const CamDev& c, float tx, float ty, float tz) {
  M3 J = m3_zero();
  J.a[0][0] = c.fx / tz;
  J.a[0][2] = -c.fx * tx / (tz * tz);
  J.a[1][1] = c.fy / tz;
  J.a[1][2] = -c.fy * ty / (tz * tz);
  M3 W;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) W.a[i][j] = c.view[i * 4 + j];
  return m3_mul(J, W);
}

}  // namespace rgs

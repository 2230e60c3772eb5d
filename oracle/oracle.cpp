// oracle.cpp -- CPU restatement of rgsplat's differentiable training step.
//
// TEST INFRASTRUCTURE ONLY. This file is the correctness oracle for the HIP
// path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// load it, and only as the checker. Nothing in rgsplat_amd/ links or calls it.
//
// Each function follows the reference file:line it cites, operation by
// operation, so that an fp32 instance reproduces the reference's rounding
// (compile with -O2 -ffp-contract=off and glibc libm; no fast-math).
// Third-party arithmetic the reference delegates to nalgebra 0.33.3 is
// restated from that library's published algorithms:
//   * Matrix * Matrix / Matrix * Vector: column-wise gemv, products summed in
//     index order k = 0, 1, 2(, 3).
//   * Vector4 norm (UnitQuaternion::from_quaternion): the unrolled dot product
//     (v0^2 + v2^2) + (v1^2 + v3^2) over storage order [i, j, k, w].
//   * UnitQuaternion::to_rotation_matrix: the ww+ii-jj-kk form.
//   * UnitQuaternion * Vector3: t = 2 (q_v x v); v' = t*w + q_v x t + v.
//   * Matrix4::try_inverse: the loop-unrolled cofactor inverse (MESA
//     gluInvertMatrix), evaluated on column-major storage.
//   * Matrix::sum: column-major fold from zero.
// rayon only schedules work; the parallel loops here use OpenMP exactly where
// the reference uses par_iter, and stay serial where the reference is serial.
// Depth ties in the sort are ordered by Gaussian index (the reference's
// sort_unstable_by leaves them unspecified).
//
// Parity pins (tests/test_oracle_*.py): the reference's own unit tests
// (sh.rs:145-156, quaternion.rs:77-106, covariance.rs:91-121,
// backward.rs:848-982, tests/training_reduces_loss.rs:38-84), fp64 central
// finite differences of every gradient, brute-force binning, a closed-form
// on-axis isotropic image, direct 2-D convolution SSIM and textbook Adam;
// for the model files, the reference's io/ply.rs tests and hand-derived
// records of crafted files; for the colour conversions, the reference's
// images.rs/rasterizer.rs tests and closed forms; for the SfM initialisation
// (pipeline/sfm.rs:365-455), points on a line, brute-force 3-NN and a numpy
// restatement of the window search. Rust's std sort and binary search are
// restated from the std library's documented behaviour (stable sort_by;
// binary_search_by as shipped since Rust 1.82, ending on the last equal key).

#include <strings.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- constants
// backward.rs:21-24, rasterizer.rs:17-19
constexpr int TILE = 16;
constexpr float ALPHA_THRESHOLD = 1.0f / 255.0f;
constexpr float TRANSMITTANCE_STOP = 1e-4f;
constexpr float MAX_ALPHA = 0.99f;

// Gaussian3D field offsets (math/mod.rs:29-46): position[3], opacity_logit,
// rotation[4] (w,x,y,z), log_scale[3], sh_coeffs[48] channel-major.
constexpr int G_POS = 0, G_OPL = 3, G_ROT = 4, G_LS = 8, G_SH = 11, G_N = 59;
// GaussianGrad field offsets (backward.rs:33-41).
constexpr int D_POS = 0, D_ROT = 3, D_LS = 7, D_OP = 10, D_SH = 11, D_SGN = 59, D_N = 60;

inline float EXP(float x) { return expf(x); }
inline double EXP(double x) { return exp(x); }
inline float LOG(float x) { return logf(x); }
inline double LOG(double x) { return log(x); }
inline float SQRT(float x) { return sqrtf(x); }
inline double SQRT(double x) { return sqrt(x); }
inline float FMAX(float a, float b) { return fmaxf(a, b); }
inline double FMAX(double a, double b) { return fmax(a, b); }
inline float FMIN(float a, float b) { return fminf(a, b); }
inline double FMIN(double a, double b) { return fmin(a, b); }

// Rust `f as i32`: truncate toward zero, saturate, NaN -> 0.
template <class F> inline int32_t as_i32(F x) {
  if (x != x) return 0;
  if (x >= (F)2147483648.0) return INT32_MAX;
  if (x <= (F)-2147483648.0) return INT32_MIN;
  return (int32_t)x;
}
// Rust f32::clamp (NaN passes through).
template <class F> inline F clampf(F x, F lo, F hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

template <class F> struct M3 {
  F a[3][3];
};
template <class F> inline M3<F> zero3() {
  M3<F> m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.a[i][j] = F(0);
  return m;
}
// nalgebra gemm: C[:,j] = sum_k A[:,k] B[k,j], accumulated in k order.
template <class F> inline M3<F> mul(const M3<F>& A, const M3<F>& B) {
  M3<F> C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      F acc = A.a[i][0] * B.a[0][j];
      acc = acc + A.a[i][1] * B.a[1][j];
      acc = acc + A.a[i][2] * B.a[2][j];
      C.a[i][j] = acc;
    }
  return C;
}
template <class F> inline M3<F> tr(const M3<F>& A) {
  M3<F> T;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T.a[i][j] = A.a[j][i];
  return T;
}
template <class F> inline void mulv3(const M3<F>& A, const F v[3], F out[3]) {
  for (int i = 0; i < 3; ++i) {
    F acc = A.a[i][0] * v[0];
    acc = acc + A.a[i][1] * v[1];
    acc = acc + A.a[i][2] * v[2];
    out[i] = acc;
  }
}

// ------------------------------------------------------------------ camera
struct Camera {
  uint32_t width, height;
  float fx, fy, cx, cy;
  float c2w[16];  // row-major
  float distortion[3];
};

// math/mod.rs:155-160 -> nalgebra Matrix4::try_inverse (cofactor form on
// column-major storage). Returns false when det == 0. Output is row-major.
template <class F> bool view_matrix(const Camera& cam, F view[16]) {
  F m[16];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) m[c * 4 + r] = (F)cam.c2w[r * 4 + c];
  F inv[16];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  F det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  if (det == F(0)) return false;
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
  F inv_det = F(1) / det;
  for (int i = 0; i < 16; ++i) inv[i] = inv[i] * inv_det;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) view[r * 4 + c] = inv[c * 4 + r];
  return true;
}

// -------------------------------------------------------------------- math
// math/mod.rs:183-185
template <class F> inline F sigmoid(F x) { return F(1) / (F(1) + EXP(-x)); }
// math/mod.rs:189-192
template <class F> inline F logit(F p) { return LOG(p / (F(1) - p)); }

// quaternion.rs:10-17
template <class F> inline void q_normalize(const F q[4], F out[4]) {
  F n = SQRT(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (n < F(1e-12)) {
    out[0] = F(1); out[1] = F(0); out[2] = F(0); out[3] = F(0);
  } else {
    out[0] = q[0] / n; out[1] = q[1] / n; out[2] = q[2] / n; out[3] = q[3] / n;
  }
}
// quaternion.rs:23-30
template <class F> inline M3<F> mat_from_quat_components(const F q[4]) {
  F w = q[0], x = q[1], y = q[2], z = q[3];
  M3<F> m;
  m.a[0][0] = F(1) - F(2) * (y * y + z * z);
  m.a[0][1] = F(2) * (x * y - w * z);
  m.a[0][2] = F(2) * (x * z + w * y);
  m.a[1][0] = F(2) * (x * y + w * z);
  m.a[1][1] = F(1) - F(2) * (x * x + z * z);
  m.a[1][2] = F(2) * (y * z - w * x);
  m.a[2][0] = F(2) * (x * z - w * y);
  m.a[2][1] = F(2) * (y * z + w * x);
  m.a[2][2] = F(1) - F(2) * (x * x + y * y);
  return m;
}
// quaternion.rs:35-37
template <class F> inline M3<F> to_rotation_matrix(const F q[4]) {
  F qn[4];
  q_normalize(q, qn);
  return mat_from_quat_components(qn);
}
// quaternion.rs:45-70
template <class F> inline void rotation_jacobian(const F q[4], M3<F> out[4]) {
  F qn[4];
  q_normalize(q, qn);
  F w = qn[0], x = qn[1], y = qn[2], z = qn[3];
  const F m[4][9] = {
      {F(0), F(-2) * z, F(2) * y, F(2) * z, F(0), F(-2) * x, F(-2) * y, F(2) * x, F(0)},
      {F(0), F(2) * y, F(2) * z, F(2) * y, F(-4) * x, F(-2) * w, F(2) * z, F(2) * w, F(-4) * x},
      {F(-4) * y, F(2) * x, F(2) * w, F(2) * x, F(0), F(2) * z, F(-2) * w, F(2) * z, F(-4) * y},
      {F(-4) * z, F(-2) * w, F(2) * x, F(2) * w, F(-4) * z, F(2) * y, F(2) * x, F(2) * y, F(0)}};
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < 9; ++i) out[k].a[i / 3][i % 3] = m[k][i];
}

// nalgebra UnitQuaternion::from_quaternion(w,x,y,z): coords stored [x,y,z,w];
// norm^2 = (x^2 + z^2) + (y^2 + w^2); each coordinate divided by the norm.
template <class F> inline void nalgebra_unit_quat(const F q[4], F& w, F& x, F& y, F& z) {
  F a = q[1] * q[1], b = q[2] * q[2], c = q[3] * q[3], d = q[0] * q[0];
  a = a + c;
  b = b + d;
  F n = SQRT(a + b);
  x = q[1] / n; y = q[2] / n; z = q[3] / n; w = q[0] / n;
}
// nalgebra UnitQuaternion::to_rotation_matrix.
template <class F> inline M3<F> nalgebra_rot(F w, F i, F j, F k) {
  F ww = w * w, ii = i * i, jj = j * j, kk = k * k;
  F ij = i * j * F(2), wk = w * k * F(2), wj = w * j * F(2);
  F ik = i * k * F(2), jk = j * k * F(2), wi = w * i * F(2);
  M3<F> r;
  r.a[0][0] = ww + ii - jj - kk; r.a[0][1] = ij - wk; r.a[0][2] = wj + ik;
  r.a[1][0] = ij + wk; r.a[1][1] = ww - ii + jj - kk; r.a[1][2] = jk - wi;
  r.a[2][0] = ik - wj; r.a[2][1] = wi + jk; r.a[2][2] = ww - ii - jj + kk;
  return r;
}
// math/mod.rs:89-96  Gaussian3D::covariance_3d
template <class F> inline M3<F> covariance_3d(const F* g) {
  F w, x, y, z;
  nalgebra_unit_quat(g + G_ROT, w, x, y, z);
  M3<F> r = nalgebra_rot(w, x, y, z);
  M3<F> s = zero3<F>();
  s.a[0][0] = EXP(g[G_LS + 0]);
  s.a[1][1] = EXP(g[G_LS + 1]);
  s.a[2][2] = EXP(g[G_LS + 2]);
  M3<F> rs = mul(r, s);
  return mul(rs, tr(rs));
}

// covariance.rs:16-24
template <class F> inline M3<F> rs_matrix(const F s[3], const F q[4]) {
  M3<F> r = to_rotation_matrix(q);
  M3<F> m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.a[i][j] = r.a[i][j] * s[j];
  return m;
}
// covariance.rs:28-31
template <class F> inline M3<F> build_cov3d(const F s[3], const F q[4]) {
  M3<F> m = rs_matrix(s, q);
  return mul(m, tr(m));
}
// covariance.rs:42-84
template <class F>
inline void backprop_cov3d(const F s[3], const F q[4], const M3<F>& g, F d_ls[3], F d_rot[4]) {
  M3<F> r = to_rotation_matrix(q);
  M3<F> m = rs_matrix(s, q);
  M3<F> gs;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) gs.a[i][j] = g.a[i][j] + g.a[j][i];
  M3<F> dl_dm = mul(gs, m);
  F dl_ds[3];
  M3<F> dl_dr;
  for (int j = 0; j < 3; ++j) {
    F acc = F(0);
    for (int i = 0; i < 3; ++i) {
      acc += dl_dm.a[i][j] * r.a[i][j];
      dl_dr.a[i][j] = dl_dm.a[i][j] * s[j];
    }
    dl_ds[j] = acc;
  }
  for (int j = 0; j < 3; ++j) d_ls[j] = dl_ds[j] * s[j];
  M3<F> jac[4];
  rotation_jacobian(q, jac);
  for (int k = 0; k < 4; ++k) {
    F acc = F(0);  // component_mul(..).sum(): column-major fold
    for (int c = 0; c < 3; ++c)
      for (int rr = 0; rr < 3; ++rr) acc = acc + jac[k].a[rr][c] * dl_dr.a[rr][c];
    d_rot[k] = acc;
  }
}

// sh.rs:9-32 (f32 constants rounded from the f64 literals)
template <class F> struct SHC {
  static constexpr F C0 = (F)0.28209479177387814;
  static constexpr F C1 = (F)0.4886025119029199;
  static constexpr F C2[5] = {(F)1.0925484305920792, (F)-1.0925484305920792, (F)0.31539156525252005,
                              (F)-1.0925484305920792, (F)0.5462742152960396};
  static constexpr F C3[7] = {(F)-0.5900435899266435, (F)2.890611442640554, (F)-0.4570457994644658,
                              (F)0.3731763325901154, (F)-0.4570457994644658, (F)1.445305721320277,
                              (F)-0.5900435899266435};
};
// sh.rs:39-61
template <class F> inline void sh_basis(const F d[3], F o[16]) {
  using C = SHC<F>;
  F x = d[0], y = d[1], z = d[2];
  F xx = x * x, yy = y * y, zz = z * z;
  F xy = x * y, xz = x * z, yz = y * z;
  o[0] = C::C0;
  o[1] = C::C1 * -y;
  o[2] = C::C1 * z;
  o[3] = C::C1 * -x;
  o[4] = C::C2[0] * xy;
  o[5] = C::C2[1] * yz;
  o[6] = C::C2[2] * (F(2) * zz - xx - yy);
  o[7] = C::C2[3] * xz;
  o[8] = C::C2[4] * (xx - yy);
  o[9] = C::C3[0] * y * (F(3) * xx - yy);
  o[10] = C::C3[1] * xy * z;
  o[11] = C::C3[2] * y * (F(4) * zz - xx - yy);
  o[12] = C::C3[3] * z * (F(2) * zz - F(3) * xx - F(3) * yy);
  o[13] = C::C3[4] * x * (F(4) * zz - xx - yy);
  o[14] = C::C3[5] * z * (xx - yy);
  o[15] = C::C3[6] * x * (xx - F(3) * yy);
}
// sh.rs:68-89
template <class F> inline void sh_basis_grad(const F d[3], F o[16][3]) {
  using C = SHC<F>;
  F x = d[0], y = d[1], z = d[2];
  F xx = x * x, yy = y * y, zz = z * z;
  const F v[16][3] = {
      {F(0), F(0), F(0)},
      {F(0), -C::C1, F(0)},
      {F(0), F(0), C::C1},
      {-C::C1, F(0), F(0)},
      {C::C2[0] * y, C::C2[0] * x, F(0)},
      {F(0), C::C2[1] * z, C::C2[1] * y},
      {C::C2[2] * F(-2) * x, C::C2[2] * F(-2) * y, C::C2[2] * F(4) * z},
      {C::C2[3] * z, F(0), C::C2[3] * x},
      {C::C2[4] * F(2) * x, C::C2[4] * F(-2) * y, F(0)},
      {C::C3[0] * F(6) * x * y, C::C3[0] * F(3) * (xx - yy), F(0)},
      {C::C3[1] * y * z, C::C3[1] * x * z, C::C3[1] * x * y},
      {C::C3[2] * F(-2) * x * y, C::C3[2] * (F(4) * zz - xx - F(3) * yy), C::C3[2] * F(8) * y * z},
      {C::C3[3] * F(-6) * x * z, C::C3[3] * F(-6) * y * z, C::C3[3] * (F(6) * zz - F(3) * xx - F(3) * yy)},
      {C::C3[4] * (F(4) * zz - F(3) * xx - yy), C::C3[4] * F(-2) * x * y, C::C3[4] * F(8) * x * z},
      {C::C3[5] * F(2) * x * z, C::C3[5] * F(-2) * y * z, C::C3[5] * (xx - yy)},
      {C::C3[6] * F(3) * (xx - yy), C::C3[6] * F(-6) * x * y, F(0)}};
  for (int k = 0; k < 16; ++k)
    for (int c = 0; c < 3; ++c) o[k][c] = v[k][c];
}
// sh.rs:93-104
template <class F> inline void eval_sh_raw(const F* coeffs, const F basis[16], F out[3]) {
  for (int ch = 0; ch < 3; ++ch) {
    F acc = F(0);
    for (int k = 0; k < 16; ++k) acc += basis[k] * coeffs[ch * 16 + k];
    out[ch] = F(0.5) + acc;
  }
}
// sh.rs:122-139
template <class F> inline void view_dir(const F* pos, const F cam[3], F out[3]) {
  F v[3] = {cam[0] - pos[0], cam[1] - pos[1], cam[2] - pos[2]};
  F len = SQRT(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (len < F(1e-8)) {
    out[0] = F(0); out[1] = F(0); out[2] = F(1);
  } else {
    out[0] = v[0] / len; out[1] = v[1] / len; out[2] = v[2] / len;
  }
}

// projection.rs:24-69 (p_cam recomputed as the reference does at :31-32)
template <class F>
inline M3<F> project_covariance(const M3<F>& sigma, const F* pos, const F view[16], F fx, F fy) {
  F p[3];
  for (int i = 0; i < 3; ++i) {
    F acc = view[i * 4 + 0] * pos[0];
    acc = acc + view[i * 4 + 1] * pos[1];
    acc = acc + view[i * 4 + 2] * pos[2];
    acc = acc + view[i * 4 + 3] * F(1);
    p[i] = acc;
  }
  F tx = p[0], ty = p[1], tz = FMAX(p[2], F(0.001));
  M3<F> J = zero3<F>();
  J.a[0][0] = fx / tz;
  J.a[0][2] = -fx * tx / (tz * tz);
  J.a[1][1] = fy / tz;
  J.a[1][2] = -fy * ty / (tz * tz);
  M3<F> W;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) W.a[i][j] = view[i * 4 + j];
  M3<F> T = mul(J, W);
  M3<F> s2 = mul(mul(T, sigma), tr(T));
  s2.a[0][0] += F(0.3);
  s2.a[1][1] += F(0.3);
  return s2;
}

// projection.rs:138-165 (returns false for None)
template <class F>
inline bool gaussian_aabb_2d(F mx, F my, const M3<F>& s2, F n_sigma, uint32_t w, uint32_t h,
                             int32_t out[4]) {
  F a = s2.a[0][0], b = s2.a[0][1], c = s2.a[1][1];
  F mid = (a + c) * F(0.5);
  F disc = SQRT((a - c) * (a - c) * F(0.25) + b * b);
  F l1 = FMAX(mid + disc, F(0));
  F l2 = FMAX(mid - disc, F(0));
  F r = n_sigma * FMAX(SQRT(l1), SQRT(l2));
  int32_t x0 = std::max(as_i32(mx - r), 0);
  int32_t y0 = std::max(as_i32(my - r), 0);
  // `as i32 + 1` wraps in a release build when the cast saturated.
  int32_t x1 = std::min((int32_t)((uint32_t)as_i32(mx + r) + 1u), (int32_t)w);
  int32_t y1 = std::min((int32_t)((uint32_t)as_i32(my + r) + 1u), (int32_t)h);
  if (x0 >= x1 || y0 >= y1) return false;
  out[0] = x0; out[1] = y0; out[2] = x1; out[3] = y1;
  return true;
}

// backward.rs:74-90
template <class F> struct Splat {
  size_t gi;
  F sx, sy;
  F conic[3];
  F cov2d[3];
  F color[3];
  F raw_color[3];
  F opacity;
  F depth;
  F basis[16];
  F pcam[3];
  int32_t tx0, ty0, tx1, ty1;
};

template <class F> struct View {
  F view[16];
  F fx, fy, cx, cy;
  F cam_pos[3];
  uint32_t w, h;
};

template <class F> bool make_view(const Camera& cam, View<F>& v) {
  if (!view_matrix<F>(cam, v.view)) return false;
  v.fx = (F)cam.fx; v.fy = (F)cam.fy; v.cx = (F)cam.cx; v.cy = (F)cam.cy;
  v.cam_pos[0] = (F)cam.c2w[3]; v.cam_pos[1] = (F)cam.c2w[7]; v.cam_pos[2] = (F)cam.c2w[11];
  v.w = cam.width; v.h = cam.height;
  return true;
}

// backward.rs:210-280
template <class F> bool project(const F* g, size_t gi, const View<F>& v, Splat<F>& s) {
  F opacity = sigmoid(g[G_OPL]);
  if (opacity < F(ALPHA_THRESHOLD)) return false;
  F p[3];
  for (int i = 0; i < 3; ++i) {
    F acc = v.view[i * 4 + 0] * g[0];
    acc = acc + v.view[i * 4 + 1] * g[1];
    acc = acc + v.view[i * 4 + 2] * g[2];
    acc = acc + v.view[i * 4 + 3] * F(1);
    p[i] = acc;
  }
  F tz = p[2];
  if (tz <= F(0.001)) return false;
  F sx = v.fx * p[0] / tz + v.cx;
  F sy = v.fy * p[1] / tz + v.cy;
  const F margin = F(128);
  if (sx < -margin || sx > (F)v.w + margin || sy < -margin || sy > (F)v.h + margin) return false;
  M3<F> s2 = project_covariance(covariance_3d(g), g, v.view, v.fx, v.fy);
  F a = s2.a[0][0], b = s2.a[0][1], c = s2.a[1][1];
  F det = a * c - b * b;
  if (std::fabs(det) < F(1e-8)) return false;
  F inv_det = F(1) / det;
  F dir[3];
  view_dir(g + G_POS, v.cam_pos, dir);
  sh_basis(dir, s.basis);
  eval_sh_raw(g + G_SH, s.basis, s.raw_color);
  int32_t bb[4];
  if (!gaussian_aabb_2d(sx, sy, s2, F(3), v.w, v.h, bb)) return false;
  s.gi = gi;
  s.sx = sx; s.sy = sy;
  s.conic[0] = c * inv_det; s.conic[1] = -b * inv_det; s.conic[2] = a * inv_det;
  s.cov2d[0] = a; s.cov2d[1] = b; s.cov2d[2] = c;
  for (int ch = 0; ch < 3; ++ch) s.color[ch] = clampf(s.raw_color[ch], F(0), F(1));
  s.opacity = opacity;
  s.depth = tz;
  s.pcam[0] = p[0]; s.pcam[1] = p[1]; s.pcam[2] = p[2];
  s.tx0 = bb[0] / TILE;
  s.ty0 = bb[1] / TILE;
  s.tx1 = (bb[2] + TILE - 1) / TILE;
  s.ty1 = (bb[3] + TILE - 1) / TILE;
  return true;
}

template <class F> struct Binned {
  std::vector<Splat<F>> splats;  // depth order
  std::vector<std::vector<uint32_t>> lists;  // per tile, indices into splats
  size_t ntx, nty;
};

// backward.rs:172-208. Returns false if a depth is NaN (the reference panics).
template <class F> bool project_and_bin(const F* gs, size_t n, const View<F>& v, Binned<F>& out) {
  out.ntx = (v.w + TILE - 1) / TILE;
  out.nty = (v.h + TILE - 1) / TILE;
  std::vector<Splat<F>> tmp(n);
  std::vector<uint8_t> ok(n);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) ok[i] = project(gs + (size_t)i * G_N, (size_t)i, v, tmp[i]);
  out.splats.clear();
  for (size_t i = 0; i < n; ++i)
    if (ok[i]) out.splats.push_back(tmp[i]);
  for (auto& s : out.splats)
    if (s.depth != s.depth) return false;
  std::sort(out.splats.begin(), out.splats.end(), [](const Splat<F>& a, const Splat<F>& b) {
    if (a.depth < b.depth) return true;
    if (b.depth < a.depth) return false;
    return a.gi < b.gi;
  });
  out.lists.assign(out.ntx * out.nty, {});
  for (size_t idx = 0; idx < out.splats.size(); ++idx) {
    const auto& s = out.splats[idx];
    size_t tx0 = (size_t)std::max(s.tx0, 0), ty0 = (size_t)std::max(s.ty0, 0);
    size_t tx1 = std::min((size_t)(uint32_t)s.tx1, out.ntx), ty1 = std::min((size_t)(uint32_t)s.ty1, out.nty);
    for (size_t ty = ty0; ty < ty1; ++ty)
      for (size_t tx = tx0; tx < tx1; ++tx) out.lists[ty * out.ntx + tx].push_back((uint32_t)idx);
  }
  return true;
}

// backward.rs:296-385 (pixels of empty tiles keep rgb = 0, T = 1, last = -1)
template <class F>
void forward_pass(const Binned<F>& b, uint32_t w, uint32_t h, const F bg[3], F* rgb, F* final_t,
                  int32_t* last_contrib, uint64_t* pairs_evaluated) {
  size_t nt = b.lists.size();
  for (size_t i = 0; i < (size_t)w * h; ++i) {
    rgb[3 * i] = rgb[3 * i + 1] = rgb[3 * i + 2] = F(0);
    final_t[i] = F(1);
    last_contrib[i] = -1;
  }
  uint64_t evaluated = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : evaluated)
  for (long long ti = 0; ti < (long long)nt; ++ti) {
    const auto& L = b.lists[ti];
    if (L.empty()) continue;
    size_t tx = ti % b.ntx, ty = ti / b.ntx;
    size_t px0 = tx * TILE, py0 = ty * TILE;
    size_t px1 = std::min(px0 + TILE, (size_t)w), py1 = std::min(py0 + TILE, (size_t)h);
    for (size_t py = py0; py < py1; ++py)
      for (size_t px = px0; px < px1; ++px) {
        F t = F(1);
        F col[3] = {F(0), F(0), F(0)};
        int32_t last = -1;
        size_t li = 0;
        for (; li < L.size(); ++li) {
          const auto& s = b.splats[L[li]];
          F dx = (F)px + F(0.5) - s.sx;
          F dy = (F)py + F(0.5) - s.sy;
          F a = s.conic[0], bb = s.conic[1], c = s.conic[2];
          F power = F(-0.5) * (a * dx * dx + F(2) * bb * dx * dy + c * dy * dy);
          if (power > F(0)) continue;
          F alpha = FMIN(s.opacity * EXP(power), F(MAX_ALPHA));
          if (alpha < F(ALPHA_THRESHOLD)) continue;
          F weight = alpha * t;
          col[0] += s.color[0] * weight;
          col[1] += s.color[1] * weight;
          col[2] += s.color[2] * weight;
          t *= F(1) - alpha;
          last = (int32_t)li;
          if (t < F(TRANSMITTANCE_STOP)) { ++li; break; }
        }
        evaluated += li;
        size_t pix = py * w + px;
        rgb[pix * 3 + 0] = col[0] + t * bg[0];
        rgb[pix * 3 + 1] = col[1] + t * bg[1];
        rgb[pix * 3 + 2] = col[2] + t * bg[2];
        final_t[pix] = t;
        last_contrib[pix] = last;
      }
  }
  if (pairs_evaluated) *pairs_evaluated = evaluated;
}

template <class F> struct Accum {
  F color[3] = {F(0), F(0), F(0)};
  F opacity = F(0);
  F conic[3] = {F(0), F(0), F(0)};
  F sx = F(0), sy = F(0);
};

// backward.rs:391-512
template <class F>
std::vector<Accum<F>> backward_pass(const Binned<F>& b, uint32_t w, uint32_t h, const F bg[3],
                                    const F* final_t, const int32_t* last_contrib, const F* dl_dpix) {
  size_t nt = b.lists.size();
  std::vector<std::vector<Accum<F>>> per_tile(nt);
#pragma omp parallel for schedule(dynamic, 4)
  for (long long ti = 0; ti < (long long)nt; ++ti) {
    const auto& L = b.lists[ti];
    if (L.empty()) continue;
    size_t tx = ti % b.ntx, ty = ti / b.ntx;
    size_t px0 = tx * TILE, py0 = ty * TILE;
    size_t px1 = std::min(px0 + TILE, (size_t)w), py1 = std::min(py0 + TILE, (size_t)h);
    std::vector<Accum<F>> local(L.size());
    for (size_t py = py0; py < py1; ++py)
      for (size_t px = px0; px < px1; ++px) {
        size_t pix = py * w + px;
        int32_t lc = last_contrib[pix];
        if (lc < 0) continue;
        F dl[3] = {dl_dpix[pix * 3], dl_dpix[pix * 3 + 1], dl_dpix[pix * 3 + 2]};
        F t = final_t[pix];
        F behind[3] = {t * bg[0], t * bg[1], t * bg[2]};
        for (long long li = lc; li >= 0; --li) {
          const auto& s = b.splats[L[li]];
          F dx = (F)px + F(0.5) - s.sx;
          F dy = (F)py + F(0.5) - s.sy;
          F a = s.conic[0], bb = s.conic[1], c = s.conic[2];
          F power = F(-0.5) * (a * dx * dx + F(2) * bb * dx * dy + c * dy * dy);
          if (power > F(0)) continue;
          F g_exp = EXP(power);
          F alpha_raw = s.opacity * g_exp;
          F alpha = FMIN(alpha_raw, F(MAX_ALPHA));
          if (alpha < F(ALPHA_THRESHOLD)) continue;
          t /= F(1) - alpha;
          F weight = alpha * t;
          auto& g = local[li];
          g.color[0] += dl[0] * weight;
          g.color[1] += dl[1] * weight;
          g.color[2] += dl[2] * weight;
          F inv_1ma = F(1) / (F(1) - alpha);
          F dl_dalpha = F(0);
          for (int ch = 0; ch < 3; ++ch) dl_dalpha += dl[ch] * (t * s.color[ch] - behind[ch] * inv_1ma);
          F dalpha_draw = alpha_raw > F(MAX_ALPHA) ? F(0) : F(1);
          F dl_draw = dl_dalpha * dalpha_draw;
          g.opacity += dl_draw * g_exp;
          F dl_dpower = dl_draw * s.opacity * g_exp;
          g.conic[0] += dl_dpower * (F(-0.5) * dx * dx);
          g.conic[1] += dl_dpower * (-dx * dy);
          g.conic[2] += dl_dpower * (F(-0.5) * dy * dy);
          g.sx += dl_dpower * (a * dx + bb * dy);
          g.sy += dl_dpower * (bb * dx + c * dy);
          behind[0] += weight * s.color[0];
          behind[1] += weight * s.color[1];
          behind[2] += weight * s.color[2];
        }
      }
    per_tile[ti] = std::move(local);
  }
  std::vector<Accum<F>> merged(b.splats.size());
  for (size_t ti = 0; ti < nt; ++ti) {
    const auto& L = b.lists[ti];
    for (size_t li = 0; li < per_tile[ti].size(); ++li) {
      auto& m = merged[L[li]];
      const auto& a = per_tile[ti][li];
      for (int ch = 0; ch < 3; ++ch) {
        m.color[ch] += a.color[ch];
        m.conic[ch] += a.conic[ch];
      }
      m.opacity += a.opacity;
      m.sx += a.sx;
      m.sy += a.sy;
    }
  }
  return merged;
}

// backward.rs:518-638
template <class F>
void geometry_backprop(const Splat<F>& s, const Accum<F>& acc, const F* g, const View<F>& v, F* out) {
  for (int i = 0; i < D_N; ++i) out[i] = F(0);
  F op = s.opacity;
  out[D_OP] = acc.opacity * op * (F(1) - op);
  F masked[3];
  for (int ch = 0; ch < 3; ++ch)
    masked[ch] = (s.raw_color[ch] > F(0) && s.raw_color[ch] < F(1)) ? acc.color[ch] : F(0);
  for (int ch = 0; ch < 3; ++ch)
    for (int k = 0; k < 16; ++k) out[D_SH + ch * 16 + k] = masked[ch] * s.basis[k];

  F ga = acc.conic[0], gb = acc.conic[1], gc = acc.conic[2];
  F a2 = s.cov2d[0], b2 = s.cov2d[1], c2 = s.cov2d[2];
  F det = a2 * c2 - b2 * b2;
  F det2 = det * det;
  F dl_da = (-ga * c2 * c2 + gb * b2 * c2 - gc * b2 * b2) / det2;
  F dl_db = (F(2) * ga * b2 * c2 - gb * (det + F(2) * b2 * b2) + F(2) * gc * a2 * b2) / det2;
  F dl_dc = (-ga * b2 * b2 + gb * a2 * b2 - gc * a2 * a2) / det2;

  F fx = v.fx, fy = v.fy;
  F tx = s.pcam[0], ty = s.pcam[1], tz = s.pcam[2];
  M3<F> J = zero3<F>();
  J.a[0][0] = fx / tz;
  J.a[0][2] = -fx * tx / (tz * tz);
  J.a[1][1] = fy / tz;
  J.a[1][2] = -fy * ty / (tz * tz);
  M3<F> W;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) W.a[i][j] = v.view[i * 4 + j];
  M3<F> T = mul(J, W);
  M3<F> sigma = covariance_3d(g);
  M3<F> G = zero3<F>();
  G.a[0][0] = dl_da; G.a[0][1] = F(0.5) * dl_db;
  G.a[1][0] = F(0.5) * dl_db; G.a[1][1] = dl_dc;
  M3<F> dsig = mul(mul(tr(T), G), T);
  F scale[3] = {EXP(g[G_LS]), EXP(g[G_LS + 1]), EXP(g[G_LS + 2])};
  backprop_cov3d(scale, g + G_ROT, dsig, out + D_LS, out + D_ROT);

  F dpc[3] = {acc.sx * fx / tz, acc.sy * fy / tz,
              acc.sx * (-fx * tx / (tz * tz)) + acc.sy * (-fy * ty / (tz * tz))};
  M3<F> dj[3] = {zero3<F>(), zero3<F>(), zero3<F>()};
  dj[0].a[0][2] = -fx / (tz * tz);
  dj[1].a[1][2] = -fy / (tz * tz);
  dj[2].a[0][0] = -fx / (tz * tz);
  dj[2].a[0][2] = F(2) * fx * tx / (tz * tz * tz);
  dj[2].a[1][1] = -fy / (tz * tz);
  dj[2].a[1][2] = F(2) * fy * ty / (tz * tz * tz);
  M3<F> sig_tt = mul(sigma, tr(T));
  M3<F> t_sig = mul(T, sigma);
  for (int k = 0; k < 3; ++k) {
    M3<F> dt = mul(dj[k], W);
    M3<F> p1 = mul(dt, sig_tt), p2 = mul(t_sig, tr(dt));
    F d00 = p1.a[0][0] + p2.a[0][0], d01 = p1.a[0][1] + p2.a[0][1], d11 = p1.a[1][1] + p2.a[1][1];
    F contrib = dl_da * d00 + dl_db * d01 + dl_dc * d11;
    dpc[k] += contrib;
  }
  F dpos[3];
  mulv3(tr(W), dpc, dpos);

  F u[3] = {v.cam_pos[0] - g[0], v.cam_pos[1] - g[1], v.cam_pos[2] - g[2]};
  F len = SQRT(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  if (len > F(1e-8)) {
    F dir[3] = {u[0] / len, u[1] / len, u[2] / len};
    F db[16][3];
    sh_basis_grad(dir, db);
    F dd[3] = {F(0), F(0), F(0)};
    for (int k = 0; k < 16; ++k) {
      F dk = masked[0] * g[G_SH + k] + masked[1] * g[G_SH + 16 + k] + masked[2] * g[G_SH + 32 + k];
      for (int c = 0; c < 3; ++c) dd[c] += dk * db[k][c];
    }
    F dot = dir[0] * dd[0] + dir[1] * dd[1] + dir[2] * dd[2];
    for (int c = 0; c < 3; ++c) {
      F proj = dd[c] - dir[c] * dot;
      dpos[c] += -(proj / len);
    }
  }
  out[D_POS] = dpos[0]; out[D_POS + 1] = dpos[1]; out[D_POS + 2] = dpos[2];
  out[D_SGN] = SQRT(acc.sx * acc.sx + acc.sy * acc.sy);
}

// backward.rs:759-772
template <class F> void gaussian_kernel_1d(int radius, F sigma, F* k) {
  int n = 2 * radius + 1;
  F sum = F(0);
  for (int i = 0; i < n; ++i) {
    F d = (F)i - (F)radius;
    F v = EXP(F(-0.5) * d * d / (sigma * sigma));
    k[i] = v;
    sum += v;
  }
  for (int i = 0; i < n; ++i) k[i] /= sum;
}
// backward.rs:776-807 (zero padding)
template <class F> void blur(const F* src, size_t w, size_t h, const F* k, int klen, F* out) {
  int radius = klen / 2;
  std::vector<F> tmp(w * h);
  for (size_t y = 0; y < h; ++y)
    for (size_t x = 0; x < w; ++x) {
      F acc = F(0);
      for (int ki = 0; ki < klen; ++ki) {
        long long sx = (long long)x + ki - radius;
        if (sx >= 0 && (size_t)sx < w) acc += k[ki] * src[y * w + sx];
      }
      tmp[y * w + x] = acc;
    }
  for (size_t y = 0; y < h; ++y)
    for (size_t x = 0; x < w; ++x) {
      F acc = F(0);
      for (int ki = 0; ki < klen; ++ki) {
        long long sy = (long long)y + ki - radius;
        if (sy >= 0 && (size_t)sy < h) acc += k[ki] * tmp[sy * w + x];
      }
      out[y * w + x] = acc;
    }
}
// backward.rs:697-756
// exact_sum (optional): the same per-pixel SSIM terms added in double
// (the reference adds them serially in f32).
template <class F>
F ssim_channel(const F* x, const F* y, size_t w, size_t h, const F* k, int klen, F* grad,
               double* exact_sum = nullptr) {
  const F C1 = F(0.01) * F(0.01);
  const F C2 = F(0.03) * F(0.03);
  size_t n = w * h;
  std::vector<F> mu_x(n), mu_y(n), xx(n), yy(n), xy(n), mu_xx(n), mu_yy(n), mu_xy(n);
  blur(x, w, h, k, klen, mu_x.data());
  blur(y, w, h, k, klen, mu_y.data());
  for (size_t p = 0; p < n; ++p) {
    xx[p] = x[p] * x[p];
    yy[p] = y[p] * y[p];
    xy[p] = x[p] * y[p];
  }
  blur(xx.data(), w, h, k, klen, mu_xx.data());
  blur(yy.data(), w, h, k, klen, mu_yy.data());
  blur(xy.data(), w, h, k, klen, mu_xy.data());
  F ssim_sum = F(0);
  std::vector<F> pm(n), qm(n), rm(n), qmu(n), rmu(n);
  for (size_t p = 0; p < n; ++p) {
    F mx = mu_x[p], my = mu_y[p];
    F sx2 = mu_xx[p] - mx * mx;
    F sy2 = mu_yy[p] - my * my;
    F sxy = mu_xy[p] - mx * my;
    F a1 = F(2) * mx * my + C1;
    F a2 = F(2) * sxy + C2;
    F b1 = mx * mx + my * my + C1;
    F b2 = sx2 + sy2 + C2;
    F num = a1 * a2;
    F den = b1 * b2;
    ssim_sum += num / den;
    if (exact_sum) *exact_sum += (double)(num / den);
    F inv_den = F(1) / den;
    F pc = inv_den * (F(2) * my * a2) - (num / (den * den)) * (F(2) * mx * b2);
    F qc = inv_den * (F(2) * a1);
    F rc = -(num / (den * den)) * b1;
    pm[p] = pc; qm[p] = qc; rm[p] = rc;
    qmu[p] = qc * my;
    rmu[p] = rc * mx;
  }
  if (grad) {
    std::vector<F> bp(n), bq(n), br(n), bqmu(n), brmu(n);
    blur(pm.data(), w, h, k, klen, bp.data());
    blur(qm.data(), w, h, k, klen, bq.data());
    blur(rm.data(), w, h, k, klen, br.data());
    blur(qmu.data(), w, h, k, klen, bqmu.data());
    blur(rmu.data(), w, h, k, klen, brmu.data());
    for (size_t p = 0; p < n; ++p)
      grad[p] = bp[p] + y[p] * bq[p] - bqmu[p] + F(2) * x[p] * br[p] - F(2) * brmu[p];
  }
  return ssim_sum;
}
// Rust f32::signum: +1 for +0.0 and positives, -1 for -0.0 and negatives.
template <class F> inline F signum(F d) {
  if (d != d) return d;
  return std::signbit(d) ? F(-1) : F(1);
}
// backward.rs:646-693. exact_loss (optional): the loss from the same f32
// per-pixel terms added in double, i.e. without the rounding of the
// reference's serial f32 sums.
template <class F>
F photometric_loss_grad(const F* rendered, const F* target_rgba, size_t w, size_t h, F lambda, F* dl,
                        double* exact_loss = nullptr) {
  size_t n = w * h;
  F m_total = (F)(n * 3);
  std::vector<F> x[3], y[3];
  for (int ch = 0; ch < 3; ++ch) {
    x[ch].resize(n);
    y[ch].resize(n);
  }
  for (size_t p = 0; p < n; ++p)
    for (int ch = 0; ch < 3; ++ch) {
      x[ch][p] = rendered[p * 3 + ch];
      y[ch][p] = target_rgba[p * 4 + ch];
    }
  F l1 = F(0);
  double l1_exact = 0.0, ssim_exact = 0.0;
  for (size_t p = 0; p < n; ++p)
    for (int ch = 0; ch < 3; ++ch) {
      F d = rendered[p * 3 + ch] - target_rgba[p * 4 + ch];
      l1 += std::fabs(d);
      l1_exact += (double)std::fabs(d);
      dl[p * 3 + ch] = (F(1) - lambda) * signum(d) / m_total;
    }
  l1 = l1 / m_total;
  F k[11];
  gaussian_kernel_1d<F>(5, F(1.5), k);
  F ssim_sum = F(0);
  std::vector<F> grad(n);
  for (int ch = 0; ch < 3; ++ch) {
    F s = ssim_channel(x[ch].data(), y[ch].data(), w, h, k, 11, grad.data(), &ssim_exact);
    ssim_sum += s;
    for (size_t p = 0; p < n; ++p) dl[p * 3 + ch] += -lambda * grad[p] / m_total;
  }
  F ssim_mean = ssim_sum / m_total;
  if (exact_loss) {
    const double m = (double)n * 3.0, lam = (double)lambda;
    *exact_loss = (1.0 - lam) * (l1_exact / m) + lam * (1.0 - ssim_exact / m);
  }
  return (F(1) - lambda) * l1 + lambda * (F(1) - ssim_mean);
}

// ------------------------------------------------------- the render object
template <class F> struct Run {
  View<F> v;
  Binned<F> b;
  std::vector<F> rgb, final_t, dl;
  std::vector<int32_t> last;
  std::vector<F> grads;
  F loss = F(0);
  double loss_exact = 0.0;  // the same per-pixel terms added in double
  uint64_t pairs_fwd = 0;
};

// backward.rs:107-149 (render_and_backward) and :152-166 (render_color)
template <class F>
Run<F>* run(const F* g, size_t n, const Camera* cam, const F* target, const F bg[3], F lambda,
            int want_backward, int* err) {
  auto* r = new Run<F>();
  *err = 0;
  if (!make_view<F>(*cam, r->v)) { *err = 1; delete r; return nullptr; }
  if (!project_and_bin(g, n, r->v, r->b)) { *err = 2; delete r; return nullptr; }
  size_t w = cam->width, h = cam->height;
  r->rgb.assign(w * h * 3, F(0));
  r->final_t.assign(w * h, F(1));
  r->last.assign(w * h, -1);
  forward_pass(r->b, cam->width, cam->height, bg, r->rgb.data(), r->final_t.data(), r->last.data(),
               &r->pairs_fwd);
  if (!want_backward || !target) return r;
  r->dl.assign(w * h * 3, F(0));
  r->loss = photometric_loss_grad(r->rgb.data(), target, w, h, lambda, r->dl.data(), &r->loss_exact);
  auto acc = backward_pass(r->b, cam->width, cam->height, bg, r->final_t.data(), r->last.data(), r->dl.data());
  r->grads.assign(n * D_N, F(0));
  for (size_t si = 0; si < r->b.splats.size(); ++si) {
    const auto& s = r->b.splats[si];
    geometry_backprop(s, acc[si], g + s.gi * G_N, r->v, r->grads.data() + s.gi * D_N);
  }
  return r;
}

// rasterizer.rs:72-192 (RGBA; tile buffer starts at bg; empty tiles stay 0)
void rasterize_rgba(const float* gs, size_t n, const View<float>& v, const float bg[3], float* out) {
  Binned<float> b;
  project_and_bin(gs, n, v, b);
  size_t w = v.w, h = v.h;
  std::fill(out, out + w * h * 4, 0.0f);
#pragma omp parallel for schedule(dynamic, 4)
  for (long long ti = 0; ti < (long long)b.lists.size(); ++ti) {
    const auto& L = b.lists[ti];
    if (L.empty()) continue;
    size_t tx = ti % b.ntx, ty = ti / b.ntx;
    size_t px0 = tx * TILE, py0 = ty * TILE;
    size_t px1 = std::min(px0 + TILE, w), py1 = std::min(py0 + TILE, h);
    float buf[TILE * TILE][4];
    float trn[TILE * TILE];
    for (int i = 0; i < TILE * TILE; ++i) {
      buf[i][0] = bg[0]; buf[i][1] = bg[1]; buf[i][2] = bg[2]; buf[i][3] = 1.0f;
      trn[i] = 1.0f;
    }
    for (uint32_t si : L) {
      const auto& s = b.splats[si];
      for (size_t py = py0; py < py1; ++py)
        for (size_t px = px0; px < px1; ++px) {
          size_t local = (py - py0) * TILE + (px - px0);
          float t = trn[local];
          if (t < TRANSMITTANCE_STOP) continue;
          float dx = (float)px + 0.5f - s.sx;
          float dy = (float)py + 0.5f - s.sy;
          float a = s.conic[0], bb = s.conic[1], c = s.conic[2];
          float power = -0.5f * (a * dx * dx + 2.0f * bb * dx * dy + c * dy * dy);
          if (power > 0.0f) continue;
          float alpha = fminf(s.opacity * expf(power), 0.99f);
          if (alpha < ALPHA_THRESHOLD) continue;
          float weight = alpha * t;
          buf[local][0] += s.color[0] * weight;
          buf[local][1] += s.color[1] * weight;
          buf[local][2] += s.color[2] * weight;
          buf[local][3] -= weight;
          trn[local] *= 1.0f - alpha;
        }
    }
    for (size_t py = py0; py < py1; ++py)
      for (size_t px = px0; px < px1; ++px) {
        size_t local = (py - py0) * TILE + (px - px0);
        size_t gidx = (py * w + px) * 4;
        out[gidx] = buf[local][0];
        out[gidx + 1] = buf[local][1];
        out[gidx + 2] = buf[local][2];
        out[gidx + 3] = 1.0f - trn[local];
      }
  }
}

}  // namespace

// =================================================================== C API
extern "C" {

typedef struct {
  float lr_position, lr_rotation, lr_scale, lr_opacity, lr_sh_dc, lr_sh_rest;
  float beta1, beta2, eps;
  float densify_grad_threshold, prune_opacity_thresh;
  uint32_t densify_from_iter, densify_until_iter, opacity_reset_iter;
  uint64_t max_gaussians, num_iterations;
} orc_train_cfg;

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

int orc_view_matrix(const Camera* cam, float out[16]) { return view_matrix<float>(*cam, out) ? 0 : 1; }

// ---- render/backward handles (fp32 and fp64 instances)
void* orc_run_f32(const float* g, size_t n, const Camera* cam, const float* target, const float* bg,
                  float lambda, int want_backward, int* err) {
  return run<float>(g, n, cam, target, bg, lambda, want_backward, err);
}
void* orc_run_f64(const double* g, size_t n, const Camera* cam, const double* target, const double* bg,
                  double lambda, int want_backward, int* err) {
  return run<double>(g, n, cam, target, bg, lambda, want_backward, err);
}
// backward_pass + geometry_backprop (backward.rs:131-146) of a forward
// handle with a given dL/dpixel (w*h*3) instead of the photometric loss's:
// the reference's backward of another loss gradient, e.g. the one the GPU's
// own image implies. g, n and bg are the handle's inputs; grads is n x 60.
void orc_backward_with_dl_f32(void* h, const float* g, size_t n, const float* bg, const float* dl, float* grads) {
  auto* r = (Run<float>*)h;
  auto acc = backward_pass(r->b, r->v.w, r->v.h, bg, r->final_t.data(), r->last.data(), dl);
  std::fill(grads, grads + n * D_N, 0.0f);
  for (size_t si = 0; si < r->b.splats.size(); ++si) {
    const auto& s = r->b.splats[si];
    geometry_backprop(s, acc[si], g + s.gi * G_N, r->v, grads + s.gi * D_N);
  }
}

void orc_free(void* h, int is_f64) {
  if (is_f64) delete (Run<double>*)h;
  else delete (Run<float>*)h;
}

#define ORC_DISPATCH(h, f64, EXPR) \
  do {                             \
    if (f64) {                     \
      auto* r = (Run<double>*)h;   \
      using F = double;            \
      EXPR;                        \
    } else {                       \
      auto* r = (Run<float>*)h;    \
      using F = float;             \
      EXPR;                        \
    }                              \
  } while (0)

size_t orc_nsplats(void* h, int f64) {
  size_t v = 0;
  ORC_DISPATCH(h, f64, v = r->b.splats.size());
  return v;
}
size_t orc_npairs(void* h, int f64) {
  size_t v = 0;
  ORC_DISPATCH(h, f64, for (auto& L : r->b.lists) v += L.size());
  return v;
}
uint64_t orc_pairs_fwd(void* h, int f64) {
  uint64_t v = 0;
  ORC_DISPATCH(h, f64, v = r->pairs_fwd);
  return v;
}
// Depth-sorted splat records: gi, and per splat 2+3+3+3+3+1+1+3 floats
// (sx, sy, conic[3], cov2d[3], color[3], raw_color[3], opacity, depth, pcam[3])
// and the tile rect [tx0, ty0, tx1, ty1].
void orc_get_splats(void* h, int f64, uint64_t* gi, void* fields /*ns x 22*/, int32_t* rects /*ns x 4*/) {
  ORC_DISPATCH(h, f64, {
    F* f = (F*)fields;
    for (size_t i = 0; i < r->b.splats.size(); ++i) {
      const auto& s = r->b.splats[i];
      gi[i] = s.gi;
      F* o = f + i * 22;
      o[0] = s.sx; o[1] = s.sy;
      for (int k = 0; k < 3; ++k) {
        o[2 + k] = s.conic[k]; o[5 + k] = s.cov2d[k]; o[8 + k] = s.color[k]; o[11 + k] = s.raw_color[k];
        o[19 + k] = s.pcam[k];
      }
      o[14] = s.opacity; o[15] = s.depth;
      o[16] = o[17] = o[18] = F(0);
      rects[i * 4] = s.tx0; rects[i * 4 + 1] = s.ty0; rects[i * 4 + 2] = s.tx1; rects[i * 4 + 3] = s.ty1;
    }
  });
}
// Per-tile lists as CSR of Gaussian indices: offsets[n_tiles + 1], gis[K].
void orc_get_tiles(void* h, int f64, uint64_t* offsets, uint32_t* gis) {
  ORC_DISPATCH(h, f64, {
    size_t k = 0;
    for (size_t t = 0; t < r->b.lists.size(); ++t) {
      offsets[t] = k;
      for (uint32_t si : r->b.lists[t]) gis[k++] = (uint32_t)r->b.splats[si].gi;
    }
    offsets[r->b.lists.size()] = k;
  });
}
void orc_get_image(void* h, int f64, void* rgb, void* final_t, int32_t* last) {
  ORC_DISPATCH(h, f64, {
    std::memcpy(rgb, r->rgb.data(), r->rgb.size() * sizeof(F));
    std::memcpy(final_t, r->final_t.data(), r->final_t.size() * sizeof(F));
    std::memcpy(last, r->last.data(), r->last.size() * sizeof(int32_t));
  });
}
double orc_get_loss_exact(void* h, int f64) {
  double v = 0.0;
  ORC_DISPATCH(h, f64, v = r->loss_exact);
  return v;
}
double orc_get_loss(void* h, int f64) {
  double v = 0;
  ORC_DISPATCH(h, f64, v = (double)r->loss);
  return v;
}
void orc_get_dl(void* h, int f64, void* dl) {
  ORC_DISPATCH(h, f64, std::memcpy(dl, r->dl.data(), r->dl.size() * sizeof(F)));
}
void orc_get_grads(void* h, int f64, void* grads) {
  ORC_DISPATCH(h, f64, std::memcpy(grads, r->grads.data(), r->grads.size() * sizeof(F)));
}

// ---- rasterizer.rs:72-192
int orc_rasterize_rgba(const float* g, size_t n, const Camera* cam, const float* bg, float* out) {
  View<float> v;
  if (!make_view<float>(*cam, v)) return 1;
  rasterize_rgba(g, n, v, bg, out);
  return 0;
}

// ---- backward.rs:646-693
double orc_loss_grad_f32(const float* rgb, const float* target, size_t w, size_t h, float lambda, float* dl) {
  return photometric_loss_grad<float>(rgb, target, w, h, lambda, dl);
}
double orc_loss_grad_f32_exact(const float* rgb, const float* target, size_t w, size_t h, float lambda, float* dl,
                               double* exact) {
  return photometric_loss_grad<float>(rgb, target, w, h, lambda, dl, exact);
}
double orc_loss_grad_f64(const double* rgb, const double* target, size_t w, size_t h, double lambda, double* dl) {
  return photometric_loss_grad<double>(rgb, target, w, h, lambda, dl);
}
double orc_ssim_channel(const float* x, const float* y, size_t w, size_t h, float* grad) {
  float k[11];
  gaussian_kernel_1d<float>(5, 1.5f, k);
  return ssim_channel<float>(x, y, w, h, k, 11, grad);
}
void orc_gaussian_kernel(float* k11) { gaussian_kernel_1d<float>(5, 1.5f, k11); }
void orc_blur(const float* src, size_t w, size_t h, float* out) {
  float k[11];
  gaussian_kernel_1d<float>(5, 1.5f, k);
  blur<float>(src, w, h, k, 11, out);
}

// glibc expf over n values (the libm call behind Rust's f32::exp).
void orc_expf_batch(const float* x, float* out, size_t n) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) out[i] = expf(x[i]);
}

// ---- math unit-level exports (math/*.rs)
float orc_sigmoid(float x) { return sigmoid(x); }
float orc_logit(float p) { return logit(p); }
void orc_sh_basis(const float d[3], float out[16]) { sh_basis(d, out); }
void orc_sh_basis_grad(const float d[3], float out[48]) { sh_basis_grad(d, (float(*)[3])out); }
void orc_eval_sh_deg3(const float* coeffs, const float d[3], float out[3]) {
  float b[16], raw[3];
  sh_basis(d, b);
  eval_sh_raw(coeffs, b, raw);
  for (int c = 0; c < 3; ++c) out[c] = clampf(raw[c], 0.0f, 1.0f);
}
void orc_q_normalize(const float q[4], float out[4]) { q_normalize(q, out); }
void orc_mat_from_quat(const float q[4], float out[9]) {
  M3<float> m = mat_from_quat_components(q);
  std::memcpy(out, m.a, sizeof(m.a));
}
void orc_to_rotation_matrix(const float q[4], float out[9]) {
  M3<float> m = to_rotation_matrix(q);
  std::memcpy(out, m.a, sizeof(m.a));
}
void orc_rotation_jacobian(const float q[4], float out[36]) {
  M3<float> j[4];
  rotation_jacobian(q, j);
  for (int k = 0; k < 4; ++k) std::memcpy(out + 9 * k, j[k].a, sizeof(j[k].a));
}
void orc_build_cov3d(const float s[3], const float q[4], float out[9]) {
  M3<float> m = build_cov3d(s, q);
  std::memcpy(out, m.a, sizeof(m.a));
}
void orc_backprop_cov3d(const float s[3], const float q[4], const float g[9], float d_ls[3], float d_rot[4]) {
  M3<float> G;
  std::memcpy(G.a, g, sizeof(G.a));
  backprop_cov3d(s, q, G, d_ls, d_rot);
}
void orc_covariance_3d(const float* g, float out[9]) {
  M3<float> m = covariance_3d(g);
  std::memcpy(out, m.a, sizeof(m.a));
}
void orc_project_covariance(const float sigma[9], const float pos[3], const float view[16], float fx, float fy,
                            float out[9]) {
  M3<float> s;
  std::memcpy(s.a, sigma, sizeof(s.a));
  M3<float> o = project_covariance(s, pos, view, fx, fy);
  std::memcpy(out, o.a, sizeof(o.a));
}
int orc_aabb(float mx, float my, const float s2[9], float n_sigma, uint32_t w, uint32_t h, int32_t out[4]) {
  M3<float> s;
  std::memcpy(s.a, s2, sizeof(s.a));
  return gaussian_aabb_2d(mx, my, s, n_sigma, w, h, out) ? 1 : 0;
}
// rasterizer.rs:273-279 and io/images.rs:76-82
float orc_linear_to_srgb(float x) { return x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f; }
float orc_srgb_to_linear(float x) { return x <= 0.04045f ? x / 12.92f : powf((x + 0.055f) / 1.055f, 2.4f); }
// load_target's loop (io/images.rs:63-69) over decoded RGBA8 pixels
void orc_rgba8_to_linear(const uint8_t* px, size_t n, float* out) {
  for (size_t i = 0; i < 4 * n; i += 4) {
    for (int c = 0; c < 3; ++c) out[i + c] = orc_srgb_to_linear((float)px[i + c] / 255.0f);
    out[i + 3] = (float)px[i + 3] / 255.0f;
  }
}
// Framebuffer::to_u8_srgb (rasterizer.rs:38-49): `as u8` truncates and saturates
void orc_to_u8_srgb(const float* v, size_t n, uint8_t* out) {
  auto as_u8 = [](float x) -> uint8_t { return (x != x || x <= 0.0f) ? 0 : (x >= 255.0f ? 255 : (uint8_t)x); };
  auto clamp01 = [](float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); };
  for (size_t i = 0; i < 4 * n; ++i)
    out[i] = (i % 4 == 3) ? as_u8(clamp01(v[i]) * 255.0f) : as_u8(orc_linear_to_srgb(clamp01(v[i])) * 255.0f);
}

// ---- optimizer.rs:116-163 (+ :165-195). Parameters AoS (Gaussian3D), grads
// AoS (GaussianGrad), moments n x 59 in the flat order
// [pos3, rot4, log_scale3, opacity1, sh48].
static float powi_f32(float a, int b) {  // compiler-rt __powisf2
  const int recip = b < 0;
  float r = 1;
  while (1) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1 / r : r;
}
float orc_powi(float a, int b) { return powi_f32(a, b); }

void orc_adam_step(float* g, const float* grads, size_t n, float* m1, float* m2, float* accum, uint32_t* denom,
                   uint64_t* t, const orc_train_cfg* cfg) {
  *t += 1;
  float bc1 = 1.0f - powi_f32(cfg->beta1, (int)*t);
  float bc2 = 1.0f - powi_f32(cfg->beta2, (int)*t);
  float lrs[59];
  for (int j = 0; j < 3; ++j) lrs[j] = cfg->lr_position;
  for (int j = 3; j < 7; ++j) lrs[j] = cfg->lr_rotation;
  for (int j = 7; j < 10; ++j) lrs[j] = cfg->lr_scale;
  lrs[10] = cfg->lr_opacity;
  for (int j = 11; j < 14; ++j) lrs[j] = cfg->lr_sh_dc;
  for (int j = 14; j < 59; ++j) lrs[j] = cfg->lr_sh_rest;
  // flat index -> Gaussian3D offset (apply_update_to_gaussian, :186-195)
  static const int dst[59] = {0, 1, 2, 4, 5, 6, 7, 8, 9, 10, 3, 11, 12, 13, 14, 15, 16, 17, 18, 19,
                              20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39,
                              40, 41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57, 58};
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    float* gp = g + (size_t)i * G_N;
    const float* gr = grads + (size_t)i * D_N;
    float* a = m1 + (size_t)i * 59;
    float* b = m2 + (size_t)i * 59;
    for (int j = 0; j < 59; ++j) {
      float gj = gr[j];
      a[j] = cfg->beta1 * a[j] + (1.0f - cfg->beta1) * gj;
      b[j] = cfg->beta2 * b[j] + (1.0f - cfg->beta2) * gj * gj;
      float m1h = a[j] / bc1;
      float m2h = b[j] / bc2;
      float upd = lrs[j] * m1h / (sqrtf(m2h) + cfg->eps);
      gp[dst[j]] -= upd;
    }
    accum[i] += gr[D_SGN];
    denom[i] += 1;
  }
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    float* q = g + (size_t)i * G_N + G_ROT;
    float nrm = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (nrm > 1e-8f) {
      q[0] /= nrm; q[1] /= nrm; q[2] /= nrm; q[3] /= nrm;
    }
  }
}

// ---- optimizer.rs:314-319
void orc_reset_opacities(float* g, size_t n) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    float* p = g + (size_t)i * G_N;
    p[G_OPL] = logit(fminf(0.01f, sigmoid(p[G_OPL])));
  }
}

// ---- pipeline/mod.rs:299-313
float orc_scene_extent(const float* g, size_t n) {
  if (n == 0) return 1.0f;
  float mn[3] = {3.40282347e+38f, 3.40282347e+38f, 3.40282347e+38f};
  float mx[3] = {-3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f};
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      mn[k] = fminf(mn[k], g[i * G_N + k]);
      mx[k] = fmaxf(mx[k], g[i * G_N + k]);
    }
  float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

// ---- optimizer.rs:206-311 (+ AdamState::resize :97-105)
// g: capacity `cap` Gaussians, *n in/out. m1/m2 (rows x 59), accum/denom
// (rows) have capacity `cap` rows; *rows is AdamState's current row count
// (it only grows, so moments stay aligned to row indices, not Gaussians).
// Returns 1 if the result would exceed `cap`.
// split_div (an extension, not the reference's): 0 for split_gaussian; > 0
// sets every child log-scale to ln(max(s_j / split_div, 1e-6)).
int orc_densify_and_prune_split(float* g, size_t* n_io, size_t cap, float* m1, float* m2, float* accum,
                                uint32_t* denom, size_t* rows, const orc_train_cfg* cfg, float extent,
                                float split_div);
int orc_densify_and_prune(float* g, size_t* n_io, size_t cap, float* m1, float* m2, float* accum, uint32_t* denom,
                          size_t* rows, const orc_train_cfg* cfg, float extent) {
  return orc_densify_and_prune_split(g, n_io, cap, m1, m2, accum, denom, rows, cfg, extent, 0.0f);
}
int orc_densify_and_prune_split(float* g, size_t* n_io, size_t cap, float* m1, float* m2, float* accum,
                                uint32_t* denom, size_t* rows, const orc_train_cfg* cfg, float extent,
                                float split_div) {
  size_t n = *n_io;
  std::vector<float> to_add;
  std::vector<uint8_t> to_remove(n, 0);
  std::vector<float> avg(n);
  for (size_t i = 0; i < n; ++i) avg[i] = denom[i] > 0 ? accum[i] / (float)denom[i] : 0.0f;
  for (size_t i = 0; i < n; ++i) {
    const float* gp = g + i * G_N;
    float alpha = sigmoid(gp[G_OPL]);
    float s[3] = {expf(gp[G_LS]), expf(gp[G_LS + 1]), expf(gp[G_LS + 2])};
    float max_scale = fmaxf(fmaxf(s[0], s[1]), s[2]);
    if (alpha < cfg->prune_opacity_thresh || max_scale > 0.1f * extent) {
      to_remove[i] = 1;
      continue;
    }
    size_t n_added = to_add.size() / G_N;
    if (avg[i] >= cfg->densify_grad_threshold && n + n_added < cfg->max_gaussians) {
      if (max_scale > 0.01f * extent) {
        // split_gaussian (:274-301)
        to_remove[i] = 1;
        float w, x, y, z;
        nalgebra_unit_quat(gp + G_ROT, w, x, y, z);
        // UnitQuaternion * Vector3(s0, 0, 0)
        float v[3] = {s[0], 0.0f, 0.0f};
        float two = 2.0f;
        float t[3] = {(y * v[2] - z * v[1]) * two, (z * v[0] - x * v[2]) * two, (x * v[1] - y * v[0]) * two};
        float cr[3] = {y * t[2] - z * t[1], z * t[0] - x * t[2], x * t[1] - y * t[0]};
        float ax[3];
        for (int k = 0; k < 3; ++k) ax[k] = t[k] * w + cr[k] + v[k];
        float a0 = ax[0] * ax[0], a1 = ax[1] * ax[1], a2 = ax[2] * ax[2];
        float nrm = sqrtf(a0 + a1 + a2);
        float off[3];
        for (int k = 0; k < 3; ++k) off[k] = ax[k] / nrm * s[0];
        float new_ls[3] = {logf(fmaxf(s[0] / 1.41421356237309504880f, 1e-6f)), gp[G_LS + 1], gp[G_LS + 2]};
        if (split_div > 0.0f)
          for (int k = 0; k < 3; ++k) new_ls[k] = logf(fmaxf(s[k] / split_div, 1e-6f));
        float new_op = clampf(alpha * 0.8f, 1e-6f, 0.999f);
        float new_logit = logit(new_op);
        for (int c = 0; c < 2; ++c) {
          size_t base = to_add.size();
          to_add.insert(to_add.end(), gp, gp + G_N);
          for (int k = 0; k < 3; ++k) {
            if (c == 0) to_add[base + k] += off[k];
            else to_add[base + k] -= off[k];
          }
          for (int k = 0; k < 3; ++k) to_add[base + G_LS + k] = new_ls[k];
          to_add[base + G_OPL] = new_logit;
        }
      } else {
        // clone_gaussian (:303-311)
        size_t base = to_add.size();
        to_add.insert(to_add.end(), gp, gp + G_N);
        float jitter = avg[i] * 0.01f;
        to_add[base + 0] += jitter;
        to_add[base + 1] += jitter;
      }
    }
  }
  size_t kept = 0;
  for (size_t i = 0; i < n; ++i)
    if (!to_remove[i]) {
      if (kept != i) std::memmove(g + kept * G_N, g + i * G_N, G_N * sizeof(float));
      ++kept;
    }
  size_t n_add = to_add.size() / G_N;
  if (kept + n_add > cap) return 1;
  if (n_add) std::memcpy(g + kept * G_N, to_add.data(), to_add.size() * sizeof(float));
  size_t new_n = kept + n_add;
  if (new_n > *rows) {
    std::fill(m1 + *rows * 59, m1 + new_n * 59, 0.0f);
    std::fill(m2 + *rows * 59, m2 + new_n * 59, 0.0f);
    std::fill(accum + *rows, accum + new_n, 0.0f);
    std::fill(denom + *rows, denom + new_n, 0u);
    *rows = new_n;
  }
  std::fill(accum, accum + *rows, 0.0f);
  std::fill(denom, denom + *rows, 0u);
  *n_io = new_n;
  return 0;
}

// ---- pipeline/sfm.rs:365-455: Gaussians from an SfM point cloud.
// Rust's slice::binary_search_by as in the std of edition-2024 toolchains
// (no early exit; on equal keys it ends on the last of them), for the search
// of a point's own x in the x-sorted points (:425-427).
static size_t rust_binary_search_x(const std::vector<std::pair<size_t, double>>& s, double key) {
  size_t size = s.size();
  if (size == 0) return 0;  // Err(0) -> unwrap_or(0)
  size_t base = 0;
  while (size > 1) {
    const size_t half = size / 2, mid = base + half;
    base = (s[mid].second > key) ? base : mid;  // cmp == Greater keeps base
    size -= half;
  }
  const double v = s[base].second;
  if (v == key) return base;                  // Ok(base)
  return 0;                                   // Err(_) -> unwrap_or(0)
}
// Returns 0, or 1 when the reference panics (a NaN in partial_cmp().unwrap()).
int orc_gaussians_from_sfm(const double* xyz, const uint8_t* rgb, size_t n, float* out, float* scales_out) {
  // estimate_initial_scales (:410-455)
  std::vector<std::pair<size_t, double>> sorted(n);
  for (size_t i = 0; i < n; ++i) {
    if (xyz[3 * i] != xyz[3 * i]) return 1;
    sorted[i] = {i, xyz[3 * i]};
  }
  std::stable_sort(sorted.begin(), sorted.end(), [](const auto& a, const auto& b) { return a.second < b.second; });
  std::vector<float> scales(n, 0.1f);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
  for (long ii = 0; ii < (long)n; ++ii) {
    const size_t i = (size_t)ii;
    const double* p = xyz + 3 * i;
    const size_t window = std::min<size_t>(50, n);
    const size_t idx = rust_binary_search_x(sorted, p[0]);
    const size_t start = idx >= window / 2 ? idx - window / 2 : 0;
    const size_t end = std::min(start + window, n);
    std::vector<double> d;
    for (size_t k = start; k < end; ++k) {
      if (sorted[k].first == i) continue;
      const double* q = xyz + 3 * sorted[k].first;
      const double dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
      d.push_back(std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    // sort_by(partial_cmp().unwrap()) panics on a NaN as soon as it compares
    // anything: from two distances on
    if (d.size() >= 2) {
      bool nan = false;
      for (double v : d) nan |= v != v;
      if (nan) {
        bad = 1;
        continue;
      }
    }
    std::sort(d.begin(), d.end());
    if (d.empty()) {
      scales[i] = 0.01f;
    } else {
      const size_t k = std::min<size_t>(3, d.size());
      double sum = -0.0;  // f64 Sum
      for (size_t q = 0; q < k; ++q) sum += d[q];
      scales[i] = (float)(sum / (double)k);
    }
  }
  if (bad) return 1;
  // gaussians_from_sfm (:365-407)
  const float SH_C0_INV = 1.0f / 0.28209479177387814f;
  const float op = logf(0.1f / (1.0f - 0.1f));  // math::logit(0.1)
  for (size_t i = 0; i < n; ++i) {
    float* g = out + i * G_N;
    std::fill(g, g + G_N, 0.0f);
    for (int d = 0; d < 3; ++d) g[G_POS + d] = (float)xyz[3 * i + d];
    g[G_ROT] = 1.0f;
    for (int ch = 0; ch < 3; ++ch) {
      const float c = (float)rgb[3 * i + ch] / 255.0f;
      g[G_SH + 16 * ch] = (powf(c, 2.2f) - 0.5f) * SH_C0_INV;
    }
    const float s = logf(fmaxf(scales[i], 1e-4f));
    g[G_LS + 0] = g[G_LS + 1] = g[G_LS + 2] = s;
    g[G_OPL] = op;
    if (scales_out) scales_out[i] = scales[i];
  }
  return 0;
}

// ---- io/images.rs:57-61: img.resize_exact(w, h, FilterType::Triangle) of
// the image crate 0.25.10 (Cargo.lock), which is not vendored in the
// reference. Restated from the crate's published imageops/sample.rs:
// `resize` runs vertical_sample (rows to an f32 image) and then
// horizontal_sample (columns, then clamp to [0, 255] and round half away
// from zero). Each pass maps output index o to the centre (o + 0.5) * ratio
// (ratio = src / dst), takes the source indices [left, right) with left =
// clamp(floor(centre - support), 0, src - 1), right = clamp(ceil(centre +
// support), left + 1, src), support = sratio = max(ratio, 1), weights
// triangle((i - (centre - 0.5)) / sratio) normalised by their f32 sum, and
// sums value * weight over the window in order, in f32.
static void resize_pass_weights(uint32_t src, uint32_t dst, uint32_t o, int64_t* left_out,
                                std::vector<float>* w) {
  const float ratio = (float)src / (float)dst;
  const float sratio = ratio < 1.0f ? 1.0f : ratio;
  const float support = 1.0f * sratio;
  const float centre = ((float)o + 0.5f) * ratio;
  int64_t left = (int64_t)floorf(centre - support);
  left = std::min<int64_t>(std::max<int64_t>(left, 0), (int64_t)src - 1);
  int64_t right = (int64_t)ceilf(centre + support);
  right = std::min<int64_t>(std::max<int64_t>(right, left + 1), (int64_t)src);
  const float c = centre - 0.5f;
  w->clear();
  float sum = 0.0f;
  for (int64_t i = left; i < right; ++i) {
    const float x = fabsf(((float)i - c) / sratio);
    const float v = x < 1.0f ? 1.0f - x : 0.0f;
    w->push_back(v);
    sum += v;
  }
  for (float& v : *w) v /= sum;
  *left_out = left;
}

int orc_resize_rgba8(const uint8_t* src, uint32_t sw, uint32_t sh, uint8_t* dst, uint32_t dw, uint32_t dh) {
  if (!sw || !sh || !dw || !dh) return 1;
  if (sw == dw && sh == dh) {
    memcpy(dst, src, (size_t)sw * sh * 4);
    return 0;
  }
  std::vector<float> tmp((size_t)sw * dh * 4);
  std::vector<float> w;
  for (uint32_t y = 0; y < dh; ++y) {  // vertical_sample
    int64_t left;
    resize_pass_weights(sh, dh, y, &left, &w);
    for (uint32_t x = 0; x < sw; ++x)
      for (int ch = 0; ch < 4; ++ch) {
        float t = 0.0f;
        for (size_t i = 0; i < w.size(); ++i) t += (float)src[((size_t)(left + i) * sw + x) * 4 + ch] * w[i];
        tmp[((size_t)y * sw + x) * 4 + ch] = t;
      }
  }
  for (uint32_t x = 0; x < dw; ++x) {  // horizontal_sample
    int64_t left;
    resize_pass_weights(sw, dw, x, &left, &w);
    for (uint32_t y = 0; y < dh; ++y)
      for (int ch = 0; ch < 4; ++ch) {
        float t = 0.0f;
        for (size_t i = 0; i < w.size(); ++i) t += tmp[((size_t)y * sw + left + i) * 4 + ch] * w[i];
        const float cl = t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t);
        dst[((size_t)y * dw + x) * 4 + ch] = (uint8_t)roundf(cl);
      }
  }
  return 0;
}

// ---- gpu/compute.rs + gpu/splat.wgsl: the preview renderer (GpuRenderer).
// SplatInfo2D (compute.rs:65-78): sx, sy, conic_a, conic_b, conic_c, color
// r, g, b, alpha, depth.
//
// splat.wgsl `project` (:120-190) with every WGSL operation as one IEEE f32
// operation in source order (WGSL allows exp 3 + 2|x| ulp and division
// 2.5 ulp; this restatement takes the correctly rounded / glibc results).
// WGSL matrix constructors take columns: quat_to_mat3 is the usual rotation
// matrix of the unnormalised quaternion, the Jacobian built as
// mat3x3(vec3(fx/tz, 0, -fx tx/tz^2), vec3(0, fy/tz, -fy ty/tz^2), 0) is
// the transpose of the perspective Jacobian (so the 2-D covariance keeps no
// perspective terms), and matrix products sum over k in order.
static void wgsl_project(const float* g, const float view[16], const Camera& cam, float out[10]) {
  for (int k = 0; k < 10; ++k) out[k] = 0.0f;  // wgpu output buffers start zeroed
  float pc[3];
  for (int r = 0; r < 3; ++r) {  // camera.view * vec4(position, 1)
    float acc = view[r * 4 + 0] * g[G_POS + 0];
    acc = acc + view[r * 4 + 1] * g[G_POS + 1];
    acc = acc + view[r * 4 + 2] * g[G_POS + 2];
    acc = acc + view[r * 4 + 3] * 1.0f;
    pc[r] = acc;
  }
  const float tx = pc[0], ty = pc[1], tz = pc[2];
  if (tz <= 0.001f) return;
  const float alpha = 1.0f / (1.0f + expf(-g[G_OPL]));
  if (alpha < 1.0f / 255.0f) return;
  const float sx = cam.fx * tx / tz + cam.cx;
  const float sy = cam.fy * ty / tz + cam.cy;
  // quat_to_mat3, q = [w, x, y, z]; R[row][col]
  const float w = g[G_ROT], x = g[G_ROT + 1], y = g[G_ROT + 2], z = g[G_ROT + 3];
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
  float S[3][3] = {{expf(g[G_LS]), 0.0f, 0.0f}, {0.0f, expf(g[G_LS + 1]), 0.0f}, {0.0f, 0.0f, expf(g[G_LS + 2])}};
  auto mul = [](const float A[3][3], const float B[3][3], float C[3][3]) {
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) C[r][c] = A[r][0] * B[0][c] + A[r][1] * B[1][c] + A[r][2] * B[2][c];
  };
  auto transpose = [](const float A[3][3], float B[3][3]) {
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) B[r][c] = A[c][r];
  };
  float RS[3][3], RSt[3][3], sig3[3][3];
  mul(R, S, RS);
  transpose(RS, RSt);
  mul(RS, RSt, sig3);
  float J[3][3] = {{cam.fx / tz, 0.0f, 0.0f}, {0.0f, cam.fy / tz, 0.0f},
                   {-cam.fx * tx / (tz * tz), -cam.fy * ty / (tz * tz), 0.0f}};
  float W[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) W[r][c] = view[r * 4 + c];
  float T[3][3], Tt[3][3], TS[3][3], sig2[3][3];
  mul(J, W, T);
  transpose(T, Tt);
  mul(T, sig3, TS);
  mul(TS, Tt, sig2);
  sig2[0][0] += 0.3f;
  sig2[1][1] += 0.3f;
  const float a = sig2[0][0], b = sig2[1][0], c = sig2[1][1];  // WGSL sigma2d[0][1]: column 0, row 1
  const float det = a * c - b * b;
  if (fabsf(det) < 1e-8f) return;
  const float inv_det = 1.0f / det;
  const float SH_C0 = 0.28209479177387814f;
  auto clamp01 = [](float v) { return fminf(fmaxf(v, 0.0f), 1.0f); };
  out[0] = sx;
  out[1] = sy;
  out[2] = c * inv_det;
  out[3] = -b * inv_det;
  out[4] = a * inv_det;
  out[5] = clamp01(0.5f + g[G_SH + 0] * SH_C0);
  out[6] = clamp01(0.5f + g[G_SH + 16] * SH_C0);
  out[7] = clamp01(0.5f + g[G_SH + 32] * SH_C0);
  out[8] = alpha;
  out[9] = tz;
}

// GpuRenderer::project (compute.rs:138-207): n SplatInfo2D records.
int orc_preview_project(const float* g, size_t n, const Camera* cam, float* out) {
  float view[16];
  if (!view_matrix<float>(*cam, view)) return 1;
  for (size_t i = 0; i < n; ++i) wgsl_project(g + i * G_N, view, *cam, out + i * 10);
  return 0;
}

// Rust `f as usize` of a value already floored and >= 0 (NaN -> 0), capped
// at w: beyond it the pixel loops are empty either way.
static size_t as_index_capped(float v, size_t w) {
  if (!(v > 0.0f)) return 0;
  if (v >= (float)w) return w;
  return (size_t)v;
}

// composite (compute.rs:243-297) over depth_order (sort.rs:10-17): the
// splats with alpha >= 1/255 and depth > 0, nearest first. sort_unstable_by
// leaves equal depths unordered; they are taken in index order here.
int orc_preview_render(const float* g, size_t n, const Camera* cam, const float bg[3], float* rgb) {
  std::vector<float> sp(n * 10);
  if (orc_preview_project(g, n, cam, sp.data())) return 1;
  const size_t w = cam->width, h = cam->height;
  std::vector<size_t> order;
  for (size_t i = 0; i < n; ++i)
    if (sp[i * 10 + 8] >= 1.0f / 255.0f && sp[i * 10 + 9] > 0.0f) order.push_back(i);
  std::stable_sort(order.begin(), order.end(), [&](size_t p, size_t q) { return sp[p * 10 + 9] < sp[q * 10 + 9]; });
  std::vector<float> trans(w * h, 1.0f);
  for (size_t k = 0; k < w * h * 3; ++k) rgb[k] = 0.0f;
  for (size_t i : order) {
    const float* s = &sp[i * 10];
    const float a = s[2], b = s[3], c = s[4];
    const float mid = 0.5f * (a + c);
    const float disc = sqrtf(0.25f * (a - c) * (a - c) + b * b);
    const float lambda_min = fmaxf(mid - disc, 1e-6f);
    const float rad = 3.0f * sqrtf(1.0f / lambda_min);
    const size_t x0 = as_index_capped(fmaxf(floorf(s[0] - rad), 0.0f), w);
    const size_t y0 = as_index_capped(fmaxf(floorf(s[1] - rad), 0.0f), h);
    const float ex = ceilf(s[0] + rad), ey = ceilf(s[1] + rad);  // `as i64`, then clamp(0, w)
    const size_t x1 = ex != ex ? 0 : (size_t)fminf(fmaxf(ex, 0.0f), (float)w);
    const size_t y1 = ey != ey ? 0 : (size_t)fminf(fmaxf(ey, 0.0f), (float)h);
    for (size_t py = y0; py < y1; ++py)
      for (size_t px = x0; px < x1; ++px) {
        const size_t pix = py * w + px;
        const float t = trans[pix];
        if (t < 1e-4f) continue;
        const float dx = (float)px + 0.5f - s[0];
        const float dy = (float)py + 0.5f - s[1];
        const float power = -0.5f * (a * dx * dx + 2.0f * b * dx * dy + c * dy * dy);
        if (power > 0.0f) continue;
        const float alpha = fminf(s[8] * expf(power), 0.99f);
        if (alpha < 1.0f / 255.0f) continue;
        const float weight = alpha * t;
        rgb[pix * 3] += s[5] * weight;
        rgb[pix * 3 + 1] += s[6] * weight;
        rgb[pix * 3 + 2] += s[7] * weight;
        trans[pix] = t * (1.0f - alpha);
      }
  }
  for (size_t pix = 0; pix < w * h; ++pix) {
    const float t = trans[pix];
    rgb[pix * 3] += t * bg[0];
    rgb[pix * 3 + 1] += t * bg[1];
    rgb[pix * 3 + 2] += t * bg[2];
  }
  return 0;
}

// ---- io/ply.rs: model files. Records are Gaussian3D (AoS, 59 floats).
// save_ply (io/ply.rs:35-102): binary little-endian 3DGS layout.
int orc_save_ply(const float* g, size_t n, const char* path) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return 1;
  std::fprintf(f, "ply\nformat binary_little_endian 1.0\nelement vertex %zu\n", n);
  const char* xyz[6] = {"x", "y", "z", "nx", "ny", "nz"};
  for (int i = 0; i < 6; ++i) std::fprintf(f, "property float %s\n", xyz[i]);
  for (int i = 0; i < 3; ++i) std::fprintf(f, "property float f_dc_%d\n", i);
  for (int i = 0; i < 45; ++i) std::fprintf(f, "property float f_rest_%d\n", i);
  std::fprintf(f, "property float opacity\n");
  for (int i = 0; i < 3; ++i) std::fprintf(f, "property float scale_%d\n", i);
  for (int i = 0; i < 4; ++i) std::fprintf(f, "property float rot_%d\n", i);
  std::fprintf(f, "end_header\n");
  std::vector<float> rec(62);
  for (size_t i = 0; i < n; ++i) {
    const float* q = g + i * G_N;
    int k = 0;
    for (int d = 0; d < 3; ++d) rec[k++] = q[G_POS + d];
    for (int d = 0; d < 3; ++d) rec[k++] = 0.0f;
    for (int ch = 0; ch < 3; ++ch) rec[k++] = q[G_SH + ch * 16];
    for (int ch = 0; ch < 3; ++ch)
      for (int j = 1; j < 16; ++j) rec[k++] = q[G_SH + ch * 16 + j];
    rec[k++] = q[G_OPL];
    for (int d = 0; d < 3; ++d) rec[k++] = q[G_LS + d];
    for (int d = 0; d < 4; ++d) rec[k++] = q[G_ROT + d];
    std::fwrite(rec.data(), 4, 62, f);  // little-endian host
  }
  return std::fclose(f) == 0 ? 0 : 1;
}

// load_ply (io/ply.rs:116-247). Returns 0, or 1 I/O error, 2 bad header,
// 3 short body, 4 capacity. *n gets the vertex count (also on code 4).
static void ply_assign(float* g, const char* name, float v) {  // gaussian_from_props (:216-247)
  if (!std::strcmp(name, "x")) g[G_POS + 0] = v;
  else if (!std::strcmp(name, "y")) g[G_POS + 1] = v;
  else if (!std::strcmp(name, "z")) g[G_POS + 2] = v;
  else if (!std::strcmp(name, "f_dc_0")) g[G_SH + 0] = v;
  else if (!std::strcmp(name, "f_dc_1")) g[G_SH + 16] = v;
  else if (!std::strcmp(name, "f_dc_2")) g[G_SH + 32] = v;
  else if (!std::strcmp(name, "opacity")) g[G_OPL] = v;
  else if (!std::strcmp(name, "scale_0")) g[G_LS + 0] = v;
  else if (!std::strcmp(name, "scale_1")) g[G_LS + 1] = v;
  else if (!std::strcmp(name, "scale_2")) g[G_LS + 2] = v;
  else if (!std::strcmp(name, "rot_0")) g[G_ROT + 0] = v;
  else if (!std::strcmp(name, "rot_1")) g[G_ROT + 1] = v;
  else if (!std::strcmp(name, "rot_2")) g[G_ROT + 2] = v;
  else if (!std::strcmp(name, "rot_3")) g[G_ROT + 3] = v;
  else if (!std::strncmp(name, "f_rest_", 7)) {
    // usize::parse: decimal digits only (a leading '+' is accepted)
    const char* d = name + 7;
    if (*d == '+') ++d;
    if (!*d) return;
    unsigned long long idx = 0;
    for (const char* c = d; *c; ++c) {
      if (*c < '0' || *c > '9') return;
      idx = idx * 10 + (unsigned)(*c - '0');
      if (idx > (1ull << 40)) return;
    }
    const unsigned long long ch = idx / 15, k = idx % 15;
    if (ch < 3) g[G_SH + ch * 16 + 1 + k] = v;
  }
}
static void ply_default(float* g) {  // Gaussian3D::new([0, 0, 0]) (math/mod.rs:52-62)
  std::fill(g, g + G_N, 0.0f);
  g[G_ROT] = 1.0f;
  g[G_LS + 0] = g[G_LS + 1] = g[G_LS + 2] = -3.0f;
  g[G_OPL] = -2.0f;
}
// Rust f32::from_str on a whitespace-free token: decimal (with exponent),
// inf/infinity/nan (any case, optional sign); anything else fails (-> 0).
static float rust_parse_f32(const char* t) {
  const char* d = t;
  if (*d == '+' || *d == '-') ++d;
  if (!strcasecmp(d, "inf") || !strcasecmp(d, "infinity")) return t[0] == '-' ? -INFINITY : INFINITY;
  if (!strcasecmp(d, "nan")) return NAN;
  for (const char* c = d; *c; ++c)
    if (!((*c >= '0' && *c <= '9') || *c == '.' || *c == 'e' || *c == 'E' || *c == '+' || *c == '-')) return 0.0f;
  char* end = nullptr;
  float v = std::strtof(t, &end);
  if (end == t || *end) return 0.0f;
  return v;
}
int orc_load_ply(const char* path, float* out, size_t cap, size_t* n_out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return 1;
  auto read_line = [&](std::string& line) {
    line.clear();
    int ch;
    while ((ch = std::fgetc(f)) != EOF) {
      line.push_back((char)ch);
      if (ch == '\n') break;
    }
    return !line.empty();
  };
  auto split = [](const std::string& l) {
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
  };
  std::string line;
  auto trim = [](const std::string& l) {
    size_t a = 0, b = l.size();
    while (a < b && std::isspace((unsigned char)l[a])) ++a;
    while (b > a && std::isspace((unsigned char)l[b - 1])) --b;
    return l.substr(a, b - a);
  };
  if (!read_line(line) || trim(line) != "ply") { std::fclose(f); return 2; }
  int fmt = 2;  // 0 binary LE, 1 binary BE, 2 ascii
  size_t count = 0;
  std::vector<std::string> names;
  std::vector<int> widths;
  while (true) {
    if (!read_line(line)) { std::fclose(f); return 2; }
    auto t = split(line);
    if (t.size() == 3 && t[0] == "format") {
      if (t[1] == "binary_little_endian") fmt = 0;
      else if (t[1] == "binary_big_endian") fmt = 1;
      else if (t[1] == "ascii") fmt = 2;
      else { std::fclose(f); return 2; }
    } else if (t.size() == 3 && t[0] == "element" && t[1] == "vertex") {
      char* end = nullptr;
      const char* s0 = t[2].c_str();
      if (*s0 == '-' || !*s0) { std::fclose(f); return 2; }
      unsigned long long v = std::strtoull(s0, &end, 10);
      if (*end) { std::fclose(f); return 2; }
      count = (size_t)v;
    } else if (t.size() == 3 && t[0] == "property") {
      const std::string& ty = t[1];
      int w;
      if (ty == "float" || ty == "float32") w = 4;
      else if (ty == "double" || ty == "float64") w = 8;
      else if (ty == "uchar" || ty == "uint8" || ty == "char" || ty == "int8") w = 1;
      else if (ty == "short" || ty == "ushort" || ty == "int16" || ty == "uint16") w = 2;
      else if (ty == "int" || ty == "uint" || ty == "int32" || ty == "uint32") w = 4;
      else { std::fclose(f); return 2; }
      names.push_back(t[2]);
      widths.push_back(w);
    } else if (t.size() == 1 && t[0] == "end_header") {
      break;
    }
  }
  *n_out = count;
  if (count > cap) { std::fclose(f); return 4; }
  const size_t np = names.size();
  if (fmt == 2) {
    for (size_t i = 0; i < count; ++i) {
      read_line(line);  // EOF leaves an empty line: every value 0
      auto t = split(line);
      float* g = out + i * G_N;
      ply_default(g);
      for (size_t k = 0; k < np; ++k) ply_assign(g, names[k].c_str(), k < t.size() ? rust_parse_f32(t[k].c_str()) : 0.0f);
    }
  } else {
    size_t rb = 0;
    for (int w : widths) rb += (size_t)w;
    std::vector<unsigned char> rec(rb);
    for (size_t i = 0; i < count; ++i) {
      if (rb && std::fread(rec.data(), 1, rb, f) != rb) { std::fclose(f); return 3; }
      float* g = out + i * G_N;
      ply_default(g);
      size_t off = 0;
      for (size_t k = 0; k < np; ++k) {
        const int w = widths[k];
        unsigned char b[8];
        for (int q = 0; q < w; ++q) b[q] = rec[off + (fmt == 0 ? q : w - 1 - q)];  // to little-endian
        float v = 0.0f;
        if (w == 4) { std::memcpy(&v, b, 4); }
        else if (w == 8) { double d; std::memcpy(&d, b, 8); v = (float)d; }
        else if (w == 1) { v = (float)b[0]; }
        else if (w == 2) { uint16_t u; std::memcpy(&u, b, 2); v = (float)u; }
        ply_assign(g, names[k].c_str(), v);
        off += (size_t)w;
      }
    }
  }
  std::fclose(f);
  return 0;
}

// save_splat (io/ply.rs:258-303): 32 bytes per Gaussian.
static uint8_t rust_as_u8(float x) {  // Rust `as u8`: saturate, NaN -> 0
  if (x != x) return 0;
  if (x <= 0.0f) return 0;
  if (x >= 255.0f) return 255;
  return (uint8_t)x;
}
static float rust_clamp(float x, float lo, float hi) {  // f32::clamp (NaN stays NaN)
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
int orc_save_splat(const float* g, size_t n, const char* path) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return 1;
  const float SH_C0_SPLAT = 0.2820948f;  // io/ply.rs:25
  for (size_t i = 0; i < n; ++i) {
    const float* q = g + i * G_N;
    unsigned char rec[32];
    int off = 0;
    for (int d = 0; d < 3; ++d) { std::memcpy(rec + off, &q[G_POS + d], 4); off += 4; }
    for (int d = 0; d < 3; ++d) { float s = EXP(q[G_LS + d]); std::memcpy(rec + off, &s, 4); off += 4; }
    auto to_u8 = [](float x) { return rust_as_u8(std::round(rust_clamp(x, 0.0f, 1.0f) * 255.0f)); };
    rec[off + 0] = to_u8(0.5f + q[G_SH + 0] * SH_C0_SPLAT);
    rec[off + 1] = to_u8(0.5f + q[G_SH + 16] * SH_C0_SPLAT);
    rec[off + 2] = to_u8(0.5f + q[G_SH + 32] * SH_C0_SPLAT);
    rec[off + 3] = to_u8(sigmoid(q[G_OPL]));
    off += 4;
    float qn[4];
    q_normalize(q + G_ROT, qn);
    for (int d = 0; d < 4; ++d) rec[off++] = rust_as_u8(std::round(rust_clamp(qn[d] * 128.0f + 128.0f, 0.0f, 255.0f)));
    std::fwrite(rec, 1, 32, f);
  }
  return std::fclose(f) == 0 ? 0 : 1;
}

}  // extern "C"

// k_optim.hip -- geometry backprop, Adam, densify/prune, opacity reset and
// parameter layout transposes.
//
// geometry_grad follows backward.rs:518-638 (with covariance.rs:42-84 and
// quaternion.rs:10-70), recomputing p_cam, the 2-D covariance, the SH basis
// and the raw colour from the parameters instead of storing a per-splat
// cache. Adam follows optimizer.rs:116-195; densify/prune follows
// optimizer.rs:206-311 as a scan-and-scatter compaction whose exclusive scan
// of requested additions reproduces the order-dependent max_gaussians cap
// exactly. All HBM-bound.
#include "rgs_kernels.h"

namespace rgs {

// Gaussian3D offset of Adam flat index j (apply_update_to_gaussian).
__device__ __forceinline__ int aos_offset(int j) {
  if (j < 3) return j;
  if (j < 10) return j + 1;
  if (j == 10) return 3;
  return j;
}

// Per view: SH basis at the view direction and the clamp-masked dL/dcolour
// (backward.rs:533-544, sh.rs:39-61,122-139).
template <class Load>
__device__ __forceinline__ void view_basis(const CamDev& c, Load prm, float basis[16], float dir[3] = nullptr) {
  const float px = prm(0), py = prm(1), pz = prm(2);
  float u0 = c.cam_pos[0] - px, u1 = c.cam_pos[1] - py, u2 = c.cam_pos[2] - pz;
  float ulen = __builtin_sqrtf(u0 * u0 + u1 * u1 + u2 * u2);
  float d0, d1, d2;
  if (ulen < 1e-8f) {
    d0 = 0.0f; d1 = 0.0f; d2 = 1.0f;
  } else {
    d0 = u0 / ulen; d1 = u1 / ulen; d2 = u2 / ulen;
  }
  sh_basis(d0, d1, d2, basis);
  if (dir) dir[0] = d0, dir[1] = d1, dir[2] = d2;
}
// bits: bit ch set when channel ch's colour is inside the clamp (0, 1).
// Also dk[k] = sum_ch masked[ch] * sh[ch][k], the colour gradient contracted
// with the coefficients, which the position gradient through the view
// direction needs. One channel's 16 coefficients are live at a time.
template <class Load>
__device__ __forceinline__ void view_sh(const CamDev& c, Load prm, const float acc[9], float basis[16],
                                        float masked[3], float dk[16], uint32_t* bits = nullptr) {
  view_basis(c, prm, basis);
  uint32_t b = 0;
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    float cf[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) cf[k] = prm(P_SH + ch * 16 + k);
    float a = 0.0f;
#pragma unroll
    for (int k = 0; k < 16; ++k) a += basis[k] * cf[k];
    float raw = 0.5f + a;
    const bool in = raw > 0.0f && raw < 1.0f;
    masked[ch] = in ? acc[ch] : 0.0f;
    b |= (in ? 1u : 0u) << ch;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (ch == 0)
        dk[k] = masked[0] * cf[k];
      else
        dk[k] = dk[k] + masked[ch] * cf[k];
    }
  }
  if (bits) *bits = b;
}

// dL/d(view direction) from dk (view_sh), through the SH basis (sh.rs:68-89);
// only used where |cam - pos| > 1e-8 (geometry_nonsh).
template <class Load>
__device__ __forceinline__ void view_dir_grad(const CamDev& c, Load prm, const float dk[16], float dd[3]) {
  const float px = prm(0), py = prm(1), pz = prm(2);
  float u0 = c.cam_pos[0] - px, u1 = c.cam_pos[1] - py, u2 = c.cam_pos[2] - pz;
  float ulen = __builtin_sqrtf(u0 * u0 + u1 * u1 + u2 * u2);
  float e0 = u0 / ulen, e1 = u1 / ulen, e2 = u2 / ulen;
  sh_basis_grad_contract(e0, e1, e2, dk, dd);
}

// Gradients of everything but the SH coefficients (backward.rs:529-531,
// 546-636), split into the view-independent inputs (GeomConst), the per-view
// part (geometry_view: dL/dposition, the screen-gradient norm, and
// dL/dSigma and dL/dopacity-logit factors, summed over views) and the
// view-independent tail (geometry_finish: rotation, log-scale and opacity
// gradients from the summed dL/dSigma, which they depend on linearly, so the
// covariance backprop runs once per Gaussian rather than once per view).
// Gradient-only arithmetic of the geometry backprop and the training step's
// Adam uses hardware reciprocal, square root and exp2 (about 1 ulp) and
// fused multiply-adds: these values feed no cull or clamp decision, and the
// results stay within float rounding of the reference's (well inside the
// gradient tolerance). The reference-layout Adam (k_adam_aos) keeps IEEE
// division and square root.
__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float exp_fast(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }

struct GeomConst {
  float q[4];      // raw quaternion
  float scale[3];  // exp(log_scale)
  float op;        // sigmoid(opacity logit)
};

// Sum over views of dL/dSigma: each view's T^T G T is symmetric (G is), so
// the upper triangle is kept; doubled() gives dsig + dsig^T.
struct SymSum {
  float u[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};  // 00 01 02 11 12 22
  __device__ M3 doubled() const {
    M3 r;
    r.a[0][0] = u[0] + u[0];
    r.a[0][1] = r.a[1][0] = u[1] + u[1];
    r.a[0][2] = r.a[2][0] = u[2] + u[2];
    r.a[1][1] = u[3] + u[3];
    r.a[1][2] = r.a[2][1] = u[4] + u[4];
    r.a[2][2] = u[5] + u[5];
    return r;
  }
};

template <class Load>
__device__ __forceinline__ GeomConst geom_const(Load prm) {
  GeomConst G;
#pragma unroll
  for (int k = 0; k < 4; ++k) G.q[k] = prm(P_ROT + k);
#pragma unroll
  for (int k = 0; k < 3; ++k) G.scale[k] = exp_fast(prm(P_LS + k));
  G.op = rcp_fast(1.0f + exp_fast(-prm(P_OP)));
  return G;
}

// 3-D covariance, upper triangle (00 01 02 11 12 22): Sigma = (R S)(R S)^T
// with covariance_3d's quaternion normalisation and rotation formula
// (math/mod.rs:89-96); R S formed column by column (S is diagonal).
struct Sym3 {
  float m[6];
};
__device__ __forceinline__ Sym3 cov_sym(const GeomConst& G) {
#pragma clang fp contract(fast)
  const float qw = G.q[0], qx = G.q[1], qy = G.q[2], qz = G.q[3];
  const float in = rcp_fast(sqrt_fast((qx * qx + qz * qz) + (qy * qy + qw * qw)));
  const float i = qx * in, j = qy * in, k = qz * in, w = qw * in;
  const float ww = w * w, ii = i * i, jj = j * j, kk = k * k;
  const float ij = i * j * 2.0f, wk = w * k * 2.0f, wj = w * j * 2.0f;
  const float ik = i * k * 2.0f, jk = j * k * 2.0f, wi = w * i * 2.0f;
  const float s0 = G.scale[0], s1 = G.scale[1], s2 = G.scale[2];
  const float r0[3] = {(ww + ii - jj - kk) * s0, (ij - wk) * s1, (wj + ik) * s2};
  const float r1[3] = {(ij + wk) * s0, (ww - ii + jj - kk) * s1, (jk - wi) * s2};
  const float r2[3] = {(ik - wj) * s0, (wi + jk) * s1, (ww - ii - jj + kk) * s2};
  Sym3 S;
  S.m[0] = r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2];
  S.m[1] = r0[0] * r1[0] + r0[1] * r1[1] + r0[2] * r1[2];
  S.m[2] = r0[0] * r2[0] + r0[1] * r2[1] + r0[2] * r2[2];
  S.m[3] = r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2];
  S.m[4] = r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2];
  S.m[5] = r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2];
  return S;
}

__device__ __forceinline__ float dot3(const float* a, const float* b) {
#pragma clang fp contract(fast)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// One view: adds dL/dSigma to dsig, dL/dposition to dpos, the screen-gradient
// norm to sgn (backward.rs:546-636 without the covariance backprop). The
// EWA Jacobian J has zero third row, so only rows T0, T1 of T = J W enter:
// the 2-D covariance is (T_r . Sigma T_s) for r, s in {0, 1}, dL/dSigma is
// T^T G T with G the 2x2 loss gradient, and each p_cam component's dJ
// (projection.rs:47-61) touches at most two entries of those rows, so the
// position path needs only the six products W_r . Sigma T_s.
__device__ __forceinline__ void geometry_view(const CamDev& c, const float p[3], const Sym3& S, const float acc[9],
                                              const float dd[3], SymSum& dsig, float dpos[3], float& sgn) {
#pragma clang fp contract(fast)
  float pc[3];
  to_camera(c, p[0], p[1], p[2], pc);
  const float tx = pc[0], ty = pc[1], tz = pc[2];
  const float fx = c.fx, fy = c.fy;
  const float itz = rcp_fast(tz), itz2 = itz * itz, itz3 = itz2 * itz;
  const float* W0 = c.view;
  const float* W1 = c.view + 4;
  const float* W2 = c.view + 8;
  const float j00 = fx * itz, j02 = -fx * tx * itz2, j11 = fy * itz, j12 = -fy * ty * itz2;
  float T0[3], T1[3], u0[3], u1[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    T0[m] = j00 * W0[m] + j02 * W2[m];
    T1[m] = j11 * W1[m] + j12 * W2[m];
  }
  // u_r = Sigma T_r
  u0[0] = S.m[0] * T0[0] + S.m[1] * T0[1] + S.m[2] * T0[2];
  u0[1] = S.m[1] * T0[0] + S.m[3] * T0[1] + S.m[4] * T0[2];
  u0[2] = S.m[2] * T0[0] + S.m[4] * T0[1] + S.m[5] * T0[2];
  u1[0] = S.m[0] * T1[0] + S.m[1] * T1[1] + S.m[2] * T1[2];
  u1[1] = S.m[1] * T1[0] + S.m[3] * T1[1] + S.m[4] * T1[2];
  u1[2] = S.m[2] * T1[0] + S.m[4] * T1[1] + S.m[5] * T1[2];
  const float a2 = dot3(T0, u0) + 0.3f, b2 = dot3(T1, u0), c2 = dot3(T1, u1) + 0.3f;
  const float ga = acc[4], gb = acc[5], gc = acc[6];
  const float det = a2 * c2 - b2 * b2;
  const float idet2 = rcp_fast(det * det);
  const float dl_da = (-ga * c2 * c2 + gb * b2 * c2 - gc * b2 * b2) * idet2;
  const float dl_db = (2.0f * ga * b2 * c2 - gb * (det + 2.0f * b2 * b2) + 2.0f * gc * a2 * b2) * idet2;
  const float dl_dc = (-ga * b2 * b2 + gb * a2 * b2 - gc * a2 * a2) * idet2;

  // dL/dSigma = T^T G T, G = [[dl_da, dl_db/2], [dl_db/2, dl_dc]]:
  // entry (i, j) = T0_i A_j + T1_i B_j with (A; B) = G (T0; T1)
  const float hb = 0.5f * dl_db;
  float A[3], B[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    A[m] = dl_da * T0[m] + hb * T1[m];
    B[m] = hb * T0[m] + dl_dc * T1[m];
  }
  dsig.u[0] += T0[0] * A[0] + T1[0] * B[0];
  dsig.u[1] += T0[0] * A[1] + T1[0] * B[1];
  dsig.u[2] += T0[0] * A[2] + T1[0] * B[2];
  dsig.u[3] += T0[1] * A[1] + T1[1] * B[1];
  dsig.u[4] += T0[1] * A[2] + T1[1] * B[2];
  dsig.u[5] += T0[2] * A[2] + T1[2] * B[2];

  // screen mean and the mean-dependence of J -> p_cam -> position. For p_cam
  // component k, dT = dJ_k W and d(a, b, c) = (2 dT0.u0, dT0.u1 + dT1.u0,
  // 2 dT1.u1); dJ_x = e at (0,2); dJ_y = f at (1,2); dJ_z = e at (0,0),
  // gx at (0,2), f at (1,1), gy at (1,2).
  const float asx = acc[7], asy = acc[8];
  const float w0u0 = dot3(W0, u0), w0u1 = dot3(W0, u1), w1u0 = dot3(W1, u0), w1u1 = dot3(W1, u1);
  const float w2u0 = dot3(W2, u0), w2u1 = dot3(W2, u1);
  const float e = -fx * itz2, f = -fy * itz2, gx = 2.0f * fx * tx * itz3, gy = 2.0f * fy * ty * itz3;
  const float da2 = 2.0f * (e * w0u0 + gx * w2u0);
  const float db2 = (e * w0u1 + gx * w2u1) + (f * w1u0 + gy * w2u0);
  const float dc2 = 2.0f * (f * w1u1 + gy * w2u1);
  float dpc[3];
  dpc[0] = asx * j00 + (dl_da * (2.0f * e * w2u0) + dl_db * (e * w2u1));
  dpc[1] = asy * j11 + (dl_db * (f * w2u0) + dl_dc * (2.0f * f * w2u1));
  dpc[2] = (asx * j02 + asy * j12) + (dl_da * da2 + dl_db * db2 + dl_dc * dc2);
  float dp[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) dp[m] = W0[m] * dpc[0] + W1[m] * dpc[1] + W2[m] * dpc[2];
  // view-dependent colour: dir = (cam - pos)/|cam - pos|
  float u0d = c.cam_pos[0] - p[0], u1d = c.cam_pos[1] - p[1], u2d = c.cam_pos[2] - p[2];
  float ulen = sqrt_fast(u0d * u0d + u1d * u1d + u2d * u2d);
  if (ulen > 1e-8f) {
    const float iu = rcp_fast(ulen);
    const float ev[3] = {u0d * iu, u1d * iu, u2d * iu};
    const float dot = ev[0] * dd[0] + ev[1] * dd[1] + ev[2] * dd[2];
#pragma unroll
    for (int m = 0; m < 3; ++m) dp[m] += -((dd[m] - ev[m] * dot) * iu);
  }
  dpos[0] += dp[0];
  dpos[1] += dp[1];
  dpos[2] += dp[2];
  sgn += sqrt_fast(asx * asx + asy * asy);
}

// Rotation (3-6), log-scale (7-9) and opacity-logit (10) gradients from the
// dL/dSigma and dL/dopacity summed over views (covariance.rs:42-84).
__device__ __forceinline__ void geometry_finish(const GeomConst& G, const SymSum& dsig, float acc3, float* out) {
#pragma clang fp contract(fast)
  const float op = G.op;
  out[P_OP] = acc3 * op * (1.0f - op);
  // quaternion.rs normalisation (identity below 1e-12)
  float qn[4];
  const float qq = G.q[0] * G.q[0] + G.q[1] * G.q[1] + G.q[2] * G.q[2] + G.q[3] * G.q[3];
  const float qlen = sqrt_fast(qq);
  if (qlen < 1e-12f) {
    qn[0] = 1.0f; qn[1] = 0.0f; qn[2] = 0.0f; qn[3] = 0.0f;
  } else {
    const float iq = rcp_fast(qlen);
#pragma unroll
    for (int k = 0; k < 4; ++k) qn[k] = G.q[k] * iq;
  }
  M3 r = mat_from_quat(qn);
  M3 m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.a[i][j] = r.a[i][j] * G.scale[j];
  const M3 gs = dsig.doubled();
  M3 dl_dm = m3_mul(gs, m);
  M3 dl_dr;
  for (int j = 0; j < 3; ++j) {
    float a = 0.0f;
    for (int i = 0; i < 3; ++i) {
      a += dl_dm.a[i][j] * r.a[i][j];
      dl_dr.a[i][j] = dl_dm.a[i][j] * G.scale[j];
    }
    out[P_LS + j] = a * G.scale[j];
  }
  const float w = qn[0], x = qn[1], y = qn[2], z = qn[3];
  const float jac[4][9] = {
      {0.0f, -2.0f * z, 2.0f * y, 2.0f * z, 0.0f, -2.0f * x, -2.0f * y, 2.0f * x, 0.0f},
      {0.0f, 2.0f * y, 2.0f * z, 2.0f * y, -4.0f * x, -2.0f * w, 2.0f * z, 2.0f * w, -4.0f * x},
      {-4.0f * y, 2.0f * x, 2.0f * w, 2.0f * x, 0.0f, 2.0f * z, -2.0f * w, 2.0f * z, -4.0f * y},
      {-4.0f * z, -2.0f * w, 2.0f * x, 2.0f * w, -4.0f * z, 2.0f * y, 2.0f * x, 2.0f * y, 0.0f}};
  for (int k = 0; k < 4; ++k) {
    float a = 0.0f;  // column-major fold of the component products
    for (int cc = 0; cc < 3; ++cc)
      for (int rr = 0; rr < 3; ++rr) a = a + jac[k][rr * 3 + cc] * dl_dr.a[rr][cc];
    out[P_ROT + k] = a;
  }
}

// Position of one Gaussian held in registers, as a parameter loader for
// view_basis (rows 0-2 only).
struct ShPos {
  const float* p;
  __device__ float operator()(int j) const { return p[j]; }
};

struct SoaLoad {
  const float* P;
  size_t cap;
  size_t i;
  __device__ float operator()(int j) const { return P[(size_t)j * cap + i]; }
};

__global__ __launch_bounds__(256) void k_geometry_aos(const float* __restrict__ P, int cap, int n,
                                                      const CamDev* __restrict__ cams, int view,
                                                      const uint32_t* __restrict__ cnt,
                                                      const float* __restrict__ accum, size_t acc_stride,
                                                      float* __restrict__ grads) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  size_t vi = (size_t)view * cap + i;
  float* o = grads + (size_t)i * 60;
  if (cnt[vi] == 0) {
    for (int j = 0; j < 60; ++j) o[j] = 0.0f;
    return;
  }
  float acc[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) acc[k] = accum[(size_t)k * acc_stride + vi];
  SoaLoad L{P, (size_t)cap, (size_t)i};
  const CamDev& c = cams[view];
  float basis[16], masked[3], dk[16], dd[3];
  view_sh(c, L, acc, basis, masked, dk);
  view_dir_grad(c, L, dk, dd);
  for (int ch = 0; ch < 3; ++ch)
    for (int k = 0; k < 16; ++k) o[P_SH + ch * 16 + k] = masked[ch] * basis[k];
  const GeomConst G = geom_const(L);
  const float p[3] = {L(0), L(1), L(2)};
  SymSum dsig;
  float dpos[3] = {0.0f, 0.0f, 0.0f}, sgn = 0.0f;
  geometry_view(c, p, cov_sym(G), acc, dd, dsig, dpos, sgn);
  o[0] = dpos[0];
  o[1] = dpos[1];
  o[2] = dpos[2];
  o[59] = sgn;
  geometry_finish(G, dsig, acc[3], o);
}

void launch_geometry_aos(const float* P, int cap, int n, const CamDev* cams, int view, const uint32_t* cnt,
                         const float* accum, size_t acc_stride, float* grads, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_geometry_aos, dim3((n + 255) / 256), dim3(256), 0, st, P, cap, n, cams, view, cnt, accum,
                     acc_stride, grads);
}

__device__ __forceinline__ float lr_of(int j, const AdamConsts& k) {
  if (j < 3) return k.lr[0];
  if (j < 7) return k.lr[1];
  if (j < 10) return k.lr[2];
  if (j == 10) return k.lr[3];
  if (j < 14) return k.lr[4];
  return k.lr_sh_rest;
}

// optimizer.rs:142-147
__device__ __forceinline__ float adam_one(float p, float g, float& m, float& v, float lr, const AdamConsts& k) {
  m = k.beta1 * m + k.one_minus_beta1 * g;
  v = k.beta2 * v + k.one_minus_beta2 * g * g;
  float mh = m / k.bc1;
  float vh = v / k.bc2;
  float upd = lr * mh / (__builtin_sqrtf(vh) + k.eps);
  return p - upd;
}

// adam_one with the bias corrections as products and the hardware
// reciprocal and square root (see rcp_fast).
__device__ __forceinline__ float adam_fast(float p, float g, float& m, float& v, float lr, const AdamConsts& k) {
#pragma clang fp contract(fast)
  m = k.beta1 * m + k.one_minus_beta1 * g;
  v = k.beta2 * v + k.one_minus_beta2 * g * g;
  const float mh = m * k.inv_bc1;
  const float vh = v * k.inv_bc2;
  return p - lr * mh * rcp_fast(sqrt_fast(vh) + k.eps);
}

__device__ __forceinline__ void renorm_quat(float& a, float& b, float& c, float& d) {
  float nrm = __builtin_sqrtf(a * a + b * b + c * c + d * d);
  if (nrm > 1e-8f) {
    a /= nrm; b /= nrm; c /= nrm; d /= nrm;
  }
}

__global__ __launch_bounds__(256) void k_adam_aos(float* __restrict__ g, const float* __restrict__ grads, size_t n,
                                                  float* __restrict__ m1, float* __restrict__ m2,
                                                  float* __restrict__ accum, uint32_t* __restrict__ denom,
                                                  AdamConsts k) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float* p = g + i * 59;
  const float* gr = grads + i * 60;
  float* a = m1 + i * 59;
  float* b = m2 + i * 59;
  for (int j = 0; j < 59; ++j) {
    float mm = a[j], vv = b[j];
    int o = aos_offset(j);
    p[o] = adam_one(p[o], gr[j], mm, vv, lr_of(j, k), k);
    a[j] = mm;
    b[j] = vv;
  }
  accum[i] += gr[59];
  denom[i] += 1u;
  float q0 = p[4], q1 = p[5], q2 = p[6], q3 = p[7];
  renorm_quat(q0, q1, q2, q3);
  p[4] = q0; p[5] = q1; p[6] = q2; p[7] = q3;
}

void launch_adam_aos(float* g, const float* grads, size_t n, float* m1, float* m2, float* accum, uint32_t* denom,
                     const AdamConsts& k, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_adam_aos, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, grads, n, m1, m2, accum,
                     denom, k);
}

// Training step tail over Gaussians (SoA parameters): one 256-thread
// workgroup per 64 Gaussians, one role per wave, lane = Gaussian:
//   waves 0-2, colour channel ch: per view the Gaussian is visible in, the
//     SH basis at the view direction, channel ch's colour clamp
//     (backward.rs:533-544), the channel's share of the direction gradient
//     (to LDS), and the SH gradient summed over views; then Adam on the
//     channel's 16 coefficient rows;
//   wave 3, geometry: per visible view, the non-SH geometry backprop with
//     the direction gradient summed over the three channels from LDS, summed
//     over views; the covariance tail; Adam on rows 0-10, quaternion
//     renormalisation and the densification statistics.
// Every wave clears the raster accumulator rows it consumed. The colour
// waves read the SH coefficients and positions before any wave updates
// them (positions are updated by the geometry wave after the last barrier),
// as in the reference's order: gradients of the whole step first, then
// Adam (pipeline/mod.rs:203-265, optimizer.rs:116-163). Views go through
// LDS in chunks of TU_VCH. A view's pair count and accumulator values are
// loaded together, one view ahead of the one being computed: the raster
// accumulators are zero wherever the count is (the backward adds only to
// Gaussians with pairs, every consumed value is cleared here, and the host
// clears them all after a void step), so their loads need not wait for the
// count. The geometry wave issues its first view's loads before it waits
// for the colour waves' direction gradients, and both roles issue their
// Adam moments' loads before the last computation that precedes their use.
// One barrier per chunk, and a second only when another chunk follows (the
// chunk's LDS slots are rewritten then), so the colour waves go on to their
// Adam while the geometry wave computes. The update runs only when the
// step's status word allows it: the same word for every thread, read once
// and tested before each role's first store or barrier, so the whole grid
// updates or none of it.
constexpr int TU_G = 64;
constexpr int TU_VCH = 4;
constexpr uint32_t ST_SKIP_UPDATE = ST_PAIR_OVERFLOW | ST_NAN_DEPTH | ST_NEED_BACK | ST_SCAN_STALL;

__device__ __forceinline__ void train_colour(float* __restrict__ P, float* __restrict__ M1, float* __restrict__ M2,
                                             int cap, int i, bool live, int ch, const CamDev* __restrict__ cams,
                                             int n_views, const uint32_t* __restrict__ cnt,
                                             float* __restrict__ accum, size_t acc_stride,
                                             float (*s_dd)[3][3][TU_G], int lane, const AdamConsts* __restrict__ kd,
                                             float* __restrict__ gout, uint32_t stv) {
  float pos[3] = {0.0f, 0.0f, 0.0f}, cf[16], gs[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) cf[k] = 0.0f, gs[k] = 0.0f;
  const size_t row0 = (size_t)(P_SH + ch * 16);
  float* arow = accum + (size_t)ch * acc_stride + i;
  // the next view's count and accumulator value (zero for lanes past the scene)
  uint32_t cn = 0;
  float an = 0.0f;
  if (live) {
#pragma unroll
    for (int q = 0; q < 3; ++q) pos[q] = P[(size_t)q * cap + i];
#pragma unroll
    for (int k = 0; k < 16; ++k) cf[k] = P[(row0 + k) * cap + i];
    cn = cnt[i];
    an = arow[0];
  }
  if (stv & ST_SKIP_UPDATE) return;
  const ShPos L{pos};
  for (int v0 = 0; v0 < n_views; v0 += TU_VCH) {
    const int nv = min(TU_VCH, n_views - v0);
#pragma unroll 1
    for (int j = 0; j < nv; ++j) {
      const int v = v0 + j;
      const uint32_t cj = cn;
      const float aj = an;
      if (live && v + 1 < n_views) {
        cn = cnt[(size_t)(v + 1) * cap + i];
        an = arow[(size_t)(v + 1) * cap];
      }
      float dd[3] = {0.0f, 0.0f, 0.0f};
      if (cj != 0) {
        float basis[16], d[3];
        view_basis(cams[v], L, basis, d);
        float sum = 0.0f;
#pragma unroll
        for (int k = 0; k < 16; ++k) sum += basis[k] * cf[k];
        const float raw = 0.5f + sum;
        const float m = (raw > 0.0f && raw < 1.0f) ? aj : 0.0f;
#pragma unroll
        for (int k = 0; k < 16; ++k) gs[k] += m * basis[k];
        float t[3];
        sh_basis_grad_contract(d[0], d[1], d[2], cf, t);
#pragma unroll
        for (int q = 0; q < 3; ++q) dd[q] = m * t[q];
        arow[(size_t)v * cap] = 0.0f;  // consumed
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) s_dd[j][ch][q][lane] = dd[q];
    }
    __syncthreads();  // direction gradients of the chunk are in LDS
    __syncthreads();  // the geometry wave has read them
  }
  if (!live) return;
  const AdamConsts K = *kd;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const size_t o = (row0 + k) * cap + i;
    if (gout) gout[o] = gs[k];
    float mm = M1[o], vv = M2[o];
    P[o] = adam_fast(cf[k], gs[k], mm, vv, lr_of(P_SH + ch * 16 + k, K), K);
    M1[o] = mm;
    M2[o] = vv;
  }
}

// s_ga[j][0][lane]: view v0 + j's pair count (as bits), [1..6]: its
// accumulator values 3-8, staged by the geometry wave for itself.
__device__ __forceinline__ void train_geometry(float* __restrict__ P, float* __restrict__ M1,
                                               float* __restrict__ M2, int cap, int i, bool live,
                                               const CamDev* __restrict__ cams, int n_views,
                                               const uint32_t* __restrict__ cnt, float* __restrict__ accum,
                                               size_t acc_stride, float (*s_dd)[3][3][TU_G],
                                               float (*s_ga)[7][TU_G], int lane, float* __restrict__ gaccum,
                                               uint32_t* __restrict__ gdenom, const AdamConsts* __restrict__ kd,
                                               float* __restrict__ gout, uint32_t stv) {
  float prm[11];
  if (live) {
#pragma unroll
    for (int j = 0; j < 11; ++j) prm[j] = P[(size_t)j * cap + i];
  } else {
#pragma unroll
    for (int j = 0; j < 11; ++j) prm[j] = 0.0f;
  }
  if (stv & ST_SKIP_UPDATE) return;
  float pos[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) pos[q] = prm[q];
  const GeomConst G = geom_const([&](int j) { return prm[j]; });
  const Sym3 S = cov_sym(G);
  SymSum dsig;
  float g[11], sgn = 0.0f, acc3 = 0.0f;
#pragma unroll
  for (int j = 0; j < 11; ++j) g[j] = 0.0f;
  bool any = false;
  for (int v0 = 0; v0 < n_views; v0 += TU_VCH) {
    const int nv = min(TU_VCH, n_views - v0);
    // the chunk's counts and accumulator values, all in flight together
    // while the colour waves work, then staged in LDS
    // (loaded into registers first and stored after: a store right after
    // its load waits for it, which would serialise the loads)
    float ld[TU_VCH][7];
#pragma unroll
    for (int j = 0; j < TU_VCH; ++j) {
      const bool in = live && j < nv;
      const size_t vi = (size_t)(v0 + j) * cap + i;
      ld[j][0] = __builtin_bit_cast(float, in ? cnt[vi] : 0u);
#pragma unroll
      for (int q = 0; q < 6; ++q) ld[j][1 + q] = in ? accum[(size_t)(q + 3) * acc_stride + vi] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < TU_VCH; ++j)
#pragma unroll
      for (int q = 0; q < 7; ++q) s_ga[j][q][lane] = ld[j][q];
    __syncthreads();  // direction gradients of the chunk are in LDS
#pragma unroll 1
    for (int j = 0; j < nv; ++j) {
      if (__builtin_bit_cast(uint32_t, s_ga[j][0][lane]) == 0u) continue;
      float acc[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) acc[q] = q < 3 ? 0.0f : s_ga[j][q - 2][lane];
      float dd[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) dd[q] = (s_dd[j][0][q][lane] + s_dd[j][1][q][lane]) + s_dd[j][2][q][lane];
      geometry_view(cams[v0 + j], pos, S, acc, dd, dsig, g, sgn);
      acc3 += acc[3];
      any = true;
    }
    // the chunk's accumulators 3-8 are consumed
#pragma unroll 1
    for (int j = 0; j < nv; ++j) {
      if (__builtin_bit_cast(uint32_t, s_ga[j][0][lane]) == 0u) continue;
      const size_t vi = (size_t)(v0 + j) * cap + i;
#pragma unroll
      for (int q = 3; q < 9; ++q) accum[(size_t)q * acc_stride + vi] = 0.0f;
    }
    __syncthreads();  // the chunk's LDS slots may be rewritten
  }
  if (!live) return;
  if (any) geometry_finish(G, dsig, acc3, g);
  const AdamConsts K = *kd;
  float rot[4];
#pragma unroll
  for (int j = 0; j < 11; ++j) {
    const size_t o = (size_t)j * cap + i;
    if (gout) gout[o] = g[j];
    // positions and the quaternion are still in registers; the log-scales
    // and the opacity logit are read again
    const float old = j < 3 ? pos[j] : j < 7 ? G.q[j - 3] : P[o];
    float mm = M1[o], vv = M2[o];
    const float p = adam_fast(old, g[j], mm, vv, lr_of(j, K), K);
    M1[o] = mm;
    M2[o] = vv;
    if (j >= P_ROT && j < P_ROT + 4)
      rot[j - P_ROT] = p;
    else
      P[o] = p;
  }
  renorm_quat(rot[0], rot[1], rot[2], rot[3]);
#pragma unroll
  for (int q = 0; q < 4; ++q) P[(size_t)(P_ROT + q) * cap + i] = rot[q];
  if (gout) gout[(size_t)59 * cap + i] = sgn;
  gaccum[i] += sgn;
  gdenom[i] += (uint32_t)n_views;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_train_update(
    float* __restrict__ P, float* __restrict__ M1, float* __restrict__ M2, int cap, int n,
    const CamDev* __restrict__ cams, int n_views, const uint32_t* __restrict__ cnt, float* __restrict__ accum,
    size_t acc_stride, float* __restrict__ gaccum, uint32_t* __restrict__ gdenom,
    const AdamConsts* __restrict__ k_dev, const uint32_t* __restrict__ status, float* __restrict__ gout) {
  __shared__ float s_dd[TU_VCH][3][3][TU_G];  // [view in chunk][channel][component][Gaussian]
  __shared__ float s_ga[TU_VCH][7][TU_G];     // the geometry wave's counts and accumulator values
  const uint32_t stv = *status;
  const int lane = threadIdx.x & 63;
  // wave-uniform role (scalar), so that each wave runs one role's barriers
  const int role = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = blockIdx.x * TU_G + lane;
  const bool live = i < n;
  if (role < 3)
    train_colour(P, M1, M2, cap, i, live, role, cams, n_views, cnt, accum, acc_stride, s_dd, lane, k_dev, gout, stv);
  else
    train_geometry(P, M1, M2, cap, i, live, cams, n_views, cnt, accum, acc_stride, s_dd, s_ga, lane, gaccum, gdenom,
                   k_dev, gout, stv);
}

void launch_train_adam(float* P, float* M1, float* M2, int cap, int n, const CamDev* cams, int n_views,
                       const uint32_t* cnt, float* accum, size_t acc_stride, float* gaccum, uint32_t* gdenom,
                       const AdamConsts* k_dev, const uint32_t* status, hipStream_t st, float* grad_out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_train_update, dim3((n + TU_G - 1) / TU_G), dim3(256), 0, st, P, M1, M2, cap, n, cams, n_views,
                     cnt, accum, acc_stride, gaccum, gdenom, k_dev, status, grad_out);
}

// ln on the device: the f64 logarithm rounded once to f32 (correctly rounded
// except in double-rounding corner cases).
__device__ __forceinline__ float logf_dev(float x) { return (float)log((double)x); }
__device__ __forceinline__ float logit_dev(float p) { return logf_dev(p / (1.0f - p)); }

// optimizer.rs:314-319
__global__ void k_reset_opacity(float* p, size_t n, size_t stride) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float* x = p + i * stride;
  *x = logit_dev(fmin_(0.01f, sigmoid(*x)));
}
void launch_reset_opacity(float* p, size_t n, size_t stride, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_reset_opacity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, stride);
}

// ---- densify/prune (optimizer.rs:206-311)
// want bits: 0-1 children requested (0, 1 clone, 2 split), bit 2 pruned.
__global__ void k_densify_flags(const float* __restrict__ P, int cap, int n, const float* __restrict__ gaccum,
                                const uint32_t* __restrict__ gdenom, float grad_thresh, float prune_thresh,
                                float extent, uint32_t* __restrict__ want, uint32_t* __restrict__ cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float alpha = sigmoid(P[(size_t)P_OP * cap + i]);
  float s0 = rgs_expf(P[(size_t)(P_LS + 0) * cap + i]);
  float s1 = rgs_expf(P[(size_t)(P_LS + 1) * cap + i]);
  float s2 = rgs_expf(P[(size_t)(P_LS + 2) * cap + i]);
  float ms = fmax_(fmax_(s0, s1), s2);
  uint32_t d = gdenom[i];
  float avg = d > 0 ? gaccum[i] / (float)d : 0.0f;
  uint32_t w = 0;
  if (alpha < prune_thresh || ms > 0.1f * extent) {
    w = 4u;
  } else if (avg >= grad_thresh) {
    w = (ms > 0.01f * extent) ? 2u : 1u;
  }
  want[i] = w;
  cnt[i] = w & 3u;
}

// After the exclusive scan W of requested children: Gaussian i adds its
// children iff n + W_i < max_gaussians (the reference's running check).
__global__ void k_densify_decide(const uint32_t* __restrict__ want, const uint32_t* __restrict__ W, int n,
                                 uint64_t max_gaussians, uint64_t* __restrict__ packed) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t w = want[i];
  uint32_t c = w & 3u;
  uint32_t actual = (c > 0 && (uint64_t)n + W[i] < max_gaussians) ? c : 0u;
  uint32_t keep = (w & 4u) ? 0u : ((c == 2u && actual) ? 0u : 1u);
  packed[i] = (uint64_t)keep | ((uint64_t)actual << 32);
}

__global__ void k_densify_scatter(const float* __restrict__ P, float* __restrict__ Q, int cap, int n,
                                  const uint64_t* __restrict__ packed_in, const uint64_t* __restrict__ prefix,
                                  const uint64_t* __restrict__ total, const float* __restrict__ gaccum,
                                  const uint32_t* __restrict__ gdenom, uint32_t* __restrict__ n_out,
                                  float split_div) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t tot = *total;
  uint32_t kept_total = (uint32_t)(tot & 0xffffffffu);
  if (i == 0) *n_out = kept_total + (uint32_t)(tot >> 32);
  if (i >= n) return;
  uint64_t pk = packed_in[i];
  uint64_t pre = prefix[i];
  uint32_t keep = (uint32_t)(pk & 0xffffffffu), actual = (uint32_t)(pk >> 32);
  uint32_t kpos = (uint32_t)(pre & 0xffffffffu), apos = kept_total + (uint32_t)(pre >> 32);
  // a result larger than the capacity is reported through *n_out and
  // nothing past the capacity is written
  if (keep && kpos < (uint32_t)cap)
    for (int j = 0; j < P_N; ++j) Q[(size_t)j * cap + kpos] = P[(size_t)j * cap + i];
  if (!actual || apos + actual > (uint32_t)cap) return;
  if (actual == 1) {
    // clone_gaussian (:303-311)
    uint32_t d = gdenom[i];
    float avg = d > 0 ? gaccum[i] / (float)d : 0.0f;
    float jitter = avg * 0.01f;
    for (int j = 0; j < P_N; ++j) {
      float x = P[(size_t)j * cap + i];
      if (j == 0 || j == 1) x += jitter;
      Q[(size_t)j * cap + apos] = x;
    }
    return;
  }
  // split_gaussian (:274-301)
  float s0 = rgs_expf(P[(size_t)P_LS * cap + i]);
  float qw = P[(size_t)(P_ROT + 0) * cap + i], qx = P[(size_t)(P_ROT + 1) * cap + i];
  float qy = P[(size_t)(P_ROT + 2) * cap + i], qz = P[(size_t)(P_ROT + 3) * cap + i];
  float a = qx * qx, b = qy * qy, cq = qz * qz, d = qw * qw;
  a = a + cq;
  b = b + d;
  float nq = __builtin_sqrtf(a + b);
  float x = qx / nq, y = qy / nq, z = qz / nq, w = qw / nq;
  // UnitQuaternion * (s0, 0, 0): t = 2 (q_v x v); v' = t w + q_v x t + v
  float v0 = s0, v1 = 0.0f, v2 = 0.0f;
  float t0 = (y * v2 - z * v1) * 2.0f, t1 = (z * v0 - x * v2) * 2.0f, t2 = (x * v1 - y * v0) * 2.0f;
  float c0 = y * t2 - z * t1, c1 = z * t0 - x * t2, c2 = x * t1 - y * t0;
  float ax0 = t0 * w + c0 + v0, ax1 = t1 * w + c1 + v1, ax2 = t2 * w + c2 + v2;
  float nrm = __builtin_sqrtf(ax0 * ax0 + ax1 * ax1 + ax2 * ax2);
  float off[3] = {ax0 / nrm * s0, ax1 / nrm * s0, ax2 / nrm * s0};
  // the reference halves the variance along axis 0 only; with split_div > 0
  // every axis is divided by it (the workload table's "scale / 1.6")
  float new_ls[3];
  if (split_div > 0.0f) {
    for (int k = 0; k < 3; ++k)
      new_ls[k] = logf_dev(fmax_(rgs_expf(P[(size_t)(P_LS + k) * cap + i]) / split_div, 1e-6f));
  } else {
    new_ls[0] = logf_dev(fmax_(s0 / 1.41421356237309504880f, 1e-6f));
    new_ls[1] = P[(size_t)(P_LS + 1) * cap + i];
    new_ls[2] = P[(size_t)(P_LS + 2) * cap + i];
  }
  float alpha = sigmoid(P[(size_t)P_OP * cap + i]);
  float op = alpha * 0.8f;
  if (op < 1e-6f) op = 1e-6f;
  if (op > 0.999f) op = 0.999f;
  float new_logit = logit_dev(op);
  for (int ci = 0; ci < 2; ++ci) {
    size_t o = apos + ci;
    for (int j = 0; j < P_N; ++j) {
      float xv = P[(size_t)j * cap + i];
      if (j < 3) xv = ci == 0 ? xv + off[j] : xv - off[j];
      if (j >= P_LS && j < P_LS + 3) xv = new_ls[j - P_LS];
      if (j == P_OP) xv = new_logit;
      Q[(size_t)j * cap + o] = xv;
    }
  }
}

void launch_densify(const float* P, float* P_out, int cap, int n, const float* gaccum, const uint32_t* gdenom,
                    float grad_thresh, float prune_thresh, uint64_t max_gaussians, float extent, DensifyScratch ds,
                    uint32_t* n_out, hipStream_t st, float split_div) {
  if (n <= 0) {
    launch_fill_u32(n_out, 1, 0u, st);
    return;
  }
  dim3 g((n + 255) / 256), b(256);
  hipLaunchKernelGGL(k_densify_flags, g, b, 0, st, P, cap, n, gaccum, gdenom, grad_thresh, prune_thresh, extent,
                     ds.want, ds.cnt);
  scan_u32(ds.cnt, (size_t)n, ds.W, ds.bsum32, ds.total32, st);
  hipLaunchKernelGGL(k_densify_decide, g, b, 0, st, ds.want, ds.W, n, max_gaussians, ds.packed);
  scan_u64(ds.packed, (size_t)n, ds.prefix, ds.bsum64, ds.total64, st);
  hipLaunchKernelGGL(k_densify_scatter, g, b, 0, st, P, P_out, cap, n, ds.packed, ds.prefix, ds.total64, gaccum,
                     gdenom, n_out, split_div);
}

__global__ void k_zero_rows(float* M, size_t cap, size_t r0, size_t r1) {
  size_t len = r1 - r0;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= len * P_N) return;
  size_t j = idx / len, r = r0 + idx % len;
  M[j * cap + r] = 0.0f;
}
void launch_zero_rows(float* M, size_t cap, size_t r0, size_t r1, hipStream_t st) {
  if (r1 <= r0) return;
  size_t tot = (r1 - r0) * P_N;
  hipLaunchKernelGGL(k_zero_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, M, cap, r0, r1);
}

// Layout transposes between Gaussian3D records (59 floats) and the SoA rows,
// through LDS: a block moves 64 records, reading and writing both sides in
// contiguous runs. The record stride (59, odd) makes the column accesses to
// the LDS tile conflict-free.
constexpr int TR_G = 64;

__global__ __launch_bounds__(256) void k_aos_to_soa(const float* __restrict__ aos, size_t n, float* __restrict__ P,
                                                    size_t cap) {
  __shared__ float s[TR_G * P_N];
  const size_t i0 = (size_t)blockIdx.x * TR_G;
  const int cnt = (int)min((size_t)TR_G, n - i0);
  const float* src = aos + i0 * P_N;
  for (int t = threadIdx.x; t < cnt * P_N; t += blockDim.x) s[t] = src[t];
  __syncthreads();
  for (int t = threadIdx.x; t < P_N * TR_G; t += blockDim.x) {
    const int j = t / TR_G, l = t % TR_G;
    if (l < cnt) P[(size_t)j * cap + i0 + l] = s[l * P_N + aos_offset(j)];
  }
}

__global__ __launch_bounds__(256) void k_soa_to_aos(const float* __restrict__ P, size_t cap, size_t n,
                                                    float* __restrict__ aos) {
  __shared__ float s[TR_G * P_N];
  const size_t i0 = (size_t)blockIdx.x * TR_G;
  const int cnt = (int)min((size_t)TR_G, n - i0);
  for (int t = threadIdx.x; t < P_N * TR_G; t += blockDim.x) {
    const int j = t / TR_G, l = t % TR_G;
    if (l < cnt) s[l * P_N + aos_offset(j)] = P[(size_t)j * cap + i0 + l];
  }
  __syncthreads();
  float* dst = aos + i0 * P_N;
  for (int t = threadIdx.x; t < cnt * P_N; t += blockDim.x) dst[t] = s[t];
}

void launch_aos_to_soa(const float* aos, size_t n, float* P, size_t cap, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_aos_to_soa, dim3((unsigned)((n + TR_G - 1) / TR_G)), dim3(256), 0, st, aos, n, P, cap);
}
void launch_soa_to_aos(const float* P, size_t cap, size_t n, float* aos, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_soa_to_aos, dim3((unsigned)((n + TR_G - 1) / TR_G)), dim3(256), 0, st, P, cap, n, aos);
}

}  // namespace rgs

// k_io.hip -- model files (io/ply.rs) from and to the device parameter rows.
//
//   k_ply_encode:   SoA rows -> 3DGS PLY vertex records (save_ply,
//                   io/ply.rs:35-102): x y z, three zero normals, f_dc_0..2,
//                   f_rest_0..44 (channel-major), opacity, scale_0..2,
//                   rot_0..3, 62 floats each;
//   k_ply_decode:   PLY vertex records of any property order, widths and
//                   endianness -> SoA rows (load_ply and gaussian_from_props,
//                   io/ply.rs:116-247), starting from Gaussian3D::new's
//                   defaults (math/mod.rs:52-62);
//   k_splat_encode: SoA rows -> 32-byte .splat records (save_splat,
//                   io/ply.rs:258-303) with the reference's exp, sigmoid,
//                   quaternion normalisation and rounding.
// The header is text and is written / parsed on the host; the records move
// between the device and the file as one block. All HBM-bound.
#include "rgs_kernels.h"

namespace rgs {

// Flat (SoA) row of PLY record field r, or -1 for the zero normals.
__device__ __forceinline__ int ply_field_row(int r) {
  if (r < 3) return P_POS + r;                       // x y z
  if (r < 6) return -1;                              // nx ny nz
  if (r < 9) return P_SH + (r - 6) * 16;             // f_dc_c = sh[c * 16]
  if (r < 54) {                                      // f_rest_q = sh[(q / 15) * 16 + 1 + q % 15]
    const int q = r - 9;
    return P_SH + (q / 15) * 16 + 1 + q % 15;
  }
  if (r == 54) return P_OP;                          // opacity
  if (r < 58) return P_LS + (r - 55);                // scale_0..2
  return P_ROT + (r - 58);                           // rot_0..3
}

constexpr int IO_G = 64;  // Gaussians per block

// Records of IO_G Gaussians staged in LDS: SoA rows read in contiguous runs,
// records written in contiguous runs (the row stride 59 is odd, so the column
// accesses to the tile are free of bank conflicts).
__global__ __launch_bounds__(256) void k_ply_encode(const float* __restrict__ P, size_t cap, size_t n,
                                                    float* __restrict__ rec) {
  __shared__ float s[IO_G * P_N];
  const size_t i0 = (size_t)blockIdx.x * IO_G;
  const int cnt = (int)min((size_t)IO_G, n - i0);
  for (int t = threadIdx.x; t < P_N * IO_G; t += blockDim.x) {
    const int j = t / IO_G, l = t % IO_G;
    if (l < cnt) s[l * P_N + j] = P[(size_t)j * cap + i0 + l];
  }
  __syncthreads();
  float* dst = rec + i0 * PLY_FIELDS;
  for (int t = threadIdx.x; t < cnt * PLY_FIELDS; t += blockDim.x) {
    const int l = t / PLY_FIELDS, r = t % PLY_FIELDS;
    const int row = ply_field_row(r);
    dst[t] = row < 0 ? 0.0f : s[l * P_N + row];
  }
}

void launch_ply_encode(const float* P, size_t cap, size_t n, float* rec, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_ply_encode, dim3((unsigned)((n + IO_G - 1) / IO_G)), dim3(256), 0, st, P, cap, n, rec);
}

// One thread per Gaussian. Fields are applied in header order, so a later
// property of the same name wins, as in gaussian_from_props.
// Records from n_body on are all zero bytes: an ASCII file whose lines end
// before its vertex count (read_line at EOF gives an empty line, whose
// properties read as 0).
__global__ __launch_bounds__(256) void k_ply_decode(const uint8_t* __restrict__ body, size_t n, size_t n_body,
                                                    int rec_bytes, const PlyField* __restrict__ fields, int n_fields,
                                                    int big_endian, float* __restrict__ P, size_t cap) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // Gaussian3D::new([0, 0, 0]) with the identity rotation
  for (int j = 0; j < P_N; ++j) {
    float v = 0.0f;
    if (j == P_ROT) v = 1.0f;
    if (j >= P_LS && j < P_LS + 3) v = -3.0f;
    if (j == P_OP) v = -2.0f;
    P[(size_t)j * cap + i] = v;
  }
  const uint8_t* r = body + i * (size_t)rec_bytes;
  for (int k = 0; k < n_fields; ++k) {
    const PlyField f = fields[k];
    if (f.row < 0) continue;
    uint8_t b[8];
    for (int q = 0; q < f.width; ++q) b[q] = i < n_body ? r[f.offset + (big_endian ? f.width - 1 - q : q)] : 0;
    float v = 0.0f;  // decode_scalar (io/ply.rs:194-212)
    if (f.width == 4) {
      v = __builtin_bit_cast(float, (uint32_t)b[0] | (uint32_t)b[1] << 8 | (uint32_t)b[2] << 16 |
                                        (uint32_t)b[3] << 24);
    } else if (f.width == 8) {
      uint64_t u = 0;
      for (int q = 7; q >= 0; --q) u = u << 8 | b[q];
      v = (float)__builtin_bit_cast(double, u);
    } else if (f.width == 1) {
      v = (float)b[0];
    } else if (f.width == 2) {
      v = (float)(uint16_t)(b[0] | b[1] << 8);
    }
    P[(size_t)f.row * cap + i] = v;
  }
}

// The same through LDS for records of at most PLY_TILE_BYTES / 16 bytes: a
// block stages the raw bytes of its records (read as contiguous words), and
// each row is written for the block's records in one contiguous run. Per row
// only the last property that sets it is decoded (row_field[j]: its offset
// and width; width 0 keeps the default), which is what applying them all in
// header order leaves.
constexpr int PLY_TILE_BYTES = 32768;
__global__ __launch_bounds__(256) void k_ply_decode_tiled(const uint8_t* __restrict__ body, size_t n, size_t n_body,
                                                          int rec_bytes, const PlyField* __restrict__ row_field,
                                                          int big_endian, float* __restrict__ P, size_t cap,
                                                          int per_block) {
  __shared__ uint32_t s_w[PLY_TILE_BYTES / 4 + 2];
  const size_t i0 = (size_t)blockIdx.x * per_block;
  const int cnt = (int)min((size_t)per_block, n - i0);
  // bytes [b0, b1) of the body, from the 4-byte word holding b0
  const size_t b0 = i0 * (size_t)rec_bytes, b1 = b0 + (size_t)cnt * rec_bytes;
  const size_t w0 = b0 / 4, w1 = (b1 + 3) / 4;
  const int skew = (int)(b0 - w0 * 4);
  const size_t body_bytes = n_body * (size_t)rec_bytes;  // bytes past it read as 0
  const size_t body_words = (body_bytes + 3) / 4;
  for (size_t w = w0 + threadIdx.x; w < w1; w += blockDim.x) {
    uint32_t v = 0;
    if (w + 1 <= body_bytes / 4) {
      v = reinterpret_cast<const uint32_t*>(body)[w];  // whole word inside the body
    } else if (w < body_words) {
      for (int q = 0; q < 4; ++q) {  // the body's last, partial word
        const size_t b = w * 4 + q;
        if (b < body_bytes) v |= (uint32_t)body[b] << (8 * q);
      }
    }
    s_w[w - w0] = v;
  }
  __syncthreads();
  const uint8_t* s_b = reinterpret_cast<const uint8_t*>(s_w) + skew;
  for (int t = threadIdx.x; t < P_N * per_block; t += blockDim.x) {
    const int j = t / per_block, l = t % per_block;
    if (l >= cnt) continue;
    float v = 0.0f;  // Gaussian3D::new([0, 0, 0]) with the identity rotation
    if (j == P_ROT) v = 1.0f;
    if (j >= P_LS && j < P_LS + 3) v = -3.0f;
    if (j == P_OP) v = -2.0f;
    const PlyField f = row_field[j];
    const size_t at = (size_t)skew + (size_t)l * rec_bytes + f.offset;
    if (f.width == 4 && !big_endian && !(at & 3)) {
      v = __builtin_bit_cast(float, s_w[at >> 2]);  // an aligned little-endian float: one LDS word
    } else if (f.width) {
      const uint8_t* r = s_b + (size_t)l * rec_bytes + f.offset;
      uint8_t b[8];
      for (int q = 0; q < f.width; ++q) b[q] = r[big_endian ? f.width - 1 - q : q];
      if (f.width == 4) {
        v = __builtin_bit_cast(float, (uint32_t)b[0] | (uint32_t)b[1] << 8 | (uint32_t)b[2] << 16 |
                                          (uint32_t)b[3] << 24);
      } else if (f.width == 8) {
        uint64_t u = 0;
        for (int q = 7; q >= 0; --q) u = u << 8 | b[q];
        v = (float)__builtin_bit_cast(double, u);
      } else if (f.width == 1) {
        v = (float)b[0];
      } else {
        v = (float)(uint16_t)(b[0] | b[1] << 8);
      }
    }
    P[(size_t)j * cap + i0 + l] = v;
  }
}

void launch_ply_decode(const uint8_t* body, size_t n, size_t n_body, int rec_bytes, const PlyField* fields,
                       int n_fields, int big_endian, float* P, size_t cap, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_ply_decode, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, body, n, n_body, rec_bytes,
                     fields, n_fields, big_endian, P, cap);
}

int ply_decode_tiled_fits(int rec_bytes) { return rec_bytes > 0 && rec_bytes <= PLY_TILE_BYTES / 16; }

void launch_ply_decode_tiled(const uint8_t* body, size_t n, size_t n_body, int rec_bytes, const PlyField* row_field,
                             int big_endian, float* P, size_t cap, hipStream_t st) {
  if (!n) return;
  const int per_block = min(64, (PLY_TILE_BYTES - 8) / rec_bytes);
  hipLaunchKernelGGL(k_ply_decode_tiled, dim3((unsigned)((n + per_block - 1) / per_block)), dim3(256), 0, st, body, n,
                     n_body, rec_bytes, row_field, big_endian, P, cap, per_block);
}

// Rust `as u8`: saturating, NaN -> 0.
__device__ __forceinline__ uint32_t as_u8(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 255.0f) return 255u;
  return (uint32_t)x;
}
// f32::clamp (a NaN stays NaN).
__device__ __forceinline__ float clamp_rs(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

__global__ __launch_bounds__(256) void k_splat_encode(const float* __restrict__ P, size_t cap, size_t n,
                                                      uint4* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  auto prm = [&](int j) { return P[(size_t)j * cap + i]; };
  const float SH_C0_SPLAT = 0.2820948f;  // io/ply.rs:25
  auto to_u8 = [](float x) { return as_u8(__builtin_roundf(clamp_rs(x, 0.0f, 1.0f) * 255.0f)); };
  const uint32_t r = to_u8(0.5f + prm(P_SH + 0) * SH_C0_SPLAT);
  const uint32_t g = to_u8(0.5f + prm(P_SH + 16) * SH_C0_SPLAT);
  const uint32_t b = to_u8(0.5f + prm(P_SH + 32) * SH_C0_SPLAT);
  const uint32_t a = to_u8(sigmoid(prm(P_OP)));
  const float q[4] = {prm(P_ROT + 0), prm(P_ROT + 1), prm(P_ROT + 2), prm(P_ROT + 3)};
  float qn[4];
  quat_normalize(q, qn);
  uint32_t qb = 0;
  for (int d = 0; d < 4; ++d) qb |= as_u8(__builtin_roundf(clamp_rs(qn[d] * 128.0f + 128.0f, 0.0f, 255.0f))) << (8 * d);
  const float s0 = rgs_expf(prm(P_LS + 0)), s1 = rgs_expf(prm(P_LS + 1)), s2 = rgs_expf(prm(P_LS + 2));
  out[2 * i] = make_uint4(__builtin_bit_cast(uint32_t, prm(P_POS + 0)), __builtin_bit_cast(uint32_t, prm(P_POS + 1)),
                          __builtin_bit_cast(uint32_t, prm(P_POS + 2)), __builtin_bit_cast(uint32_t, s0));
  out[2 * i + 1] = make_uint4(__builtin_bit_cast(uint32_t, s1), __builtin_bit_cast(uint32_t, s2),
                              r | g << 8 | b << 16 | a << 24, qb);
}

void launch_splat_encode(const float* P, size_t cap, size_t n, void* out, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_splat_encode, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, P, cap, n,
                     static_cast<uint4*>(out));
}

}  // namespace rgs

namespace rgs {

// ---- Colour conversions ----------------------------------------------------
// load_target's per-pixel loop (io/images.rs:63-69): channel c of an RGBA8
// pixel -> srgb_to_linear(c / 255) for RGB, c / 255 for alpha. The 256 RGB
// values come from a table the host fills with the same libm powf the
// reference calls (srgb8_table), so the kernel is a lookup; 4 pixels per
// thread (16 bytes in, 64 bytes out).
__global__ __launch_bounds__(256) void k_rgba8_to_linear(const uint32_t* __restrict__ px, size_t n,
                                                         const float* __restrict__ lut, float4* __restrict__ out) {
  __shared__ float s_lut[512];  // [0, 256): RGB, [256, 512): alpha (c / 255)
  for (int t = threadIdx.x; t < 512; t += blockDim.x) s_lut[t] = lut[t];
  __syncthreads();
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = px[i];
  out[i] = make_float4(s_lut[p & 255u], s_lut[(p >> 8) & 255u], s_lut[(p >> 16) & 255u], s_lut[256 + (p >> 24)]);
}

void launch_rgba8_to_linear(const uint8_t* rgba8, size_t n_pixels, const float* lut, float* out, hipStream_t st) {
  if (!n_pixels) return;
  hipLaunchKernelGGL(k_rgba8_to_linear, dim3((unsigned)((n_pixels + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<const uint32_t*>(rgba8), n_pixels, lut, reinterpret_cast<float4*>(out));
}

// Framebuffer::to_u8_srgb (cpu/rasterizer.rs:38-49): alpha (clamp(v) * 255)
// as u8; RGB (linear_to_srgb(clamp(v)) * 255) as u8 with linear_to_srgb
// (rasterizer.rs:271-279). powf(x, 1/2.4) is the f64 pow of the f32 inputs
// rounded to f32 (the correctly rounded value but for double-rounding
// corner cases; glibc's powf is within 0.82 ulp of it): a byte can differ
// only where the scaled value lies within an ulp of an integer.
__device__ __forceinline__ float linear_to_srgb(float x) {
  if (x <= 0.0031308f) return 12.92f * x;
  return 1.055f * (float)pow((double)x, (double)(1.0f / 2.4f)) - 0.055f;
}
__device__ __forceinline__ uint32_t as_u8_trunc(float x) {  // Rust `as u8`
  if (!(x > 0.0f)) return 0u;
  if (x >= 255.0f) return 255u;
  return (uint32_t)x;
}
__global__ __launch_bounds__(256) void k_to_u8_srgb(const float4* __restrict__ in, size_t n, uint32_t* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 v = in[i];
  const float c[4] = {v.x, v.y, v.z, v.w};
  uint32_t r = 0;
  for (int k = 0; k < 4; ++k) {
    const float x = clamp_rs(c[k], 0.0f, 1.0f);
    r |= as_u8_trunc((k == 3 ? x : linear_to_srgb(x)) * 255.0f) << (8 * k);
  }
  out[i] = r;
}

void launch_to_u8_srgb(const float* rgba, size_t n_pixels, uint8_t* out, hipStream_t st) {
  if (!n_pixels) return;
  hipLaunchKernelGGL(k_to_u8_srgb, dim3((unsigned)((n_pixels + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<const float4*>(rgba), n_pixels, reinterpret_cast<uint32_t*>(out));
}

}  // namespace rgs

namespace rgs {

// image 0.25's resize (imageops/sample.rs: vertical_sample, then
// horizontal_sample) of RGBA8 with the Triangle filter. Each output row /
// column has a window of taps [left, left + cnt) with normalised f32
// weights, computed on the host (see resize_taps); the sums run over the
// taps in order from 0, in f32.
//
// Vertical pass: RGBA8 (sw x sh) -> f32 RGBA (sw x dh).
__global__ __launch_bounds__(256) void k_resize_v(const uchar4* __restrict__ src, uint32_t sw, uint32_t dh,
                                                  const int2* __restrict__ taps, const float* __restrict__ w, int K,
                                                  float4* __restrict__ tmp) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= sw || y >= dh) return;
  const int2 tp = taps[y];
  float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int i = 0; i < tp.y; ++i) {
    const uchar4 p = src[(size_t)(tp.x + i) * sw + x];
    const float wi = w[(size_t)y * K + i];
    t.x += (float)p.x * wi;
    t.y += (float)p.y * wi;
    t.z += (float)p.z * wi;
    t.w += (float)p.w * wi;
  }
  tmp[(size_t)y * sw + x] = t;
}

// FloatNearest(clamp(t, 0, 255)) as u8: round half away from zero.
__device__ __forceinline__ unsigned char nearest_u8(float t) {
  const float c = t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t);
  return (unsigned char)__builtin_roundf(c);
}

// Horizontal pass: f32 RGBA (sw x dh) -> RGBA8 (dw x dh).
__global__ __launch_bounds__(256) void k_resize_h(const float4* __restrict__ tmp, uint32_t sw, uint32_t dw,
                                                  uint32_t dh, const int2* __restrict__ taps,
                                                  const float* __restrict__ w, int K, uchar4* __restrict__ dst) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dw || y >= dh) return;
  const int2 tp = taps[x];
  float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int i = 0; i < tp.y; ++i) {
    const float4 p = tmp[(size_t)y * sw + tp.x + i];
    const float wi = w[(size_t)x * K + i];
    t.x += p.x * wi;
    t.y += p.y * wi;
    t.z += p.z * wi;
    t.w += p.w * wi;
  }
  dst[(size_t)y * dw + x] = make_uchar4(nearest_u8(t.x), nearest_u8(t.y), nearest_u8(t.z), nearest_u8(t.w));
}

void launch_resize_rgba8(const uint8_t* src, uint32_t sw, uint32_t sh, uint8_t* dst, uint32_t dw, uint32_t dh,
                         const int2* vtaps, const float* vw, int vK, const int2* htaps, const float* hw, int hK,
                         float* tmp, hipStream_t st) {
  (void)sh;
  if (!sw || !dw || !dh) return;
  hipLaunchKernelGGL(k_resize_v, dim3((sw + 255) / 256, dh), dim3(256), 0, st, reinterpret_cast<const uchar4*>(src),
                     sw, dh, vtaps, vw, vK, reinterpret_cast<float4*>(tmp));
  hipLaunchKernelGGL(k_resize_h, dim3((dw + 255) / 256, dh), dim3(256), 0, st, reinterpret_cast<const float4*>(tmp),
                     sw, dw, dh, htaps, hw, hK, reinterpret_cast<uchar4*>(dst));
}

}  // namespace rgs

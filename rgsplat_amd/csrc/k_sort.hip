// k_sort.hip -- depth sort, tile-key generation, tile sort and tile ranges.
//
// The reference sorts splats by depth (backward.rs:191) and pushes each into
// the lists of the tiles its rect covers, in depth order (:193-205), so every
// tile list is in global depth order. Here that is two stable sorts:
//   1. (depth bits, v*cap + i) over all views at once -> depth order, ties by
//      Gaussian index (stability over the index-ordered input);
//   2. after writing the (global tile id, splat) pairs in that order, a
//      stable sort on the tile id groups the pairs by tile without
//      disturbing depth order inside a tile.
// Both are the same LSD radix sort with 8-bit digits: per 4096-key tile a
// histogram kernel, one scan block per digit, and a scatter kernel that ranks
// keys stably with wave64 ballot matching, stages the sorted tile in LDS and
// writes each digit's run contiguously. Counts come from device memory so the
// whole sequence can be captured in a hipGraph.
#include "rgs_kernels.h"

namespace rgs {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = RS_TILE / RS_THREADS;  // 16
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_WAVE_ITEMS = RS_TILE / RS_WAVES;  // 1024

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_excl_scan_256(uint32_t v, uint32_t* s_tmp /*>= 4*/, uint32_t* total) {
  uint32_t lane = lane_id();
  uint32_t w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d) x += y;
  }
  if (lane == 63) s_tmp[w] = x;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    uint32_t t = s_tmp[k];
    if ((uint32_t)k < w) base += t;
    tot += t;
  }
  __syncthreads();
  if (total) *total = tot;
  return base + x - v;
}

// K: the key type (uint32_t, or uint16_t for the tile sort of at most 2^16
// bins: a quarter less traffic per pair and pass).
template <class K>
__global__ __launch_bounds__(RS_THREADS) void k_radix_hist(const K* __restrict__ keys,
                                                         const uint32_t* __restrict__ n_dev, int shift,
                                                         uint32_t mask, uint32_t* __restrict__ hist, int nb_max) {
  __shared__ uint32_t h[256];
  uint32_t n = *n_dev;
  const uint32_t b = xcd_chunked(blockIdx.x);
  uint32_t nb = (n + RS_TILE - 1) / RS_TILE;
  if (b >= nb) return;
  h[threadIdx.x] = 0;
  size_t base = (size_t)b * RS_TILE;
  // the tile's keys first (all loads in flight together), then the counts
  uint32_t kv[RS_ITEMS];
  const K* kp = keys + base + threadIdx.x;
  if (base + RS_TILE <= n) {
#pragma unroll
    for (int k = 0; k < RS_ITEMS; ++k) kv[k] = kp[k * RS_THREADS];
  } else {
#pragma unroll
    for (int k = 0; k < RS_ITEMS; ++k) kv[k] = base + (size_t)k * RS_THREADS + threadIdx.x < n ? kp[k * RS_THREADS] : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < RS_ITEMS; ++k)
    if (base + (size_t)k * RS_THREADS + threadIdx.x < n) atomicAdd(&h[(kv[k] >> shift) & mask], 1u);
  __syncthreads();
  hist[(size_t)threadIdx.x * nb_max + b] = h[threadIdx.x];
}

// One block per digit: in-place exclusive scan of hist[d][0..nb), row total.
// Each thread owns 8 consecutive entries per 8192-entry chunk (two 16-byte
// loads; rows are 16-byte aligned as nb_max is a multiple of 256).
constexpr int RSCAN_THREADS = 1024;
constexpr int RSCAN_ITEMS = 8;
// Rows of digits the pass does not have (d >= n_digits) hold zeros: their
// scan is zeros too and is skipped.
__global__ __launch_bounds__(RSCAN_THREADS) void k_radix_rowscan(uint32_t* __restrict__ hist,
                                                               const uint32_t* __restrict__ n_dev, int nb_max,
                                                               int n_digits, uint32_t* __restrict__ row_total) {
  __shared__ uint32_t s_w[RSCAN_THREADS / 64];
  if ((int)blockIdx.x >= n_digits) {
    if (threadIdx.x == 0) row_total[blockIdx.x] = 0u;
    return;
  }
  const uint32_t n = *n_dev;
  const uint32_t nb = (n + RS_TILE - 1) / RS_TILE;
  uint32_t* row = hist + (size_t)blockIdx.x * nb_max;
  const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
  uint32_t carry = 0;
  for (uint32_t j0 = 0; j0 < nb; j0 += RSCAN_THREADS * RSCAN_ITEMS) {
    const uint32_t j = j0 + threadIdx.x * RSCAN_ITEMS;
    uint32_t v[RSCAN_ITEMS];
    const bool full = j + RSCAN_ITEMS <= nb;
    if (full) {
      uint4 a = *reinterpret_cast<const uint4*>(row + j);
      uint4 c = *reinterpret_cast<const uint4*>(row + j + 4);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      v[4] = c.x; v[5] = c.y; v[6] = c.z; v[7] = c.w;
    } else {
#pragma unroll
      for (int k = 0; k < RSCAN_ITEMS; ++k) v[k] = j + k < nb ? row[j + k] : 0u;
    }
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < RSCAN_ITEMS; ++k) sum += v[k];
    uint32_t x = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t y = __shfl_up(x, d, 64);
      if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < RSCAN_THREADS / 64; ++k) {
      uint32_t t = s_w[k];
      if ((uint32_t)k < w) base += t;
      tot += t;
    }
    __syncthreads();  // s_w is rewritten by the next chunk
    uint32_t run = carry + base + x - sum;
#pragma unroll
    for (int k = 0; k < RSCAN_ITEMS; ++k) {
      uint32_t c = v[k];
      v[k] = run;
      run += c;
    }
    if (full) {
      *reinterpret_cast<uint4*>(row + j) = make_uint4(v[0], v[1], v[2], v[3]);
      *reinterpret_cast<uint4*>(row + j + 4) = make_uint4(v[4], v[5], v[6], v[7]);
    } else {
#pragma unroll
      for (int k = 0; k < RSCAN_ITEMS; ++k)
        if (j + k < nb) row[j + k] = v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) row_total[blockIdx.x] = carry;
}

// Stable scatter of one radix tile over the digit of `dbits` (<= 8) bits at
// `shift`. Keys are ranked with wave64 ballot matching, the tile is staged
// in LDS in digit order and each digit's run is written contiguously (keys
// first, then the values through the same LDS).
// LAST: the final pass of the tile sort; keys are not written, and the first
// and last position of every key segment inside the tile go to the ranges
// (atomicMin / atomicMax), which yields each key's global [start, end).
// D8: the digit is 8 bits wide (dbits is ignored; the matching loop is
// fully unrolled).
template <bool LAST, bool D8, class K>
__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(
    const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, K* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, const uint32_t* __restrict__ n_dev, int shift, int dbits,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ row_total, int nb_max, TileRanges tr) {
  if (D8) dbits = 8;
  const uint32_t mask = (1u << dbits) - 1u;
  __shared__ uint32_t s_cnt[RS_WAVES][256];
  __shared__ uint32_t s_blk_start[256];
  __shared__ uint32_t s_glob[256];
  __shared__ uint32_t s_stage[RS_TILE];
  __shared__ uint8_t s_dig[RS_TILE];  // digit of each staged slot, for the value pass
  __shared__ uint32_t s_tmp[8];

  const uint32_t n = *n_dev;
  const uint32_t nb = (n + RS_TILE - 1) / RS_TILE;
  const uint32_t b = xcd_chunked(blockIdx.x);
  if (b >= nb) return;
  const uint32_t t = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t w = t >> 6;
  const size_t base = (size_t)b * RS_TILE;
  const uint32_t blk_n = (uint32_t)min((size_t)RS_TILE, (size_t)n - base);

#pragma unroll
  for (int k = 0; k < RS_WAVES; ++k) s_cnt[k][t] = 0;
  // global base of digit t: exclusive scan of row totals + this block's prefix
  uint32_t dig_off = block_excl_scan_256(row_total[t], s_tmp, nullptr);
  s_glob[t] = dig_off + hist[(size_t)t * nb_max + b];
  __syncthreads();

  uint32_t key[RS_ITEMS], val[RS_ITEMS], pos[RS_ITEMS];
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // items of this wave: wave-local index r * 64 + lane < w_n
  const uint32_t w_lo = w * RS_WAVE_ITEMS;
  const uint32_t w_n = blk_n > w_lo ? min(blk_n - w_lo, (uint32_t)RS_WAVE_ITEMS) : 0u;
  const K* kin = keys_in + base + w_lo + lane;
  const uint32_t* vin = vals_in + base + w_lo + lane;
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    bool valid = (uint32_t)r * 64 + lane < w_n;
    key[r] = valid ? (uint32_t)kin[r * 64] : 0xffffffffu;
  }
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    bool valid = (uint32_t)r * 64 + lane < w_n;
    uint32_t d = (key[r] >> shift) & mask;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      if (!D8 && bit >= dbits) break;  // uniform
      uint64_t bb = __ballot((d >> bit) & 1u);
      peers &= ((d >> bit) & 1u) ? bb : ~bb;
    }
    uint32_t lrank = (uint32_t)__popcll(peers & lt_mask);
    uint32_t before = 0;
    if (valid) before = s_cnt[w][d];
    __builtin_amdgcn_wave_barrier();
    if (valid && lrank == 0) s_cnt[w][d] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    pos[r] = before + lrank;
  }
  // values are loaded only now (in flight across the block scan) to keep
  // the register footprint of the ranking loop small
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    bool valid = (uint32_t)r * 64 + lane < w_n;
    val[r] = valid ? vin[r * 64] : 0u;
  }
  __syncthreads();
  // per-digit: offsets of each wave's run inside the block, and block start
  uint32_t c0 = s_cnt[0][t], c1 = s_cnt[1][t], c2 = s_cnt[2][t], c3 = s_cnt[3][t];
  uint32_t tot = c0 + c1 + c2 + c3;
  uint32_t start = block_excl_scan_256(tot, s_tmp, nullptr);
  s_blk_start[t] = start;
  __syncthreads();
  s_cnt[0][t] = start;
  s_cnt[1][t] = start + c0;
  s_cnt[2][t] = start + c0 + c1;
  s_cnt[3][t] = start + c0 + c1 + c2;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    // recompute the digit here instead of keeping 16 LDS addresses live
    // across the scan (an empty asm makes the key opaque to CSE)
    asm volatile("" : "+v"(key[r]));
    if ((uint32_t)r * 64 + lane < w_n) {
      uint32_t d = (key[r] >> shift) & mask;
      pos[r] += s_cnt[w][d];
      s_stage[pos[r]] = key[r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    uint32_t p = (uint32_t)r * RS_THREADS + t;
    if (p < blk_n) {
      uint32_t k = s_stage[p];
      uint32_t d = (k >> shift) & mask;
      s_dig[p] = (uint8_t)d;
      uint32_t o = s_glob[d] + (p - s_blk_start[d]);
      if (!LAST) {
        keys_out[o] = (K)k;
      } else if (k < tr.n_keys) {
        // within a digit run the keys are sorted (lower digits were sorted
        // by the earlier passes), so equal keys form one segment here
        if (p == 0 || s_stage[p - 1] != k) atomicMin(&tr.rstart[k], o);
        if (p + 1 == blk_n || s_stage[p + 1] != k) atomicMax(&tr.rend[k], o + 1);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r)
    if ((uint32_t)r * 64 + lane < w_n) s_stage[pos[r]] = val[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    uint32_t p = (uint32_t)r * RS_THREADS + t;
    if (p < blk_n) {
      uint32_t d = s_dig[p];
      vals_out[s_glob[d] + (p - s_blk_start[d])] = s_stage[p];
    }
  }
}

// ------------------------------------------------- single-pass digit offsets
// The digit totals of every pass in one read of the keys (gtot[p * 256 + d],
// zero at launch): a block histogram in LDS per pass, then one global add per
// non-empty digit and block.
template <class K>
__global__ __launch_bounds__(256) void k_radix_gtot(const K* __restrict__ keys, const uint32_t* __restrict__ n_dev,
                                                    int bits, int db, uint32_t* __restrict__ gtot) {
  __shared__ uint32_t h[OS_PASSES][256];
#pragma unroll
  for (int p = 0; p < OS_PASSES; ++p) h[p][threadIdx.x] = 0;
  __syncthreads();
  const uint32_t n = *n_dev;
  const int passes = (bits + db - 1) / db;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const uint32_t k = keys[i];
    for (int p = 0; p < passes; ++p) {
      const int w = min(db, bits - p * db);
      atomicAdd(&h[p][(k >> (p * db)) & ((1u << w) - 1u)], 1u);
    }
  }
  __syncthreads();
  for (int p = 0; p < passes; ++p) {
    const uint32_t v = h[p][threadIdx.x];
    if (v) atomicAdd(&gtot[p * 256 + threadIdx.x], v);
  }
}

// One radix pass with the block's digit offsets from a decoupled look-back
// (instead of a histogram pass and a row scan): blocks take tiles in ticket
// order, rank their keys (as k_radix_scatter), publish each digit's count,
// then each thread (one digit) walks back over the earlier tiles until one
// that has published its inclusive prefix. A tile waits only for tiles
// whose workgroups already run and have published their counts before
// waiting, so the look-back ends; its wait is bounded all the same
// (OS_SPIN_LIMIT polls), past which it flags ST_SCAN_STALL and goes on with
// a wrong offset. The rest is k_radix_scatter's staging and writes, which
// keep the sort stable.
constexpr uint32_t OS_AGG = 1u << 30, OS_INC = 2u << 30, OS_VAL = OS_AGG - 1u;
constexpr uint32_t OS_SPIN_LIMIT = 1u << 16;
constexpr int OS_WIN = 8;
template <bool LAST, bool D8, class K>
__global__ __launch_bounds__(RS_THREADS) void k_radix_onesweep(
    const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, K* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, const uint32_t* __restrict__ n_dev, int shift, int dbits,
    const uint32_t* __restrict__ gtot, uint32_t* __restrict__ tw, uint32_t* __restrict__ status, TileRanges tr) {
  if (D8) dbits = 8;
  const uint32_t mask = (1u << dbits) - 1u;
  __shared__ uint32_t s_cnt[RS_WAVES][256];
  __shared__ uint32_t s_blk_start[256];
  __shared__ uint32_t s_glob[256];
  __shared__ uint32_t s_stage[RS_TILE];
  __shared__ uint8_t s_dig[RS_TILE];
  __shared__ uint32_t s_tmp[8];
  __shared__ uint32_t s_b;

  const uint32_t n = *n_dev;
  const uint32_t nb = (n + RS_TILE - 1) / RS_TILE;
  if (blockIdx.x >= nb) return;  // exactly nb workgroups take tickets 0..nb-1
  const uint32_t t = threadIdx.x;
  if (t == 0) s_b = atomicAdd(&tw[0], 1u);
#pragma unroll
  for (int k = 0; k < RS_WAVES; ++k) s_cnt[k][t] = 0;
  __syncthreads();
  const uint32_t b = s_b;
  const uint32_t lane = lane_id();
  const uint32_t w = t >> 6;
  const size_t base = (size_t)b * RS_TILE;
  const uint32_t blk_n = (uint32_t)min((size_t)RS_TILE, (size_t)n - base);

  uint32_t key[RS_ITEMS], val[RS_ITEMS], pos[RS_ITEMS];
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t w_lo = w * RS_WAVE_ITEMS;
  const uint32_t w_n = blk_n > w_lo ? min(blk_n - w_lo, (uint32_t)RS_WAVE_ITEMS) : 0u;
  const K* kin = keys_in + base + w_lo + lane;
  const uint32_t* vin = vals_in + base + w_lo + lane;
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    bool valid = (uint32_t)r * 64 + lane < w_n;
    key[r] = valid ? (uint32_t)kin[r * 64] : 0xffffffffu;
  }
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    bool valid = (uint32_t)r * 64 + lane < w_n;
    uint32_t d = (key[r] >> shift) & mask;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      if (!D8 && bit >= dbits) break;  // uniform
      uint64_t bb = __ballot((d >> bit) & 1u);
      peers &= ((d >> bit) & 1u) ? bb : ~bb;
    }
    uint32_t lrank = (uint32_t)__popcll(peers & lt_mask);
    uint32_t before = 0;
    if (valid) before = s_cnt[w][d];
    __builtin_amdgcn_wave_barrier();
    if (valid && lrank == 0) s_cnt[w][d] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    pos[r] = before + lrank;
  }
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    bool valid = (uint32_t)r * 64 + lane < w_n;
    val[r] = valid ? vin[r * 64] : 0u;
  }
  __syncthreads();
  uint32_t c0 = s_cnt[0][t], c1 = s_cnt[1][t], c2 = s_cnt[2][t], c3 = s_cnt[3][t];
  uint32_t tot = c0 + c1 + c2 + c3;
  // digit t of this tile: publish the count, look back for the tiles before
  uint32_t* words = tw + OS_HEAD;
  __hip_atomic_store(&words[(size_t)b * 256 + t], (b == 0 ? OS_INC : OS_AGG) | tot, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  uint32_t prefix = 0;
  if (b > 0) {
    // OS_WIN tiles' words in flight per step (tiles j, j - 1, ...): their
    // counts are added up to the nearest inclusive prefix, or up to the
    // first tile that has published nothing yet, from where the walk resumes
    int j = (int)b - 1;
    uint32_t polls = 0;
    while (j >= 0) {
      uint32_t wv[OS_WIN];
#pragma unroll
      for (int k = 0; k < OS_WIN; ++k)
        wv[k] = j - k >= 0 ? __hip_atomic_load(&words[(size_t)(j - k) * 256 + t], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT)
                           : OS_INC;
      int used = 0;
      bool done = false;
#pragma unroll
      for (int k = 0; k < OS_WIN; ++k) {
        const uint32_t fl = wv[k] & ~OS_VAL;
        if (done || fl == 0) break;
        prefix += wv[k] & OS_VAL;
        used = k + 1;
        done = fl == OS_INC;
      }
      if (done) break;
      j -= used;
      if (used == 0) {
        if (++polls > OS_SPIN_LIMIT) {
          atomicOr(status, ST_SCAN_STALL);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __hip_atomic_store(&words[(size_t)b * 256 + t], OS_INC | ((prefix + tot) & OS_VAL), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  // global start of digit t: the totals of the digits below it, then the
  // tiles before this one
  const uint32_t dig_off = block_excl_scan_256(t <= mask ? gtot[t] : 0u, s_tmp, nullptr);
  s_glob[t] = dig_off + prefix;
  uint32_t start = block_excl_scan_256(tot, s_tmp, nullptr);
  s_blk_start[t] = start;
  __syncthreads();
  s_cnt[0][t] = start;
  s_cnt[1][t] = start + c0;
  s_cnt[2][t] = start + c0 + c1;
  s_cnt[3][t] = start + c0 + c1 + c2;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    asm volatile("" : "+v"(key[r]));
    if ((uint32_t)r * 64 + lane < w_n) {
      uint32_t d = (key[r] >> shift) & mask;
      pos[r] += s_cnt[w][d];
      s_stage[pos[r]] = key[r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    uint32_t p = (uint32_t)r * RS_THREADS + t;
    if (p < blk_n) {
      uint32_t k = s_stage[p];
      uint32_t d = (k >> shift) & mask;
      s_dig[p] = (uint8_t)d;
      uint32_t o = s_glob[d] + (p - s_blk_start[d]);
      if (!LAST) {
        keys_out[o] = (K)k;
      } else if (k < tr.n_keys) {
        if (p == 0 || s_stage[p - 1] != k) atomicMin(&tr.rstart[k], o);
        if (p + 1 == blk_n || s_stage[p + 1] != k) atomicMax(&tr.rend[k], o + 1);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r)
    if ((uint32_t)r * 64 + lane < w_n) s_stage[pos[r]] = val[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    uint32_t p = (uint32_t)r * RS_THREADS + t;
    if (p < blk_n) {
      uint32_t d = s_dig[p];
      vals_out[s_glob[d] + (p - s_blk_start[d])] = s_stage[p];
    }
  }
}

// The sort's passes with single-pass digit offsets (rs.os set).
template <class K>
static int radix_sort_onesweep(K* keys[2], uint32_t* vals[2], const uint32_t* n_dev, size_t n_max, int bits,
                               RadixScratch rs, hipStream_t st, const TileRanges* tr) {
  const int nb = radix_nb(n_max);
  const int db = radix_digit_bits(bits);
  const int passes = (bits + db - 1) / db;
  // the totals and the tile words of the passes, zeroed in one launch
  // (unless the caller cleared them)
  if (!rs.os_zeroed) launch_fill_u32(rs.os, radix_os_used_words(nb, passes), 0u, st);
  if (tr && !tr->prefilled) {
    launch_fill_u32(tr->rstart, tr->n_keys, 0xffffffffu, st);
    launch_fill_u32(tr->rend, tr->n_keys, 0u, st);
  }
  // a modest grid: every workgroup adds its non-empty digit counts to the
  // totals with global atomics
  hipLaunchKernelGGL(k_radix_gtot<K>, dim3((unsigned)std::min(nb, 128)), dim3(256), 0, st, keys[0], n_dev, bits, db,
                     rs.os);
  int cur = 0;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * db;
    const int w = std::min(db, bits - shift);
    const bool last = p + 1 == passes;
    const uint32_t* gt = rs.os + (size_t)256 * p;
    uint32_t* tw = rs.os + (size_t)256 * OS_PASSES + (size_t)p * radix_os_pass_words(nb);
    const TileRanges none{nullptr, nullptr, 0u};
    const dim3 grid((unsigned)nb);
    if (last && tr && w == 8)
      hipLaunchKernelGGL((k_radix_onesweep<true, true, K>), grid, dim3(RS_THREADS), 0, st, keys[cur], vals[cur],
                         keys[cur ^ 1], vals[cur ^ 1], n_dev, shift, w, gt, tw, rs.status, *tr);
    else if (last && tr)
      hipLaunchKernelGGL((k_radix_onesweep<true, false, K>), grid, dim3(RS_THREADS), 0, st, keys[cur], vals[cur],
                         keys[cur ^ 1], vals[cur ^ 1], n_dev, shift, w, gt, tw, rs.status, *tr);
    else if (w == 8)
      hipLaunchKernelGGL((k_radix_onesweep<false, true, K>), grid, dim3(RS_THREADS), 0, st, keys[cur], vals[cur],
                         keys[cur ^ 1], vals[cur ^ 1], n_dev, shift, w, gt, tw, rs.status, none);
    else
      hipLaunchKernelGGL((k_radix_onesweep<false, false, K>), grid, dim3(RS_THREADS), 0, st, keys[cur], vals[cur],
                         keys[cur ^ 1], vals[cur ^ 1], n_dev, shift, w, gt, tw, rs.status, none);
    cur ^= 1;
  }
  return cur;
}

template <class K>
static int radix_sort_impl(K* keys[2], uint32_t* vals[2], const uint32_t* n_dev, size_t n_max, int bits,
                           RadixScratch rs, hipStream_t st, bool first_hist_done, const TileRanges* tr) {
  int cur = 0;
  int nb = radix_nb(n_max);
  if (nb == 0) return 0;
  // single-pass offsets for the depth sorts only: over the tile sort's
  // thousand-odd tiles the look-back's round trips to device-coherent memory
  // cost more than the histogram and row-scan launches they replace
  if (rs.os && rs.status && !tr && sizeof(K) == 4 && n_max < ((size_t)1 << 30) &&
      radix_digit_bits(bits) * OS_PASSES >= bits)
    return radix_sort_onesweep(keys, vals, n_dev, n_max, bits, rs, st, tr);
  const unsigned grid = xcd_grid((size_t)nb);
  if (tr && !tr->prefilled) {
    launch_fill_u32(tr->rstart, tr->n_keys, 0xffffffffu, st);
    launch_fill_u32(tr->rend, tr->n_keys, 0u, st);
  }
  const int db = radix_digit_bits(bits);
  for (int shift = 0; shift < bits; shift += db) {
    const bool last = shift + db >= bits;
    const int w = std::min(db, bits - shift);
    if (!(shift == 0 && first_hist_done))
      hipLaunchKernelGGL(k_radix_hist<K>, dim3(grid), dim3(RS_THREADS), 0, st, keys[cur], n_dev, shift,
                         (1u << w) - 1u, rs.hist, rs.nb_max);
    hipLaunchKernelGGL(k_radix_rowscan, dim3(256), dim3(RSCAN_THREADS), 0, st, rs.hist, n_dev, rs.nb_max, 1 << w,
                       rs.row_total);
    const TileRanges none{nullptr, nullptr, 0u};
    if (last && tr && w == 8)
      hipLaunchKernelGGL((k_radix_scatter<true, true, K>), dim3(grid), dim3(RS_THREADS), 0, st, keys[cur], vals[cur],
                         keys[cur ^ 1], vals[cur ^ 1], n_dev, shift, w, rs.hist, rs.row_total, rs.nb_max, *tr);
    else if (last && tr)
      hipLaunchKernelGGL((k_radix_scatter<true, false, K>), dim3(grid), dim3(RS_THREADS), 0, st, keys[cur], vals[cur],
                         keys[cur ^ 1], vals[cur ^ 1], n_dev, shift, w, rs.hist, rs.row_total, rs.nb_max, *tr);
    else if (w == 8)
      hipLaunchKernelGGL((k_radix_scatter<false, true, K>), dim3(grid), dim3(RS_THREADS), 0, st, keys[cur], vals[cur],
                         keys[cur ^ 1], vals[cur ^ 1], n_dev, shift, w, rs.hist, rs.row_total, rs.nb_max, none);
    else
      hipLaunchKernelGGL((k_radix_scatter<false, false, K>), dim3(grid), dim3(RS_THREADS), 0, st, keys[cur], vals[cur],
                         keys[cur ^ 1], vals[cur ^ 1], n_dev, shift, w, rs.hist, rs.row_total, rs.nb_max, none);
    cur ^= 1;
  }
  return cur;
}

int radix_sort_pairs(uint32_t* keys[2], uint32_t* vals[2], const uint32_t* n_dev, size_t n_max, int bits,
                     RadixScratch rs, hipStream_t st) {
  return radix_sort_impl(keys, vals, n_dev, n_max, bits, rs, st, false, nullptr);
}
int radix_sort_tiles(uint32_t* keys[2], uint32_t* vals[2], const uint32_t* n_dev, size_t n_max, int bits,
                     RadixScratch rs, TileRanges tr, hipStream_t st) {
  if (tile_keys_16(bits)) {
    uint16_t* k16[2] = {reinterpret_cast<uint16_t*>(keys[0]), reinterpret_cast<uint16_t*>(keys[1])};
    return radix_sort_impl(k16, vals, n_dev, n_max, bits, rs, st, true, &tr);
  }
  return radix_sort_impl(keys, vals, n_dev, n_max, bits, rs, st, true, &tr);
}
int radix_sort_tiles_full(uint32_t* keys[2], uint32_t* vals[2], const uint32_t* n_dev, size_t n_max, int bits,
                          RadixScratch rs, TileRanges tr, hipStream_t st) {
  return radix_sort_impl(keys, vals, n_dev, n_max, bits, rs, st, false, &tr);
}

// ------------------------------------------------------------------- scans
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = SCAN_TILE / SCAN_THREADS;  // 8

template <class T>
__device__ __forceinline__ T block_excl_scan_generic(T v, T* s_tmp, T* total) {
  uint32_t lane = lane_id();
  uint32_t w = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d) x += y;
  }
  if (lane == 63) s_tmp[w] = x;
  __syncthreads();
  T base = 0, tot = 0;
  const int nw = blockDim.x >> 6;
  for (int k = 0; k < nw; ++k) {
    T t = s_tmp[k];
    if ((uint32_t)k < w) base += t;
    tot += t;
  }
  __syncthreads();
  if (total) *total = tot;
  return base + x - v;
}

struct GatherCnt {
  const uint32_t* vals;
  const uint32_t* cnt;
  __device__ uint32_t operator()(size_t r) const { return cnt[vals[r]]; }
};
template <class T>
struct Plain {
  const T* in;
  __device__ T operator()(size_t r) const { return in[r]; }
};

template <class T, class F>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(F in, size_t n, T* block_sums) {
  __shared__ T s_tmp[8];
  size_t base = (size_t)blockIdx.x * SCAN_TILE;
  T sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    size_t e = base + (size_t)k * SCAN_THREADS + threadIdx.x;
    if (e < n) sum += in(e);
  }
  T tot;
  block_excl_scan_generic<T>(sum, s_tmp, &tot);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// The total of a 32-bit scan saturates: a true total of 2^32 - 1 or more is
// written as SCAN_SATURATED (the offsets have wrapped; callers treat the scan
// as failed), so a count past the 32-bit range is never taken for a small one.
template <class T>
__global__ __launch_bounds__(1024) void k_scan_top(T* block_sums, int nb, T* total) {
  __shared__ T s_tmp[16];
  __shared__ unsigned long long s_wide;
  int per = (nb + 1023) / 1024;
  int beg = threadIdx.x * per;
  int end = min(beg + per, nb);
  T sum = 0;
  unsigned long long wide = 0;
  for (int j = beg; j < end; ++j) {
    sum += block_sums[j];
    wide += (unsigned long long)block_sums[j];
  }
  if (threadIdx.x == 0) s_wide = 0;
  __syncthreads();
  if (sizeof(T) == 4) atomicAdd(&s_wide, wide);
  T tot;
  T run = block_excl_scan_generic<T>(sum, s_tmp, &tot);
  for (int j = beg; j < end; ++j) {
    T c = block_sums[j];
    block_sums[j] = run;
    run += c;
  }
  if (threadIdx.x == 0) {
    if (sizeof(T) == 4 && s_wide >= (unsigned long long)SCAN_SATURATED) tot = (T)SCAN_SATURATED;
    *total = tot;
  }
}

template <class T, class F>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(F in, size_t n, const T* block_sums, T* out) {
  __shared__ T s_tmp[8];
  size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  T v[SCAN_ITEMS];
  T sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    size_t e = base + k;
    v[k] = e < n ? in(e) : (T)0;
    sum += v[k];
  }
  T run = block_excl_scan_generic<T>(sum, s_tmp, nullptr) + block_sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    size_t e = base + k;
    if (e < n) out[e] = run;
    run += v[k];
  }
}

template <class T>
__global__ void k_zero(T* p) {
  *p = 0;
}

template <class T, class F>
static void scan_impl(F in, size_t n, T* out, T* block_sums, T* total, hipStream_t st) {
  int nb = scan_nb(n);
  if (nb == 0) {
    hipLaunchKernelGGL((k_zero<T>), dim3(1), dim3(1), 0, st, total);
    return;
  }
  hipLaunchKernelGGL((k_scan_reduce<T, F>), dim3(nb), dim3(SCAN_THREADS), 0, st, in, n, block_sums);
  hipLaunchKernelGGL((k_scan_top<T>), dim3(1), dim3(1024), 0, st, block_sums, nb, total);
  hipLaunchKernelGGL((k_scan_down<T, F>), dim3(nb), dim3(SCAN_THREADS), 0, st, in, n, block_sums, out);
}

void scan_gathered_counts(const uint32_t* vals, const uint32_t* cnt, size_t n, uint32_t* out,
                          uint32_t* block_sums, uint32_t* total, hipStream_t st) {
  scan_impl<uint32_t>(GatherCnt{vals, cnt}, n, out, block_sums, total, st);
}
// Number of flagged bins in the bin rect of splat vi.
__device__ __forceinline__ uint32_t flagged_bins(uint32_t vi, const uint32_t* cnt, const uint2* rect, int cap,
                                                 const CamDev* cams, const uint8_t* flag, bool any) {
  if (!cnt[vi]) return 0u;
  const CamDev& cm = cams[vi / (uint32_t)cap];
  const uint2 rr = bin_rect(rect[vi], cm.bin_shift);
  const uint32_t x0 = rr.x & 0xffffu, y0 = rr.x >> 16, x1 = rr.y & 0xffffu, y1 = rr.y >> 16;
  uint32_t k = 0;
  for (uint32_t by = y0; by < y1; ++by) {
    const uint8_t* row = flag + cm.bin_base + (size_t)by * cm.nbx;
    for (uint32_t bx = x0; bx < x1; ++bx) {
      k += row[bx];
      if (any && k) return 1u;
    }
  }
  return k;
}
// Back entries (first *n_dev) whose rect meets a flagged bin.
struct TouchFlagged {
  const uint32_t* vals;
  const uint32_t* n_dev;
  const uint32_t* n_flagged;
  const uint32_t* cnt;
  const uint2* rect;
  int cap;
  const CamDev* cams;
  const uint8_t* flag;
  __device__ uint32_t operator()(size_t r) const {
    if (*n_flagged == 0 || r >= *n_dev) return 0u;
    return flagged_bins(vals[r], cnt, rect, cap, cams, flag, true);
  }
};

void scan_touch_flagged(const uint32_t* vals, size_t n, const uint32_t* n_dev, const uint32_t* n_flagged,
                        const SplatArrays S, int cap, const CamDev* cams, const uint8_t* flag, uint32_t* out,
                        uint32_t* block_sums, uint32_t* total, hipStream_t st) {
  scan_impl<uint32_t>(TouchFlagged{vals, n_dev, n_flagged, S.cnt, S.rect, cap, cams, flag}, n, out, block_sums,
                      total, st);
}

void scan_u32(const uint32_t* in, size_t n, uint32_t* out, uint32_t* block_sums, uint32_t* total, hipStream_t st) {
  scan_impl<uint32_t>(Plain<uint32_t>{in}, n, out, block_sums, total, st);
}
void scan_u64(const uint64_t* in, size_t n, uint64_t* out, uint64_t* block_sums, uint64_t* total, hipStream_t st) {
  scan_impl<uint64_t>(Plain<uint64_t>{in}, n, out, block_sums, total, st);
}

// ------------------------------------------- single-pass scans (look-back)
// An exclusive scan in one launch: each workgroup takes the next tile of
// LB_TILE entries from a ticket counter, publishes its tile's sum, then one
// wave looks back over the tiles before it (256 per round, four per lane)
// for the sum of everything in front of it, publishes its inclusive prefix
// and the workgroup writes its entries. Tiles are numbered in the order
// workgroups start, so a tile only waits for tiles whose workgroups are
// already running, and every running workgroup publishes its own sum before
// it waits for anything: the look-back always ends. Its wait is bounded all
// the same (LB_SPIN_LIMIT polls); past the bound the workgroup flags
// ST_SCAN_STALL in the status word and finishes with a wrong prefix, and the
// caller turns the flag into an error. Only the workgroups whose tiles hold
// entries take a ticket. Tile words: bits 62-63 the flag (0: nothing yet,
// LB_AGG: the tile's own sum, LB_INC: the inclusive prefix through the
// tile), bits 0-61 the value. The state (LB_HEAD words, the first holding
// the ticket, then a word per tile) must be zero when the launch starts.
constexpr uint64_t LB_AGG = 1ull << 62, LB_INC = 2ull << 62, LB_VAL = (1ull << 62) - 1;
constexpr uint32_t LB_SPIN_LIMIT = 1u << 16;
constexpr int LB_WIN = 4;  // look-back words per lane and round

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// The sum of the tiles before `tile` (tw: the tile words), by one whole wave.
__device__ __forceinline__ uint64_t look_back(uint64_t* tw, int tile, uint32_t* status) {
  const int lane = (int)(threadIdx.x & 63);
  uint64_t prefix = 0;
  int hi = tile - 1;  // the nearest tile not yet added
  uint32_t polls = 0;
  while (hi >= 0) {
    // word j of this lane: tile hi - (64 j + lane), i.e. distance 64 j + lane
    uint64_t w[LB_WIN];
#pragma unroll
    for (int j = 0; j < LB_WIN; ++j) {
      const int idx = hi - (64 * j + lane);
      w[j] = idx >= 0 ? __hip_atomic_load(&tw[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : LB_INC;
    }
    // the nearest inclusive prefix (word jf, lane kf); everything nearer it
    // must have published at least its own sum
    int jf = LB_WIN, kf = 63;
    bool missing = false;
#pragma unroll
    for (int j = 0; j < LB_WIN; ++j) {
      if (jf < LB_WIN) break;
      const uint64_t fl = w[j] & ~LB_VAL;
      const uint64_t inc = __ballot(fl == LB_INC);
      const uint64_t none = __ballot(fl == 0);
      const int k = inc ? __builtin_ctzll(inc) : 63;
      const uint64_t upto = k == 63 ? ~0ull : ((2ull << k) - 1ull);
      if (none & upto) {
        missing = true;
        break;
      }
      if (inc) jf = j, kf = k;
    }
    if (missing) {
      if (++polls > LB_SPIN_LIMIT) {
        if (lane == 0) atomicOr(status, ST_SCAN_STALL);
        return prefix;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint64_t x = 0;
#pragma unroll
    for (int j = 0; j < LB_WIN; ++j)
      if (j < jf || (j == jf && lane <= kf)) x += w[j] & LB_VAL;
    prefix += wave_sum_u64(x);
    if (jf < LB_WIN) break;
    hi -= 64 * LB_WIN;
  }
  return prefix;
}

// The entries' counts and what is written for them. F::Item holds what an
// entry's output needs; F::count gives its count; F::put writes entry e
// given the exclusive prefix of the counts; F::last gets the total. With
// F::kTally, F::tally(item) is also summed over the entries into
// *F::tally_out (one atomic per workgroup). TILE:
// entries per workgroup (a multiple of SCAN_THREADS; fewer, larger tiles
// shorten the look-back, smaller ones need fewer registers).
template <class F, int TILE>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_lookback(F f, size_t n_max, const uint32_t* __restrict__ n_dev,
                                                                uint64_t* __restrict__ state,
                                                                uint32_t* __restrict__ total,
                                                                uint32_t* __restrict__ status) {
  __shared__ uint64_t s_tmp[8];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const size_t n = n_dev ? min((size_t)*n_dev, n_max) : n_max;
  constexpr int LB_ITEMS = TILE / SCAN_THREADS;
  const uint32_t needed = max((uint32_t)((n + TILE - 1) / TILE), 1u);
  if (blockIdx.x >= needed) return;  // exactly `needed` workgroups take tickets 0..needed-1
  if (threadIdx.x == 0) s_tile = atomicAdd(reinterpret_cast<uint32_t*>(state), 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  uint64_t* tw = state + LB_HEAD;
  const size_t base = (size_t)tile * TILE;
  typename F::Item it[LB_ITEMS];
  // thread t holds entries base + k * SCAN_THREADS + t (row k): coalesced
  // loads and stores; entries past n hold the item of count 0. A full tile
  // loads unguarded, so that all its loads are in flight together.
  if (base + TILE <= n) {
#pragma unroll
    for (int k = 0; k < LB_ITEMS; ++k) it[k] = f.item(base + (size_t)k * SCAN_THREADS + threadIdx.x);
  } else {
#pragma unroll
    for (int k = 0; k < LB_ITEMS; ++k) {
      const size_t e = base + (size_t)k * SCAN_THREADS + threadIdx.x;
      it[k] = e < n ? f.item(e) : F::empty();
    }
  }
  // exclusive prefix within the tile: within the row's wave, then the row's
  // earlier waves, then the earlier rows
  uint32_t rank[LB_ITEMS];
#pragma unroll
  for (int k = 0; k < LB_ITEMS; ++k) {
    const uint32_t c = f.count(it[k]);
    uint32_t x = c;
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    rank[k] = x - c;  // exclusive within the wave
  }
  // per (row, wave) totals -> offsets
  __shared__ uint32_t s_wt[LB_ITEMS][SCAN_THREADS / 64];
  if ((threadIdx.x & 63) == 63) {
#pragma unroll
    for (int k = 0; k < LB_ITEMS; ++k) s_wt[k][threadIdx.x >> 6] = rank[k] + f.count(it[k]);
  }
  __syncthreads();
  uint64_t tsum = 0;
  {
    const int w = (int)(threadIdx.x >> 6);
    uint64_t run = 0;
#pragma unroll
    for (int k = 0; k < LB_ITEMS; ++k) {
      uint64_t before = 0, row = 0;
#pragma unroll
      for (int u = 0; u < SCAN_THREADS / 64; ++u) {
        const uint32_t t = s_wt[k][u];
        if (u < w) before += t;
        row += t;
      }
      rank[k] = (uint32_t)(run + before) + rank[k];  // exclusive within the tile (< 2^32 per tile)
      run += row;
    }
    tsum = run;
  }
  if (threadIdx.x < 64) {
    if (threadIdx.x == 0)
      __hip_atomic_store(&tw[tile], (tile == 0 ? LB_INC : LB_AGG) | (tsum & LB_VAL), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    uint64_t prefix = 0;
    if (tile > 0) {
      prefix = look_back(tw, (int)tile, status);
      if (threadIdx.x == 0)
        __hip_atomic_store(&tw[tile], LB_INC | ((prefix + tsum) & LB_VAL), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x == 0) s_prefix = prefix;
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
#pragma unroll
  for (int k = 0; k < LB_ITEMS; ++k) {
    const size_t e = base + (size_t)k * SCAN_THREADS + threadIdx.x;
    if (e < n) f.put(e, it[k], prefix + rank[k]);
  }
  if constexpr (F::kTally) {
    uint32_t t = 0;
#pragma unroll
    for (int k = 0; k < LB_ITEMS; ++k) t += f.tally(it[k]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o, 64);
    __shared__ uint32_t s_tally[SCAN_THREADS / 64];
    if ((threadIdx.x & 63) == 0) s_tally[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t sum = 0;
#pragma unroll
      for (int w = 0; w < SCAN_THREADS / 64; ++w) sum += s_tally[w];
      if (sum) atomicAdd(f.tally_out, sum);
    }
  }
  if (threadIdx.x == 0 && tile == needed - 1) {
    const uint64_t t = prefix + tsum;
    *total = t >= (uint64_t)SCAN_SATURATED ? SCAN_SATURATED : (uint32_t)t;
    f.last(t);
  }
}

// The depth-cut partition: entry r (key, value) goes to the front arrays at
// its exclusive count of front entries, or (with back arrays) to the back
// arrays at r minus that count; the totals and the cut switch go to n_back
// and cut_on, the number of visible entries behind the cut (culled entries
// carry the key 0xffffffff) to tally_out.
struct PartitionAtCut {
  static constexpr bool kTally = true;
  const uint32_t* keys;
  const uint32_t* vals;
  const CutState* cs;
  uint32_t *kf, *vf, *kb, *vb;
  size_t n;
  uint32_t* n_back;
  uint32_t* cut_on;
  uint32_t* tally_out;
  struct Item {
    uint32_t k, v;
  };
  __device__ Item item(size_t r) const { return {keys[r], vals[r]}; }
  __device__ static Item empty() { return {0xffffffffu, 0u}; }  // at or past any cut
  __device__ uint32_t count(const Item& it) const { return it.k < cs->key ? 1u : 0u; }
  __device__ uint32_t tally(const Item& it) const { return (it.k >= cs->key && it.k != 0xffffffffu) ? 1u : 0u; }
  __device__ void put(size_t r, const Item& it, uint64_t o) const {
    if (it.k < cs->key) {
      kf[o] = it.k;
      vf[o] = it.v;
    } else if (kb) {
      kb[r - o] = it.k;
      vb[r - o] = it.v;
    }
  }
  __device__ void last(uint64_t t) const {
    *n_back = (uint32_t)(n - t);
    *cut_on = cs->key != 0xffffffffu ? 1u : 0u;
  }
};
// Pair offsets: the exclusive scan of cnt[vals[r]] over the first *n_dev
// sorted entries; with first_splat set, also the entry holding the first
// pair of every RS_TILE-pair block below k_cap (entry r's pairs are
// [o, o + c)), which pair generation starts from.
struct PairOffsets {
  static constexpr bool kTally = false;
  const uint32_t* vals;
  const uint32_t* cnt;
  uint32_t* out;
  uint32_t* first_splat;
  size_t k_cap;
  struct Item {
    uint32_t c;
  };
  __device__ Item item(size_t r) const { return {cnt[vals[r]]}; }
  __device__ static Item empty() { return {0u}; }
  __device__ uint32_t count(const Item& it) const { return it.c; }
  __device__ void put(size_t r, const Item& it, uint64_t o) const {
    out[r] = (uint32_t)o;
    if (first_splat) {
      const uint64_t hi = min(o + it.c, (uint64_t)k_cap);
      for (uint64_t b = (o + RS_TILE - 1) / RS_TILE; b * RS_TILE < hi; ++b) first_splat[b] = (uint32_t)r;
    }
  }
  __device__ void last(uint64_t) const {}
  __device__ uint32_t tally(const Item&) const { return 0u; }
  uint32_t* tally_out = nullptr;
};

// The partition in two launches, with no look-back: each tile's count of
// front entries (and the visible back tally), then every workgroup sums the
// counts of the tiles before its own (a few KB, from the cache), ranks its
// entries by ballot and writes them. Over the thousand-odd tiles of a step a
// look-back's round trips to device-coherent memory cost more than the
// second read of the keys.
constexpr int PT_ITEMS = 16;
constexpr int PT_TILE = PT_ITEMS * SCAN_THREADS;  // 4096 entries per tile

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t x, uint32_t* s_tmp /* >= SCAN_THREADS / 64 */) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
  if ((threadIdx.x & 63) == 0) s_tmp[threadIdx.x >> 6] = x;
  __syncthreads();
  uint32_t r = 0;
#pragma unroll
  for (int w = 0; w < SCAN_THREADS / 64; ++w) r += s_tmp[w];
  __syncthreads();
  return r;
}

// The tile's keys (row k: base + k * SCAN_THREADS + t; ~0u past n), loaded
// all at once: a full tile unguarded, the last one guarded.
__device__ __forceinline__ void pt_load_keys(const uint32_t* __restrict__ keys, size_t n, size_t base,
                                             uint32_t (&key)[PT_ITEMS]) {
  const uint32_t* kp = keys + base + threadIdx.x;
  if (base + PT_TILE <= n) {
#pragma unroll
    for (int k = 0; k < PT_ITEMS; ++k) key[k] = kp[k * SCAN_THREADS];
  } else {
#pragma unroll
    for (int k = 0; k < PT_ITEMS; ++k)
      key[k] = base + (size_t)k * SCAN_THREADS + threadIdx.x < n ? kp[k * SCAN_THREADS] : 0xffffffffu;
  }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_part_count(const uint32_t* __restrict__ keys, size_t n,
                                                            const CutState* __restrict__ cs,
                                                            uint32_t* __restrict__ tcount,
                                                            uint32_t* __restrict__ tally_out) {
  __shared__ uint32_t s_tmp[SCAN_THREADS / 64];
  const uint32_t cut = cs->key;
  const size_t base = (size_t)blockIdx.x * PT_TILE;
  uint32_t key[PT_ITEMS];
  pt_load_keys(keys, n, base, key);
  uint32_t f = 0, vis = 0;
#pragma unroll
  for (int k = 0; k < PT_ITEMS; ++k) {
    f += key[k] < cut ? 1u : 0u;
    vis += (key[k] >= cut && key[k] != 0xffffffffu) ? 1u : 0u;
  }
  f = block_sum_u32(f, s_tmp);
  vis = block_sum_u32(vis, s_tmp);
  if (threadIdx.x == 0) {
    tcount[blockIdx.x] = f;
    if (vis) atomicAdd(tally_out, vis);
  }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_part_scatter(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, size_t n, const CutState* __restrict__ cs,
    const uint32_t* __restrict__ tcount, uint32_t* __restrict__ kf, uint32_t* __restrict__ vf,
    uint32_t* __restrict__ kb, uint32_t* __restrict__ vb, uint32_t* __restrict__ n_front,
    uint32_t* __restrict__ n_back, uint32_t* __restrict__ cut_on) {
  __shared__ uint32_t s_tmp[SCAN_THREADS / 64];
  __shared__ uint32_t s_wt[PT_ITEMS][SCAN_THREADS / 64];
  const uint32_t cut = cs->key;
  const uint32_t b = blockIdx.x;
  const size_t base = (size_t)b * PT_TILE;
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  // the front entries of the tiles before this one
  uint32_t pre = 0;
  for (uint32_t i = threadIdx.x; i < b; i += SCAN_THREADS) pre += tcount[i];
  pre = block_sum_u32(pre, s_tmp);
  // row k: entries base + k * SCAN_THREADS + t; rank of a front entry within
  // its wave by ballot, then the waves and rows before it
  uint32_t key[PT_ITEMS], rank[PT_ITEMS];
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  pt_load_keys(keys, n, base, key);
#pragma unroll
  for (int k = 0; k < PT_ITEMS; ++k) {
    const uint64_t bal = __builtin_amdgcn_ballot_w64(key[k] < cut);
    rank[k] = (uint32_t)__popcll(bal & lt);
    if (lane == 0) s_wt[k][w] = (uint32_t)__popcll(bal);
  }
  __syncthreads();
  uint32_t run = 0;
#pragma unroll
  for (int k = 0; k < PT_ITEMS; ++k) {
    uint32_t before = 0, row = 0;
#pragma unroll
    for (int u = 0; u < SCAN_THREADS / 64; ++u) {
      const uint32_t c = s_wt[k][u];
      if (u < w) before += c;
      row += c;
    }
    rank[k] += run + before;
    run += row;
  }
  // the values of the entries written, loaded together before the writes
  uint32_t val[PT_ITEMS];
#pragma unroll
  for (int k = 0; k < PT_ITEMS; ++k) {
    const size_t e = base + (size_t)k * SCAN_THREADS + threadIdx.x;
    val[k] = (e < n && (key[k] < cut || kb)) ? vals[e] : 0u;
  }
#pragma unroll
  for (int k = 0; k < PT_ITEMS; ++k) {
    const size_t e = base + (size_t)k * SCAN_THREADS + threadIdx.x;
    if (e >= n) continue;
    const uint32_t o = pre + rank[k];  // front entries before e
    if (key[k] < cut) {
      kf[o] = key[k];
      vf[o] = val[k];
    } else if (kb) {
      kb[e - o] = key[k];
      vb[e - o] = val[k];
    }
  }
  if (threadIdx.x == 0 && b + 1 == gridDim.x) {
    *n_front = pre + run;
    *n_back = (uint32_t)(n - (pre + run));
    *cut_on = cut != 0xffffffffu ? 1u : 0u;
  }
}

void partition_at_cut(const uint32_t* keys, const uint32_t* vals, size_t n, const CutState* cs, uint32_t* kf,
                      uint32_t* vf, uint32_t* kb, uint32_t* vb, uint32_t* n_front, uint32_t* n_back, uint32_t* cut_on,
                      uint32_t* n_back_visible, uint64_t* state, uint32_t* status, hipStream_t st) {
  (void)status;
  const int nb = std::max((int)((n + PT_TILE - 1) / PT_TILE), 1);
  // the tile counts in the look-back state's words (lookback_words(n) u64
  // words hold more than nb counts)
  uint32_t* tcount = reinterpret_cast<uint32_t*>(state);
  hipLaunchKernelGGL(k_part_count, dim3(nb), dim3(SCAN_THREADS), 0, st, keys, n, cs, tcount, n_back_visible);
  hipLaunchKernelGGL(k_part_scatter, dim3(nb), dim3(SCAN_THREADS), 0, st, keys, vals, n, cs, tcount, kf, vf, kb, vb,
                     n_front, n_back, cut_on);
}

void scan_pair_offsets(const uint32_t* vals, const uint32_t* cnt, size_t n_max, const uint32_t* n_dev,
                       uint32_t* out, uint32_t* total, uint64_t* state, uint32_t* status, hipStream_t st,
                       uint32_t* first_splat, size_t k_cap) {
  const int nb = std::max(lookback_tiles(n_max), 1);
  PairOffsets f{vals, cnt, out, first_splat, k_cap, nullptr};
  hipLaunchKernelGGL((k_scan_lookback<PairOffsets, LB_TILE>), dim3(nb), dim3(SCAN_THREADS), 0, st, f, n_max, n_dev, state,
                     total, status);
}

// ------------------------------------------------------------ pairs/ranges
// Pair generation. Block b writes pairs [b*RS_TILE, (b+1)*RS_TILE) of the
// depth-ordered pair sequence (a splat's pairs follow the reference's push
// order: tile rows, then columns, backward.rs:200-204), so the writes are
// coalesced, every block has the same amount of work, and the block's keys
// are exactly the radix sort's tile b, whose low-digit histogram it counts in
// LDS on the way. first_splat[b] (the splat holding pair b*RS_TILE) comes
// from k_block_first.
constexpr int GP_THREADS = 256;
constexpr int GP_PER = RS_TILE / GP_THREADS;  // window slots scanned per thread

__global__ __launch_bounds__(256) void k_block_first(const uint32_t* __restrict__ offsets, size_t n_sorted,
                                                     const uint32_t* __restrict__ total, size_t k_cap,
                                                     uint32_t* __restrict__ first_splat) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_sorted) return;
  const uint32_t tot = *total;
  uint64_t lo = offsets[r];
  uint64_t hi = r + 1 < n_sorted ? offsets[r + 1] : tot;
  hi = hi < k_cap ? hi : k_cap;
  // blocks whose first pair lies in [lo, hi)
  for (uint64_t b = (lo + RS_TILE - 1) / RS_TILE; b * RS_TILE < hi; ++b) first_splat[b] = (uint32_t)r;
}

template <class KT>
__global__ __launch_bounds__(GP_THREADS) void k_gen_pairs(const uint32_t* __restrict__ sorted_vals, size_t n_sorted,
                                                          const uint32_t* __restrict__ offsets,
                                                          const uint32_t* __restrict__ total,
                                                          const uint32_t* __restrict__ cnt,
                                                          const uint2* __restrict__ rect, int cap,
                                                          const CamDev* __restrict__ cams,
                                                          KT* __restrict__ pair_keys,
                                                          uint32_t* __restrict__ pair_vals, size_t k_cap,
                                                          uint32_t* __restrict__ k_out, uint32_t* __restrict__ status,
                                                          const uint32_t* __restrict__ first_splat,
                                                          uint32_t* __restrict__ hist, int nb_max,
                                                          const uint32_t* __restrict__ n_dev, uint32_t hmask) {
  __shared__ uint32_t s_off[GP_THREADS + 1];
  __shared__ uint32_t s_vi[GP_THREADS];
  __shared__ uint32_t s_org[GP_THREADS];  // tx0 | ty0 << 16
  __shared__ uint32_t s_w[GP_THREADS];    // rect width in tiles
  __shared__ uint32_t s_tb[GP_THREADS];   // bin base of the splat's view
  __shared__ uint32_t s_ntx[GP_THREADS];
  __shared__ float s_rw[GP_THREADS];       // 1 / rect width
  __shared__ uint32_t s_map[RS_TILE];     // window slot -> splat of the pair
  __shared__ uint32_t s_wmax[GP_THREADS / 64];
  __shared__ uint32_t s_h[256];
  const int t = threadIdx.x;
  const uint32_t tot = *total;
  const size_t K = (size_t)tot < k_cap ? (size_t)tot : k_cap;
  if (blockIdx.x == 0 && t == 0) {
    if ((size_t)tot > k_cap) atomicOr(status, ST_PAIR_OVERFLOW);
    *k_out = (uint32_t)K;
  }
  const size_t b = xcd_chunked(blockIdx.x);
  const size_t p_lo = b * RS_TILE;
  if (p_lo >= K) return;
  const size_t p_hi = p_lo + RS_TILE < K ? p_lo + RS_TILE : K;
  // sorted entries past *n_dev hold no valid splat (their offsets are total)
  if (n_dev && (size_t)*n_dev < n_sorted) n_sorted = *n_dev;
  s_h[t] = 0;
  size_t rc = first_splat[b];
  for (;;) {
    __syncthreads();
    const size_t r = rc + t;
    if (r < n_sorted) {
      uint32_t vi = sorted_vals[r];
      s_off[t] = offsets[r];
      s_vi[t] = vi;
      const CamDev& cm = cams[vi / (uint32_t)cap];
      const uint2 rr = bin_rect(rect[vi], cm.bin_shift);
      s_org[t] = rr.x;
      s_w[t] = cnt[vi] ? (rr.y & 0xffffu) - (rr.x & 0xffffu) : 1u;
      s_rw[t] = 1.0f / (float)s_w[t];
      s_tb[t] = (uint32_t)cm.bin_base;
      s_ntx[t] = (uint32_t)cm.nbx;
    } else {
      s_off[t] = tot;
    }
    if (t == 0) s_off[GP_THREADS] = (rc + GP_THREADS < n_sorted) ? offsets[rc + GP_THREADS] : tot;
    __syncthreads();
    const size_t c_lo = (size_t)s_off[0] > p_lo ? (size_t)s_off[0] : p_lo;
    const size_t c_end = (size_t)s_off[GP_THREADS];
    const size_t c_hi = c_end < p_hi ? c_end : p_hi;
    // the window's splat of every pair in [c_lo, c_hi) (at most RS_TILE):
    // each splat marks the slot of its first pair there (the splat holding
    // c_lo, the last to start at or before it, marks slot 0; of the splats
    // that start at one offset the last, which has the pairs, wins), and an
    // inclusive max-scan carries the marks over the following slots
    const uint32_t m = c_hi > c_lo ? (uint32_t)(c_hi - c_lo) : 0u;
    for (uint32_t k = (uint32_t)t; k < m; k += GP_THREADS) s_map[k] = 0u;
    __syncthreads();
    if ((size_t)s_off[t] < c_hi) {
      const size_t o = (size_t)s_off[t];
      atomicMax(&s_map[o > c_lo ? (uint32_t)(o - c_lo) : 0u], (uint32_t)t);
    }
    __syncthreads();
    {
      // thread t scans slots [GP_PER t, GP_PER (t + 1)) in registers, then
      // the threads' maxima are scanned across the block
      uint32_t v[GP_PER], run = 0;
#pragma unroll
      for (int k = 0; k < GP_PER; ++k) {
        const uint32_t slot = (uint32_t)t * GP_PER + k;
        v[k] = slot < m ? s_map[slot] : 0u;
        run = max(run, v[k]);
        v[k] = run;
      }
      uint32_t x = run;
      const int lane = t & 63;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
        if (lane >= d) x = max(x, y);
      }
      if (lane == 63) s_wmax[t >> 6] = x;
      __syncthreads();
      uint32_t before = (uint32_t)__shfl_up((int)x, 1, 64);
      if (lane == 0) before = 0u;
      for (int w = 0; w < (t >> 6); ++w) before = max(before, s_wmax[w]);
#pragma unroll
      for (int k = 0; k < GP_PER; ++k) {
        const uint32_t slot = (uint32_t)t * GP_PER + k;
        if (slot < m) s_map[slot] = max(before, v[k]);
      }
    }
    __syncthreads();
    for (uint32_t k = (uint32_t)t; k < m; k += GP_THREADS) {
      const size_t p = c_lo + k;
      const uint32_t j = s_map[k];
      const uint32_t local = (uint32_t)(p - s_off[j]);
      const uint32_t w = s_w[j];
      // local / w through the float reciprocal, corrected to the exact
      // quotient (local < 2^24: a bin rect holds fewer pairs)
      uint32_t q = (uint32_t)((float)local * s_rw[j]);
      if (q * w > local) --q;
      else if ((q + 1) * w <= local) ++q;
      const uint32_t org = s_org[j];
      const uint32_t ty = (org >> 16) + q;
      const uint32_t tx = (org & 0xffffu) + (local - q * w);
      const uint32_t key = s_tb[j] + ty * s_ntx[j] + tx;
      pair_keys[p] = (KT)key;
      pair_vals[p] = s_vi[j];
      atomicAdd(&s_h[key & hmask], 1u);
    }
    if (c_end >= p_hi) break;  // uniform: every thread read the same s_off
    rc += GP_THREADS;
  }
  __syncthreads();
  hist[(size_t)t * nb_max + b] = s_h[t];
}

void launch_gen_pairs(const uint32_t* sorted_vals, size_t n_sorted, const uint32_t* offsets, const uint32_t* total,
                      const SplatArrays S, int cap, const CamDev* cams, uint32_t* pair_keys, uint32_t* pair_vals,
                      size_t k_cap, uint32_t* k_out, uint32_t* status, uint32_t* first_splat, RadixScratch rs,
                      int key_bits, hipStream_t st, const uint32_t* n_dev) {
  // with n_dev the pair-offset scan has recorded the blocks' first entries
  // (it sees only the first *n_dev entries; the offsets past them are stale)
  if (n_sorted && !n_dev)
    hipLaunchKernelGGL(k_block_first, dim3((unsigned)((n_sorted + 255) / 256)), dim3(256), 0, st, offsets, n_sorted,
                       total, k_cap, first_splat);
  size_t nb = (size_t)radix_nb(k_cap);
  if (nb == 0) nb = 1;
  const uint32_t hmask = (1u << radix_digit_bits(key_bits)) - 1u;
  if (tile_keys_16(key_bits))
    hipLaunchKernelGGL(k_gen_pairs<uint16_t>, dim3(xcd_grid(nb)), dim3(GP_THREADS), 0, st, sorted_vals, n_sorted,
                       offsets, total, S.cnt, S.rect, cap, cams, reinterpret_cast<uint16_t*>(pair_keys), pair_vals,
                       k_cap, k_out, status, first_splat, rs.hist, rs.nb_max, n_dev, hmask);
  else
    hipLaunchKernelGGL(k_gen_pairs<uint32_t>, dim3(xcd_grid(nb)), dim3(GP_THREADS), 0, st, sorted_vals, n_sorted,
                       offsets, total, S.cnt, S.rect, cap, cams, pair_keys, pair_vals, k_cap, k_out, status,
                       first_splat, rs.hist, rs.nb_max, n_dev, hmask);
}

// ------------------------------------------------------- depth-cut binning
// Keeps the entries that the scan of 0/1 selections off[] (total *total)
// selected: entry r is kept iff off[r + 1] > off[r].
__global__ __launch_bounds__(256) void k_compact(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                 size_t n, const uint32_t* __restrict__ off,
                                                 const uint32_t* __restrict__ total, uint32_t* __restrict__ kc,
                                                 uint32_t* __restrict__ vc) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || *total == 0) return;
  const uint32_t o = off[r];
  const uint32_t o1 = r + 1 < n ? off[r + 1] : *total;
  if (o1 == o) return;
  kc[o] = keys[r];
  vc[o] = vals[r];
}

void launch_compact(const uint32_t* keys, const uint32_t* vals, size_t n, const uint32_t* off, const uint32_t* total,
                    uint32_t* kc, uint32_t* vc, hipStream_t st) {
  hipLaunchKernelGGL(k_compact, dim3((unsigned)(((n ? n : 1) + 255) / 256)), dim3(256), 0, st, keys, vals, n, off,
                     total, kc, vc);
}

__global__ __launch_bounds__(256) void k_back_count(const uint32_t* __restrict__ sorted_vals, size_t n,
                                                    const uint32_t* __restrict__ n_dev,
                                                    const uint32_t* __restrict__ n_flagged,
                                                    const uint32_t* __restrict__ cnt, const uint2* __restrict__ rect,
                                                    int cap, const CamDev* __restrict__ cams,
                                                    const uint8_t* __restrict__ flag, uint32_t* __restrict__ count) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint32_t k = 0;
  if (*n_flagged != 0 && r < *n_dev) k = flagged_bins(sorted_vals[r], cnt, rect, cap, cams, flag, false);
  count[r] = k;
}

void launch_back_count(const uint32_t* sorted_vals, size_t n, const uint32_t* n_dev, const uint32_t* n_flagged,
                       const SplatArrays S, int cap, const CamDev* cams, const uint8_t* flag, uint32_t* count,
                       hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_back_count, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sorted_vals, n, n_dev,
                     n_flagged, S.cnt, S.rect, cap, cams, flag, count);
}

// One thread per sorted back entry with back pairs: its flagged bins in
// row-major order at its offset.
__global__ __launch_bounds__(256) void k_back_gen(const uint32_t* __restrict__ sorted_vals, size_t n,
                                                  const uint32_t* __restrict__ n_dev,
                                                  const uint32_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ total,
                                                  const uint32_t* __restrict__ n_flagged,
                                                  const uint2* __restrict__ rect, int cap,
                                                  const CamDev* __restrict__ cams, const uint8_t* __restrict__ flag,
                                                  uint32_t* __restrict__ pair_keys, uint32_t* __restrict__ pair_vals,
                                                  size_t k_cap, uint32_t* __restrict__ k_out,
                                                  uint32_t* __restrict__ status) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t tot = *total;
  if (r == 0) {
    *k_out = (size_t)tot < k_cap ? tot : (uint32_t)k_cap;
    if ((size_t)tot > k_cap) atomicOr(status, ST_PAIR_OVERFLOW);
  }
  if (r >= n || *n_flagged == 0 || r >= *n_dev) return;
  size_t o = offsets[r];
  const size_t o_end = r + 1 < n ? offsets[r + 1] : tot;
  if (o == o_end) return;
  const uint32_t vi = sorted_vals[r];
  const CamDev& cm = cams[vi / (uint32_t)cap];
  const uint2 rr = bin_rect(rect[vi], cm.bin_shift);
  const uint32_t x0 = rr.x & 0xffffu, y0 = rr.x >> 16, x1 = rr.y & 0xffffu, y1 = rr.y >> 16;
  for (uint32_t ty = y0; ty < y1; ++ty) {
    const uint32_t rowbase = (uint32_t)cm.bin_base + ty * (uint32_t)cm.nbx;
    for (uint32_t tx = x0; tx < x1; ++tx) {
      if (!flag[rowbase + tx]) continue;
      if (o < k_cap) {
        pair_keys[o] = rowbase + tx;
        pair_vals[o] = vi;
      }
      ++o;
    }
  }
}

void launch_back_gen(const uint32_t* sorted_vals, size_t n, const uint32_t* n_dev, const uint32_t* offsets,
                     const uint32_t* total, const uint32_t* n_flagged, const SplatArrays S, int cap,
                     const CamDev* cams, const uint8_t* flag, uint32_t* pair_keys, uint32_t* pair_vals, size_t k_cap,
                     uint32_t* k_out, uint32_t* status, hipStream_t st) {
  hipLaunchKernelGGL(k_back_gen, dim3((unsigned)(((n ? n : 1) + 255) / 256)), dim3(256), 0, st, sorted_vals, n, n_dev,
                     offsets, total, n_flagged, S.rect, cap, cams, flag, pair_keys, pair_vals, k_cap, k_out, status);
}

__global__ void k_need_back(const uint32_t* n_flagged, uint32_t* status) {
  if (*n_flagged) atomicOr(status, ST_NEED_BACK);
}

void launch_need_back(const uint32_t* n_flagged, uint32_t* status, hipStream_t st) {
  hipLaunchKernelGGL(k_need_back, dim3(1), dim3(1), 0, st, n_flagged, status);
}

__global__ void k_cut_end(CutState* cs) {
  if (cs->mode == 0) cs->key = cs->zneed ? f2u(__uint_as_float(cs->zneed) * 1.05f) : 0xffffffffu;
  cs->zneed = 0;
}

void launch_cut_end(CutState* cs, hipStream_t st) { hipLaunchKernelGGL(k_cut_end, dim3(1), dim3(1), 0, st, cs); }

// k_need_back then k_cut_end, in one launch.
__global__ void k_need_back_cut_end(const uint32_t* n_flagged, uint32_t* status, CutState* cs) {
  if (*n_flagged) atomicOr(status, ST_NEED_BACK);
  if (cs->mode == 0) cs->key = cs->zneed ? f2u(__uint_as_float(cs->zneed) * 1.05f) : 0xffffffffu;
  cs->zneed = 0;
}

void launch_need_back_cut_end(const uint32_t* n_flagged, uint32_t* status, CutState* cs, hipStream_t st) {
  hipLaunchKernelGGL(k_need_back_cut_end, dim3(1), dim3(1), 0, st, n_flagged, status, cs);
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f32(float* p, size_t n, float v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// Streaming copy, one 16-byte element per thread (on MI355X this flat form
// streams faster than grid-stride loops of any unrolling: 6.2 against 4.5 to
// 5.2 TB/s for 1 GiB).
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
void launch_copy16(const void* a, void* b, size_t n16, hipStream_t st) {
  if (!n16) return;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)((n16 + 255) / 256)), dim3(256), 0, st, static_cast<const uint4*>(a),
                     static_cast<uint4*>(b), n16);
}

void launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

// Several fills in one launch (segments in turn; block-strided in each).
__global__ __launch_bounds__(256) void k_fill_multi(FillSegs f) {
  for (int s = 0; s < f.count; ++s) {
    uint32_t* const p = f.p[s];
    const uint32_t n = f.n[s], v = f.v[s];
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) p[i] = v;
  }
}

void launch_fill_multi(const FillSegs& f, hipStream_t st) {
  uint32_t most = 0;
  for (int s = 0; s < f.count; ++s) most = std::max(most, f.n[s]);
  if (!most) return;
  const unsigned grid = std::min(1024u, (most + 255u) / 256u);
  hipLaunchKernelGGL(k_fill_multi, dim3(grid), dim3(256), 0, st, f);
}
void launch_fill_f32(float* p, size_t n, float v, hipStream_t st) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

}  // namespace rgs

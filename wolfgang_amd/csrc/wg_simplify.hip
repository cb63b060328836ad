// Sort-reduce simplify on the device.
//
// Semantics (reference src/simplify.cpp:147-430): terms with equal packed keys
// (x, z) are summed in construction order, sums with |s| <= atol + rtol*max|c|
// are dropped, survivors come out in canonical lexicographic order over
// (x0, z0, x1, z1, ...) compared as unsigned words.
//
// Pipeline (all on one stream):
//   1. key statistics: bitwise OR / AND of every key word over all terms, so
//      that 8-bit digits that are constant across the operator are skipped;
//   2. sort records: one key word plus a uint32 construction index. Keys
//      that fit one word (x<<32|z when every mask fits in 32 bits) sort in
//      one LSD radix sort; wider keys sort word-major: all items by x0, then
//      only the items whose leading words tie by the next word (records
//      [run rank, word]), written back into the tied positions;
//   3. each LSD radix pass uses 8-bit digits (constant digits skipped) and is
//      histogram -> exclusive scan -> stable scatter. Stability within a
//      4096-key tile comes from wave-level digit matching (ballots), each
//      wave ranking a contiguous quarter of the tile, so equal keys keep
//      construction order; the tile is reordered in LDS and written out in
//      digit runs;
//   4. segment heads, one thread per segment summing its coefficients in
//      sorted (= construction) order, tolerance test with hypot;
//   5. survivor compaction by scan and a final gather into a new operator
//      (masks of up to 4 words first interleaved into one record per term).
//
// The accumulation seed follows the reference exactly: its hash path starts
// from the first coefficient and its sort path from 0.0, which differ only for
// a component whose every contribution is -0.0. The device sums from the first
// coefficient and, when the reference would have taken the sort path, adds
// +0.0 (which maps -0.0 to +0.0 and leaves every other value unchanged).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_set>
#include <vector>

#include "wg_internal.hpp"
#include "wg_product_term.hpp"
#include "wg_simplify.hpp"

namespace wg {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr std::size_t kTile = static_cast<std::size_t>(kBlock) * kItems;

// ---------------------------------------------------------------- key stats

__global__ __launch_bounds__(256) void key_stats_kernel(const u64* __restrict__ x,
                                                        const u64* __restrict__ z, std::size_t T,
                                                        std::size_t W,
                                                        unsigned long long* or_out,
                                                        unsigned long long* and_out) {
  // or_out / and_out hold 2W words: [x0, z0, x1, z1, ...].
  for (std::size_t k = 0; k < 2 * W; ++k) {
    const u64* src = (k & 1) ? z : x;
    const std::size_t w = k >> 1;
    u64 o = 0, a = ~0ULL;
    for (std::size_t t = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < T;
         t += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
      const u64 v = src[t * W + w];
      o |= v;
      a &= v;
    }
    for (int d = 32; d > 0; d >>= 1) {
      o |= __shfl_xor(o, d, 64);
      a &= __shfl_xor(a, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicOr(&or_out[k], static_cast<unsigned long long>(o));
      atomicAnd(&and_out[k], static_cast<unsigned long long>(a));
    }
  }
}

__global__ __launch_bounds__(256) void max_abs_kernel(const double2* __restrict__ c,
                                                      std::size_t T, double* block_max) {
  __shared__ double lds[4];
  double m = 0.0;
  for (std::size_t t = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < T;
       t += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    m = fmax(m, hypot(c[t].x, c[t].y));
  }
  for (int d = 32; d > 0; d >>= 1) m = fmax(m, __shfl_xor(m, d, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    block_max[blockIdx.x] = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
  }
}

// ---------------------------------------------------------------- records

__global__ __launch_bounds__(256) void build_records_kernel(const u64* __restrict__ x,
                                                            const u64* __restrict__ z,
                                                            std::size_t T, std::size_t W, int KW,
                                                            u64* __restrict__ keys,
                                                            u32* __restrict__ idx) {
  for (std::size_t t = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < T;
       t += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    if (W == 0) {
      keys[t] = 0;
    } else if (KW == 1) {
      keys[t] = (x[t] << 32) | z[t];
    } else {
      for (std::size_t w = 0; w < W; ++w) {
        keys[t * KW + 2 * w] = x[t * W + w];
        keys[t * KW + 2 * w + 1] = z[t * W + w];
      }
    }
    idx[t] = static_cast<u32>(t);
  }
}

// ---------------------------------------------------------------- radix pass

__global__ __launch_bounds__(kBlock) void radix_hist_kernel(const u64* __restrict__ keys,
                                                            std::size_t T, int KW, int word,
                                                            int shift, std::size_t nblocks,
                                                            u32* __restrict__ hist) {
  // One histogram per wave (fewer colliding LDS atomics), summed at the end.
  __shared__ u32 counts[kBlock / 64][256];
  static_assert(kBlock == 256 && kItems % 2 == 0, "histogram layout");
  const u32 wave = threadIdx.x >> 6;
  for (int w = 0; w < kBlock / 64; ++w) counts[w][threadIdx.x] = 0;
  __syncthreads();
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kTile;
  if (KW == 1 && base + kTile <= T && (reinterpret_cast<std::uintptr_t>(keys) & 15U) == 0) {
    // Full tile of one-word keys: two keys per 16-byte load.
    const ulonglong2* k2 = reinterpret_cast<const ulonglong2*>(keys + base);
    for (int r = 0; r < kItems / 2; ++r) {
      const ulonglong2 v = k2[r * kBlock + threadIdx.x];
      atomicAdd(&counts[wave][static_cast<u32>(v.x >> shift) & 0xFFU], 1U);
      atomicAdd(&counts[wave][static_cast<u32>(v.y >> shift) & 0xFFU], 1U);
    }
  } else {
    for (int r = 0; r < kItems; ++r) {
      const std::size_t i = base + static_cast<std::size_t>(r) * kBlock + threadIdx.x;
      if (i < T) {
        const u32 d = static_cast<u32>(keys[i * KW + word] >> shift) & 0xFFU;
        atomicAdd(&counts[wave][d], 1U);
      }
    }
  }
  __syncthreads();
  u32 c = 0;
  for (int w = 0; w < kBlock / 64; ++w) c += counts[w][threadIdx.x];
  hist[static_cast<std::size_t>(threadIdx.x) * nblocks + blockIdx.x] = c;
}

// Stable scatter of one 8-bit digit pass over a tile of kTile records.
// Wave w owns the tile's records [w * kTile / 4, (w + 1) * kTile / 4) in order
// (rows of 64), so the tile order is (wave, row, lane) and a wave ranks its
// records without block barriers: per row, the lanes holding the same digit
// are found with ballots and the wave's running count per digit lives in its
// own LDS row. The block then forms, per digit, the waves' offsets and the
// digit's start in the sorted tile, places every record at its sorted position
// in LDS, and writes the sorted tile out: consecutive threads write
// consecutive positions of a digit's run (coalesced), where scattering each
// record from its original slot would touch one cache line per record.
//
// With kLook (one-sweep passes, one-word keys) the block's tile is the next
// number from look.tile_counter and the digits' output offsets come from the
// upfront digit starts plus a decoupled look-back over the earlier tiles'
// published digit counts (LookBack below), instead of a per-pass histogram and
// scan (offsets).
struct LookBack {
  const u32* starts;  // the pass's 256 digit starts
  u32* status;        // [tiles][256] published counts: flag bits 30..31, value 0..29
  u32* tile_counter;
  u32* error;
};
constexpr u32 kLookAggregate = 1U << 30;  // value = this tile's digit count
constexpr u32 kLookInclusive = 2U << 30;  // value = digit count of tiles 0..this
constexpr u32 kLookValue = (1U << 30) - 1;
// Polls of a predecessor's status word before the pass gives up and reports
// an error (the wait ends on its own: every tile publishes its counts before
// it waits, and tiles are numbered in the order their blocks start).
constexpr u32 kLookSpins = 1U << 16;

template <int KW, bool kLook>
__global__ __launch_bounds__(kBlock) void radix_scatter_kernel(
    const u64* __restrict__ keys_in, const u32* __restrict__ idx_in, u64* __restrict__ keys_out,
    u32* __restrict__ idx_out, std::size_t T, int word, int shift, std::size_t nblocks,
    const u32* __restrict__ offsets, LookBack look) {
  constexpr int kWaves = kBlock / 64;
  constexpr int kRows = kTile / kBlock;
  static_assert(kBlock == 256, "one thread per digit");
  // (16-bit counts and tile positions keep the block under a third of the LDS,
  // three blocks per CU.)
  __shared__ unsigned short counts[kWaves][256];  // running counts per wave, then the waves' offsets
  __shared__ unsigned short tstart[256];          // digit's first position in the sorted tile
  __shared__ u32 gbase[256];           // digit's first output position for this tile
  __shared__ u32 wsum[kWaves];
  __shared__ u64 skeys[kTile * KW];
  __shared__ u32 sidx[kTile];
  const u32 tid = threadIdx.x, lane = tid & 63U, wave = tid >> 6;
  __shared__ u32 tile_s;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) counts[w][tid] = 0;
  if constexpr (kLook) {
    if (tid == 0) tile_s = atomicAdd(look.tile_counter, 1U);
  } else {
    gbase[tid] = offsets[static_cast<std::size_t>(tid) * nblocks + blockIdx.x];
    if (tid == 0) tile_s = blockIdx.x;
  }
  __syncthreads();
  const u32 tile_id = tile_s;
  const std::size_t tile = static_cast<std::size_t>(tile_id) * kTile;
  const u32 n_tile = static_cast<u32>(std::min<std::size_t>(kTile, T - tile));
  const u64 lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
  u64 key[kRows][KW];
  u32 idx[kRows], dig[kRows], rank[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const u32 j = wave * (kTile / kWaves) + static_cast<u32>(r) * 64U + lane;
    const std::size_t i = tile + j;
    if (j < n_tile) {
#pragma unroll
      for (int k = 0; k < KW; ++k) key[r][k] = keys_in[i * KW + k];
      idx[r] = idx_in[i];
    } else {
#pragma unroll
      for (int k = 0; k < KW; ++k) key[r][k] = 0;
      idx[r] = 0;
    }
  }
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const u32 j = wave * (kTile / kWaves) + static_cast<u32>(r) * 64U + lane;
    const bool valid = j < n_tile;
    // (word is 0 or KW - 1; a select keeps the key array in registers.)
    const u64 kword = word == 0 ? key[r][0] : key[r][KW - 1];
    const u32 d = static_cast<u32>(kword >> shift) & 0xFFU;
    u64 peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1U;
      const u64 m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    dig[r] = d;
    const u32 before = valid ? counts[wave][d] : 0U;
    rank[r] = before + static_cast<u32>(__popcll(peers & lt));
    if (valid && (peers & lt) == 0) {
      counts[wave][d] = static_cast<unsigned short>(before + static_cast<u32>(__popcll(peers)));
    }
  }
  __syncthreads();
  // Thread tid handles digit tid: the waves' offsets, then the tile's
  // exclusive scan over digits.
  u32 total = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const u32 c = counts[w][tid];
    counts[w][tid] = static_cast<unsigned short>(total);
    total += c;
  }
  u32* my_status = nullptr;
  if constexpr (kLook) {
    my_status = look.status + static_cast<std::size_t>(tile_id) * 256 + tid;
    __hip_atomic_store(my_status, (tile_id == 0 ? kLookInclusive : kLookAggregate) | total,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  u32 incl = total;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const u32 t = static_cast<u32>(__shfl_up(static_cast<int>(incl), off, 64));
    if (lane >= static_cast<u32>(off)) incl += t;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  u32 before_wave = 0;
  for (u32 w = 0; w < wave; ++w) before_wave += wsum[w];
  tstart[tid] = static_cast<unsigned short>(before_wave + incl - total);
  if constexpr (kLook) {
    // Digit tid's count in the tiles before this one: walk back until a tile
    // that has published its inclusive count.
    u32 prefix = 0;
    for (u32 t = tile_id; t-- > 0;) {
      const u32* st = look.status + static_cast<std::size_t>(t) * 256 + tid;
      u32 v = __hip_atomic_load(st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (u32 spin = 0; (v >> 30) == 0; ++spin) {
        if (spin == kLookSpins) {
          atomicOr(look.error, 1U);
          v = kLookInclusive;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        v = __hip_atomic_load(st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      prefix += v & kLookValue;
      if ((v >> 30) == 2) break;
    }
    if (tile_id > 0) {
      __hip_atomic_store(my_status, kLookInclusive | (prefix + total), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    gbase[tid] = look.starts[tid] + prefix;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const u32 j = wave * (kTile / kWaves) + static_cast<u32>(r) * 64U + lane;
    if (j < n_tile) {
      const u32 pos = tstart[dig[r]] + counts[wave][dig[r]] + rank[r];
#pragma unroll
      for (int k = 0; k < KW; ++k) skeys[pos * KW + k] = key[r][k];
      sidx[pos] = idx[r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const u32 j = static_cast<u32>(r) * kBlock + tid;
    if (j < n_tile) {
      const u32 d = static_cast<u32>(skeys[j * KW + word] >> shift) & 0xFFU;
      const std::size_t g = static_cast<std::size_t>(gbase[d]) + (j - tstart[d]);
#pragma unroll
      for (int k = 0; k < KW; ++k) keys_out[g * KW + k] = skeys[j * KW + k];
      idx_out[g] = sidx[j];
    }
  }
}

// Digit counts of every pass of a one-sweep sort in one read of the keys
// (counts do not depend on the order the passes leave the keys in):
// hist[p * 256 + d] for the digit at shifts.s[p].
struct DigitShifts {
  int s[8];
};

__global__ __launch_bounds__(256) void digit_hist_all_kernel(const u64* __restrict__ keys,
                                                             std::size_t n, DigitShifts shifts,
                                                             int passes, u32* __restrict__ hist) {
  // One histogram set per wave: the waves' atomics never collide (the biased
  // digits of product keys make one bin hot).
  __shared__ u32 h[4][8][256];
  const u32 wave = threadIdx.x >> 6;
  for (int w = 0; w < 4; ++w) {
    for (int p = 0; p < 8; ++p) h[w][p][threadIdx.x] = 0;
  }
  __syncthreads();
  // Four keys per thread per step, loaded before any of them is counted.
  const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
  std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    u64 k[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) k[q] = keys[i + q * stride];
    for (int p = 0; p < passes; ++p) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        atomicAdd(&h[wave][p][static_cast<u32>(k[q] >> shifts.s[p]) & 0xFFU], 1U);
      }
    }
  }
  for (; i < n; i += stride) {
    const u64 k = keys[i];
    for (int p = 0; p < passes; ++p) {
      atomicAdd(&h[wave][p][static_cast<u32>(k >> shifts.s[p]) & 0xFFU], 1U);
    }
  }
  __syncthreads();
  for (int p = 0; p < passes; ++p) {
    const u32 c = h[0][p][threadIdx.x] + h[1][p][threadIdx.x] + h[2][p][threadIdx.x] +
                  h[3][p][threadIdx.x];
    if (c) atomicAdd(&hist[p * 256 + threadIdx.x], c);
  }
}

// starts[p * 256 + d] = exclusive prefix over digits of hist[p * 256 + .]
// (one block per pass).
__global__ __launch_bounds__(256) void digit_starts_kernel(const u32* __restrict__ hist,
                                                           u32* __restrict__ starts) {
  __shared__ u32 wsum[4];
  const u32 tid = threadIdx.x, lane = tid & 63U, wave = tid >> 6;
  const u32 c = hist[blockIdx.x * 256 + tid];
  u32 incl = c;
  for (int off = 1; off < 64; off <<= 1) {
    const u32 t = static_cast<u32>(__shfl_up(static_cast<int>(incl), off, 64));
    if (lane >= static_cast<u32>(off)) incl += t;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  u32 before = 0;
  for (u32 w = 0; w < wave; ++w) before += wsum[w];
  starts[blockIdx.x * 256 + tid] = before + incl - c;
}

// Digit counts of the products' one-word keys without reading them: a
// product's key word is the XOR of its operands' key words (x0, or
// x0 << 32 | z0 when packed), so the count of products whose digit at shift
// s is b is sum_a L_s[a] R_s[a ^ b], a XOR convolution of the operands' digit
// counts (L * R < 2^32 products, so the sums fit).
__global__ __launch_bounds__(256) void operand_digit_hist_kernel(const u64* __restrict__ rec,
                                                                 u32 count, u32 RW, u32 W,
                                                                 int packed, DigitShifts shifts,
                                                                 int passes,
                                                                 u32* __restrict__ hist) {
  __shared__ u32 h[8][256];
  for (int q = 0; q < 8; ++q) h[q][threadIdx.x] = 0;
  __syncthreads();
  for (u32 r = blockIdx.x * blockDim.x + threadIdx.x; r < count; r += gridDim.x * blockDim.x) {
    const u64* e = rec + static_cast<std::size_t>(r) * RW;
    const u64 key = packed ? ((e[0] << 32) | e[W]) : e[0];
    for (int q = 0; q < passes; ++q) atomicAdd(&h[q][static_cast<u32>(key >> shifts.s[q]) & 0xFFU], 1U);
  }
  __syncthreads();
  for (int q = 0; q < passes; ++q) {
    if (h[q][threadIdx.x]) atomicAdd(&hist[q * 256 + threadIdx.x], h[q][threadIdx.x]);
  }
}

__global__ __launch_bounds__(256) void xor_convolve_kernel(const u32* __restrict__ lh,
                                                           const u32* __restrict__ rh,
                                                           u32* __restrict__ hist) {
  __shared__ u32 l[256], r[256];
  const u32 q = blockIdx.x, b = threadIdx.x;
  l[b] = lh[q * 256 + b];
  r[b] = rh[q * 256 + b];
  __syncthreads();
  u64 c = 0;
  for (u32 a = 0; a < 256; ++a) c += static_cast<u64>(l[a]) * r[a ^ b];
  hist[q * 256 + b] = static_cast<u32>(c);
}

__global__ __launch_bounds__(256) void clear_u32_kernel(u32* __restrict__ p, std::size_t n) {
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    p[i] = 0;
  }
}

// ---------------------------------------------------------------- segments

// Where the segments of equal keys lie in sorted order. With `start` the
// table: segment s is [start[s], start[s + 1]). Without it, the keys of the
// sorted order that repeat their predecessor are the m positions `nonheads`
// (ascending; m = 0: every key distinct): the head of segment s is the
// position s + c, c the number of non-heads before it, i.e. the number of j
// with nonheads[j] - j <= s, and its run goes on through the non-heads right
// after it.
struct SegMap {
  const u32* start;
  const u32* nonheads;
  u32 m;
};

__device__ inline void segment_bounds(const SegMap& map, std::size_t s, u32& b, u32& e) {
  if (map.start) {
    b = map.start[s];
    e = map.start[s + 1];
    return;
  }
  u32 lo = 0, hi = map.m;  // first j with nonheads[j] - j > s
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    if (map.nonheads[mid] - mid <= s) lo = mid + 1; else hi = mid;
  }
  b = static_cast<u32>(s) + lo;
  e = b + 1;
  for (u32 j = lo; j < map.m && map.nonheads[j] == e; ++j) ++e;
}

// Contiguous ranges of sorted one-word keys, one per block: head flags
// (head[i] = 1 where a run of equal keys starts), and per block the number of
// non-heads and of tied items (items whose run has two or more), for an
// ordered list of the tied items (tied_list_kernel).
__device__ inline bool key_tied(const u64* __restrict__ keys, std::size_t n, std::size_t i) {
  const u64 k = keys[i];
  return (i > 0 && keys[i - 1] == k) || (i + 1 < n && keys[i + 1] == k);
}

__global__ __launch_bounds__(256) void key_heads_kernel(const u64* __restrict__ keys,
                                                        std::size_t n, std::size_t chunk,
                                                        u32* __restrict__ head,
                                                        u32* __restrict__ block_nonheads,
                                                        u32* __restrict__ block_tied) {
  __shared__ u32 lds[2][4];
  u32 nonheads = 0, tied = 0;
  const std::size_t begin = static_cast<std::size_t>(blockIdx.x) * chunk;
  const std::size_t end = std::min(n, begin + chunk);
  for (std::size_t i = begin + threadIdx.x; i < end; i += blockDim.x) {
    const u64 k = keys[i];
    const bool h = i == 0 || keys[i - 1] != k;
    head[i] = h ? 1U : 0U;
    nonheads += h ? 0U : 1U;
    tied += (!h || (i + 1 < n && keys[i + 1] == k)) ? 1U : 0U;
  }
  for (int d = 32; d > 0; d >>= 1) {
    nonheads += static_cast<u32>(__shfl_xor(static_cast<int>(nonheads), d, 64));
    tied += static_cast<u32>(__shfl_xor(static_cast<int>(tied), d, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    lds[0][threadIdx.x >> 6] = nonheads;
    lds[1][threadIdx.x >> 6] = tied;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    block_nonheads[blockIdx.x] = lds[0][0] + lds[0][1] + lds[0][2] + lds[0][3];
    block_tied[blockIdx.x] = lds[1][0] + lds[1][1] + lds[1][2] + lds[1][3];
  }
}

// The tied items' positions in ascending order: block b writes its range's
// tied positions from offset[b] on, in order (ballots within each 256-item
// step, the waves' counts through LDS).
__global__ __launch_bounds__(256) void tied_list_kernel(const u64* __restrict__ keys,
                                                        std::size_t n, std::size_t chunk,
                                                        const u32* __restrict__ block_tied,
                                                        const u32* __restrict__ offset,
                                                        u32* __restrict__ list) {
  __shared__ u32 wcount[4];
  if (block_tied[blockIdx.x] == 0) return;  // (uniform: the whole block leaves)
  const u32 lane = threadIdx.x & 63U, wave = threadIdx.x >> 6;
  const u64 lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
  u32 out = offset[blockIdx.x];
  const std::size_t begin = static_cast<std::size_t>(blockIdx.x) * chunk;
  const std::size_t end = std::min(n, begin + chunk);
  for (std::size_t base = begin; base < end; base += blockDim.x) {
    const std::size_t i = base + threadIdx.x;
    const bool t = i < end && key_tied(keys, n, i);
    const u64 ball = __ballot(t);
    if (lane == 0) wcount[wave] = static_cast<u32>(__popcll(ball));
    __syncthreads();
    u32 before = 0;
    for (u32 w = 0; w < wave; ++w) before += wcount[w];
    if (t) list[out + before + static_cast<u32>(__popcll(ball & lt))] = static_cast<u32>(i);
    out += wcount[0] + wcount[1] + wcount[2] + wcount[3];
    __syncthreads();
  }
}

// First refinement level from the list of tied items: records [run rank, key
// word w] and the items' term indices (run rank = heads before the item in
// the list - 1; runs are whole in the list).
template <class Src>
__global__ __launch_bounds__(256) void list_records_kernel(
    const u32* __restrict__ list, std::size_t A, const u32* __restrict__ head,
    const u32* __restrict__ lhead_scan, const u32* __restrict__ idx, Src src, int w,
    u64* __restrict__ keys2, u32* __restrict__ idx2) {
  for (std::size_t j = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < A;
       j += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    const u32 pos = list[j];
    const u32 t = idx[pos];
    keys2[2 * j] = lhead_scan[j] + head[pos] - 1U;
    keys2[2 * j + 1] = src.word(t, w);
    idx2[j] = t;
  }
}

// Head flags of the listed items: list_head[j] = head[list[j]].
__global__ __launch_bounds__(256) void gather_heads_kernel(const u32* __restrict__ list,
                                                           std::size_t A,
                                                           const u32* __restrict__ head,
                                                           u32* __restrict__ out) {
  for (std::size_t j = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < A;
       j += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    out[j] = head[list[j]];
  }
}

// Full-key head flags of the listed (tied) items, after the refinement; the
// other positions' flags (heads) are final already. notes[j] = 1 when the
// item repeats its predecessor.
template <class Src>
__global__ __launch_bounds__(256) void list_full_heads_kernel(
    const u32* __restrict__ list, std::size_t A, const u64* __restrict__ k0,
    const u32* __restrict__ idx, Src src, int KW, u32* __restrict__ flags,
    u32* __restrict__ repeats) {
  for (std::size_t j = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < A;
       j += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    const u32 i = list[j];
    bool h = i == 0 || k0[i] != k0[i - 1];
    if (!h) {
      const u32 a = idx[i], b = idx[i - 1];
      for (int w = 1; w < KW && !h; ++w) h = src.word(a, w) != src.word(b, w);
    }
    flags[i] = h ? 1U : 0U;
    repeats[j] = h ? 0U : 1U;
  }
}

// nonheads[pos[j]] = list[j] for the listed items that repeat their
// predecessor (pos: exclusive scan of repeats).
__global__ __launch_bounds__(256) void compact_list_kernel(const u32* __restrict__ list,
                                                           std::size_t A,
                                                           const u32* __restrict__ repeats,
                                                           const u32* __restrict__ pos,
                                                           u32* __restrict__ nonheads) {
  for (std::size_t j = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < A;
       j += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    if (repeats[j]) nonheads[pos[j]] = list[j];
  }
}

__global__ __launch_bounds__(256) void segment_starts_kernel(const u32* __restrict__ flags,
                                                             const u32* __restrict__ seg_pos,
                                                             std::size_t T,
                                                             u32* __restrict__ seg_start) {
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < T;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    if (flags[i]) seg_start[seg_pos[i]] = static_cast<u32>(i);
    if (i == T - 1) seg_start[seg_pos[i] + flags[i]] = static_cast<u32>(T);
  }
}

__global__ __launch_bounds__(256) void segment_sum_kernel(
    SegMap map, std::size_t S, const u32* __restrict__ idx,
    const double2* __restrict__ c, int zero_seed, double threshold, double2* __restrict__ sums,
    u32* __restrict__ keep) {
  for (std::size_t s = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < S;
       s += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    u32 b, e;
    segment_bounds(map, s, b, e);
    double2 acc = c[idx[b]];
    for (u32 j = b + 1; j < e; ++j) {
      const double2 v = c[idx[j]];
      acc.x += v.x;
      acc.y += v.y;
    }
    if (zero_seed) {
      acc.x += 0.0;
      acc.y += 0.0;
    }
    sums[s] = acc;
    keep[s] = hypot(acc.x, acc.y) > threshold ? 1U : 0U;
  }
}

__global__ __launch_bounds__(256) void compact_kernel(
    SegMap map, std::size_t S, const u32* __restrict__ keep,
    const u32* __restrict__ out_pos, const u64* __restrict__ keys, int KW, std::size_t W,
    const double2* __restrict__ sums, u64* __restrict__ ox, u64* __restrict__ oz,
    double2* __restrict__ oc) {
  for (std::size_t s = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < S;
       s += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    if (!keep[s]) continue;
    const std::size_t o = out_pos[s];
    u32 kb, ke;
    segment_bounds(map, s, kb, ke);
    const std::size_t k = kb;
    if (W == 0) {
      // Zero-qubit operators have no mask words (their buffers are null).
    } else if (KW == 1) {
      ox[o] = keys[k] >> 32;
      oz[o] = keys[k] & 0xFFFFFFFFULL;
    } else {
      for (std::size_t w = 0; w < W; ++w) {
        ox[o * W + w] = keys[k * KW + 2 * w];
        oz[o * W + w] = keys[k * KW + 2 * w + 1];
      }
    }
    oc[o] = sums[s];
  }
}

// ---------------------------------------------------------------- word-major sort
//
// Keys wider than one word are sorted most significant word first: a stable
// LSD radix sort of (word 0, index) records, then, only for the items whose
// word-0 values tie, a stable sort of (run, next word) records written back
// into the tied positions, and so on word by word. Records carry one key word
// instead of the whole key, and for distinct-ish keys (random products) the
// later words touch few items.

// Materialized terms: key word w of term t in the canonical order
// [x0, z0, x1, z1, ...]. (ProductTerms reads the same words through the
// operands of a product.)
struct ArrayTerms {
  const u64* x;
  const u64* z;
  std::size_t W;
  __device__ u64 word(u32 t, int w) const {
    const std::size_t at = static_cast<std::size_t>(t) * W + static_cast<std::size_t>(w >> 1);
    return (w & 1) ? z[at] : x[at];
  }
};

__global__ __launch_bounds__(256) void word_records_kernel(const u64* __restrict__ x,
                                                           const u64* __restrict__ z,
                                                           std::size_t T, std::size_t W,
                                                           u64* __restrict__ keys,
                                                           u32* __restrict__ idx) {
  for (std::size_t t = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < T;
       t += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    keys[t] = x[t * W];
    idx[t] = static_cast<u32>(t);
  }
}

// head[i] = 1 where a run of equal keys starts (KW-word records).
__global__ __launch_bounds__(256) void run_heads_kernel(const u64* __restrict__ keys,
                                                        std::size_t n, int KW,
                                                        u32* __restrict__ head) {
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    bool h = i == 0;
    if (!h) {
      for (int k = 0; k < KW; ++k) h = h || keys[i * KW + k] != keys[(i - 1) * KW + k];
    }
    head[i] = h ? 1U : 0U;
  }
}

// tied[i] = 1 when item i's run has at least two items.
__global__ __launch_bounds__(256) void tied_kernel(const u32* __restrict__ head, std::size_t n,
                                                   u32* __restrict__ tied) {
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    tied[i] = (!head[i] || (i + 1 < n && !head[i + 1])) ? 1U : 0U;
  }
}

// Refinement records for the tied items of the current list (list == null:
// the identity over all positions): global position, run rank (monotone in
// position) and word w of the item's term.
template <class Src>
__global__ __launch_bounds__(256) void refine_records_kernel(
    const u32* __restrict__ list, std::size_t n, const u32* __restrict__ tied,
    const u32* __restrict__ tpos, const u32* __restrict__ head, const u32* __restrict__ hscan,
    const u32* __restrict__ idx, Src src, int w, u32* __restrict__ plist,
    u64* __restrict__ keys2, u32* __restrict__ idx2) {
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    if (!tied[i]) continue;
    const u32 j = tpos[i];
    const u32 pos = list ? list[i] : static_cast<u32>(i);
    const u32 t = idx[pos];
    plist[j] = pos;
    keys2[2 * static_cast<std::size_t>(j)] = hscan[i] + head[i] - 1U;
    keys2[2 * static_cast<std::size_t>(j) + 1] = src.word(t, w);
    idx2[j] = t;
  }
}

__global__ __launch_bounds__(256) void refine_scatter_kernel(const u32* __restrict__ plist,
                                                             const u32* __restrict__ idx2,
                                                             std::size_t n, u32* __restrict__ idx) {
  for (std::size_t j = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < n;
       j += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    idx[plist[j]] = idx2[j];
  }
}

// Segment heads of the fully sorted order: word 0 from the sorted word-0 keys,
// the other words compared through the terms only where word 0 ties.
template <class Src>
__global__ __launch_bounds__(256) void full_heads_kernel(const u64* __restrict__ k0,
                                                         const u32* __restrict__ idx, Src src,
                                                         int KW, std::size_t T,
                                                         u32* __restrict__ flags) {
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < T;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    bool h = i == 0 || k0[i] != k0[i - 1];
    if (!h) {
      const u32 a = idx[i], b = idx[i - 1];
      for (int w = 1; w < KW && !h; ++w) h = src.word(a, w) != src.word(b, w);
    }
    flags[i] = h ? 1U : 0U;
  }
}

__global__ __launch_bounds__(256) void compact_gather_kernel(
    SegMap map, std::size_t S, const u32* __restrict__ keep,
    const u32* __restrict__ out_pos, const u32* __restrict__ idx, const u64* __restrict__ x,
    const u64* __restrict__ z, std::size_t W, const double2* __restrict__ sums,
    u64* __restrict__ ox, u64* __restrict__ oz, double2* __restrict__ oc) {
  for (std::size_t s = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < S;
       s += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    if (!keep[s]) continue;
    const std::size_t o = out_pos[s];
    u32 b, e;
    segment_bounds(map, s, b, e);
    const std::size_t t = idx[b];
    for (std::size_t w = 0; w < W; ++w) {
      ox[o * W + w] = x[t * W + w];
      oz[o * W + w] = z[t * W + w];
    }
    oc[o] = sums[s];
  }
}

// Masks of every term interleaved into one record of RW words (x words, z
// words, zero padding; RW a power of two), so that the compaction's random
// gather touches one memory sector per term instead of one per mask array.
__global__ __launch_bounds__(256) void mask_records_kernel(const u64* __restrict__ x,
                                                           const u64* __restrict__ z,
                                                           std::size_t T, std::size_t W,
                                                           std::size_t RW, u64* __restrict__ rec) {
  for (std::size_t t = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < T;
       t += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    u64* r = rec + t * RW;
    for (std::size_t w = 0; w < W; ++w) {
      r[w] = x[t * W + w];
      r[W + w] = z[t * W + w];
    }
    for (std::size_t w = 2 * W; w < RW; ++w) r[w] = 0;
  }
}

constexpr std::size_t record_words(std::size_t W) { return W == 1 ? 2 : W == 2 ? 4 : 8; }

// Records written by the prologue (2 <= W <= 4): 64 bytes. For W <= 3 the
// term's coefficient rides in words 2W and 2W + 1, so the final gather reads
// masks and coefficient from one record.
constexpr bool record_has_coeff(std::size_t W) { return W >= 2 && W <= 3; }
constexpr std::size_t prologue_record_words(std::size_t W) {
  return record_has_coeff(W) ? 8 : record_words(W);
}

// One pass over the terms of a multi-word operator (2 <= W <= 4, whose keys
// never pack into one word): OR / AND of every key word ([x0, z0, x1, z1,
// ...], combined with atomics), the block's max |c|, the word-0 sort records
// and the interleaved mask records of the final gather. Replaces separate
// passes that each read the masks again.
template <int W>
__global__ __launch_bounds__(256) void prologue_kernel(
    const u64* __restrict__ x, const u64* __restrict__ z, const double2* __restrict__ c,
    std::size_t T, unsigned long long* or_out, unsigned long long* and_out,
    double* __restrict__ block_max, u64* __restrict__ keys, u32* __restrict__ idx,
    u64* __restrict__ rec) {
  static_assert(W >= 2 && W <= 4, "prologue for multi-word keys");
  constexpr int RW = static_cast<int>(prologue_record_words(W));
  __shared__ double lds[4];
  u64 o[2 * W], a[2 * W];
#pragma unroll
  for (int k = 0; k < 2 * W; ++k) {
    o[k] = 0;
    a[k] = ~0ULL;
  }
  double m = 0.0;
  for (std::size_t t = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < T;
       t += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    u64 r[RW];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      r[w] = x[t * W + w];
      r[W + w] = z[t * W + w];
    }
#pragma unroll
    for (int w = 2 * W; w < RW; ++w) r[w] = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      o[2 * w] |= r[w];
      a[2 * w] &= r[w];
      o[2 * w + 1] |= r[W + w];
      a[2 * w + 1] &= r[W + w];
    }
    const double2 ct = c[t];
    m = fmax(m, hypot(ct.x, ct.y));
    if constexpr (record_has_coeff(W)) {
      r[2 * W] = static_cast<u64>(__double_as_longlong(ct.x));
      r[2 * W + 1] = static_cast<u64>(__double_as_longlong(ct.y));
    }
    keys[t] = r[0];
    idx[t] = static_cast<u32>(t);
#pragma unroll
    for (int k = 0; k < RW; k += 2) {
      *reinterpret_cast<ulonglong2*>(rec + t * RW + k) = make_ulonglong2(r[k], r[k + 1]);
    }
  }
#pragma unroll
  for (int k = 0; k < 2 * W; ++k) {
    u64 ok = o[k], ak = a[k];
    for (int d = 32; d > 0; d >>= 1) {
      ok |= __shfl_xor(ok, d, 64);
      ak &= __shfl_xor(ak, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicOr(&or_out[k], static_cast<unsigned long long>(ok));
      atomicAnd(&and_out[k], static_cast<unsigned long long>(ak));
    }
  }
  for (int d = 32; d > 0; d >>= 1) m = fmax(m, __shfl_xor(m, d, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    block_max[blockIdx.x] = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
  }
}

// Segment sums and output in one gather (records with the coefficient, W =
// 2, 3): segment s (sorted order) takes its head's record (masks and
// coefficient), adds the other members' coefficients in sorted (=
// construction) order, the reference's +0.0 seed when it applies, and writes
// masks and sum at output position s; keep[s] flags |sum| > threshold and the
// survivors are counted (one atomic per block). When every segment survives
// the output is final; otherwise the survivors are compacted afterwards.
template <int W>
__global__ __launch_bounds__(256) void segment_emit_kernel(
    SegMap map, std::size_t S, const u32* __restrict__ idx,
    const u64* __restrict__ rec, int zero_seed, double threshold, u64* __restrict__ ox,
    u64* __restrict__ oz, double2* __restrict__ oc, u32* __restrict__ keep,
    u32* __restrict__ survivors) {
  static_assert(record_has_coeff(W), "records carry the coefficient");
  constexpr int RW = static_cast<int>(prologue_record_words(W));
  __shared__ u32 lds[4];
  u32 kept = 0;
  for (std::size_t s = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < S;
       s += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    u32 b, e;
    segment_bounds(map, s, b, e);
    const u64* head = rec + static_cast<std::size_t>(idx[b]) * RW;
    u64 r[2 * W + 2];
#pragma unroll
    for (int k = 0; k < 2 * W + 2; k += 2) {
      const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(head + k);
      r[k] = v.x;
      r[k + 1] = v.y;
    }
    double2 acc = make_double2(__longlong_as_double(static_cast<long long>(r[2 * W])),
                               __longlong_as_double(static_cast<long long>(r[2 * W + 1])));
    for (u32 j = b + 1; j < e; ++j) {
      const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(
          rec + static_cast<std::size_t>(idx[j]) * RW + 2 * W);
      acc.x += __longlong_as_double(static_cast<long long>(v.x));
      acc.y += __longlong_as_double(static_cast<long long>(v.y));
    }
    if (zero_seed) {
      acc.x += 0.0;
      acc.y += 0.0;
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
      ox[s * W + w] = r[w];
      oz[s * W + w] = r[W + w];
    }
    oc[s] = acc;
    const u32 k = hypot(acc.x, acc.y) > threshold ? 1U : 0U;
    keep[s] = k;
    kept += k;
  }
  for (int d = 32; d > 0; d >>= 1) kept += static_cast<u32>(__shfl_xor(static_cast<int>(kept), d, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = kept;
  __syncthreads();
  if (threadIdx.x == 0) {
    const u32 total = lds[0] + lds[1] + lds[2] + lds[3];
    if (total) atomicAdd(survivors, total);
  }
}

__global__ void clear_count_kernel(u32* count) { *count = 0; }

// Survivors of segment_emit_kernel's output (every segment at its position)
// moved to their compacted positions.
__global__ __launch_bounds__(256) void compact_emitted_kernel(
    const u32* __restrict__ keep, const u32* __restrict__ out_pos, std::size_t S, std::size_t W,
    const u64* __restrict__ sx, const u64* __restrict__ sz, const double2* __restrict__ sc,
    u64* __restrict__ ox, u64* __restrict__ oz, double2* __restrict__ oc) {
  for (std::size_t s = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < S;
       s += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    if (!keep[s]) continue;
    const std::size_t o = out_pos[s];
    for (std::size_t w = 0; w < W; ++w) {
      ox[o * W + w] = sx[s * W + w];
      oz[o * W + w] = sz[s * W + w];
    }
    oc[o] = sc[s];
  }
}

template <int W>
__global__ __launch_bounds__(256) void compact_records_kernel(
    SegMap map, std::size_t S, const u32* __restrict__ keep,
    const u32* __restrict__ out_pos, const u32* __restrict__ idx, const u64* __restrict__ rec,
    const double2* __restrict__ sums, u64* __restrict__ ox, u64* __restrict__ oz,
    double2* __restrict__ oc) {
  constexpr int RW = static_cast<int>(record_words(W));
  for (std::size_t s = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < S;
       s += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    if (!keep[s]) continue;
    const std::size_t o = out_pos[s];
    u32 b, e;
    segment_bounds(map, s, b, e);
    const std::size_t t = idx[b];
    u64 r[RW];
#pragma unroll
    for (int k = 0; k < RW; k += 2) {
      const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(rec + t * RW + k);
      r[k] = v.x;
      r[k + 1] = v.y;
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
      ox[o * W + w] = r[w];
      oz[o * W + w] = r[W + w];
    }
    oc[o] = sums[s];
  }
}

// ---------------------------------------------------------------- products
//
// Simplify of lhs @ rhs without materializing the L*R products. One pass
// computes every product's sort key (word 0; x << 32 | z when the masks fit
// 32 bits), the key statistics and max |c|, and writes only the (key, index)
// records; the sort runs on those. The tie refinement, the segment heads and
// the segment sums read masks and coefficients through the operands
// (ProductTerms), which stay in L2, instead of gathering materialized products
// from HBM in sorted (random) order. The coefficient arithmetic is the product
// kernel's, so the sums are bit-identical to simplifying materialized products.


// Operand records (product_record_words(W) words per term: x words, z words,
// coefficient bits, zero padding).
__global__ __launch_bounds__(256) void operand_records_kernel(const u64* __restrict__ x,
                                                              const u64* __restrict__ z,
                                                              const double2* __restrict__ c,
                                                              std::size_t T, u32 W,
                                                              u64* __restrict__ rec) {
  const std::size_t RW = product_record_words(W);
  for (std::size_t t = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < T;
       t += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    u64* r = rec + t * RW;
    for (std::size_t w = 0; w < W; ++w) {
      r[w] = x[t * W + w];
      r[W + w] = z[t * W + w];
    }
    r[2 * W] = static_cast<u64>(__double_as_longlong(c[t].x));
    r[2 * W + 1] = static_cast<u64>(__double_as_longlong(c[t].y));
    for (std::size_t w = 2 * W + 2; w < RW; ++w) r[w] = 0;
  }
}

// One block per (lhs term, kProductChunk = 2,048 rhs terms): twice the product
// kernel's 1,024, which halves the per-block statistics reductions (0.30 ->
// 0.19 ms for 8,192^2 products).
template <int W, bool kPacked, bool kMax>
__global__ __launch_bounds__(kProductKeyBlock) void product_keys_kernel(
    ProductTerms P, u32 chunks, u64* __restrict__ block_stats, u64* __restrict__ keys,
    u32* __restrict__ idx) {
  __shared__ u64 lds[kProductKeyBlock / 64][kProductMaxStats];
  const u32 i = blockIdx.x / chunks;
  const u32 j0 = (blockIdx.x - i * chunks) * kProductChunk;
  product_keys_unit<W, kPacked, kMax>(P, i, j0, keys, idx, block_stats, gridDim.x, blockIdx.x,
                                      lds);
}

// Combines the per-block statistics of product_keys_kernel (word k of block b
// at k * blocks + b): block k of this kernel reduces word k (OR for k < n_or,
// AND for k < 2 n_or, max after that). Same-address atomics from every block
// would serialize.
__global__ __launch_bounds__(1024) void stats_reduce_kernel(const u64* __restrict__ block_stats,
                                                            u32 blocks, u32 n_or,
                                                            u64* __restrict__ out) {
  __shared__ u64 part[16];
  const u32 k = blockIdx.x;
  const int op = k < n_or ? 0 : k < 2 * n_or ? 1 : 2;
  auto combine = [op](u64 v, u64 u) { return op == 0 ? (v | u) : op == 1 ? (v & u) : (u > v ? u : v); };
  u64 v = op == 1 ? ~0ULL : 0ULL;
  const u64* row = block_stats + static_cast<std::size_t>(k) * blocks;
  for (u32 b = threadIdx.x; b < blocks; b += blockDim.x) v = combine(v, row[b]);
  for (int d = 32; d > 0; d >>= 1) v = combine(v, __shfl_xor(v, d, 64));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 16; ++w) v = combine(v, part[w]);
    out[k] = v;
  }
}

// Masks of the first `count` products (the duplicate-rate sample of the
// simplify seed rule).
__global__ __launch_bounds__(256) void product_sample_kernel(ProductTerms P, u32 count,
                                                             u64* __restrict__ x,
                                                             u64* __restrict__ z) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  for (u32 w = 0; w < P.W; ++w) {
    x[static_cast<std::size_t>(t) * P.W + w] = P.word(t, static_cast<int>(2 * w));
    z[static_cast<std::size_t>(t) * P.W + w] = P.word(t, static_cast<int>(2 * w + 1));
  }
}

// Segment sums and output of the products: segment s (sorted order) takes its
// head's masks, adds its members' coefficients in sorted (= construction)
// order with the reference's +0.0 seed when it applies, and writes masks and
// sum at output position s; keep[s] flags |sum| > threshold and the survivors
// are counted (one atomic per block).
typedef __attribute__((address_space(3))) void lds_void;

template <int W>
__global__ __launch_bounds__(256) void product_emit_kernel(
    SegMap map, std::size_t S, const u32* __restrict__ idx, ProductTerms P,
    int zero_seed, double threshold, u64* __restrict__ ox, u64* __restrict__ oz,
    double2* __restrict__ oc, u32* __restrict__ keep, u32* __restrict__ survivors) {
  constexpr u32 RW = product_record_words(W);
  constexpr u32 kChunks = RW / 2;         // 16-byte pieces of a record
  constexpr u32 kPerLoad = 64 / kChunks;  // records one wave-wide load moves
  // Per wave: the operand records of the first product of its 64 segments.
  __shared__ u64 stage[4][2][64 * RW];
  __shared__ u32 lds[4];
  const u32 lane = threadIdx.x & 63U, wave = threadIdx.x >> 6;
  u64* const lst = stage[wave][0];
  u64* const rst = stage[wave][1];
  u32 kept = 0;
  const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
  for (std::size_t base = static_cast<std::size_t>(blockIdx.x) * blockDim.x + 64U * wave; base < S;
       base += stride) {
    const std::size_t s = base + lane;
    const bool valid = s < S;
    u32 b = 0, e = 0, t = 0;
    if (valid) {
      segment_bounds(map, s, b, e);
      t = idx[b];
    }
    // Records in sorted order are scattered lines: loaded by the lane that
    // needs them they cost 2 (W + 1) 16-byte requests per segment. Here each
    // wave-wide load (gfx950's 16-byte LDS DMA) gives kChunks consecutive
    // lanes one record, a single line, and each lane then reads its segment's
    // records from LDS.
#pragma unroll
    for (u32 g = 0; g < kChunks; ++g) {
      const u32 src = g * kPerLoad + lane / kChunks;
      const u32 ts = static_cast<u32>(__shfl(static_cast<int>(t), static_cast<int>(src), 64));
      if (base + src < S) {
        u32 i, j;
        P.split(ts, i, j);
        const u32 c = 2U * (lane % kChunks);
        __builtin_amdgcn_global_load_lds(P.lrec + static_cast<std::size_t>(i) * RW + c,
                                         (lds_void*)(lst + g * kPerLoad * RW), 16, 0, 0);
        __builtin_amdgcn_global_load_lds(P.rrec + static_cast<std::size_t>(j) * RW + c,
                                         (lds_void*)(rst + g * kPerLoad * RW), 16, 0, 0);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    if (valid) {
      u64 mx[W], mz[W];
      double2 acc = ProductTerms::combine<W>(lst + lane * RW, rst + lane * RW, mx, mz);
#pragma unroll
      for (int w = 0; w < W; ++w) {
        ox[s * W + w] = mx[w];
        oz[s * W + w] = mz[w];
      }
      for (u32 q = b + 1; q < e; ++q) {
        const double2 v = P.coeff<W>(idx[q]);
        acc.x += v.x;
        acc.y += v.y;
      }
      if (zero_seed) {
        acc.x += 0.0;
        acc.y += 0.0;
      }
      oc[s] = acc;
      const u32 k = hypot(acc.x, acc.y) > threshold ? 1U : 0U;
      keep[s] = k;
      kept += k;
    }
  }
  for (int d = 32; d > 0; d >>= 1) kept += static_cast<u32>(__shfl_xor(static_cast<int>(kept), d, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = kept;
  __syncthreads();
  if (threadIdx.x == 0) {
    const u32 total = lds[0] + lds[1] + lds[2] + lds[3];
    if (total) atomicAdd(survivors, total);
  }
}

unsigned stride_grid(std::size_t work) {
  const std::size_t g = (work + 255) / 256;
  return static_cast<unsigned>(std::max<std::size_t>(1, std::min<std::size_t>(g, 65536)));
}

// Sampled duplicate rate of the first `sample` keys (reference
// src/simplify.cpp:67-97).
double duplicate_rate(const std::vector<u64>& hx, const std::vector<u64>& hz, std::size_t W,
                      std::size_t sample) {
  if (sample == 0) return 0.0;
  struct H {
    std::size_t W;
    const std::vector<u64>* x;
    const std::vector<u64>* z;
    std::size_t operator()(std::size_t t) const {
      std::size_t h = 0x9E3779B97F4A7C15ULL;
      for (std::size_t w = 0; w < W; ++w) {
        h ^= (*x)[t * W + w] + 0x632BE59BD9B4E019ULL + (h << 6) + (h >> 2);
        h ^= (*z)[t * W + w] + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
      }
      return h;
    }
  };
  struct E {
    std::size_t W;
    const std::vector<u64>* x;
    const std::vector<u64>* z;
    bool operator()(std::size_t a, std::size_t b) const {
      for (std::size_t w = 0; w < W; ++w) {
        if ((*x)[a * W + w] != (*x)[b * W + w] || (*z)[a * W + w] != (*z)[b * W + w]) return false;
      }
      return true;
    }
  };
  std::unordered_set<std::size_t, H, E> seen(sample, H{W, &hx, &hz}, E{W, &hx, &hz});
  for (std::size_t t = 0; t < sample; ++t) seen.insert(t);
  return 1.0 - static_cast<double>(seen.size()) / static_cast<double>(sample);
}

}  // namespace

void copy_sample(const u64* x, const u64* z, std::size_t W, std::size_t count,
                 std::vector<u64>& hx, std::vector<u64>& hz, hipStream_t stream) {
  hx.resize(count * W);
  hz.resize(count * W);
  if (count * W == 0) return;
  check_hip(hipMemcpyAsync(hx.data(), x, count * W * sizeof(u64), hipMemcpyDeviceToHost, stream),
            "copy device buffer to host");
  check_hip(hipMemcpyAsync(hz.data(), z, count * W * sizeof(u64), hipMemcpyDeviceToHost, stream),
            "copy device buffer to host");
  sync(stream);
}

bool simplify_takes_hash_path(const u64* x, const u64* z, std::size_t T, std::size_t W,
                              hipStream_t stream) {
  // Reference src/simplify.cpp:107-143.
  if (W != 1 && W != 2) return false;
  if (T < 4096) return false;
  std::vector<u64> hx, hz;
  copy_sample(x, z, W, 4096, hx, hz, stream);
  const double first = duplicate_rate(hx, hz, W, 512);
  if (first >= 0.25) return true;
  if (first <= 0.05) return false;
  return duplicate_rate(hx, hz, W, 4096) >= 0.25;
}

bool operand_duplicate_heavy(const wg_pauli_sum* op, hipStream_t stream) {
  // Reference src/multiply.cpp:29-69: >= 128 terms and a duplicate rate of at
  // least 0.25 among the first min(T, 512) terms.
  if (op->num_terms < 128) return false;
  const std::size_t sample = std::min<std::size_t>(op->num_terms, 512);
  std::vector<u64> hx, hz;
  copy_sample(op->x, op->z, op->words, sample, hx, hz, stream);
  return duplicate_rate(hx, hz, op->words, sample) >= 0.25;
}

namespace {

struct DigitPass {
  int word, shift;
};

// 8-bit digit passes over KW-word records, least significant first; a digit
// is skipped when none of its bits varies (varying[k]: bits of word k that
// differ somewhere; word 0 is the most significant).
std::vector<DigitPass> digit_passes(const std::vector<u64>& varying) {
  std::vector<DigitPass> passes;
  for (int k = static_cast<int>(varying.size()) - 1; k >= 0; --k) {
    for (int shift = 0; shift < 64; shift += 8) {
      if ((varying[k] >> shift) & 0xFFULL) passes.push_back({k, shift});
    }
  }
  return passes;
}

// One-word sorts of at least this many records (and fewer than 2^30) run
// their digit passes one-sweep.
constexpr std::size_t kOneSweepMin = std::size_t{1} << 20;

// Stable LSD radix sort of (KW-word key, index) records; the sorted records
// end up in (keys_a, idx_a). Large one-word sorts take one histogram of every
// pass's digits up front and then one kernel per pass (look-back offsets);
// the others run histogram -> scan -> scatter per pass.
//
// With `products` (the keys are the one-word keys of those products, packed
// when `packed`), the one-sweep digit counts come from the operands.
void radix_sort(u64*& keys_a, u32*& idx_a, u64*& keys_b, u32*& idx_b, std::size_t n, int KW,
                const std::vector<DigitPass>& passes, Scratch& scratch, hipStream_t stream,
                const char* prof_name = "radix_scatter", const ProductTerms* products = nullptr,
                bool packed = false) {
  if (passes.empty() || n < 2) return;
  const std::size_t nblocks = (n + kTile - 1) / kTile;
  if (KW == 1 && n >= kOneSweepMin && n < (std::size_t{1} << 30) && passes.size() <= 8) {
    const std::size_t P = passes.size();
    // One allocation: status words of every pass, tile counters, error flag,
    // digit counts and starts; cleared by a kernel.
    const std::size_t status_words = P * nblocks * 256;
    const std::size_t words = status_words + P + 1 + 2 * P * 256;
    u32* mem = scratch.take<u32>(words);
    u32* status = mem;
    u32* counters = status + status_words;
    u32* error = counters + P;
    u32* hist = error + 1;
    u32* starts = hist + P * 256;
    clear_u32_kernel<<<stride_grid(words), 256, 0, stream>>>(mem, words);
    check_launch("clear");
    DigitShifts sh{};
    for (std::size_t p = 0; p < P; ++p) sh.s[p] = passes[p].shift;
    if (products) {
      const ProductTerms& PT = *products;
      const u32 L = static_cast<u32>(n / PT.R), RW = product_record_words(PT.W);
      u32* oh = scratch.take<u32>(2 * P * 256);
      clear_u32_kernel<<<stride_grid(2 * P * 256), 256, 0, stream>>>(oh, 2 * P * 256);
      check_launch("clear");
      operand_digit_hist_kernel<<<std::min<unsigned>(stride_grid(L), 64), 256, 0, stream>>>(
          PT.lrec, L, RW, PT.W, packed ? 1 : 0, sh, static_cast<int>(P), oh);
      check_launch("operand_digit_hist");
      operand_digit_hist_kernel<<<std::min<unsigned>(stride_grid(PT.R), 64), 256, 0, stream>>>(
          PT.rrec, PT.R, RW, PT.W, packed ? 1 : 0, sh, static_cast<int>(P), oh + P * 256);
      check_launch("operand_digit_hist");
      xor_convolve_kernel<<<static_cast<unsigned>(P), 256, 0, stream>>>(oh, oh + P * 256, hist);
      check_launch("xor_convolve");
    } else {
      digit_hist_all_kernel<<<std::min<unsigned>(stride_grid(n), 2048), 256, 0, stream>>>(
          keys_a, n, sh, static_cast<int>(P), hist);
      check_launch("digit_hist_all");
    }
    digit_starts_kernel<<<static_cast<unsigned>(P), 256, 0, stream>>>(hist, starts);
    check_launch("digit_starts");
    for (std::size_t p = 0; p < P; ++p) {
      const LookBack look{starts + p * 256, status + p * nblocks * 256, counters + p, error};
      ProfileScope prof(prof_name, stream);
      radix_scatter_kernel<1, true><<<grid_for(nblocks, 1), kBlock, 0, stream>>>(
          keys_a, idx_a, keys_b, idx_b, n, passes[p].word, passes[p].shift, nblocks, nullptr,
          look);
      check_launch("radix_scatter");
      std::swap(keys_a, keys_b);
      std::swap(idx_a, idx_b);
    }
    u32 err = 0;
    check_hip(hipMemcpyAsync(&err, error, sizeof(u32), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
    if (err) fail_runtime("radix sort look-back did not complete");
    return;
  }
  u32* hist = scratch.take<u32>(256 * nblocks);
  u32* offs = scratch.take<u32>(256 * nblocks);
  for (const DigitPass& p : passes) {
    radix_hist_kernel<<<grid_for(nblocks, 1), kBlock, 0, stream>>>(keys_a, n, KW, p.word,
                                                                   p.shift, nblocks, hist);
    check_launch("radix_hist");
    exclusive_scan_u32(hist, offs, 256 * nblocks, nullptr, scratch, stream);
    ProfileScope prof(prof_name, stream);
    const LookBack none{};
    if (KW == 1) {
      radix_scatter_kernel<1, false><<<grid_for(nblocks, 1), kBlock, 0, stream>>>(
          keys_a, idx_a, keys_b, idx_b, n, p.word, p.shift, nblocks, offs, none);
    } else if (KW == 2) {
      radix_scatter_kernel<2, false><<<grid_for(nblocks, 1), kBlock, 0, stream>>>(
          keys_a, idx_a, keys_b, idx_b, n, p.word, p.shift, nblocks, offs, none);
    } else {
      fail_runtime("radix sort records hold one or two key words");
    }
    check_launch("radix_scatter");
    std::swap(keys_a, keys_b);
    std::swap(idx_a, idx_b);
  }
}

// Head flags of sorted one-word keys (flags[i] = 1 where a run of equal keys
// starts) and the counts behind them; each block takes a contiguous range.
struct HeadInfo {
  u32 nonheads = 0;  // items equal to their predecessor
  u32 tied = 0;      // items whose run has two or more
  unsigned grid = 0;
  std::size_t chunk = 0;
  u32* block_tied = nullptr;   // per block: its tied items
  u32* tied_offset = nullptr;  // per block: its first slot in the ordered tied list
};

HeadInfo key_heads(const u64* keys, std::size_t T, u32* flags, Scratch& scratch,
                   hipStream_t stream) {
  HeadInfo h;
  h.grid = static_cast<unsigned>(std::min<std::size_t>((T + 255) / 256, 2048));
  h.chunk = ((T + h.grid - 1) / h.grid + 255) / 256 * 256;
  u32* nonheads = scratch.take<u32>(h.grid);
  u32* tied = scratch.take<u32>(h.grid);
  h.block_tied = tied;
  u32* scanned = scratch.take<u32>(h.grid);
  u32* totals = scratch.take<u32>(2);
  h.tied_offset = scratch.take<u32>(h.grid);
  key_heads_kernel<<<h.grid, 256, 0, stream>>>(keys, T, h.chunk, flags, nonheads, tied);
  check_launch("key_heads");
  exclusive_scan_u32(nonheads, scanned, h.grid, totals, scratch, stream);
  exclusive_scan_u32(tied, h.tied_offset, h.grid, totals + 1, scratch, stream);
  u32 host[2] = {0, 0};
  check_hip(hipMemcpyAsync(host, totals, sizeof(host), hipMemcpyDeviceToHost, stream),
            "copy device buffer to host");
  sync(stream);
  h.nonheads = host[0];
  h.tied = host[1];
  return h;
}

// Segment starts of the sorted order from its head flags (seg_start[s], s <=
// S; seg_start[S] = T); returns S.
u32 segment_table(const u32* flags, std::size_t T, u32* seg_start, u32* d_count, Scratch& scratch,
                  hipStream_t stream) {
  u32* seg_pos = scratch.take<u32>(T);
  exclusive_scan_u32(flags, seg_pos, T, d_count, scratch, stream);
  segment_starts_kernel<<<stride_grid(T), 256, 0, stream>>>(flags, seg_pos, T, seg_start);
  check_launch("segment_starts");
  u32 S = 0;
  check_hip(hipMemcpyAsync(&S, d_count, sizeof(u32), hipMemcpyDeviceToHost, stream),
            "copy device buffer to host");
  sync(stream);
  return S;
}

// Refinement levels w_begin .. KW - 1 over the tied items of the current list
// (list == null: all of [0, n_cur)) whose run heads are `head`: the items
// whose previous words tie are re-sorted by (run rank, word w) and written
// back into their positions of idx, word by word.
template <class Src>
void refine_levels(Src src, std::size_t w_begin, std::size_t KW, const u32* list,
                   std::size_t n_cur, u32* head, u32* idx,
                   const std::vector<unsigned long long>& ors,
                   const std::vector<unsigned long long>& ands, Scratch& scratch,
                   hipStream_t stream) {
  if (w_begin >= KW || n_cur == 0) return;
  u32* tied = scratch.take<u32>(n_cur);
  u32* tpos = scratch.take<u32>(n_cur);
  u32* hscan = scratch.take<u32>(n_cur);
  u32* d_tied = scratch.take<u32>(1);
  for (std::size_t w = w_begin; w < KW; ++w) {
    tied_kernel<<<stride_grid(n_cur), 256, 0, stream>>>(head, n_cur, tied);
    check_launch("tied");
    exclusive_scan_u32(tied, tpos, n_cur, d_tied, scratch, stream);
    u32 A = 0;
    check_hip(hipMemcpyAsync(&A, d_tied, sizeof(u32), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
    if (A == 0) break;
    exclusive_scan_u32(head, hscan, n_cur, nullptr, scratch, stream);
    u32* plist = scratch.take<u32>(A);
    u64* r_keys_a = scratch.take<u64>(2 * static_cast<std::size_t>(A));
    u64* r_keys_b = scratch.take<u64>(2 * static_cast<std::size_t>(A));
    u32* r_idx_a = scratch.take<u32>(A);
    u32* r_idx_b = scratch.take<u32>(A);
    refine_records_kernel<<<stride_grid(n_cur), 256, 0, stream>>>(
        list, n_cur, tied, tpos, head, hscan, idx, src, static_cast<int>(w), plist, r_keys_a,
        r_idx_a);
    check_launch("refine_records");
    // Records [run, word w]: word w's varying bits, then the run ranks
    // (all bits below the bit length of n_cur).
    int run_bits = 0;
    while (run_bits < 64 && (u64{1} << run_bits) < n_cur) ++run_bits;
    const u64 run_varying = run_bits >= 64 ? ~0ULL : ((u64{1} << run_bits) - 1);
    radix_sort(r_keys_a, r_idx_a, r_keys_b, r_idx_b, A, 2,
               digit_passes({run_varying, ors[w] ^ ands[w]}), scratch, stream, "radix_refine");
    refine_scatter_kernel<<<stride_grid(A), 256, 0, stream>>>(plist, r_idx_a, A, idx);
    check_launch("refine_scatter");
    // The next level works on the tied items, in their new order.
    u32* head2 = scratch.take<u32>(A);
    run_heads_kernel<<<stride_grid(A), 256, 0, stream>>>(r_keys_a, A, 2, head2);
    check_launch("run_heads");
    head = head2;
    list = plist;
    n_cur = A;
  }
}

// At most this many tied items take the list path of sorted_segments: its
// segment map is searched per output term (log2 of the non-head count steps).
constexpr u32 kTiedListMax = 8192;

// Segments of equal keys, given the items sorted stably by key word 0 (k0,
// idx; KW = 1 when k0 is the whole key): the items whose word-0 values tie are
// re-sorted word by word (refine_levels), so equal keys end up adjacent in
// construction order, and flags[i] = 1 where a segment starts. src reads the
// KW key words of a term; ors / ands hold each word's bitwise OR / AND over all
// terms (their XOR: the varying bits). Three outcomes, by the number of ties:
//  * none: every key its own segment (no map, S = T);
//  * a few (at most kTiedListMax tied items): the ordered list of tied items
//    is refined alone, and the map holds the positions that repeat their
//    predecessor;
//  * more: every level and the head flags run over all positions, and the
//    map is the segment start table.
// *S receives the segment count.
template <class Src>
SegMap sorted_segments(Src src, std::size_t KW, std::size_t T, u64* k0, u32* idx,
                       const std::vector<unsigned long long>& ors,
                       const std::vector<unsigned long long>& ands, u32* flags, u32* d_count,
                       u32* S, Scratch& scratch, hipStream_t stream) {
  const HeadInfo h = key_heads(k0, T, flags, scratch, stream);
  if (h.nonheads == 0) {
    *S = static_cast<u32>(T);
    return SegMap{nullptr, nullptr, 0};
  }
  if (h.tied <= kTiedListMax) {
    const std::size_t A = h.tied;
    u32* plist = scratch.take<u32>(A);
    tied_list_kernel<<<h.grid, 256, 0, stream>>>(k0, T, h.chunk, h.block_tied, h.tied_offset,
                                                 plist);
    check_launch("tied_list");
    if (KW > 1) {
      // Level 1 over the listed items, then the later levels as usual.
      u32* lhead = scratch.take<u32>(A);
      u32* lscan = scratch.take<u32>(A);
      gather_heads_kernel<<<stride_grid(A), 256, 0, stream>>>(plist, A, flags, lhead);
      check_launch("gather_heads");
      exclusive_scan_u32(lhead, lscan, A, nullptr, scratch, stream);
      u64* r_keys_a = scratch.take<u64>(2 * A);
      u64* r_keys_b = scratch.take<u64>(2 * A);
      u32* r_idx_a = scratch.take<u32>(A);
      u32* r_idx_b = scratch.take<u32>(A);
      list_records_kernel<<<stride_grid(A), 256, 0, stream>>>(plist, A, flags, lscan, idx, src, 1,
                                                             r_keys_a, r_idx_a);
      check_launch("list_records");
      int run_bits = 0;
      while (run_bits < 64 && (u64{1} << run_bits) < A) ++run_bits;
      const u64 run_varying = run_bits >= 64 ? ~0ULL : ((u64{1} << run_bits) - 1);
      radix_sort(r_keys_a, r_idx_a, r_keys_b, r_idx_b, A, 2,
                 digit_passes({run_varying, ors[1] ^ ands[1]}), scratch, stream, "radix_refine");
      refine_scatter_kernel<<<stride_grid(A), 256, 0, stream>>>(plist, r_idx_a, A, idx);
      check_launch("refine_scatter");
      u32* head2 = scratch.take<u32>(A);
      run_heads_kernel<<<stride_grid(A), 256, 0, stream>>>(r_keys_a, A, 2, head2);
      check_launch("run_heads");
      refine_levels(src, 2, KW, plist, A, head2, idx, ors, ands, scratch, stream);
    }
    u32* repeats = scratch.take<u32>(A);
    u32* rpos = scratch.take<u32>(A);
    list_full_heads_kernel<<<stride_grid(A), 256, 0, stream>>>(plist, A, k0, idx, src,
                                                               static_cast<int>(KW), flags, repeats);
    check_launch("list_full_heads");
    exclusive_scan_u32(repeats, rpos, A, d_count, scratch, stream);
    u32 m = 0;
    check_hip(hipMemcpyAsync(&m, d_count, sizeof(u32), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
    u32* nonheads = scratch.take<u32>(std::max<u32>(m, 1));
    if (m) {
      compact_list_kernel<<<stride_grid(A), 256, 0, stream>>>(plist, A, repeats, rpos, nonheads);
      check_launch("compact_list");
    }
    *S = static_cast<u32>(T - m);
    return SegMap{nullptr, m ? nonheads : nullptr, m};
  }
  refine_levels(src, 1, KW, nullptr, T, flags, idx, ors, ands, scratch, stream);
  if (KW > 1) {
    full_heads_kernel<<<stride_grid(T), 256, 0, stream>>>(k0, idx, src, static_cast<int>(KW), T,
                                                          flags);
    check_launch("full_heads");
  }
  u32* seg_start = scratch.take<u32>(T + 1);
  *S = segment_table(flags, T, seg_start, d_count, scratch, stream);
  return SegMap{seg_start, nullptr, 0};
}

}  // namespace

wg_pauli_sum* simplify_device(const u64* x, const u64* z, const double2* c, std::size_t n,
                              std::size_t T, int device, double atol, double rtol, int seed_mode,
                              Scratch& scratch, hipStream_t stream) {
  const std::size_t W = words_for(n);
  if (T == 0) return new_pauli_sum(n, 0, device);
  if (T >= 0xFFFFFFFFULL) fail_invalid("simplify supports at most 2^32 - 2 terms per call");

  // Threshold atol + rtol * max|c| (reference src/simplify.cpp:156-160), key
  // statistics and, for multi-word keys, the sort and mask records in the
  // same pass.
  const bool fused = W >= 2 && W <= 4;
  u32* flags = scratch.take<u32>(T);
  u64* keys_a = scratch.take<u64>(T);
  u32* idx_a = scratch.take<u32>(T);
  u64* keys_b = scratch.take<u64>(T);
  u32* idx_b = scratch.take<u32>(T);
  u64* rec = fused ? scratch.take<u64>(checked_mul(T, prologue_record_words(W), "mask records"))
                   : nullptr;
  const unsigned mgrid = std::min<unsigned>(stride_grid(T), 1024);
  double* dmax = scratch.take<double>(mgrid);
  std::vector<unsigned long long> ors(2 * W, 0ULL), ands(2 * W, ~0ULL);
  unsigned long long* dstat = W ? scratch.take<unsigned long long>(4 * W) : nullptr;
  if (W) {
    check_hip(hipMemcpyAsync(dstat, ors.data(), 2 * W * sizeof(u64), hipMemcpyHostToDevice, stream),
              "copy host buffer to device");
    check_hip(hipMemcpyAsync(dstat + 2 * W, ands.data(), 2 * W * sizeof(u64),
                             hipMemcpyHostToDevice, stream),
              "copy host buffer to device");
  }
  if (fused) {
    auto launch = [&](auto kernel) {
      kernel<<<mgrid, 256, 0, stream>>>(x, z, c, T, dstat, dstat + 2 * W, dmax, keys_a, idx_a, rec);
    };
    switch (W) {
      case 2: launch(prologue_kernel<2>); break;
      case 3: launch(prologue_kernel<3>); break;
      default: launch(prologue_kernel<4>); break;
    }
    check_launch("prologue");
  } else {
    max_abs_kernel<<<mgrid, 256, 0, stream>>>(c, T, dmax);
    check_launch("max_abs");
    if (W) {
      key_stats_kernel<<<mgrid, 256, 0, stream>>>(x, z, T, W, dstat, dstat + 2 * W);
      check_launch("key_stats");
    }
  }
  std::vector<double> hmax(mgrid);
  check_hip(hipMemcpyAsync(hmax.data(), dmax, mgrid * sizeof(double), hipMemcpyDeviceToHost,
                           stream),
            "copy device buffer to host");
  if (W) {
    check_hip(hipMemcpyAsync(ors.data(), dstat, 2 * W * sizeof(u64), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    check_hip(hipMemcpyAsync(ands.data(), dstat + 2 * W, 2 * W * sizeof(u64),
                             hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
  }
  sync(stream);
  double max_abs = 0.0;
  for (double v : hmax) max_abs = std::max(max_abs, v);
  const double threshold = atol + rtol * max_abs;

  bool hash_seed;
  if (seed_mode < 0) {
    hash_seed = simplify_takes_hash_path(x, z, T, W, stream);
  } else {
    hash_seed = seed_mode != 0;
  }

  // Segments of equal keys in sorted order, as positions into idx; the
  // word-0 keys (one-word keys: the whole key) in sorted order.
  const bool packed = W == 0 || (W == 1 && (ors[0] >> 32) == 0 && (ors[1] >> 32) == 0);
  u32* d_count = scratch.take<u32>(2);
  u32 S = 0;
  SegMap map;
  if (packed) {
    // One-word keys: x << 32 | z (or 0 for zero-qubit operators).
    const u64 kor = W ? (ors[0] << 32) | ors[1] : 0;
    const u64 kand = W ? (ands[0] << 32) | ands[1] : 0;
    build_records_kernel<<<stride_grid(T), 256, 0, stream>>>(x, z, T, W, 1, keys_a, idx_a);
    check_launch("build_records");
    radix_sort(keys_a, idx_a, keys_b, idx_b, T, 1, digit_passes({kor ^ kand}), scratch, stream);
    map = sorted_segments(ArrayTerms{x, z, W}, 1, T, keys_a, idx_a, ors, ands, flags, d_count, &S,
                          scratch, stream);
  } else {
    // Word 0 (x0) for every item, then the tied items word by word.
    if (!fused) {
      word_records_kernel<<<stride_grid(T), 256, 0, stream>>>(x, z, T, W, keys_a, idx_a);
      check_launch("word_records");
    }
    radix_sort(keys_a, idx_a, keys_b, idx_b, T, 1, digit_passes({ors[0] ^ ands[0]}), scratch,
               stream);
    map = sorted_segments(ArrayTerms{x, z, W}, 2 * W, T, keys_a, idx_a, ors, ands, flags, d_count,
                          &S, scratch, stream);
  }
  clear_count_kernel<<<1, 1, 0, stream>>>(d_count + 1);
  check_launch("clear_count");

  u32* keep = scratch.take<u32>(S);
  u32* out_pos = scratch.take<u32>(S);
  if (fused && record_has_coeff(W)) {
    // Sums and output in one gather of the 64-byte records, written at each
    // segment's position; compacted only when some segment is dropped.
    std::unique_ptr<wg_pauli_sum> staged(new_pauli_sum(n, S, device));
    clear_count_kernel<<<1, 1, 0, stream>>>(d_count + 1);
    check_launch("clear_count");
    auto emit = [&](auto kernel) {
      kernel<<<stride_grid(S), 256, 0, stream>>>(map, S, idx_a, rec, hash_seed ? 0 : 1,
                                                 threshold, staged->x, staged->z, staged->coeffs,
                                                 keep, d_count + 1);
    };
    if (W == 2) emit(segment_emit_kernel<2>); else emit(segment_emit_kernel<3>);
    check_launch("segment_emit");
    u32 U = 0;
    check_hip(hipMemcpyAsync(&U, d_count + 1, sizeof(u32), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
    if (U == S) return staged.release();
    exclusive_scan_u32(keep, out_pos, S, nullptr, scratch, stream);
    std::unique_ptr<wg_pauli_sum> out(new_pauli_sum(n, U, device));
    if (U) {
      compact_emitted_kernel<<<stride_grid(S), 256, 0, stream>>>(
          keep, out_pos, S, W, staged->x, staged->z, staged->coeffs, out->x, out->z, out->coeffs);
      check_launch("compact_emitted");
    }
    sync(stream);
    return out.release();
  }
  double2* sums = scratch.take<double2>(S);
  segment_sum_kernel<<<stride_grid(S), 256, 0, stream>>>(map, S, idx_a, c,
                                                         hash_seed ? 0 : 1, threshold, sums, keep);
  check_launch("segment_sum");
  exclusive_scan_u32(keep, out_pos, S, d_count + 1, scratch, stream);
  u32 U = 0;
  check_hip(hipMemcpyAsync(&U, d_count + 1, sizeof(u32), hipMemcpyDeviceToHost, stream),
            "copy device buffer to host");
  sync(stream);

  std::unique_ptr<wg_pauli_sum> out(new_pauli_sum(n, U, device));
  if (U && packed) {
    compact_kernel<<<stride_grid(S), 256, 0, stream>>>(map, S, keep, out_pos, keys_a, 1, W,
                                                       sums, out->x, out->z, out->coeffs);
    check_launch("compact");
  } else if (U && W <= 4) {
    // Interleaved mask records (16, 32 or 64 bytes per term), then the gather.
    if (!fused) {
      const std::size_t RW = record_words(W);
      rec = scratch.take<u64>(checked_mul(T, RW, "mask records"));
      mask_records_kernel<<<stride_grid(T), 256, 0, stream>>>(x, z, T, W, RW, rec);
      check_launch("mask_records");
    }
    auto gather = [&](auto kernel) {
      kernel<<<stride_grid(S), 256, 0, stream>>>(map, S, keep, out_pos, idx_a, rec, sums,
                                                 out->x, out->z, out->coeffs);
    };
    // (Two and three words took segment_emit_kernel above.)
    if (W == 1) gather(compact_records_kernel<1>); else gather(compact_records_kernel<4>);
    check_launch("compact");
  } else if (U) {
    compact_gather_kernel<<<stride_grid(S), 256, 0, stream>>>(
        map, S, keep, out_pos, idx_a, x, z, W, sums, out->x, out->z, out->coeffs);
    check_launch("compact");
  }
  sync(stream);
  return out.release();
}

void build_operand_records(const wg_pauli_sum* op, u64* rec, hipStream_t stream) {
  if (op->num_terms == 0) return;
  operand_records_kernel<<<stride_grid(op->num_terms), 256, 0, stream>>>(
      op->x, op->z, op->coeffs, op->num_terms, static_cast<u32>(op->words), rec);
  check_launch("operand_records");
}

wg_pauli_sum* simplify_products(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, double atol,
                                double rtol, int seed_mode, Scratch& scratch,
                                hipStream_t stream) {
  const std::size_t n = lhs->num_qubits, W = lhs->words;
  const std::size_t L = lhs->num_terms, R = rhs->num_terms;
  const int device = lhs->device;
  if (W < 1 || W > 4) fail_runtime("product simplify handles 1 to 4 mask words");
  const std::size_t T = checked_mul(L, R, "matmul products");
  if (T == 0) return new_pauli_sum(n, 0, device);
  if (T >= 0xFFFFFFFFULL) fail_invalid("simplify supports at most 2^32 - 2 terms per call");
  const bool packed = W == 1 && n <= 32;
  const std::size_t RW = product_record_words(static_cast<u32>(W));
  u64* lrec = scratch.take<u64>(checked_mul(L, RW, "operand records"));
  u64* rrec = scratch.take<u64>(checked_mul(R, RW, "operand records"));
  build_operand_records(lhs, lrec, stream);
  build_operand_records(rhs, rrec, stream);
  const ProductTerms P{lrec, rrec, static_cast<u32>(R), static_cast<u32>(W)};

  u64* keys = scratch.take<u64>(T);
  u32* idx = scratch.take<u32>(T);
  // Statistics: OR of the 2W key words, AND of them, then max |c| (bits),
  // per block; max |c| only matters when rtol != 0.
  const u32 chunks = static_cast<u32>((R + kProductChunk - 1) / kProductChunk);
  const std::size_t units = checked_mul(L, chunks, "matmul grid");
  const unsigned grid = grid_for(units, 1);
  const u32 ns = static_cast<u32>(4 * W + 1);
  u64* block_stats = scratch.take<u64>(checked_mul(units, ns, "product statistics"));
  {
    ProfileScope prof("product_keys", stream);
    auto launch = [&](auto kernel) {
      kernel<<<grid, kProductKeyBlock, 0, stream>>>(P, chunks, block_stats, keys, idx);
    };
    const bool need_max = rtol != 0.0;
    if (packed) {
      if (need_max) launch(product_keys_kernel<1, true, true>);
      else launch(product_keys_kernel<1, true, false>);
    } else if (need_max) {
      switch (W) {
        case 1: launch(product_keys_kernel<1, false, true>); break;
        case 2: launch(product_keys_kernel<2, false, true>); break;
        case 3: launch(product_keys_kernel<3, false, true>); break;
        default: launch(product_keys_kernel<4, false, true>); break;
      }
    } else {
      switch (W) {
        case 1: launch(product_keys_kernel<1, false, false>); break;
        case 2: launch(product_keys_kernel<2, false, false>); break;
        case 3: launch(product_keys_kernel<3, false, false>); break;
        default: launch(product_keys_kernel<4, false, false>); break;
      }
    }
    check_launch("product_keys");
  }
  return simplify_product_keys(P, n, T, device, keys, idx, block_stats, units, atol, rtol,
                               seed_mode, scratch, stream);
}

wg_pauli_sum* simplify_product_keys(const ProductTerms& P, std::size_t n, std::size_t T,
                                    int device, u64* keys, u32* idx, const u64* block_stats,
                                    std::size_t units, double atol, double rtol, int seed_mode,
                                    Scratch& scratch, hipStream_t stream) {
  const std::size_t W = P.W;
  const bool packed = W == 1 && n <= 32;
  const u32 ns = static_cast<u32>(4 * W + 1);
  std::vector<unsigned long long> stat(ns);
  u64* dstat = scratch.take<u64>(ns);
  u64* keys_a = keys;
  u32* idx_a = idx;
  u32* flags = scratch.take<u32>(T);
  u64* keys_b = scratch.take<u64>(T);
  u32* idx_b = scratch.take<u32>(T);
  stats_reduce_kernel<<<ns, 1024, 0, stream>>>(block_stats, static_cast<u32>(units),
                                               static_cast<u32>(2 * W), dstat);
  check_launch("stats_reduce");
  check_hip(hipMemcpyAsync(stat.data(), dstat, ns * sizeof(u64), hipMemcpyDeviceToHost, stream),
            "copy device buffer to host");
  sync(stream);
  const std::vector<unsigned long long> ors(stat.begin(), stat.begin() + 2 * W);
  const std::vector<unsigned long long> ands(stat.begin() + 2 * W, stat.begin() + 4 * W);
  double max_abs = 0.0;
  std::memcpy(&max_abs, &stat[4 * W], sizeof(double));
  const double threshold = atol + rtol * max_abs;

  bool hash_seed;
  if (seed_mode < 0) {
    // The reference's rule samples the first 4,096 products.
    const u32 count = static_cast<u32>(std::min<std::size_t>(T, 4096));
    u64* sx = scratch.take<u64>(static_cast<std::size_t>(count) * W);
    u64* sz = scratch.take<u64>(static_cast<std::size_t>(count) * W);
    product_sample_kernel<<<(count + 255) / 256, 256, 0, stream>>>(P, count, sx, sz);
    check_launch("product_sample");
    hash_seed = simplify_takes_hash_path(sx, sz, T, W, stream);
  } else {
    hash_seed = seed_mode != 0;
  }

  u32* d_count = scratch.take<u32>(2);
  u32 S = 0;
  SegMap map;
  if (packed) {
    const u64 kor = (ors[0] << 32) | ors[1];
    const u64 kand = (ands[0] << 32) | ands[1];
    radix_sort(keys_a, idx_a, keys_b, idx_b, T, 1, digit_passes({kor ^ kand}), scratch, stream,
               "radix_scatter", &P, true);
    map = sorted_segments(P, 1, T, keys_a, idx_a, ors, ands, flags, d_count, &S, scratch, stream);
  } else {
    radix_sort(keys_a, idx_a, keys_b, idx_b, T, 1, digit_passes({ors[0] ^ ands[0]}), scratch,
               stream, "radix_scatter", &P, false);
    map = sorted_segments(P, 2 * W, T, keys_a, idx_a, ors, ands, flags, d_count, &S, scratch,
                          stream);
  }
  clear_count_kernel<<<1, 1, 0, stream>>>(d_count + 1);
  check_launch("clear_count");

  // Sums and output at each segment's position; compacted only when some
  // segment is dropped.
  u32* keep = scratch.take<u32>(S);
  std::unique_ptr<wg_pauli_sum> staged(new_pauli_sum(n, S, device));
  {
    ProfileScope prof("product_emit", stream);
    auto emit = [&](auto kernel) {
      kernel<<<stride_grid(S), 256, 0, stream>>>(map, S, idx_a, P, hash_seed ? 0 : 1,
                                                 threshold, staged->x, staged->z, staged->coeffs,
                                                 keep, d_count + 1);
    };
    switch (W) {
      case 1: emit(product_emit_kernel<1>); break;
      case 2: emit(product_emit_kernel<2>); break;
      case 3: emit(product_emit_kernel<3>); break;
      default: emit(product_emit_kernel<4>); break;
    }
    check_launch("product_emit");
  }
  u32 U = 0;
  check_hip(hipMemcpyAsync(&U, d_count + 1, sizeof(u32), hipMemcpyDeviceToHost, stream),
            "copy device buffer to host");
  sync(stream);
  if (U == S) return staged.release();
  u32* out_pos = scratch.take<u32>(S);
  exclusive_scan_u32(keep, out_pos, S, nullptr, scratch, stream);
  std::unique_ptr<wg_pauli_sum> out(new_pauli_sum(n, U, device));
  if (U) {
    compact_emitted_kernel<<<stride_grid(S), 256, 0, stream>>>(
        keep, out_pos, S, W, staged->x, staged->z, staged->coeffs, out->x, out->z, out->coeffs);
    check_launch("compact_emitted");
  }
  sync(stream);
  return out.release();
}

}  // namespace wg

using namespace wg;

extern "C" {

int wg_simplify(const wg_pauli_sum* op, double atol, double rtol, void* stream_ptr,
                wg_pauli_sum** out) {
  return guard([&] {
    *out = nullptr;
    require_op(op);
    // Reference src/hip/device_pauli_sum.hip.hpp:47-54.
    if (!std::isfinite(atol) || !std::isfinite(rtol) || atol < 0.0 || rtol < 0.0) {
      fail_invalid("simplify tolerances must be finite and non-negative");
    }
    DeviceScope scope(op->device);
    hipStream_t stream = as_stream(stream_ptr);
    Scratch scratch(op->device);
    *out = simplify_device(op->x, op->z, op->coeffs, op->num_qubits, op->num_terms, op->device,
                           atol, rtol, -1, scratch, stream);
  });
}

}  // extern "C"

// Statevector expectation E = <psi|H|psi> and its adjoint.
//
// Reference semantics (src/expectation.cpp:37-155): for P = i^{|x&z|} X^x Z^z,
//   <P> = sum_{b : b & pivot = 0} 2 Re( i^{|x&z|} (-1)^{|z&b|} conj(psi[b^x]) psi[b] )
// with pivot any set bit of x (the pair {b, b^x} contributes a value and its
// conjugate), and <P> = sum_b (-1)^{|z&b|} |psi[b]|^2 for x = 0.
//
// Device plan (built once per operator on the host, cached on the handle):
//   * distinct (x, z) keys, with every term mapped to its key;
//   * keys grouped by x-mask: a group's keys share the amplitude products
//     a_b = conj(psi[b^x]) psi[b], so one sweep serves all of its z-masks;
//     "items" hold up to kZ keys of one group;
//   * partial sums S_z = sum_b (-1)^{|z&b|} a_b per (item, block) go to a
//     partial buffer; a second kernel reduces them in a fixed order and a third
//     forms E = sum_t c_t <P_key(t)> in a fixed tree order, so repeated calls
//     give identical bits.
//
// Two sweeps over psi:
//   * direct (small states): every item streams psi[b], psi[b^x] from global
//     memory; the whole state is cache resident.
//   * tiled (large states): psi is cut into tiles that are cosets of a subspace
//     U of the index space (the low 3 bits, some further coordinate bits and
//     some generator masks); a block stages one tile in LDS and each of its
//     waves applies one item whose x lies in U. A pass covers every x-group in
//     U with one read of psi from HBM, so the HBM traffic is (#passes) x |psi|
//     instead of (#x-groups) x |psi|. Passes are chosen greedily on the host.
//
// Adjoint: sigma = H psi = dE/dpsi* with (P psi)[b ^ x] = i^{|x&z|}
// (-1)^{|z&b|} psi[b], and dE/dc_k = <P_k>, in one fused sweep: each thread
// owns a few output amplitudes, gathers psi[o ^ x] once per x-group and applies
// all of the group's keys to sigma, sigma^dagger and the per-key partial
// sums <P> = sum_o conj(psi[o]) (P psi)[o].

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <utility>
#include <numeric>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "wg_expect.hpp"
#include "wg_internal.hpp"

namespace wg {

namespace {

constexpr int kBlock = 256;
constexpr int kTileBlock = 512;
constexpr int kTileWaves = kTileBlock / 64;
// States with at least this many qubits use the tiled sweep, unless the plan
// needs more passes (launches) than this.
constexpr std::size_t kTiledMinQubits = 18;
constexpr std::size_t kMaxTiledPasses = 256;
// A pass (one launch, one read of psi) is worth it while its launch cost
// stays small against the work it replaces: allow up to one pass per
// 2^kPassPairsLog amplitude pairs of direct-sweep work (G_x * 2^(n-1) pairs),
// and never fewer than kMaxTiledPasses.
constexpr int kPassPairsLog = 23;

std::size_t max_tiled_passes(std::size_t n, std::size_t groups) {
  const int shift = static_cast<int>(n) - 1 - kPassPairsLog;
  const std::size_t by_work = shift >= 0 ? (groups << std::min(shift, 20)) : (groups >> -shift);
  return std::max(kMaxTiledPasses, by_work);
}
// Diagonal keys use the Walsh-Hadamard sweep beyond this many items' worth.
constexpr int kDiagWhtMinItems = 4;
constexpr u32 kDiagWhtBlocks = 1024;
// Off-diagonal items of tiled passes use the owner sweep (one LDS read per
// amplitude pair) instead of the per-wave item sweep.
constexpr bool kOwnerSweep = true;

__device__ inline u64 insert_zero(u64 v, u64 pivot) {
  const u64 low = pivot - 1;
  return ((v & ~low) << 1) | (v & low);
}

// ------------------------------------------------------------ direct sweep

__global__ __launch_bounds__(kBlock) void expect_direct_kernel(const double2* __restrict__ psi,
                                                               u64 len,
                                                               const ExpItem* __restrict__ items,
                                                               u32 blocks_per_item,
                                                               double2* __restrict__ partials) {
  __shared__ double2 lds[kBlock / 64][kZ];
  const u32 item_id = blockIdx.x / blocks_per_item;
  const u32 chunk = blockIdx.x % blocks_per_item;
  const ExpItem& it = items[item_id];
  const u64 x = it.x;
  const int nz = it.nz;
  const u64 range = x ? (len >> 1) : len;
  const u64 per = (range + blocks_per_item - 1) / blocks_per_item;
  const u64 begin = static_cast<u64>(chunk) * per;
  const u64 end = begin + per < range ? begin + per : range;

  double ar[kZ], ai[kZ];
#pragma unroll
  for (int j = 0; j < kZ; ++j) {
    ar[j] = 0.0;
    ai[j] = 0.0;
  }
  if (x) {
    const u64 pivot = x & (0 - x);
    for (u64 v = begin + threadIdx.x; v < end; v += kBlock) {
      const u64 b = insert_zero(v, pivot);
      const double2 a = psi[b];
      const double2 t = psi[b ^ x];
      // conj(t) * a
      const double re = fma(t.x, a.x, t.y * a.y);
      const double im = fma(t.x, a.y, -(t.y * a.x));
#pragma unroll
      for (int j = 0; j < kZ; ++j) {
        if (j < nz) {
          const bool odd = __popcll(it.z[j] & b) & 1;
          ar[j] += odd ? -re : re;
          ai[j] += odd ? -im : im;
        }
      }
    }
  } else {
    for (u64 b = begin + threadIdx.x; b < end; b += kBlock) {
      const double2 a = psi[b];
      const double p = fma(a.x, a.x, a.y * a.y);
#pragma unroll
      for (int j = 0; j < kZ; ++j) {
        if (j < nz) {
          const bool odd = __popcll(it.z[j] & b) & 1;
          ar[j] += odd ? -p : p;
        }
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kZ; ++j) {
    if (j < nz) {
      double r = ar[j], i = ai[j];
      for (int d = 32; d > 0; d >>= 1) {
        r += __shfl_xor(r, d, 64);
        i += __shfl_xor(i, d, 64);
      }
      if (lane == 0) lds[wave][j] = make_double2(r, i);
    }
  }
  __syncthreads();
  if (threadIdx.x < static_cast<unsigned>(nz)) {
    double r = 0.0, i = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) {
      r += lds[w][threadIdx.x].x;
      i += lds[w][threadIdx.x].y;
    }
    partials[(static_cast<std::size_t>(item_id) * blocks_per_item + chunk) * kZ + threadIdx.x] =
        make_double2(r, i);
  }
}

// ------------------------------------------------------------ tiled sweep

// Every tiled pass uses tiles of kTilePoints = 2^kTileMaxLog amplitudes: the
// planner pads each subspace with coordinate bits up to that dimension.
constexpr u32 kTilePoints = 1U << kTileMaxLog;
constexpr int kOffSteps = kTilePoints / 2 / 64;  // pair steps per wave (b & pivot = 0)
constexpr int kDiagSteps = kTilePoints / 64;     // point steps per wave (x = 0)

// The bits of v at the positions of `mask`, packed into the low bits in
// ascending order (software pext); parity(z & tile_rep(t, mask)) equals
// parity(compress_bits(z, mask) & t).
__device__ inline u64 compress_bits(u64 v, u64 mask) {
  u64 out = 0;
  int k = 0;
  while (mask) {
    const u64 low = mask & (0 - mask);
    if (v & low) out |= 1ULL << k;
    ++k;
    mask ^= low;
  }
  return out;
}

// Coset representative of tile t: the bits of t deposited into the free
// positions (ascending), i.e. a software pdep over the pass's free mask.
__device__ inline u64 tile_rep(u64 t, u64 free_mask) {
  u64 rep = 0;
  while (free_mask) {
    const u64 low = free_mask & (0 - free_mask);
    if (t & 1ULL) rep |= low;
    t >>= 1;
    free_mask ^= low;
  }
  return rep;
}

__device__ __forceinline__ double2 lds_at(const double2* tile, u32 byte_off) {
  return *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(tile) + byte_off);
}

// Thread t stages points u = t + i * kTileBlock (i < 8) of a tile, i.e. the
// segments (t >> 3) + 64 i at offset t & 7. Segment offsets are linear over
// GF(2) in the segment index, so seg_off[(t >> 3) + 64 i] = seg_off[t >> 3] ^
// (the offsets of the set bits of 64 i): one per-thread value and three
// uniform ones instead of eight.
constexpr u32 kSegStride = kTileBlock >> kTileLowLog;
static_assert(kTilePoints / kTileBlock == 8 && kSegStride == 64,
              "staging is written for 8 points per thread and 64 segments per block");

// The staged points live in eight named registers (an array or a struct
// passed by reference may be left in scratch memory by the compiler).
#define WG_FETCH_TILE(rep)                                                  \
  do {                                                                      \
    const u64 r0_ = (rep) ^ sb_w, r1_ = r0_ ^ sb_e0, r2_ = r0_ ^ sb_e1;      \
    const u64 r3_ = r1_ ^ sb_e1;                                            \
    st0 = psi[r0_ | lo];                                                    \
    st1 = psi[r1_ | lo];                                                    \
    st2 = psi[r2_ | lo];                                                    \
    st3 = psi[r3_ | lo];                                                    \
    st4 = psi[(r0_ ^ sb_e2) | lo];                                          \
    st5 = psi[(r1_ ^ sb_e2) | lo];                                          \
    st6 = psi[(r2_ ^ sb_e2) | lo];                                          \
    st7 = psi[(r3_ ^ sb_e2) | lo];                                          \
  } while (0)

// The block's tile loop: stage tile t in LDS, prefetch tile t + tile_blocks
// into registers, let this wave's work consume the staged tile. Every wave of
// the block runs the same number of iterations (same barriers), whatever its
// work type.
template <class Work>
__device__ __forceinline__ void sweep_tiles(const double2* __restrict__ psi,
                                            const PassDesc* __restrict__ pd, u32 wave, u64 t0,
                                            u32 tile_blocks, double2* __restrict__ tile,
                                            Work& work) {
  const u64 free_mask = pd->free_mask;
  const u64 ntiles = 1ULL << (pd->n - kTileMaxLog);
  (void)wave;
  const u64 sb_w = pd->seg_off[threadIdx.x >> kTileLowLog], sb_e0 = pd->seg_off[kSegStride],
            sb_e1 = pd->seg_off[2 * kSegStride], sb_e2 = pd->seg_off[4 * kSegStride];
  const u32 lo = threadIdx.x & ((1U << kTileLowLog) - 1U);
  double2 st0, st1, st2, st3, st4, st5, st6, st7;
  u64 t = t0;
  if (t < ntiles) WG_FETCH_TILE(tile_rep(t, free_mask));
  for (; t < ntiles; t += tile_blocks) {
    __syncthreads();
    {
      const u32 u = threadIdx.x;
      tile[u + 0 * kTileBlock] = st0;
      tile[u + 1 * kTileBlock] = st1;
      tile[u + 2 * kTileBlock] = st2;
      tile[u + 3 * kTileBlock] = st3;
      tile[u + 4 * kTileBlock] = st4;
      tile[u + 5 * kTileBlock] = st5;
      tile[u + 6 * kTileBlock] = st6;
      tile[u + 7 * kTileBlock] = st7;
    }
    __syncthreads();
    const u64 tn = t + tile_blocks;
    // Prefetch the next tile into registers while this one is consumed.
    if (tn < ntiles) WG_FETCH_TILE(tile_rep(tn, free_mask));
    work(tile, t);
  }
}
#undef WG_FETCH_TILE

struct IdleWork {
  __device__ void operator()(const double2*, u64) {}
};
// (Work functors get the tile index t; the tile's coset representative is
// tile_rep(t, free_mask).)

// Writes a wave's NZ (sum over lanes) partial values for one tile block; the
// per-lane sign parity(zl_j & f_lane) is applied first.
template <int NZ>
__device__ __forceinline__ void write_partials(const double (&ar)[NZ], const double (&ai)[NZ],
                                               u32 lane_par, u32 lane, double2* __restrict__ row) {
#pragma unroll
  for (int j = 0; j < NZ; ++j) {
    const bool neg = (lane_par >> j) & 1U;
    double r = neg ? -ar[j] : ar[j];
    double i = neg ? -ai[j] : ai[j];
    for (int d = 32; d > 0; d >>= 1) {
      r += __shfl_xor(r, d, 64);
      i += __shfl_xor(i, d, 64);
    }
    if (lane == 0) row[j] = make_double2(r, i);
  }
}

// One off-diagonal item on one wave. Lane points u = f ^ g(s), s < kOffSteps,
// with f = insert_zero(lane, pivot) and g(s) = insert_zero(s << 6, pivot):
// they enumerate the pair representatives b & pivot = 0 of the tile. Since
// parity(zl & u) = parity(zl & f) ^ parity(zl & g(s)), the lane part is a sign
// applied once at the end and the step part (bit s of smask) is wave-uniform,
// so each (pair, key) update is one FMA by a uniform +-1 per component. Steps
// go in batches of 4: g(s0 + s) = g(s0) ^ g(s) for s < 4 (disjoint bits).
template <int NZ>
struct OffDiagWork {
  static constexpr int B = NZ <= 4 ? 4 : 2;
  u64 zt[NZ];       // z compressed to the free (tile index) bits: parity(z & rep(t)) = parity(zt & t)
  u32 smask[NZ];
  u32 a16[B], b16[B];  // (f ^ g(s)) * 16 and (f ^ xl ^ g(s)) * 16 for s < B
  u64 pbit;
  u32 lane_par;
  double ar[NZ], ai[NZ];

  __device__ OffDiagWork(const ExpItem& it, u64 free_mask, u32 lane) {
    const u32 xl = static_cast<u32>(it.xl);
    pbit = it.pbit;
    const u32 f = static_cast<u32>(insert_zero(lane, pbit));
#pragma unroll
    for (int s = 0; s < B; ++s) {
      const u32 g = static_cast<u32>(insert_zero(static_cast<u64>(s) << 6, pbit));
      a16[s] = (f ^ g) << 4;
      b16[s] = (f ^ g ^ xl) << 4;
    }
    lane_par = 0;
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
      zt[j] = compress_bits(it.z[j], free_mask);
      const u32 zlj = static_cast<u32>(it.zl[j]);
      lane_par |= static_cast<u32>(__popc(zlj & f) & 1) << j;
      u32 m = 0;
      for (int s = 0; s < kOffSteps; ++s) {
        const u32 g = static_cast<u32>(insert_zero(static_cast<u64>(s) << 6, pbit));
        m |= static_cast<u32>(__popc(zlj & g) & 1) << s;
      }
      smask[j] = m;
      ar[j] = 0.0;
      ai[j] = 0.0;
    }
  }

  __device__ __forceinline__ void operator()(const double2* __restrict__ tile, u64 t) {
    u32 sm[NZ];
#pragma unroll
    for (int j = 0; j < NZ; ++j) sm[j] = (__popcll(zt[j] & t) & 1) ? ~smask[j] : smask[j];
    constexpr int kUnroll = NZ <= 2 ? 2 : 1;
#pragma unroll kUnroll
    for (int s0 = 0; s0 < kOffSteps; s0 += B) {
      const u32 g0 = static_cast<u32>(insert_zero(static_cast<u64>(s0) << 6, pbit)) << 4;
      double2 a[B], b[B];
#pragma unroll
      for (int s = 0; s < B; ++s) {
        a[s] = lds_at(tile, a16[s] ^ g0);
        b[s] = lds_at(tile, b16[s] ^ g0);
      }
#pragma unroll
      for (int s = 0; s < B; ++s) {
        const double re = fma(b[s].x, a[s].x, b[s].y * a[s].y);
        const double im = fma(b[s].x, a[s].y, -(b[s].y * a[s].x));
#pragma unroll
        for (int j = 0; j < NZ; ++j) {
          const double sg = ((sm[j] >> (s0 + s)) & 1U) ? -1.0 : 1.0;
          ar[j] = fma(sg, re, ar[j]);
          ai[j] = fma(sg, im, ai[j]);
        }
      }
    }
  }
};

// One off-diagonal item whose step signs take 2^D classes (ExpItem::cls_log):
// steps are enumerated class by class, so a step only adds its pair product
// to its class sum Q[class] (no per-step sign work); at the end of a tile,
// key j gains (-1)^{parity(z_j & rep)} sum_a (-1)^{lam_j . a} Q[a]. Lane
// points and the lane sign are as in OffDiagWork. Reads: own point at
// f * 16 + G(e), partner at (f ^ x_lane) * 16 + (G(e) ^ xh16), where G(e) is
// the uniform step offset (XOR of the enumeration basis offsets) and xh16 the
// non-lane bits of x (pivot included).
template <int D, int NZK>
struct OffDiagClassWork {
  static constexpr int NC = 1 << D;
  static constexpr int B = NZK <= 2 ? 4 : 2;
  int nz;
  u32 zt[NZK];  // tile indices fit 32 bits on any state a GPU holds
  u32 lam[NZK];
  u32 gb[5];
  u32 xh16;
  u32 f16, fx16;
  u32 lane_par;
  double ar[NZK], ai[NZK];

  __device__ OffDiagClassWork(const ExpItem& it, u64 free_mask, u32 lane) {
    nz = it.nz;
    const u64 pbit = it.pbit;
    const u32 xl = static_cast<u32>(it.xl);
    const u32 lanepos = static_cast<u32>(insert_zero(63, pbit));
    const u32 f = static_cast<u32>(insert_zero(lane, pbit));
    f16 = f << 4;
    fx16 = (f ^ (xl & lanepos)) << 4;
    xh16 = (xl & ~lanepos) << 4;
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      gb[l] = static_cast<u32>(insert_zero(static_cast<u64>(it.ebasis[l]) << 6, pbit)) << 4;
    }
    lane_par = 0;
#pragma unroll
    for (int j = 0; j < NZK; ++j) {
      const bool on = j < nz;
      zt[j] = on ? static_cast<u32>(compress_bits(it.z[j], free_mask)) : 0;
      lam[j] = on ? it.lam[j] : 0;
      lane_par |= static_cast<u32>(on && (__popc(static_cast<u32>(it.zl[j]) & f) & 1)) << j;
      ar[j] = 0.0;
      ai[j] = 0.0;
    }
  }

  __device__ __forceinline__ void operator()(const double2* __restrict__ tile, u64 t64) {
    // Steps of class c are e = (c << (5 - D)) | k, k < KS: offset
    // G(e) = G_class(c) ^ G_kernel(k), G_kernel linear in the bits of k. The
    // k loop is a runtime loop so that the 32 step offsets are formed per
    // batch instead of being held in scalar registers. Each class sum goes to
    // the keys as soon as it is complete.
    constexpr int KS = kOffSteps >> D;
    constexpr int kLogB = B == 4 ? 2 : 1;
    const u32 t = static_cast<u32>(t64);
    u32 flip = 0;  // bit j: parity(z_j & rep)
#pragma unroll
    for (int j = 0; j < NZK; ++j) flip |= static_cast<u32>(__popc(zt[j] & t) & 1) << j;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      u32 gc = 0;
#pragma unroll
      for (int i = 0; i < D; ++i) {
        if ((c >> i) & 1) gc ^= gb[5 - D + i];
      }
      double cr[B], ci[B];
#pragma unroll
      for (int s = 0; s < B; ++s) {
        cr[s] = 0.0;
        ci[s] = 0.0;
      }
#pragma unroll 1
      for (int k = 0; k < KS; k += B) {
        u32 gk = gc;
        for (int l = kLogB; l < 5 - D; ++l) {
          if ((k >> l) & 1) gk ^= gb[l];
        }
        double2 a[B], b[B];
#pragma unroll
        for (int s = 0; s < B; ++s) {
          u32 g = gk;
#pragma unroll
          for (int l = 0; l < kLogB; ++l) {
            if ((s >> l) & 1) g ^= gb[l];
          }
          a[s] = lds_at(tile, f16 + g);
          b[s] = lds_at(tile, fx16 + (g ^ xh16));
        }
#pragma unroll
        for (int s = 0; s < B; ++s) {
          cr[s] = fma(b[s].x, a[s].x, fma(b[s].y, a[s].y, cr[s]));
          ci[s] = fma(b[s].x, a[s].y, fma(-b[s].y, a[s].x, ci[s]));
        }
      }
#pragma unroll
      for (int s = 1; s < B; ++s) {
        cr[0] += cr[s];
        ci[0] += ci[s];
      }
#pragma unroll
      for (int j = 0; j < NZK; ++j) {
        if (j < nz) {
          const bool neg = ((flip >> j) ^ static_cast<u32>(__popc(lam[j] & static_cast<u32>(c)))) & 1U;
          ar[j] += neg ? -cr[0] : cr[0];
          ai[j] += neg ? -ci[0] : ci[0];
        }
      }
    }
  }
};

// One diagonal item (x = 0) on one wave: lane points u = lane + 64 s.
template <int NZ>
struct DiagWork {
  u64 zt[NZ];
  u64 smask[NZ];
  u32 lane_par;
  u32 f16;
  double ar[NZ], ai[NZ];

  __device__ DiagWork(const ExpItem& it, u64 free_mask, u32 lane) {
    f16 = lane << 4;
    lane_par = 0;
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
      zt[j] = compress_bits(it.z[j], free_mask);
      const u32 zlj = static_cast<u32>(it.zl[j]);
      lane_par |= static_cast<u32>(__popc(zlj & lane) & 1) << j;
      u64 m = 0;
      for (int s = 0; s < kDiagSteps; ++s) m |= static_cast<u64>(__popc(zlj & (static_cast<u32>(s) << 6)) & 1) << s;
      smask[j] = m;
      ar[j] = 0.0;
      ai[j] = 0.0;
    }
  }

  __device__ __forceinline__ void operator()(const double2* __restrict__ tile, u64 t) {
    u64 sm[NZ];
#pragma unroll
    for (int j = 0; j < NZ; ++j) sm[j] = (__popcll(zt[j] & t) & 1) ? ~smask[j] : smask[j];
    constexpr int B = NZ <= 4 ? 4 : 2;
    constexpr int kUnroll = NZ <= 2 ? 2 : 1;
#pragma unroll kUnroll
    for (int s0 = 0; s0 < kDiagSteps; s0 += B) {
      double2 a[B];
#pragma unroll
      for (int s = 0; s < B; ++s) a[s] = lds_at(tile, f16 + (static_cast<u32>(s0 + s) << 10));
#pragma unroll
      for (int s = 0; s < B; ++s) {
        const double p = fma(a[s].x, a[s].x, a[s].y * a[s].y);
#pragma unroll
        for (int j = 0; j < NZ; ++j) {
          const double sg = ((sm[j] >> (s0 + s)) & 1ULL) ? -1.0 : 1.0;
          ar[j] = fma(sg, p, ar[j]);
        }
      }
    }
  }
};

template <int D, int NZK>
__device__ __forceinline__ void run_class_item(const double2* __restrict__ psi,
                                               const PassDesc* __restrict__ pd, const ExpItem& it,
                                               u32 wave, u32 lane, u64 t0, u32 tile_blocks,
                                               double2* __restrict__ tile,
                                               double2* __restrict__ row) {
  OffDiagClassWork<D, NZK> w(it, pd->free_mask, lane);
  sweep_tiles(psi, pd, wave, t0, tile_blocks, tile, w);
#pragma unroll
  for (int j = 0; j < NZK; ++j) {
    if (j < w.nz) {
      const bool neg = (w.lane_par >> j) & 1U;
      double r = neg ? -w.ar[j] : w.ar[j];
      double i = neg ? -w.ai[j] : w.ai[j];
      for (int d = 32; d > 0; d >>= 1) {
        r += __shfl_xor(r, d, 64);
        i += __shfl_xor(i, d, 64);
      }
      if (lane == 0) row[j] = make_double2(r, i);
    }
  }
}

template <int NZ, bool kDiag>
__device__ __forceinline__ void run_item(const double2* __restrict__ psi,
                                         const PassDesc* __restrict__ pd, const ExpItem& it,
                                         u32 wave, u32 lane, u64 t0, u32 tile_blocks,
                                         double2* __restrict__ tile, double2* __restrict__ row) {
  if constexpr (kDiag) {
    DiagWork<NZ> w(it, pd->free_mask, lane);
    sweep_tiles(psi, pd, wave, t0, tile_blocks, tile, w);
    write_partials<NZ>(w.ar, w.ai, w.lane_par, lane, row);
  } else {
    OffDiagWork<NZ> w(it, pd->free_mask, lane);
    sweep_tiles(psi, pd, wave, t0, tile_blocks, tile, w);
    write_partials<NZ>(w.ar, w.ai, w.lane_par, lane, row);
  }
}

// Sweep variants: one kernel instance per kind of item, so that each
// instance's registers are allocated for one work type.
//   [0, 8):   off-diagonal, per-step signs, nz = v + 1
//   [8, 16):  diagonal, nz = v - 7
//   [16, 24): off-diagonal with step classes, D = (v - 16) / 2, up to 2
//             (even v) or kZ (odd v) keys
constexpr int kVariants = 24;

int sweep_variant(const ExpItem& it) { return it.variant; }

// Grid: item_groups * tile_blocks blocks. Block -> (item group, tile block)
// such that the blocks sharing a tile block differ by multiples of 8 in their
// linear id: they are dispatched to the same XCD close together, so the later
// reads of a tile hit that XCD's L2 instead of HBM. Wave w of a block applies
// item (group * 8 + w) to every tile of its tile block; every item of a launch
// is of sweep variant V.
// Two blocks per CU (LDS 2 x 64 KiB) need 4 waves per SIMD: <= 128 VGPRs.
template <int V>
__global__ __launch_bounds__(kTileBlock) __attribute__((amdgpu_waves_per_eu(4)))
void expect_tiled_kernel(
    const double2* __restrict__ psi, const PassDesc* __restrict__ pd,
    const ExpItem* __restrict__ items,
    u32 num_items, u32 item_groups, u32 tile_blocks, u32 partial_base,
    double2* __restrict__ partials) {
  __shared__ double2 tile[kTilePoints];
  const u32 b = blockIdx.x;
  const u32 xcd = b & 7U;
  const u32 slot = b >> 3;
  const u32 ig = slot % item_groups;
  const u32 tb = (slot / item_groups) * 8U + xcd;
  const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const u32 lane = threadIdx.x & 63U;
  const u32 item_id = ig * kTileWaves + wave;
  if (item_id >= num_items) {
    IdleWork w;
    sweep_tiles(psi, pd, wave, tb, tile_blocks, tile, w);
    return;
  }
  const ExpItem& it = items[item_id];
  double2* row = partials + (static_cast<std::size_t>(partial_base) +
                             static_cast<std::size_t>(item_id) * tile_blocks + tb) * kZ;
  if constexpr (V < 8) {
    run_item<V + 1, false>(psi, pd, it, wave, lane, tb, tile_blocks, tile, row);
  } else if constexpr (V < 16) {
    run_item<V - 7, true>(psi, pd, it, wave, lane, tb, tile_blocks, tile, row);
  } else {
    run_class_item<(V - 16) / 2, ((V - 16) & 1) ? kZ : 2>(psi, pd, it, wave, lane, tb,
                                                           tile_blocks, tile, row);
  }
}

using TiledKernel = void (*)(const double2*, const PassDesc*, const ExpItem*, u32, u32, u32, u32,
                             double2*);

template <int... V>
constexpr std::array<TiledKernel, sizeof...(V)> tiled_kernel_table(std::integer_sequence<int, V...>) {
  return {&expect_tiled_kernel<V>...};
}

const std::array<TiledKernel, kVariants> kTiledKernels =
    tiled_kernel_table(std::make_integer_sequence<int, kVariants>{});

// ------------------------------------------------------------ owner sweep

// Off-diagonal items of a pass in "bundles" of item slots (2 or 4 keys per
// slot); one launch per pass runs all its bundles, so the bundles that share
// a tile read it from L2. A 512-thread block owns one bundle; thread tid keeps
// its points u = tid + 512 i (i < 8) of each tile in registers and, for every
// slot, evaluates the 4 pairs whose representative it owns, with one LDS read
// (the partner) per pair:
//  * register pivot (local bit 11 of x set in the bundle's layout):
//    representatives are registers i = j < 4 and partners
//    (tid ^ x[0..8], j ^ x[9..11]) lie in registers 4..7, so only that half
//    of the tile goes to LDS (owner_half_sweep);
//  * lane pivot (x inside local bits 0..2): representatives are registers
//    i = j + 4 h, h = bit `hbit` of tid, partners (tid ^ x, i); the whole tile
//    is staged (owner_lane_sweep).
// For representative j and key k the pair's sign is (-1)^{parity(z & rep(t))}
// (-1)^{c_k . j} times a per-thread sign applied once at the end (c_k: z's
// local bits 9, 10). Each key needs one component of conj(psi[b ^ x]) psi[b]
// (real for even |x & z|, imaginary for odd), which is all 2 Re(i^y S) uses.
// The bundle's per-slot and per-key words live in scalar registers for the
// whole sweep, so the tile loop issues no scalar loads (their completion
// counter is shared with the LDS reads).
//
// Bundle kinds: bit 0 odd |x & z|, bit 1 lane pivot, bits 2.. the bundle
// shape (slots, keys per slot).
constexpr int kOwnerBlock = kTileBlock;
constexpr int kOwnerWaves = kOwnerBlock / 64;
// Shapes 0..6 are register-pivot bundles (half-tile sweep), 7..10 lane-pivot
// bundles (full-tile sweep: their partners may sit in any register, and they
// also hold all 8 points and the representatives' selects, so they are smaller).
// The 11-slot shape forms its partner addresses at every tile to fit in 128
// VGPRs: it pays where each tile read costs most, states beyond the
// Infinity Cache (kOwnerWideMinQubits and up); smaller states start from the
// 9-slot shape, which keeps its addresses in registers.
constexpr int kOwnerShapes = 11;
constexpr int kOwnerShapeSlots[kOwnerShapes] = {11, 9, 4, 2, 4, 2, 1, 4, 2, 2, 1};
constexpr int kOwnerShapeKeys[kOwnerShapes] = {2, 2, 2, 2, 4, 4, 4, 2, 2, 4, 4};
constexpr int kOwnerKinds = 4 * kOwnerShapes;
constexpr int kOwnerMaxAcc = 22;
constexpr int kOwnerWideMinQubits = 25;
// Shapes available to a group of slots, largest first.
constexpr int owner_first_shape(bool wide, bool lanex, bool big) {
  return lanex ? (wide ? 9 : 7) : (wide ? 4 : (big ? 0 : 1));
}
constexpr int owner_last_shape(bool wide, bool lanex) {
  return lanex ? (wide ? 10 : 8) : (wide ? 6 : 3);
}
constexpr bool owner_kind_used(int kind) {
  const int shape = kind >> 2;
  const bool lanex = (kind >> 1) & 1;
  return lanex ? shape >= 7 : shape < 7;
}
// Owner passes give each block about kOwnerTilesPerBlock tiles: the blocks of
// the bundles that share a tile start together and stay close, so all but the
// first read of each tile hit L2 (long runs drift apart by hundreds of tiles).
// A launch keeps at least kOwnerMinBlocks blocks (two per CU).
constexpr u64 kOwnerTilesPerBlock = 64;
constexpr u64 kOwnerMinBlocks = 512;
// Partial rows of one owner pass (items x tile blocks, kZ values each) stay
// within this many rows.
constexpr u64 kOwnerMaxPassRows = u64{1} << 20;
static_assert(kTilePoints / kOwnerBlock == 8 && kOwnerWaves == 8,
              "owner staging is written for 8 points per thread and 8 waves");
static_assert(kOwnerMaxSlots == 11, "bundle slots match the largest shape");

__device__ __forceinline__ double unit_sign(u32 hi) { return __hiloint2double(static_cast<int>(hi), 0); }

// Lane-pivot bundles: the whole tile is staged in LDS (partners may lie in
// either half), the next tile prefetched into registers; two barriers per tile.
template <int M, int NZ, bool kOdd>
__device__ __forceinline__ void owner_lane_sweep(const double2* __restrict__ psi,
                                                   const PassDesc* __restrict__ pd,
                                                   const OwnerBundle& B, u32 tb, u32 tile_blocks,
                                                   u32 partial_base, double2* __restrict__ lo,
                                                   double2* __restrict__ hi,
                                                   double (*red)[kOwnerMaxAcc],
                                                   u64* __restrict__ step_mask,
                                                   double2* __restrict__ partials) {
  static_assert(M * NZ <= kOwnerMaxAcc && M <= kOwnerMaxSlots && NZ <= kOwnerMaxKeys,
                "owner bundle shape");
  const u32 tid = threadIdx.x;
  const u32 wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const u32 lane = tid & 63U;
  const u64 free_mask = pd->free_mask;
  const u64 ntiles = 1ULL << (pd->n - kTileMaxLog);
  // Segment part of this thread's points: bits 0..2 of the segment index are
  // lane bits 3..5, bits 3..5 the wave; bits 6..8 (the register index) are
  // the uniform sb_e0..sb_e2.
  u64 sb_w = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if ((tid >> (kTileLowLog + i)) & 1U) sb_w ^= B.seg_basis[i];
  }
  const u64 sb_e0 = B.seg_basis[6], sb_e1 = B.seg_basis[7], sb_e2 = B.seg_basis[8];
  const u32 lane_lo = lane & ((1U << kTileLowLog) - 1U);
  const u32 tid16 = tid << 4;
  u32 addr0[M], kw[M][NZ];
  u32 hmask = 0;  // bit m = bit hbit_m of tid
  double acc[M][NZ];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    addr0[m] = __builtin_amdgcn_readfirstlane(B.slot[m].addr0);
    hmask |= ((tid >> B.slot[m].hbit) & 1U) << m;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      kw[m][k] = __builtin_amdgcn_readfirstlane(B.slot[m].key[k].w);
      acc[m][k] = 0.0;
    }
  }
  // The block's tiles t_s = tb + tile_blocks * gray(s), s < run, in Gray-code
  // order: the coset representative of t_{s+1} is that of t_s XOR one free
  // index bit, step_mask[ctz(s + 1)] (a software pdep per tile would cost
  // ~70 scalar instructions).
  const u64 run = ntiles / tile_blocks;
  const int log_tb = __builtin_ctz(tile_blocks);
  if (tid < 32U) step_mask[tid] = tile_rep(static_cast<u64>(tile_blocks) << tid, free_mask);
  double2 st0, st1, st2, st3, st4, st5, st6, st7;
  u64 rep = tile_rep(tb, free_mask);
#define WG_OWNER_FETCH(rep)                                                 \
  do {                                                                      \
    const u64 r0_ = (rep) ^ sb_w, r1_ = r0_ ^ sb_e0, r2_ = r0_ ^ sb_e1;      \
    const u64 r3_ = r1_ ^ sb_e1;                                            \
    st0 = psi[r0_ | lane_lo];                                               \
    st1 = psi[r1_ | lane_lo];                                               \
    st2 = psi[r2_ | lane_lo];                                               \
    st3 = psi[r3_ | lane_lo];                                               \
    st4 = psi[(r0_ ^ sb_e2) | lane_lo];                                     \
    st5 = psi[(r1_ ^ sb_e2) | lane_lo];                                     \
    st6 = psi[(r2_ ^ sb_e2) | lane_lo];                                     \
    st7 = psi[(r3_ ^ sb_e2) | lane_lo];                                     \
  } while (0)
  WG_OWNER_FETCH(rep);
  for (u64 step = 0; step < run; ++step) {
    __syncthreads();
    lo[tid + 0 * kOwnerBlock] = st0;
    lo[tid + 1 * kOwnerBlock] = st1;
    lo[tid + 2 * kOwnerBlock] = st2;
    lo[tid + 3 * kOwnerBlock] = st3;
    hi[tid + 0 * kOwnerBlock] = st4;
    hi[tid + 1 * kOwnerBlock] = st5;
    hi[tid + 2 * kOwnerBlock] = st6;
    hi[tid + 3 * kOwnerBlock] = st7;
    const double2 own0 = st0, own1 = st1, own2 = st2, own3 = st3;
    const double2 own4 = st4, own5 = st5, own6 = st6, own7 = st7;
    __syncthreads();
    const u32 t32 = tb | (static_cast<u32>(step ^ (step >> 1)) << log_tb);
    if (step + 1 < run) {
      const u64 m = step_mask[__builtin_ctzll(step + 1)];
      rep ^= (static_cast<u64>(__builtin_amdgcn_readfirstlane(static_cast<u32>(m >> 32))) << 32) |
             __builtin_amdgcn_readfirstlane(static_cast<u32>(m));
      WG_OWNER_FETCH(rep);
    }
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const bool h = (hmask >> m) & 1U;
      const double2 a[4] = {h ? own4 : own0, h ? own5 : own1, h ? own6 : own2, h ? own7 : own3};
      // The partners share the representatives' half (x is inside the low bits).
      const double2* half = h ? hi : lo;
      const u32 base = tid16 ^ addr0[m];
      double2 q[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) q[j] = lds_at(half, base + (static_cast<u32>(j) << 13));
      double v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v[j] = kOdd ? fma(q[j].x, a[j].y, -(q[j].y * a[j].x)) : fma(q[j].x, a[j].x, q[j].y * a[j].y);
      }
#pragma unroll
      for (int k = 0; k < NZ; ++k) {
        const u32 w = kw[m][k];
        const u32 h0 = 0x3FF00000U ^ (static_cast<u32>(__popc(w & t32)) << 31);
        const u32 c1 = w & 0x80000000U, c0 = (w << 1) & 0x80000000U;
        double s = acc[m][k];
        s = fma(unit_sign(h0), v[0], s);
        s = fma(unit_sign(h0 ^ c0), v[1], s);
        s = fma(unit_sign(h0 ^ c1), v[2], s);
        s = fma(unit_sign(h0 ^ c0 ^ c1), v[3], s);
        acc[m][k] = s;
      }
    }
  }
#undef WG_OWNER_FETCH
  // Per-thread sign, then a fixed-order block sum per slot key.
#pragma unroll
  for (int m = 0; m < M; ++m) {
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      const OwnerKey& K = B.slot[m].key[k];
      u32 par = static_cast<u32>(__popc(K.zlt & tid)) & 1U;
      par ^= (K.info >> 2) & (hmask >> m) & 1U;
      double r = par ? -acc[m][k] : acc[m][k];
      for (int d = 32; d > 0; d >>= 1) r += __shfl_xor(r, d, 64);
      if (lane == 0) red[wave][m * NZ + k] = r;
    }
  }
  __syncthreads();
  if (tid < static_cast<u32>(M * NZ)) {
    const int m = static_cast<int>(tid) / NZ, k = static_cast<int>(tid) % NZ;
    const OwnerSlot& S = B.slot[m];
    const u32 info = S.key[k].info;
    if ((info >> 15) & 1U) {
      double r = 0.0;
      for (int w = 0; w < kOwnerWaves; ++w) r += red[w][tid];
      const std::size_t row = static_cast<std::size_t>(partial_base) +
                              static_cast<std::size_t>(S.item) * tile_blocks + tb;
      partials[row * kZ + ((info >> 4) & 15U)] = make_double2(r, r);
    }
  }
}

// Register-pivot bundles: representatives are registers 0..3 of every thread
// and every partner lies in registers 4..7 (local bit 11 set), so LDS holds
// only that half of the tile (32 KiB), double-buffered: while tile s is
// evaluated, global_load_lds fills the other half-buffer with tile s + 1 and
// the own half of tile s + 1 is loaded into registers. One barrier per tile.
// The two half-buffers are distinct LDS arrays and the tile loop is unrolled
// by two, so each step's buffers are known at compile time: the compiler then
// sees that the reads of tile s cannot alias the loads of tile s + 1 and does
// not wait for those loads before the reads (with one array and a computed
// buffer it inserts that wait, and loads would not overlap evaluation).
typedef __attribute__((address_space(3))) void lds_void;

template <int M, int NZ, bool kOdd>
__device__ __forceinline__ void owner_half_sweep(const double2* __restrict__ psi,
                                                 const PassDesc* __restrict__ pd,
                                                 const OwnerBundle& B, u32 tb, u32 tile_blocks,
                                                 u32 partial_base, double2* __restrict__ buf0,
                                                 double2* __restrict__ buf1,
                                                 double (*red)[kOwnerMaxAcc],
                                                 u64* __restrict__ step_mask,
                                                 double2* __restrict__ partials) {
  static_assert(M * NZ <= kOwnerMaxAcc && M <= kOwnerMaxSlots && NZ <= kOwnerMaxKeys,
                "owner bundle shape");
  const u32 tid = threadIdx.x;
  const u32 wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const u32 lane = tid & 63U;
  const u64 free_mask = pd->free_mask;
  const u64 ntiles = 1ULL << (pd->n - kTileMaxLog);
  // Segment part of this thread's points: bits 0..2 of the segment index are
  // lane bits 3..5, bits 3..5 the wave; bits 6..8 (the register index) are
  // the uniform sb_e0..sb_e2.
  u64 sb_w = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if ((tid >> (kTileLowLog + i)) & 1U) sb_w ^= B.seg_basis[i];
  }
  const u64 sb_e0 = B.seg_basis[6], sb_e1 = B.seg_basis[7], sb_e2 = B.seg_basis[8];
  const u32 lane_lo = lane & ((1U << kTileLowLog) - 1U);
  const u32 tid16 = tid << 4;
  // Partner of representative j: byte (tid << 4) ^ haddr[m] ^ (j << 13) of the
  // half-buffer (register j ^ x[9..11] minus 4; x bit 11 is set). Above 9
  // slots the 4 M partner addresses are formed again at every tile (5 VALU per
  // slot): held across the tile loop they would take 4 M VGPRs, and 11 slots
  // fit in 128 VGPRs only without them (each slot a bundle holds saves a tile
  // read).
  u32 haddr[M], kw[M][NZ];
  double acc[M][NZ];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    haddr[m] = __builtin_amdgcn_readfirstlane(B.slot[m].addr0 ^ (1U << 15));
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      kw[m][k] = __builtin_amdgcn_readfirstlane(B.slot[m].key[k].w);
      acc[m][k] = 0.0;
    }
  }
  const u64 run = ntiles / tile_blocks;
  const int log_tb = __builtin_ctz(tile_blocks);
  if (tid < 32U) step_mask[tid] = tile_rep(static_cast<u64>(tile_blocks) << tid, free_mask);
  u64 rep = tile_rep(tb, free_mask);
  double2 nx0, nx1, nx2, nx3;
  auto issue = [&](u64 r, double2* buf) {
    const u64 r0 = r ^ sb_w, r1 = r0 ^ sb_e0, r2 = r0 ^ sb_e1, r3 = r1 ^ sb_e1;
    nx0 = psi[r0 | lane_lo];
    nx1 = psi[r1 | lane_lo];
    nx2 = psi[r2 | lane_lo];
    nx3 = psi[r3 | lane_lo];
    double2* half = buf + 64U * wave;
    __builtin_amdgcn_global_load_lds(psi + ((r0 ^ sb_e2) | lane_lo), (lds_void*)(half + 0 * kOwnerBlock), 16, 0, 0);
    __builtin_amdgcn_global_load_lds(psi + ((r1 ^ sb_e2) | lane_lo), (lds_void*)(half + 1 * kOwnerBlock), 16, 0, 0);
    __builtin_amdgcn_global_load_lds(psi + ((r2 ^ sb_e2) | lane_lo), (lds_void*)(half + 2 * kOwnerBlock), 16, 0, 0);
    __builtin_amdgcn_global_load_lds(psi + ((r3 ^ sb_e2) | lane_lo), (lds_void*)(half + 3 * kOwnerBlock), 16, 0, 0);
  };
  // One step: this wave's loads of tile `step` have landed; the barrier makes
  // every wave's half-buffer writes visible and retires tile step - 1's reads
  // of `next`, which then receives tile step + 1 while `cur` is evaluated.
  auto step_once = [&](u64 step, const double2* cur, double2* next) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const double2 own0 = nx0, own1 = nx1, own2 = nx2, own3 = nx3;
    if (step + 1 < run) {
      const u64 msk = step_mask[__builtin_ctzll(step + 1)];
      rep ^= (static_cast<u64>(__builtin_amdgcn_readfirstlane(static_cast<u32>(msk >> 32))) << 32) |
             __builtin_amdgcn_readfirstlane(static_cast<u32>(msk));
      issue(rep, next);
    }
    const u32 t32 = tb | (static_cast<u32>(step ^ (step >> 1)) << log_tb);
    u32 t16 = tid16;
    if constexpr (M > 9) {
      asm volatile("" : "+v"(t16));  // opaque per tile: keeps the addresses out of the loop invariants
    }
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const u32 base = t16 ^ haddr[m];
      double2 q[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) q[j] = lds_at(cur, base ^ (static_cast<u32>(j) << 13));
      const double2 a[4] = {own0, own1, own2, own3};
      double v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v[j] = kOdd ? fma(q[j].x, a[j].y, -(q[j].y * a[j].x)) : fma(q[j].x, a[j].x, q[j].y * a[j].y);
      }
#pragma unroll
      for (int k = 0; k < NZ; ++k) {
        const u32 w = kw[m][k];
        const u32 h0 = 0x3FF00000U ^ (static_cast<u32>(__popc(w & t32)) << 31);
        const u32 c1 = w & 0x80000000U, c0 = (w << 1) & 0x80000000U;
        double sacc = acc[m][k];
        sacc = fma(unit_sign(h0), v[0], sacc);
        sacc = fma(unit_sign(h0 ^ c0), v[1], sacc);
        sacc = fma(unit_sign(h0 ^ c1), v[2], sacc);
        sacc = fma(unit_sign(h0 ^ c0 ^ c1), v[3], sacc);
        acc[m][k] = sacc;
      }
    }
  };
  issue(rep, buf0);
  for (u64 step = 0; step < run; step += 2) {
    step_once(step, buf0, buf1);
    if (step + 1 == run) break;
    step_once(step + 1, buf1, buf0);
  }
  // Per-thread sign, then a fixed-order block sum per slot key.
#pragma unroll
  for (int m = 0; m < M; ++m) {
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      const OwnerKey& K = B.slot[m].key[k];
      const u32 par = static_cast<u32>(__popc(K.zlt & tid)) & 1U;
      double r = par ? -acc[m][k] : acc[m][k];
      for (int d = 32; d > 0; d >>= 1) r += __shfl_xor(r, d, 64);
      if (lane == 0) red[wave][m * NZ + k] = r;
    }
  }
  __syncthreads();
  if (tid < static_cast<u32>(M * NZ)) {
    const int m = static_cast<int>(tid) / NZ, k = static_cast<int>(tid) % NZ;
    const OwnerSlot& S = B.slot[m];
    const u32 info = S.key[k].info;
    if ((info >> 15) & 1U) {
      double r = 0.0;
      for (int w = 0; w < kOwnerWaves; ++w) r += red[w][tid];
      const std::size_t row = static_cast<std::size_t>(partial_base) +
                              static_cast<std::size_t>(S.item) * tile_blocks + tb;
      partials[row * kZ + ((info >> 4) & 15U)] = make_double2(r, r);
    }
  }
}

template <int K>
__device__ __forceinline__ void owner_dispatch_one(int kind, const double2* __restrict__ psi,
                                                   const PassDesc* __restrict__ pd,
                                                   const OwnerBundle& B, u32 tb, u32 tile_blocks,
                                                   u32 partial_base, double2* __restrict__ lo,
                                                   double2* __restrict__ hi,
                                                   double (*red)[kOwnerMaxAcc],
                                                   u64* __restrict__ step_mask,
                                                   double2* __restrict__ partials) {
  if constexpr (owner_kind_used(K)) {
    if (kind != K) return;
    constexpr int shape = K >> 2;
    if constexpr ((K >> 1) & 1) {
      owner_lane_sweep<kOwnerShapeSlots[shape], kOwnerShapeKeys[shape], (K & 1) != 0>(
          psi, pd, B, tb, tile_blocks, partial_base, lo, hi, red, step_mask, partials);
    } else {
      owner_half_sweep<kOwnerShapeSlots[shape], kOwnerShapeKeys[shape], (K & 1) != 0>(
          psi, pd, B, tb, tile_blocks, partial_base, lo, hi, red, step_mask, partials);
    }
  }
}

template <int... K>
__device__ __forceinline__ void owner_dispatch(int kind, std::integer_sequence<int, K...>,
                                               const double2* __restrict__ psi,
                                               const PassDesc* __restrict__ pd,
                                               const OwnerBundle& B, u32 tb, u32 tile_blocks,
                                               u32 partial_base, double2* __restrict__ lo,
                                               double2* __restrict__ hi,
                                               double (*red)[kOwnerMaxAcc],
                                               u64* __restrict__ step_mask,
                                               double2* __restrict__ partials) {
  (owner_dispatch_one<K>(kind, psi, pd, B, tb, tile_blocks, partial_base, lo, hi, red, step_mask,
                         partials),
   ...);
}

// One launch for the bundles of every pass: the pass of a block comes from
// the table (its blocks are a contiguous range starting at a multiple of 8);
// within a pass, bundles * tile_blocks blocks, and the blocks of one tile
// block are consecutive multiples of 8 apart (one XCD, close in time). The
// passes are independent (each writes its own partial rows), so one pass's
// tail overlaps the next pass's start. 128 VGPRs per lane: two blocks per CU.
__global__ __launch_bounds__(kOwnerBlock) __attribute__((amdgpu_waves_per_eu(4)))
void expect_owner_kernel(const double2* __restrict__ psi, const PassDesc* __restrict__ passes,
                         const OwnerBundle* __restrict__ bundles,
                         const OwnerLaunch* __restrict__ table, const u32* __restrict__ entry,
                         double2* __restrict__ partials) {
  // Lane-pivot bundles stage points 0..3 of each thread in `lo` and 4..7 in
  // `hi`; register-pivot bundles use them as the two partner half-buffers.
  __shared__ double2 lo[kTilePoints / 2];
  __shared__ double2 hi[kTilePoints / 2];
  __shared__ double red[kOwnerWaves][kOwnerMaxAcc];
  __shared__ u64 step_mask[32];
  const OwnerLaunch& L = table[entry[blockIdx.x >> 3]];
  const u32 groups = L.groups, tile_blocks = L.tile_blocks, partial_base = L.partial_base;
  const PassDesc* pd = passes + L.pass;
  const u32 b = blockIdx.x - L.block_begin;
  const u32 xcd = b & 7U;
  const u32 slot_id = b >> 3;
  const u32 ig = slot_id % groups;
  const u32 tb = (slot_id / groups) * 8U + xcd;
  const OwnerBundle& B = bundles[L.owner_first + ig];
  const int kind = static_cast<int>(__builtin_amdgcn_readfirstlane(B.kind));
  owner_dispatch(kind, std::make_integer_sequence<int, kOwnerKinds>{}, psi, pd, B, tb, tile_blocks,
                 partial_base, lo, hi, red, step_mask, partials);
}

// ------------------------------------------------------------ tiled sigma

// sigma = H psi (and H^dagger psi) over the tiled passes: for output o and
// x-group g, (P_k psi)[o] = i^{y_k} (-1)^{|z_k & (o ^ x)|} psi[o ^ x], so
// sigma[o] = sum_g psi[o ^ x_g] * sum_{k in g} cy_k (-1)^{|z_k & (o ^ x_g)|} with
// cy_k = c_k i^{y_k} (H^dagger: conj(c_k) i^{y_k}).
//
// One 1024-thread block per CU walks its tiles in Gray-code order (as the
// owner sweep). The tile is double-buffered in LDS (2 x 64 KiB): global_load_lds
// fills the other buffer with the next tile while this one is evaluated, so a
// tile costs one barrier. Thread tid owns the points u = tid + 1024 I (I < 4),
// which are segment (tid >> 3) + 128 I, offset tid & 7: the same addresses it copies into
// LDS, reads sigma from and writes sigma back to (one read-modify-write per
// point per pass; the first launched pass writes without reading). The current
// sigma values are loaded at the start of a tile and the sums are stored after
// the next tile's barrier, so neither latency sits in front of a barrier.
//
// Keys are a flat list (SigmaKey), read through scalar loads a chunk of
// records at a time. The sign of key j at point u is
// parity(z_j & rep) ^ parity(zl_j & (u ^ xl)), i.e. a uniform part, a lane part
// parity(zl_j & tid) and the register part parity(((zl_j >> 10) & 3) & I).
constexpr int kSigmaBlock = 1024;
constexpr int kSigmaPoints = kTilePoints / kSigmaBlock;

static_assert(kSigmaPoints == 4 && (kSigmaBlock >> kTileLowLog) == 128,
              "sigma sweep is written for 4 points per thread and 128 segments per block");

__device__ __forceinline__ double flip_sign(double v, u32 m) {
  return __hiloint2double(__double2hiint(v) ^ static_cast<int>(m), __double2loint(v));
}

// c += s * k for a uniform sign s (hi word 0x3FF00000 or 0xBFF00000).
__device__ __forceinline__ void add_signed(double2& c, u32 s_hi, double2 k) {
  const double s = unit_sign(s_hi);
  c.x = fma(s, k.x, c.x);
  c.y = fma(s, k.y, c.y);
}

// a += c * b (complex).
__device__ __forceinline__ void cmac(double2& a, double2 c, double2 b) {
  a.x = fma(c.x, b.x, a.x);
  a.x = fma(-c.y, b.y, a.x);
  a.y = fma(c.x, b.y, a.y);
  a.y = fma(c.y, b.x, a.y);
}

template <bool kPair>
__device__ __forceinline__ void hold_key(const SigmaKey& k) {
  asm volatile("" ::"s"(k.z), "s"(k.zl), "s"(k.xl), "s"(k.sh[0]), "s"(k.sh[1]), "s"(k.sh[2]),
               "s"(k.c.x), "s"(k.c.y));
  if (kPair) asm volatile("" ::"s"(k.cd.x), "s"(k.cd.y));
}

template <bool kPair>
__global__ __launch_bounds__(kSigmaBlock) void sigma_tiled_kernel(
    const double2* __restrict__ psi, const PassDesc* __restrict__ pd,
    const SigmaKey* __restrict__ keys, u32 key_begin, u32 key_end, u32 log_blocks, u32 init,
    double2* __restrict__ out, double2* __restrict__ out_dag) {
  __shared__ double2 tile[2 * kTilePoints];
  __shared__ u64 step_mask[32];
  const u32 tid = threadIdx.x;
  const u32 wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const u32 lane = tid & 63U;
  const u64 free_mask = pd->free_mask;
  const u64 run = (1ULL << (pd->n - kTileMaxLog)) >> log_blocks;
  // Point I of thread tid is local point tid + 1024 I: segment (tid >> 3) +
  // 128 I at offset tid & 7.
  const u64 sb_w = pd->seg_off[tid >> kTileLowLog], sb_e0 = pd->seg_off[128], sb_e1 = pd->seg_off[256];
  const u32 lane_lo = lane & ((1U << kTileLowLog) - 1U);
  if (tid < 32U) step_mask[tid] = tile_rep(1ULL << (tid + log_blocks), free_mask);
  u64 rep = tile_rep(blockIdx.x, free_mask);
  // Global index of point I of this thread in the tile with representative r.
#define WG_PT(r, I) \
  ((((r) ^ sb_w ^ (((I) & 1) ? sb_e0 : 0) ^ (((I) & 2) ? sb_e1 : 0))) | lane_lo)
  auto issue = [&](u64 r, u32 buf) {
    double2* seg = tile + buf * kTilePoints + 64U * wave;
    __builtin_amdgcn_global_load_lds(psi + WG_PT(r, 0), (lds_void*)(seg + 0 * kSigmaBlock), 16, 0, 0);
    __builtin_amdgcn_global_load_lds(psi + WG_PT(r, 1), (lds_void*)(seg + 1 * kSigmaBlock), 16, 0, 0);
    __builtin_amdgcn_global_load_lds(psi + WG_PT(r, 2), (lds_void*)(seg + 2 * kSigmaBlock), 16, 0, 0);
    __builtin_amdgcn_global_load_lds(psi + WG_PT(r, 3), (lds_void*)(seg + 3 * kSigmaBlock), 16, 0, 0);
  };
  const double2 zero = make_double2(0.0, 0.0);
  double2 o0 = zero, o1 = zero, o2 = zero, o3 = zero;  // sums waiting to be stored
  double2 q0 = zero, q1 = zero, q2 = zero, q3 = zero;  // sigma^dagger sums
  u64 prev = 0;
  issue(rep, 0);
  for (u64 step = 0; step < run; ++step) {
    // This wave's copies of tile `step` have landed; the barrier makes every
    // wave's copies visible and retires the reads of the other buffer.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (step > 0) {
      out[WG_PT(prev, 0)] = o0;
      out[WG_PT(prev, 1)] = o1;
      out[WG_PT(prev, 2)] = o2;
      out[WG_PT(prev, 3)] = o3;
      if (kPair) {
        out_dag[WG_PT(prev, 0)] = q0;
        out_dag[WG_PT(prev, 1)] = q1;
        out_dag[WG_PT(prev, 2)] = q2;
        out_dag[WG_PT(prev, 3)] = q3;
      }
    }
    const u64 cur = rep;
    double2 s0 = zero, s1 = zero, s2 = zero, s3 = zero;
    double2 t0 = zero, t1 = zero, t2 = zero, t3 = zero;
    if (!init) {
      s0 = out[WG_PT(cur, 0)];
      s1 = out[WG_PT(cur, 1)];
      s2 = out[WG_PT(cur, 2)];
      s3 = out[WG_PT(cur, 3)];
      if (kPair) {
        t0 = out_dag[WG_PT(cur, 0)];
        t1 = out_dag[WG_PT(cur, 1)];
        t2 = out_dag[WG_PT(cur, 2)];
        t3 = out_dag[WG_PT(cur, 3)];
      }
    }
    if (step + 1 < run) {
      const u64 msk = step_mask[__builtin_ctzll(step + 1)];
      rep ^= (static_cast<u64>(__builtin_amdgcn_readfirstlane(static_cast<u32>(msk >> 32))) << 32) |
             __builtin_amdgcn_readfirstlane(static_cast<u32>(msk));
      issue(rep, static_cast<u32>(step + 1) & 1U);
    }
    const double2* tl = tile + (static_cast<u32>(step) & 1U) * kTilePoints;
    double2 a0 = zero, a1 = zero, a2 = zero, a3 = zero;
    double2 d0 = zero, d1 = zero, d2 = zero, d3 = zero;
    double2 c0 = zero, c1 = zero, c2 = zero, c3 = zero;
    double2 e0 = zero, e1 = zero, e2 = zero, e3 = zero;
    auto key_step = [&](const SigmaKey& k) {
      const u32 xl = k.xl & ~kSigmaLast;
      // Bit 31 of m: parity(z & rep) ^ parity(zl & tid) (the parity of zl & xl
      // is folded into the coefficients).
      const u32 m = (static_cast<u32>(__popc(k.zl & tid)) + static_cast<u32>(__popcll(k.z & cur)))
                    << 31;
      const u32 s1h = k.sh[0], s2h = k.sh[1], s3h = k.sh[2];
      const double2 kp = make_double2(flip_sign(k.c.x, m), flip_sign(k.c.y, m));
      c0.x += kp.x;
      c0.y += kp.y;
      add_signed(c1, s1h, kp);
      add_signed(c2, s2h, kp);
      add_signed(c3, s3h, kp);
      if (kPair) {
        const double2 kq = make_double2(flip_sign(k.cd.x, m), flip_sign(k.cd.y, m));
        e0.x += kq.x;
        e0.y += kq.y;
        add_signed(e1, s1h, kq);
        add_signed(e2, s2h, kq);
        add_signed(e3, s3h, kq);
      }
      if (k.xl & kSigmaLast) {
        const double2 b0 = tl[(tid + 0 * kSigmaBlock) ^ xl];
        const double2 b1 = tl[(tid + 1 * kSigmaBlock) ^ xl];
        const double2 b2 = tl[(tid + 2 * kSigmaBlock) ^ xl];
        const double2 b3 = tl[(tid + 3 * kSigmaBlock) ^ xl];
        cmac(a0, c0, b0);
        cmac(a1, c1, b1);
        cmac(a2, c2, b2);
        cmac(a3, c3, b3);
        c0 = c1 = c2 = c3 = zero;
        if (kPair) {
          cmac(d0, e0, b0);
          cmac(d1, e1, b1);
          cmac(d2, e2, b2);
          cmac(d3, e3, b3);
          e0 = e1 = e2 = e3 = zero;
        }
      }
    };
    // Key records are loaded kSigmaChunk at a time (one wait on the scalar
    // loads per chunk); every pass's list is padded with zero records (zero
    // coefficient, not an item's last key) to a whole number of chunks.
    // (Pairs of records when sigma^dagger rides along: four would not fit
    // the scalar registers.)
    constexpr u32 kChunk = kPair ? kSigmaChunk / 2 : kSigmaChunk;
    for (u32 r = key_begin; r < key_end; r += kChunk) {
      SigmaKey k[kChunk];
#pragma unroll
      for (u32 i = 0; i < kChunk; ++i) k[i] = keys[r + i];
      // Empty asm statements that take the records' fields as scalar
      // operands keep the compiler from moving a record's loads down to its
      // own step (which would wait on the scalar loads once per key again).
#pragma unroll
      for (u32 i = 0; i < kChunk; ++i) hold_key<kPair>(k[i]);
#pragma unroll
      for (u32 i = 0; i < kChunk; ++i) key_step(k[i]);
    }
    o0 = make_double2(s0.x + a0.x, s0.y + a0.y);
    o1 = make_double2(s1.x + a1.x, s1.y + a1.y);
    o2 = make_double2(s2.x + a2.x, s2.y + a2.y);
    o3 = make_double2(s3.x + a3.x, s3.y + a3.y);
    if (kPair) {
      q0 = make_double2(t0.x + d0.x, t0.y + d0.y);
      q1 = make_double2(t1.x + d1.x, t1.y + d1.y);
      q2 = make_double2(t2.x + d2.x, t2.y + d2.y);
      q3 = make_double2(t3.x + d3.x, t3.y + d3.y);
    }
    prev = cur;
  }
  if (run > 0) {
    out[WG_PT(prev, 0)] = o0;
    out[WG_PT(prev, 1)] = o1;
    out[WG_PT(prev, 2)] = o2;
    out[WG_PT(prev, 3)] = o3;
    if (kPair) {
      out_dag[WG_PT(prev, 0)] = q0;
      out_dag[WG_PT(prev, 1)] = q1;
      out_dag[WG_PT(prev, 2)] = q2;
      out_dag[WG_PT(prev, 3)] = q3;
    }
  }
#undef WG_PT
}

// Per-call sigma key records: the plan's records with c = cy[key] (or
// cy_dag[key] when only sigma^dagger is wanted) and cd = cy_dag[key], both
// times (-1)^{parity(zl & xl)}; padding records keep zero coefficients.
__global__ __launch_bounds__(256) void sigma_keys_kernel(const SigmaKey* __restrict__ plan_keys,
                                                         std::size_t count,
                                                         const double2* __restrict__ cy,
                                                         const double2* __restrict__ cy_dag,
                                                         int dagger_only,
                                                         SigmaKey* __restrict__ keys) {
  const std::size_t r = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= count) return;
  SigmaKey k = plan_keys[r];
  if (k.key != kSigmaPadKey) {
    const bool neg = __popc(k.zl & k.xl & ~kSigmaLast) & 1;
    const double2 a = dagger_only ? cy_dag[k.key] : cy[k.key], b = cy_dag[k.key];
    k.c = neg ? make_double2(-a.x, -a.y) : a;
    k.cd = neg ? make_double2(-b.x, -b.y) : b;
  }
  keys[r] = k;
}

// cy[id] = c_key * i^{y_key}, cy_dag[id] = conj(c_key) * i^{y_key}, by distinct-key id.
__global__ __launch_bounds__(256) void key_cy_kernel(const double2* __restrict__ key_c_flat,
                                                     const int* __restrict__ key_y_flat,
                                                     const u32* __restrict__ flat_key,
                                                     std::size_t num_keys,
                                                     double2* __restrict__ cy,
                                                     double2* __restrict__ cy_dag) {
  const std::size_t p = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= num_keys) return;
  const double2 c = key_c_flat[p];
  const double2 cc = make_double2(c.x, -c.y);
  const int y = key_y_flat[p] & 3;
  auto times_iy = [y](double2 v) {
    switch (y) {
      case 0: return v;
      case 1: return make_double2(-v.y, v.x);
      case 2: return make_double2(-v.x, -v.y);
      default: return make_double2(v.y, -v.x);
    }
  };
  cy[flat_key[p]] = times_iy(c);
  cy_dag[flat_key[p]] = times_iy(cc);
}

// ------------------------------------------------------------ diagonal keys

// All diagonal keys (x = 0) at once: for a tile of kTilePoints consecutive
// amplitudes (index = t * kTilePoints + u) the block forms p(u) = |psi|^2 and
// its Walsh-Hadamard transform P(v) = sum_u (-1)^{|v & u|} p(u); key z then
// gains (-1)^{|(z >> kTileMaxLog) & t|} P(z mod kTilePoints). The transform is
// kTileMaxLog butterfly stages, 3 in registers and 9 through LDS, so a tile
// costs O(k 2^k) whatever the number of keys.
constexpr int kWhtKeysPerThread = 4;
constexpr u32 kWhtMaxKeys = kTileBlock * kWhtKeysPerThread;

// LDS slot of transform index u: one pad double every 16 spreads the strided
// accesses of the butterfly rounds over the banks.
__device__ __forceinline__ u32 wht_slot(u32 u) { return u + (u >> 4); }

__device__ __forceinline__ void wht8(double (&v)[8]) {
#pragma unroll
  for (int h = 1; h < 8; h <<= 1) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (!(i & h)) {
        const double a = v[i], c = v[i | h];
        v[i] = a + c;
        v[i | h] = a - c;
      }
    }
  }
}

// Keys [key_begin, key_begin + num_keys) of the diagonal items (kZ keys per
// item, all items full but the last); item i's partial row for block c is
// partial_base + i * gridDim.x + c.
__global__ __launch_bounds__(kTileBlock) void expect_diag_wht_kernel(
    const double2* __restrict__ psi, u32 log_tiles, const ExpItem* __restrict__ items,
    u32 key_begin, u32 num_keys, u32 partial_base, double2* __restrict__ partials) {
  __shared__ double w[kTilePoints + kTilePoints / 16];
  const u32 tid = threadIdx.x;
  u64 zh[kWhtKeysPerThread];
  u32 zslot[kWhtKeysPerThread];
  double acc[kWhtKeysPerThread];
#pragma unroll
  for (int r = 0; r < kWhtKeysPerThread; ++r) {
    const u32 kk = tid + r * kTileBlock;
    acc[r] = 0.0;
    zh[r] = 0;
    zslot[r] = 0;
    if (kk < num_keys) {
      const u32 g = key_begin + kk;
      const u64 z = items[g / kZ].z[g % kZ];
      zh[r] = z >> kTileMaxLog;
      zslot[r] = wht_slot(static_cast<u32>(z) & (kTilePoints - 1));
    }
  }
  const u64 ntiles = 1ULL << log_tiles;
  double2 s[8];
  u64 t = blockIdx.x;
  if (t < ntiles) {
#pragma unroll
    for (int i = 0; i < 8; ++i) s[i] = psi[(t << kTileMaxLog) + tid + i * kTileBlock];
  }
  for (; t < ntiles; t += gridDim.x) {
    double v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = fma(s[i].x, s[i].x, s[i].y * s[i].y);
    const u64 tn = t + gridDim.x;
    if (tn < ntiles) {
#pragma unroll
      for (int i = 0; i < 8; ++i) s[i] = psi[(tn << kTileMaxLog) + tid + i * kTileBlock];
    }
    wht8(v);  // bits 9..11 (register index)
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) w[wht_slot(tid + i * kTileBlock)] = v[i];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 3; ++r) {  // bits 3r .. 3r+2
      const u32 lowbits = tid & ((1U << (3 * r)) - 1U);
      const u32 high = (tid >> (3 * r)) << (3 * r + 3);
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = w[wht_slot(high | (static_cast<u32>(j) << (3 * r)) | lowbits)];
      wht8(v);
#pragma unroll
      for (int j = 0; j < 8; ++j) w[wht_slot(high | (static_cast<u32>(j) << (3 * r)) | lowbits)] = v[j];
      __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < kWhtKeysPerThread; ++r) {
      if (tid + r * kTileBlock < num_keys) {
        const double val = w[zslot[r]];
        acc[r] += (__popcll(zh[r] & t) & 1) ? -val : val;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kWhtKeysPerThread; ++r) {
    const u32 kk = tid + r * kTileBlock;
    if (kk < num_keys) {
      const u32 g = key_begin + kk;
      partials[(static_cast<std::size_t>(partial_base) +
                static_cast<std::size_t>(g / kZ) * gridDim.x + blockIdx.x) * kZ + g % kZ] =
          make_double2(acc[r], 0.0);
    }
  }
}

// One block per item: thread r sums the item's partial rows r, r + 256, ...
// for every key column, a fixed-shape tree adds the 256 sums, and the key
// value is <P> = 2 Re(i^y S) (x != 0) or Re S (x == 0).
constexpr int kReduceBlock = 256;

__global__ __launch_bounds__(kReduceBlock) void expect_reduce_kernel(
    const ExpItem* __restrict__ items, const u32* __restrict__ offset,
    const u32* __restrict__ count, const double2* __restrict__ partials,
    double* __restrict__ key_vals) {
  __shared__ double2 lds[kReduceBlock / 64][kZ];
  const u32 item_id = blockIdx.x;
  const ExpItem& it = items[item_id];
  const int nz = it.nz;
  const std::size_t base = offset[item_id];
  const u32 rows = count[item_id];
  double r[kZ], i[kZ];
#pragma unroll
  for (int j = 0; j < kZ; ++j) r[j] = i[j] = 0.0;
  for (u32 c = threadIdx.x; c < rows; c += kReduceBlock) {
    const double2* row = partials + (base + c) * kZ;
#pragma unroll
    for (int j = 0; j < kZ; ++j) {
      if (j < nz) {
        const double2 p = row[j];
        r[j] += p.x;
        i[j] += p.y;
      }
    }
  }
  const u32 wave = threadIdx.x >> 6, lane = threadIdx.x & 63U;
#pragma unroll
  for (int j = 0; j < kZ; ++j) {
    if (j < nz) {
      for (int d = 32; d > 0; d >>= 1) {
        r[j] += __shfl_xor(r[j], d, 64);
        i[j] += __shfl_xor(i[j], d, 64);
      }
      if (lane == 0) lds[wave][j] = make_double2(r[j], i[j]);
    }
  }
  __syncthreads();
  if (threadIdx.x < static_cast<u32>(nz)) {
    const int j = static_cast<int>(threadIdx.x);
    double sr = 0.0, si = 0.0;
    for (int w = 0; w < kReduceBlock / 64; ++w) {
      sr += lds[w][j].x;
      si += lds[w][j].y;
    }
    double val;
    if (it.x == 0) {
      val = sr;
    } else {
      switch (it.y[j] & 3) {
        case 0: val = 2.0 * sr; break;
        case 1: val = -2.0 * si; break;
        case 2: val = -2.0 * sr; break;
        default: val = 2.0 * si; break;
      }
    }
    key_vals[it.key[j]] = val;
  }
}

// E = sum_t c_t <P_key(t)>: one block, strided per-thread sums in term order,
// then a fixed-shape tree. Optionally writes per-term values (complex, imag 0).
__global__ __launch_bounds__(1024) void expect_finalize_kernel(const double2* __restrict__ coeffs,
                                                               const u32* __restrict__ term_key,
                                                               std::size_t T,
                                                               const double* __restrict__ key_vals,
                                                               double2* __restrict__ term_values,
                                                               double2* __restrict__ out) {
  __shared__ double2 lds[16];
  double r = 0.0, i = 0.0;
  for (std::size_t t = threadIdx.x; t < T; t += blockDim.x) {
    const double v = key_vals[term_key[t]];
    const double2 c = coeffs[t];
    r = fma(c.x, v, r);
    i = fma(c.y, v, i);
    if (term_values) term_values[t] = make_double2(v, 0.0);
  }
  for (int d = 32; d > 0; d >>= 1) {
    r += __shfl_xor(r, d, 64);
    i += __shfl_xor(i, d, 64);
  }
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = make_double2(r, i);
  __syncthreads();
  if (threadIdx.x == 0 && out) {
    double sr = 0.0, si = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x / 64); ++w) {
      sr += lds[w].x;
      si += lds[w].y;
    }
    *out = make_double2(sr, si);
  }
}

__global__ __launch_bounds__(256) void widen_c64_kernel(const float2* __restrict__ in,
                                                        std::size_t len,
                                                        double2* __restrict__ out) {
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < len;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    out[i] = make_double2(static_cast<double>(in[i].x), static_cast<double>(in[i].y));
  }
}

// ------------------------------------------------------------ sigma = H psi

__device__ inline double2 times_i_pow(double2 v, int k) {
  switch (k & 3) {
    case 0: return v;
    case 1: return make_double2(-v.y, v.x);
    case 2: return make_double2(-v.x, -v.y);
    default: return make_double2(v.y, -v.x);
  }
}

// Direct sigma sweep (states swept without tiles): sigma = H psi and
// sigma_dag = H^dagger psi, with
//   sigma[o] = sum_g psi[o ^ x_g] sum_{k in g} c_k i^{y_k} (-1)^{|z_k & (o ^ x_g)|}
// (sigma_dag: conj(c_k)). The work is split two ways so that small states with
// many terms still fill the GPU: blockIdx.x picks a chunk of kSigOut *
// kSigBlock consecutive outputs, blockIdx.y a slice of consecutive x-groups.
// Each block writes its slice's partial sums to its own row (row = slice,
// row_stride = len); sigma_rows_sum_kernel then adds the rows in slice order,
// so the result does not depend on scheduling. With one slice the block writes
// sigma directly.
constexpr int kSigBlock = 256;
constexpr int kSigOut = 8;
constexpr u64 kSigChunk = static_cast<u64>(kSigBlock) * kSigOut;

__global__ __launch_bounds__(kSigBlock) void sigma_direct_kernel(
    const double2* __restrict__ psi, u64 len, const u64* __restrict__ group_x,
    const u32* __restrict__ group_begin, u32 num_groups, u32 groups_per_slice,
    const u64* __restrict__ key_z, const int* __restrict__ key_y, const double2* __restrict__ key_c,
    double2* __restrict__ rows_s, double2* __restrict__ rows_d) {
  const u64 base = static_cast<u64>(blockIdx.x) * kSigChunk;
  const u32 g0 = blockIdx.y * groups_per_slice;
  const u32 g1 = min(num_groups, g0 + groups_per_slice);
  double2 sg[kSigOut], sd[kSigOut];
#pragma unroll
  for (int i = 0; i < kSigOut; ++i) {
    sg[i] = make_double2(0.0, 0.0);
    sd[i] = make_double2(0.0, 0.0);
  }
  for (u32 g = g0; g < g1; ++g) {
    const u64 x = group_x[g];
    double2 tv[kSigOut];
#pragma unroll
    for (int i = 0; i < kSigOut; ++i) {
      const u64 o = base + threadIdx.x + static_cast<u64>(i) * kSigBlock;
      tv[i] = o < len ? psi[o ^ x] : make_double2(0.0, 0.0);
    }
    const u32 kb = group_begin[g], ke = group_begin[g + 1];
    for (u32 k = kb; k < ke; ++k) {
      const u64 z = key_z[k];
      // c i^y and conj(c) i^y, uniform per key.
      const double2 c = key_c[k];
      const double2 cy = times_i_pow(c, key_y[k]);
      const double2 cdy = times_i_pow(make_double2(c.x, -c.y), key_y[k]);
#pragma unroll
      for (int i = 0; i < kSigOut; ++i) {
        const u64 o = base + threadIdx.x + static_cast<u64>(i) * kSigBlock;
        double2 w = tv[i];
        if (__popcll(z & (o ^ x)) & 1) w = make_double2(-w.x, -w.y);
        sg[i].x = fma(cy.x, w.x, fma(-cy.y, w.y, sg[i].x));
        sg[i].y = fma(cy.x, w.y, fma(cy.y, w.x, sg[i].y));
        sd[i].x = fma(cdy.x, w.x, fma(-cdy.y, w.y, sd[i].x));
        sd[i].y = fma(cdy.x, w.y, fma(cdy.y, w.x, sd[i].y));
      }
    }
  }
  const std::size_t row = static_cast<std::size_t>(blockIdx.y) * len;
#pragma unroll
  for (int i = 0; i < kSigOut; ++i) {
    const u64 o = base + threadIdx.x + static_cast<u64>(i) * kSigBlock;
    if (o < len) {
      if (rows_s) rows_s[row + o] = sg[i];
      if (rows_d) rows_d[row + o] = sd[i];
    }
  }
}

// out[o] = sum over the rows of rows[r * len + o], rows in order.
__global__ __launch_bounds__(256) void sigma_rows_sum_kernel(const double2* __restrict__ rows,
                                                             u32 num_rows, u64 len,
                                                             double2* __restrict__ out) {
  for (u64 o = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; o < len;
       o += static_cast<u64>(gridDim.x) * blockDim.x) {
    double2 acc = rows[o];
    for (u32 r = 1; r < num_rows; ++r) {
      const double2 v = rows[static_cast<std::size_t>(r) * len + o];
      acc.x += v.x;
      acc.y += v.y;
    }
    out[o] = acc;
  }
}

// Per-key coefficient sums from per-term coefficients, each key's terms in term
// order (the same order the host plan uses for the operator's own coefficients).
__global__ __launch_bounds__(256) void key_coeff_kernel(const double2* __restrict__ coeffs,
                                                        const u32* __restrict__ begin,
                                                        const u32* __restrict__ terms,
                                                        std::size_t num_keys,
                                                        double2* __restrict__ key_c) {
  const std::size_t k = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= num_keys) return;
  const u32 b = begin[k], e = begin[k + 1];
  double2 acc = coeffs[terms[b]];
  for (u32 i = b + 1; i < e; ++i) {
    acc.x += coeffs[terms[i]].x;
    acc.y += coeffs[terms[i]].y;
  }
  key_c[k] = acc;
}

unsigned stride_grid(std::size_t work, unsigned block) {
  const std::size_t g = (work + block - 1) / block;
  return static_cast<unsigned>(std::max<std::size_t>(1, std::min<std::size_t>(g, 1u << 20)));
}

// ------------------------------------------------------------ pass planner

// Subspace U = span(coordinate bits in `coord`, generators `gens`), generators
// kept in reduced echelon form: each has a distinct pivot (its highest bit),
// no bit inside `coord`, and no generator has another generator's pivot set.
struct Span {
  u64 coord = 0;
  std::vector<u64> gens;

  int dim() const { return __builtin_popcountll(coord) + static_cast<int>(gens.size()); }

  static int top(u64 v) { return 63 - __builtin_clzll(v); }

  void normalize() {
    std::vector<u64> out;
    for (u64 v : gens) {
      v &= ~coord;
      for (u64 e : out) {
        if ((v >> top(e)) & 1ULL) v ^= e;
      }
      if (v) {
        for (u64& e : out) {
          if ((e >> top(v)) & 1ULL) e ^= v;
        }
        out.push_back(v);
      }
    }
    gens = out;
  }

  // Reduces x against the generators; returns the residual (inside `coord`
  // iff x lies in U) and the generator selection.
  u64 reduce(u64 x, u32* sel) const {
    u32 s = 0;
    for (std::size_t i = 0; i < gens.size(); ++i) {
      if ((x >> top(gens[i])) & 1ULL) {
        x ^= gens[i];
        s |= 1U << i;
      }
    }
    if (sel) *sel = s;
    return x;
  }

  bool contains(u64 x) const { return (reduce(x, nullptr) & ~coord) == 0; }

  void add_coords(u64 bits) {
    coord |= bits;
    normalize();
  }

  void add_gen(u64 v) {
    gens.push_back(v);
    normalize();
  }
};

}  // namespace

namespace {

u64 host_insert_zero(u64 v, u64 pivot) {
  const u64 low = pivot - 1;
  return ((v & ~low) << 1) | (v & low);
}

int host_parity(u64 v) { return __builtin_popcountll(v) & 1; }

// Step classes of an off-diagonal tiled item (see ExpItem): the step
// characters c_j (bit i = parity(zl_j & g(2^i)), g(s) = insert_zero(s << 6,
// pivot)) span a space of dimension d; with a basis b_1..b_d, class(s) =
// (b_i . s)_i and c_j = sum_i lam_j[i] b_i. The enumeration basis is a basis
// of the kernel of s -> class(s) (low 5 - d positions) followed by vectors v_i
// with class(v_i) = e_i (top d positions).
void assign_step_classes(ExpItem& e, int min_d) {
  e.cls_log = kNoClasses;
  if (e.xl == 0) return;
  constexpr int kStepBits = 5;
  static_assert((1 << kStepBits) == kOffSteps, "five step bits per off-diagonal tile sweep");
  u32 c[kZ] = {};
  for (int j = 0; j < e.nz; ++j) {
    for (int i = 0; i < kStepBits; ++i) {
      const u64 g = host_insert_zero(u64{1} << (i + 6), e.pbit);
      c[j] |= static_cast<u32>(host_parity(e.zl[j] & g)) << i;
    }
  }
  // Basis of span{c_j} (in the order first met) and each c_j's coordinates.
  std::vector<u32> basis;
  auto coords = [&](u32 v, u32* lam) -> bool {
    for (u32 m = 0; m < (1U << basis.size()); ++m) {
      u32 acc = 0;
      for (std::size_t i = 0; i < basis.size(); ++i) {
        if ((m >> i) & 1U) acc ^= basis[i];
      }
      if (acc == v) {
        *lam = m;
        return true;
      }
    }
    return false;
  };
  u32 lam[kZ] = {};
  for (int j = 0; j < e.nz; ++j) {
    if (!coords(c[j], &lam[j])) {
      if (static_cast<int>(basis.size()) == kMaxClassLog) return;  // too many classes
      basis.push_back(c[j]);
      coords(c[j], &lam[j]);
    }
  }
  // Extra class dimensions (zero in every lam) up to min_d, so that all items
  // of a pass can share one kernel instance.
  for (u32 unit = 1; static_cast<int>(basis.size()) < std::min(min_d, kMaxClassLog) &&
                     unit < (1U << kStepBits);
       unit <<= 1) {
    u32 probe = 0;
    if (!coords(unit, &probe)) basis.push_back(unit);
  }
  const int d = static_cast<int>(basis.size());
  auto cls = [&](u32 s) {
    u32 a = 0;
    for (int i = 0; i < d; ++i) a |= static_cast<u32>(host_parity(basis[i] & s)) << i;
    return a;
  };
  std::vector<u32> kern;  // basis of the kernel, by elimination on the top bit
  for (u32 s = 1; s < (1U << kStepBits); ++s) {
    if (cls(s)) continue;
    u32 w = s;
    for (u32 b : kern) {
      if ((w >> (31 - __builtin_clz(b))) & 1U) w ^= b;
    }
    if (w) kern.push_back(w);
  }
  if (static_cast<int>(kern.size()) != kStepBits - d) fail_runtime("step class kernel has the wrong dimension");
  for (int l = 0; l < kStepBits - d; ++l) e.ebasis[l] = static_cast<unsigned char>(kern[l]);
  for (int i = 0; i < d; ++i) {
    u32 v = 0;
    for (u32 s = 1; s < (1U << kStepBits) && !v; ++s) {
      if (cls(s) == (1U << i)) v = s;
    }
    if (!v) fail_runtime("step class has no representative");
    e.ebasis[kStepBits - d + i] = static_cast<unsigned char>(v);
  }
  for (int j = 0; j < kZ; ++j) e.lam[j] = static_cast<unsigned char>(j < e.nz ? lam[j] : 0);
  e.cls_log = static_cast<unsigned char>(d);
}

// ------------------------------------------------------------ owner slots

constexpr u32 kSegMask = (1U << kSegBits) - 1;

// One slot of the owner sweep: the keys of an item whose |x & z| has the given
// parity (a slot computes one component of the pair product), 2 per slot, or
// 4 when the parity class holds more than 2 keys ("wide").
struct OwnerCand {
  u32 item;  // the item (relative to the pass in plan_owner_bundles)
  u32 xl;    // x in pass-local coordinates
  int odd;
  int ncols;
  int cols[kOwnerMaxKeys];
  int wide;
  int lanex;  // x inside the contiguous low bits: lane-pivot bundle
};

void owner_cands(const ExpItem& e, u32 item, u32 xl, std::vector<OwnerCand>& out) {
  for (int odd = 0; odd < 2; ++odd) {
    int cols[kZ];
    int nc = 0;
    for (int j = 0; j < e.nz; ++j) {
      if ((e.y[j] & 1) == odd) cols[nc++] = j;
    }
    const int wide = nc > 2 ? 1 : 0;
    const int per = wide ? 4 : 2;
    for (int s = 0; s < nc; s += per) {
      OwnerCand c{};
      c.item = item;
      c.xl = xl;
      c.odd = odd;
      c.ncols = std::min(per, nc - s);
      for (int q = 0; q < c.ncols; ++q) c.cols[q] = cols[s + q];
      c.wide = wide;
      c.lanex = ((xl >> kTileLowLog) & kSegMask) == 0 ? 1 : 0;
      out.push_back(c);
    }
  }
}

// Register-pivot slots of one pool grouped by bundle layout, the functional
// lam with lam . s = 1 for the segment part s of every member's x: repeatedly
// the lam covering the most remaining slots (the lowest on ties); while it
// covers a full bundle only whole bundles are taken (its other slots may fill
// bundles of other layouts), and the last groups take everything their lam
// covers. The counts come from a Walsh-Hadamard transform of the pool's
// histogram of segment parts: #{i : lam . s_i = 1} = (k - W(lam)) / 2.
struct OwnerGroup {
  u32 lam;
  std::vector<u32> members;  // indices into the pool, in pool order
};

std::vector<OwnerGroup> owner_layout_groups(const std::vector<u32>& segs, std::size_t full) {
  std::vector<OwnerGroup> out;
  std::vector<u32> pool(segs.size()), rest;
  for (u32 i = 0; i < pool.size(); ++i) pool[i] = i;
  std::array<int, std::size_t{1} << kSegBits> w;
  while (!pool.empty()) {
    w.fill(0);
    for (u32 i : pool) ++w[segs[i] & kSegMask];
    for (std::size_t len = 1; len < w.size(); len <<= 1) {
      for (std::size_t b = 0; b < w.size(); b += 2 * len) {
        for (std::size_t i = b; i < b + len; ++i) {
          const int u = w[i], v = w[i + len];
          w[i] = u + v;
          w[i + len] = u - v;
        }
      }
    }
    const int k = static_cast<int>(pool.size());
    u32 best = 1;
    std::size_t best_n = 0;
    for (u32 lam = 1; lam <= kSegMask; ++lam) {
      const std::size_t cnt = static_cast<std::size_t>((k - w[lam]) / 2);
      if (cnt > best_n) {
        best_n = cnt;
        best = lam;
      }
    }
    if (best_n == 0) fail_runtime("owner bundle slot has no register pivot");
    const std::size_t quota = best_n >= full ? best_n / full * full : best_n;
    OwnerGroup g{best, {}};
    rest.clear();
    for (u32 i : pool) {
      const bool covered = host_parity(best & segs[i]);
      (covered && g.members.size() < quota ? g.members : rest).push_back(i);
    }
    out.push_back(std::move(g));
    pool.swap(rest);
  }
  return out;
}

// Shape of each bundle of a group of `size` slots: bundles of the largest
// shape, then the smallest shape that holds the rest.
template <class F>
void owner_group_shapes(std::size_t size, int first, int last, F&& f) {
  for (std::size_t s = 0; s < size;) {
    const std::size_t left = size - s;
    int shape = first;
    for (int q = last; q >= first; --q) {
      if (static_cast<std::size_t>(kOwnerShapeSlots[q]) >= left) {
        shape = q;
        break;
      }
    }
    f(shape, s);
    s += static_cast<std::size_t>(kOwnerShapeSlots[shape]);
  }
}

// The owner slots of a pass in bundle pools (width, parity, pivot kind) and
// layout groups: calls f(pool, group lam (0 for lane pivots), member indices
// into cands, wide, base kind) in the order the bundles are emitted.
template <class F>
void owner_pass_groups(const std::vector<OwnerCand>& cands, bool big, F&& f) {
  for (int wide = 0; wide < 2; ++wide) {
    for (int base_kind = 0; base_kind < 4; ++base_kind) {
      std::vector<u32> pool;
      for (u32 i = 0; i < cands.size(); ++i) {
        const OwnerCand& c = cands[i];
        if (c.odd == (base_kind & 1) && c.lanex == ((base_kind >> 1) & 1) && c.wide == wide) {
          pool.push_back(i);
        }
      }
      if (pool.empty()) continue;
      if ((base_kind >> 1) & 1) {
        f(0U, pool, wide, base_kind);
        continue;
      }
      std::vector<u32> segs(pool.size());
      for (std::size_t i = 0; i < pool.size(); ++i) {
        segs[i] = (cands[pool[i]].xl >> kTileLowLog) & kSegMask;
      }
      const std::size_t full =
          static_cast<std::size_t>(kOwnerShapeSlots[owner_first_shape(wide != 0, false, big)]);
      for (const OwnerGroup& g : owner_layout_groups(segs, full)) {
        std::vector<u32> members(g.members.size());
        for (std::size_t i = 0; i < members.size(); ++i) members[i] = pool[g.members[i]];
        f(g.lam, members, wide, base_kind);
      }
    }
  }
}

// Bundles the owner sweep needs for these slots, and the sum over layout
// groups of (slots in the group's partial bundle)^2, which grows as partial
// bundles concentrate.
struct OwnerCost {
  std::size_t bundles = 0;
  std::size_t spread = 0;
};

OwnerCost owner_cost(const std::vector<OwnerCand>& cands, bool big) {
  OwnerCost c;
  owner_pass_groups(cands, big, [&](u32, const std::vector<u32>& members, int wide, int base_kind) {
    const bool lanex = (base_kind >> 1) & 1;
    const int first = owner_first_shape(wide != 0, lanex, big);
    owner_group_shapes(members.size(), first, owner_last_shape(wide != 0, lanex),
                       [&](int, std::size_t) { ++c.bundles; });
    const std::size_t r = members.size() % static_cast<std::size_t>(kOwnerShapeSlots[first]);
    c.spread += r * r;
  });
  return c;
}

// Segment part (pass-local bits kTileLowLog.. of x) of x in U.
u32 span_seg(const Span& sp, u64 x, int l) {
  u32 sel = 0;
  const u64 r = sp.reduce(x, &sel);
  u32 out = 0;
  int a = 0;
  for (int q = l; q < 64; ++q) {
    if (!((sp.coord >> q) & 1ULL)) continue;
    if ((r >> q) & 1ULL) out |= 1U << a;
    ++a;
  }
  return out | (sel << a);
}

// Moves x-groups between the passes whose subspace holds them so that the
// owner sweep needs fewer bundles (each bundle reads every tile of its pass):
// local search over single moves, taking a move when it lowers the bundle
// count, or keeps it and concentrates the partial bundles. Passes left
// without an x-group are dropped (pass 0 stays: it carries the diagonal
// items). Items keep their input order within a pass.
// Perturbation rounds of the packing search, for states of kOwnerWideMinQubits
// and up (where a bundle reads gigabytes) and up to this many x-groups
// (structured operators; each round is a full local search).
constexpr int kPackRounds = 100;
constexpr std::size_t kPackRoundsMaxGroups = 1024;

void pack_owner_passes(const std::vector<ExpItem>& items, const std::vector<u64>& xs,
                       std::vector<Span>& spans, std::vector<std::vector<u64>>& pass_xs, int l,
                       bool big) {
  const std::size_t P = spans.size();
  if (P < 2) return;
  std::unordered_map<u64, u32> group_of;
  std::unordered_map<u64, std::vector<u32>> by_high;  // x >> l -> x-groups
  for (u32 g = 0; g < xs.size(); ++g) {
    group_of.emplace(xs[g], g);
    by_high[xs[g] >> l].push_back(g);
  }
  // Each x-group's items (consecutive in the input) and the passes holding it
  // with its pass-local x. Every pass's subspace holds the low l bits, so x
  // lies in it exactly when x >> l lies in the span of its coordinate bits
  // and generators above them: 2^(k - l) lookups per pass.
  std::vector<u32> first_item(xs.size(), 0), num_items(xs.size(), 0);
  for (u32 i = 0; i < items.size(); ++i) {
    if (items[i].x == 0) continue;
    const u32 g = group_of.at(items[i].x);
    if (num_items[g]++ == 0) first_item[g] = i;
  }
  std::vector<std::vector<std::pair<int, u32>>> cover(xs.size());
  for (std::size_t p = 0; p < P; ++p) {
    const Span& sp = spans[p];
    std::vector<u64> basis;
    for (int q = l; q < 64; ++q) {
      if ((sp.coord >> q) & 1ULL) basis.push_back(1ULL << (q - l));
    }
    for (u64 gv : sp.gens) basis.push_back(gv >> l);
    std::vector<u64> elems{0};
    for (u64 b : basis) {
      const std::size_t n0 = elems.size();
      for (std::size_t e = 0; e < n0; ++e) elems.push_back(elems[e] ^ b);
    }
    for (u64 h : elems) {
      const auto it = by_high.find(h);
      if (it == by_high.end()) continue;
      for (u32 g : it->second) {
        if (!sp.contains(xs[g])) fail_runtime("owner pass packing: span membership differs");
        cover[g].push_back({static_cast<int>(p), span_seg(sp, xs[g], l) << kTileLowLog});
      }
    }
  }
  std::vector<int> at(xs.size(), -1);
  std::vector<std::vector<u32>> members(P);  // ascending x-group indices (= item order)
  for (std::size_t p = 0; p < P; ++p) {
    for (u64 x : pass_xs[p]) {
      const u32 g = group_of.at(x);
      at[g] = static_cast<int>(p);
      members[p].push_back(g);
    }
    std::sort(members[p].begin(), members[p].end());
  }
  auto seg_in = [&](u32 g, int p) -> u32 {
    for (const auto& [q, xl] : cover[g]) {
      if (q == p) return xl;
    }
    fail_runtime("owner pass packing lost an x-group");
  };
  // Cost of pass p with x-group `add` moved in and `drop` moved out.
  constexpr u32 kNone = 0xFFFFFFFFU;
  std::vector<OwnerCand> cands;
  auto cost = [&](int p, u32 add, u32 drop) {
    cands.clear();
    auto put = [&](u32 g) {
      for (u32 i = 0; i < num_items[g]; ++i) owner_cands(items[first_item[g] + i], 0, seg_in(g, p), cands);
    };
    bool added = add == kNone;
    for (u32 g : members[p]) {
      if (!added && add < g) {
        put(add);
        added = true;
      }
      if (g != drop) put(g);
    }
    if (!added) put(add);
    return owner_cost(cands, big);
  };
  std::vector<OwnerCost> pc(P);
  for (std::size_t p = 0; p < P; ++p) pc[p] = cost(static_cast<int>(p), kNone, kNone);
  // Most moves happen in the first sweep; large unstructured sets (each
  // evaluation is a layout grouping of two passes) stop after it.
  auto relocate = [&](u32 g, int q, const OwnerCost& cp, const OwnerCost& cq) {
    const int p = at[g];
    at[g] = q;
    members[p].erase(std::lower_bound(members[p].begin(), members[p].end(), g));
    members[q].insert(std::lower_bound(members[q].begin(), members[q].end(), g), g);
    pc[p] = cp;
    pc[q] = cq;
  };
  auto local_search = [&](int sweeps) {
    for (int sweep = 0; sweep < sweeps; ++sweep) {
      bool moved = false;
      for (u32 g = 0; g < xs.size(); ++g) {
        if (cover[g].size() < 2 || at[g] < 0) continue;
        const int p = at[g];
        const OwnerCost cp = cost(p, kNone, g);
        int best_q = -1;
        long best_db = 0, best_ds = 0;
        OwnerCost best_cq{};
        for (const auto& [q, xl] : cover[g]) {
          if (q == p) continue;
          const OwnerCost cq = cost(q, g, kNone);
          const long db = static_cast<long>(cp.bundles + cq.bundles) -
                          static_cast<long>(pc[p].bundles + pc[q].bundles);
          const long ds = static_cast<long>(cp.spread + cq.spread) -
                          static_cast<long>(pc[p].spread + pc[q].spread);
          if ((db < best_db) || (db == best_db && ds > best_ds)) {
            best_q = q;
            best_db = db;
            best_ds = ds;
            best_cq = cq;
          }
        }
        if (best_q < 0) continue;
        relocate(g, best_q, cp, best_cq);
        moved = true;
      }
      if (!moved) break;
    }
  };
  auto total = [&]() {
    std::size_t b = 0;
    for (const OwnerCost& c : pc) b += c.bundles;
    return b;
  };
  // Most moves happen in the first sweep; large unstructured sets (each
  // evaluation is a layout grouping of two passes) stop after it.
  local_search(xs.size() > 4096 ? 1 : 4);
  if (big && xs.size() <= kPackRoundsMaxGroups) {
    // Perturbation rounds: a few x-groups move to another pass that holds
    // them at random (a fixed-seed generator, so plans are reproducible),
    // the local search runs again, and the best assignment is kept.
    std::vector<u32> movable;
    for (u32 g = 0; g < xs.size(); ++g) {
      if (cover[g].size() >= 2 && at[g] >= 0) movable.push_back(g);
    }
    u64 rng = 0x9E3779B97F4A7C15ULL ^ xs.size();
    auto next = [&rng]() {
      rng += 0x9E3779B97F4A7C15ULL;
      u64 z = rng;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      return z ^ (z >> 31);
    };
    std::vector<int> best_at = at;
    std::vector<std::vector<u32>> best_members = members;
    std::vector<OwnerCost> best_pc = pc;
    std::size_t best_total = total();
    for (int round = 0; round < kPackRounds && !movable.empty(); ++round) {
      const std::size_t kicks = 1 + movable.size() / 16;
      for (std::size_t k = 0; k < kicks; ++k) {
        const u32 g = movable[next() % movable.size()];
        const auto& cv = cover[g];
        const int q = cv[next() % cv.size()].first;
        if (q == at[g]) continue;
        relocate(g, q, cost(at[g], kNone, g), cost(q, g, kNone));
      }
      local_search(4);
      const std::size_t t = total();
      if (t < best_total) {
        best_total = t;
        best_at = at;
        best_members = members;
        best_pc = pc;
      } else {
        at = best_at;
        members = best_members;
        pc = best_pc;
      }
    }
  }
  std::vector<Span> kept_spans;
  std::vector<std::vector<u64>> kept_xs;
  for (std::size_t p = 0; p < P; ++p) {
    if (p > 0 && members[p].empty()) continue;
    kept_spans.push_back(spans[p]);
    std::vector<u64> xp;
    for (u32 g : members[p]) xp.push_back(xs[g]);
    kept_xs.push_back(std::move(xp));
  }
  spans = std::move(kept_spans);
  pass_xs = std::move(kept_xs);
}

}  // namespace

void plan_tiled_passes(std::size_t n, std::vector<ExpItem>& items, std::vector<PassLaunch>& passes,
                       bool diag_separate, u32* num_diag_front, std::size_t max_passes) {
  const int l = kTileLowLog;
  const int kmax = std::min<int>(kTileMaxLog, static_cast<int>(n) - 3);
  const u64 lowmask = (1ULL << l) - 1;
  // Distinct nonzero x in first-appearance order.
  std::vector<u64> xs;
  std::unordered_map<u64, int> seen;
  for (const ExpItem& it : items) {
    if (it.x && seen.emplace(it.x, 1).second) xs.push_back(it.x);
  }
  std::vector<char> covered(xs.size(), 0);
  std::size_t remaining = xs.size();
  std::vector<Span> spans;
  std::vector<std::vector<u64>> pass_xs;
  const bool structured = xs.size() <= 1024;
  // Unstructured sets: the uncovered x-groups keyed by their part above the l
  // contiguous low bits (those bits are coordinates of every pass, so an
  // x-group lies in U exactly when its high part lies in the span of the
  // generators' high parts), and the distinct high parts in first-appearance
  // order (the candidate generators).
  std::unordered_map<u64, std::vector<u32>> by_high;
  std::vector<u64> highs;
  if (!structured) {
    by_high.reserve(xs.size() * 2);
    for (std::size_t i = 0; i < xs.size(); ++i) {
      auto [it, fresh] = by_high.try_emplace(xs[i] >> l);
      if (fresh) highs.push_back(xs[i] >> l);
      it->second.push_back(static_cast<u32>(i));
    }
  }
  std::size_t hcursor = 0;
  std::vector<u64> span_elems;  // high parts of the current pass's generator span
  while (remaining > 0 || spans.empty()) {
    Span sp;
    sp.coord = lowmask;
    if (remaining > 0) {
      if (structured) {
        // Greedy: repeatedly extend U by the option (the coordinate bits of one
        // candidate's residual, or that residual as a generator) that covers
        // the most uncovered x-groups per added dimension.
        while (sp.dim() < kmax) {
          double best_score = 0.0;
          int best_cost = 0;
          u64 best_bits = 0;
          bool best_is_gen = false;
          std::size_t tried = 0;
          for (std::size_t i = 0; i < xs.size() && tried < 192; ++i) {
            if (covered[i] || sp.contains(xs[i])) continue;
            ++tried;
            const u64 resid = sp.reduce(xs[i], nullptr) & ~sp.coord;
            for (int option = 0; option < 2; ++option) {
              const bool as_gen = option == 1;
              const int cost = as_gen ? 1 : __builtin_popcountll(resid);
              if (as_gen && __builtin_popcountll(resid) == 1) continue;
              if (sp.dim() + cost > kmax) continue;
              Span trial = sp;
              if (as_gen) trial.add_gen(resid); else trial.add_coords(resid);
              std::size_t gain = 0;
              for (std::size_t j = 0; j < xs.size(); ++j) {
                if (!covered[j] && trial.contains(xs[j])) ++gain;
              }
              const double score = static_cast<double>(gain) / cost;
              if (score > best_score + 1e-12 || (score > best_score - 1e-12 && cost < best_cost)) {
                best_score = score;
                best_cost = cost;
                best_bits = resid;
                best_is_gen = as_gen;
              }
            }
          }
          if (best_cost == 0) break;
          if (best_is_gen) sp.add_gen(best_bits); else sp.add_coords(best_bits);
        }
      } else {
        // Unstructured x-masks: greedy generators. Each step takes, among the
        // first kUnstructuredCandidates uncovered high parts outside the
        // current span, the one whose coset of the span holds the most
        // uncovered x-groups (2^d hash lookups per candidate), so coincidences
        // x_a ^ x_b = x_c, plentiful in large random sets, share passes.
        constexpr std::size_t kUnstructuredCandidates = 48;
        span_elems.assign(1, 0);
        std::unordered_set<u64> in_span{0};
        while (sp.dim() < kmax) {
          while (hcursor < highs.size() && by_high.find(highs[hcursor]) == by_high.end()) ++hcursor;
          u64 best = 0;
          std::size_t best_gain = 0;
          std::size_t seen_c = 0;
          for (std::size_t j = hcursor; j < highs.size() && seen_c < kUnstructuredCandidates; ++j) {
            const u64 c = highs[j];
            if (by_high.find(c) == by_high.end() || in_span.count(c)) continue;
            ++seen_c;
            std::size_t gain = 0;
            for (u64 e : span_elems) {
              const auto it = by_high.find(e ^ c);
              if (it != by_high.end()) gain += it->second.size();
            }
            if (gain > best_gain) {
              best_gain = gain;
              best = c;
            }
          }
          if (best_gain == 0) break;
          const std::size_t d0 = span_elems.size();
          for (std::size_t i = 0; i < d0; ++i) {
            span_elems.push_back(span_elems[i] ^ best);
            in_span.insert(span_elems[i] ^ best);
          }
          sp.add_gen(sp.reduce(best << l, nullptr) & ~sp.coord);
        }
      }
    }
    // Pad U with the lowest unused index bits up to the full tile dimension:
    // every pass then runs with tiles of 2^kmax points (more work per tile,
    // the same single read of psi).
    for (int q = l; q < static_cast<int>(n) && sp.dim() < kmax; ++q) {
      u64 occupied = sp.coord;
      for (u64 g : sp.gens) occupied |= 1ULL << Span::top(g);
      if (!((occupied >> q) & 1ULL)) sp.add_coords(1ULL << q);
    }
    // Membership: a scan of all x-groups for structured sets; for
    // unstructured sets, lookups of every high part of U (the generator span
    // plus the padding coordinates above the low bits: 2^(kmax - l) keys).
    std::vector<u64> in_pass;
    if (structured) {
      for (std::size_t j = 0; j < xs.size(); ++j) {
        if (!covered[j] && sp.contains(xs[j])) {
          covered[j] = 1;
          --remaining;
          in_pass.push_back(xs[j]);
        }
      }
    } else if (remaining > 0) {
      std::vector<u64> keys = span_elems;
      for (int q = l; q < 64; ++q) {
        if (!((sp.coord >> q) & 1ULL)) continue;
        const u64 bit = 1ULL << (q - l);
        const std::size_t k0 = keys.size();
        for (std::size_t i = 0; i < k0; ++i) keys.push_back(keys[i] ^ bit);
      }
      for (u64 key : keys) {
        const auto it = by_high.find(key);
        if (it == by_high.end()) continue;
        for (u32 j : it->second) {
          if (!sp.contains(xs[j])) fail_runtime("tiled pass planning lost an x-group");
          covered[j] = 1;
          --remaining;
          in_pass.push_back(xs[j]);
        }
        by_high.erase(it);
      }
    }
    spans.push_back(sp);
    pass_xs.push_back(std::move(in_pass));
    if (remaining == 0) break;
    if (spans.size() > max_passes) {
      // More passes than the caller accepts: report the count only.
      passes.assign(spans.size(), PassLaunch{});
      if (num_diag_front) *num_diag_front = 0;
      return;
    }
  }

  if (kOwnerSweep) {
    pack_owner_passes(items, xs, spans, pass_xs, l, static_cast<int>(n) >= kOwnerWideMinQubits);
  }

  // Emit passes: local coordinates, segment offsets, free positions; items are
  // reordered pass by pass. Diagonal items either lead the item list (swept
  // separately) or ride in the first pass.
  std::unordered_map<u64, int> pass_of;
  for (std::size_t p = 0; p < pass_xs.size(); ++p) {
    for (u64 x : pass_xs[p]) pass_of[x] = static_cast<int>(p);
  }
  std::vector<ExpItem> ordered;
  if (diag_separate) {
    for (const ExpItem& it : items) {
      if (it.x == 0) ordered.push_back(it);
    }
  }
  if (num_diag_front) *num_diag_front = static_cast<u32>(ordered.size());
  passes.clear();
  for (std::size_t p = 0; p < spans.size(); ++p) {
    const Span& sp = spans[p];
    PassLaunch pl{};
    PassDesc& d = pl.desc;
    d.n = static_cast<int>(n);
    d.l = l;
    std::vector<int> apos;
    for (int q = l; q < 64; ++q) {
      if ((sp.coord >> q) & 1ULL) apos.push_back(q);
    }
    const int a = static_cast<int>(apos.size());
    const int m = static_cast<int>(sp.gens.size());
    d.k = l + a + m;
    u64 occupied = sp.coord;
    for (u64 g : sp.gens) occupied |= 1ULL << Span::top(g);
    d.nfree = 0;
    d.free_mask = 0;
    for (int q = 0; q < static_cast<int>(n); ++q) {
      if (!((occupied >> q) & 1ULL)) {
        d.free_pos[d.nfree++] = static_cast<unsigned char>(q);
        d.free_mask |= 1ULL << q;
      }
    }
    for (int s = 0; s < (1 << (a + m)); ++s) {
      u64 off = 0;
      for (int i = 0; i < a; ++i) {
        if ((s >> i) & 1) off |= 1ULL << apos[i];
      }
      for (int i = 0; i < m; ++i) {
        if ((s >> (a + i)) & 1) off ^= sp.gens[i];
      }
      d.seg_off[s] = off;
    }
    // x in U -> its coordinates in the basis (low bits, apos, gens).
    auto x_local = [&](u64 v) -> u64 {
      u32 sel = 0;
      const u64 r = sp.reduce(v, &sel);
      u64 out = r & lowmask;
      for (int i = 0; i < a; ++i) {
        if ((r >> apos[i]) & 1ULL) out |= 1ULL << (l + i);
      }
      return out | (static_cast<u64>(sel) << (l + a));
    };
    // z -> local sign mask: parity(z & (rep ^ b_local(u))) = parity(z & rep) ^ parity(zl & u).
    auto z_local = [&](u64 v) -> u64 {
      u64 out = v & lowmask;
      for (int i = 0; i < a; ++i) {
        if ((v >> apos[i]) & 1ULL) out |= 1ULL << (l + i);
      }
      for (int i = 0; i < m; ++i) {
        if (__builtin_popcountll(v & sp.gens[i]) & 1) out |= 1ULL << (l + a + i);
      }
      return out;
    };
    pl.first_item = static_cast<u32>(ordered.size());
    if (p == 0 && diag_separate) {
      // The separately swept diagonal items lead the list; their local masks
      // in pass 0's coordinates serve the tiled sigma sweep.
      for (u32 i = 0; i < (num_diag_front ? *num_diag_front : 0); ++i) {
        ExpItem& e = ordered[i];
        e.xl = 0;
        e.pbit = 0;
        for (int j = 0; j < e.nz; ++j) e.zl[j] = z_local(e.z[j]);
        e.cls_log = kNoClasses;
        e.variant = static_cast<unsigned char>(8 + e.nz - 1);
      }
    }
    for (int diag_round = 1; diag_round >= 0; --diag_round) {
      for (const ExpItem& it : items) {
        const bool is_diag = it.x == 0;
        if (is_diag != (diag_round == 1)) continue;
        if (is_diag && diag_separate) continue;
        if (is_diag ? p != 0 : pass_of.at(it.x) != static_cast<int>(p)) continue;
        ExpItem e = it;
        e.xl = is_diag ? 0 : x_local(it.x);
        // Pivot: the highest local bit of x. Any set bit of x selects one
        // representative per pair {b, b ^ x}; a pivot above the low 6 local
        // bits keeps each wave's LDS reads on 64 consecutive points.
        e.pbit = is_diag ? 0 : (1ULL << (63 - __builtin_clzll(e.xl)));
        for (int j = 0; j < e.nz; ++j) e.zl[j] = z_local(it.z[j]);
        assign_step_classes(e, 0);
        ordered.push_back(e);
      }
    }
    // Waves of a block wait for each other at every tile: group items of
    // similar cost (key count) into the same block.
    // Sweep variants: the off-diagonal items with step classes share the
    // pass's largest class dimension and key count, so a pass usually needs
    // one launch; others get their own instance.
    {
      int dmax = -1;
      bool wide = false;
      for (std::size_t i = pl.first_item; i < ordered.size(); ++i) {
        const ExpItem& it = ordered[i];
        if (it.xl && it.cls_log != kNoClasses) {
          dmax = std::max<int>(dmax, it.cls_log);
          wide = wide || it.nz > 2;
        }
      }
      for (std::size_t i = pl.first_item; i < ordered.size(); ++i) {
        ExpItem& it = ordered[i];
        if (it.xl == 0) {
          it.variant = static_cast<unsigned char>(8 + it.nz - 1);
        } else if (it.cls_log != kNoClasses) {
          assign_step_classes(it, dmax);
          if (it.cls_log != dmax) fail_runtime("step classes could not be widened");
          it.variant = static_cast<unsigned char>(16 + 2 * dmax + (wide ? 1 : 0));
        } else {
          it.variant = static_cast<unsigned char>(it.nz - 1);
        }
      }
    }
    // One launch per sweep variant present in the pass.
    std::stable_sort(ordered.begin() + pl.first_item, ordered.end(),
                     [](const ExpItem& a, const ExpItem& b) {
                       return sweep_variant(a) < sweep_variant(b);
                     });
    pl.num_items = static_cast<u32>(ordered.size()) - pl.first_item;
    pl.item_groups = 0;
    for (u32 i = 0; i < pl.num_items;) {
      const int v = sweep_variant(ordered[pl.first_item + i]);
      u32 e = i;
      while (e < pl.num_items && sweep_variant(ordered[pl.first_item + e]) == v) ++e;
      const u32 groups = (e - i + kTileWaves - 1) / kTileWaves;
      pl.subs.push_back({v, i, e - i, groups});
      pl.item_groups += groups;
      i = e;
    }
    pl.item_groups = std::max<u32>(1, pl.item_groups);
    const u64 ntiles = 1ULL << (n - d.k);
    u64 tb = (1024 + pl.item_groups - 1) / pl.item_groups;
    tb = std::max<u64>(8, (tb + 7) / 8 * 8);
    tb = std::min<u64>(tb, std::max<u64>(8, ntiles / 8 * 8));
    pl.tile_blocks = static_cast<u32>(tb);
    passes.push_back(pl);
  }
  items = std::move(ordered);
}

namespace {

// Bundle layout for a register-pivot bundle: new segment coordinate 5 is the
// functional lam (a set bit of every item's x lands at local bit 11); the
// others are the old coordinates except lam's top bit p, in order. With
// s = A s' the basis images are A e'_i = e_{pi(i)} (+ e_p if lam has bit
// pi(i)) for i < 5 and A e'_5 = e_p.
struct SegLayout {
  u32 img[kSegBits];  // A e'_i as old segment vectors
  u32 lam;
  int p;
  int pi[kSegBits - 1];
};

SegLayout seg_layout(u32 lam) {
  SegLayout L{};
  L.lam = lam;
  if (lam == 0) {
    for (int i = 0; i < kSegBits; ++i) L.img[i] = 1U << i;
    L.p = -1;
    return L;
  }
  L.p = 31 - __builtin_clz(lam);
  int c = 0;
  for (int i = 0; i < kSegBits; ++i) {
    if (i != L.p) L.pi[c++] = i;
  }
  for (int i = 0; i < kSegBits - 1; ++i) {
    L.img[i] = (1U << L.pi[i]) | (((lam >> L.pi[i]) & 1U) << L.p);
  }
  L.img[kSegBits - 1] = 1U << L.p;
  return L;
}

// x (12-bit pass-local) in the bundle layout: the segment part's last
// coordinate becomes lam(x).
u32 layout_x(const SegLayout& L, u32 xl) {
  if (L.p < 0) return xl;
  const u32 s = (xl >> kTileLowLog) & kSegMask;
  u32 s2 = 0;
  for (int i = 0; i < kSegBits - 1; ++i) s2 |= ((s >> L.pi[i]) & 1U) << i;
  s2 |= static_cast<u32>(host_parity(L.lam & s)) << (kSegBits - 1);
  return (xl & ((1U << kTileLowLog) - 1U)) | (s2 << kTileLowLog);
}

// z character (12-bit pass-local) in the bundle layout: bit i of the new
// segment part is parity(zl_seg & A e'_i).
u32 layout_z(const SegLayout& L, u32 zl) {
  const u32 s = (zl >> kTileLowLog) & kSegMask;
  u32 s2 = 0;
  for (int i = 0; i < kSegBits; ++i) s2 |= static_cast<u32>(host_parity(s & L.img[i])) << i;
  return (zl & ((1U << kTileLowLog) - 1U)) | (s2 << kTileLowLog);
}

u64 host_compress(u64 v, u64 mask) {
  u64 out = 0;
  int k = 0;
  while (mask) {
    const u64 low = mask & (0 - mask);
    if (v & low) out |= 1ULL << k;
    ++k;
    mask ^= low;
  }
  return out;
}

}  // namespace

void plan_owner_bundles(const std::vector<ExpItem>& items, std::vector<PassLaunch>& passes,
                        std::vector<OwnerBundle>& bundles) {
  for (PassLaunch& pl : passes) {
    const PassDesc& d = pl.desc;
    // The diagonal items keep their sweep variants; the off-diagonal ones move
    // to owner bundles.
    std::vector<SubLaunch> keep;
    u32 groups = 0;
    for (const SubLaunch& sub : pl.subs) {
      if (sub.variant >= 8 && sub.variant < 16) {
        keep.push_back(sub);
        groups += sub.groups;
      }
    }
    pl.subs = keep;
    pl.owner_first = static_cast<u32>(bundles.size());
    pl.owner_count = 0;
    std::vector<OwnerCand> cands;
    for (u32 r = 0; r < pl.num_items; ++r) {
      const ExpItem& e = items[pl.first_item + r];
      if (e.x != 0) owner_cands(e, r, static_cast<u32>(e.xl), cands);
    }
    const bool big = d.n >= kOwnerWideMinQubits;
    owner_pass_groups(cands, big, [&](u32 lam, const std::vector<u32>& group, int wide, int base_kind) {
      const bool lanex = (base_kind >> 1) & 1;
      const SegLayout L = seg_layout(lanex ? 0U : lam);
      owner_group_shapes(group.size(), owner_first_shape(wide != 0, lanex, big),
                         owner_last_shape(wide != 0, lanex), [&](int shape, std::size_t s) {
        const int M = kOwnerShapeSlots[shape];
        OwnerBundle bd{};
        for (int i = 0; i < kSegBits; ++i) bd.seg_basis[i] = d.seg_off[L.img[i]];
        bd.kind = static_cast<u32>(shape * 4 + base_kind);
        for (int m = 0; m < M; ++m) {
          OwnerSlot& S = bd.slot[m];
          if (s + m >= group.size()) {
            // Padding: in-bounds reads, no keys written.
            S.addr0 = lanex ? 0U : (4U << 13);
            continue;
          }
          const OwnerCand& c = cands[group[s + m]];
          const ExpItem& e = items[pl.first_item + c.item];
          const u32 xl = layout_x(L, c.xl);
          S.item = c.item;
          const u32 xt16 = (xl & 511U) << 4, xr = (xl >> 9) & 7U;
          if (c.lanex) {
            if (xl >= (1U << kTileLowLog)) fail_runtime("lane-pivot item has x outside the low bits");
            S.hbit = static_cast<u32>(31 - __builtin_clz(xl));
            S.addr0 = xt16;
          } else {
            if (xr < 4U) fail_runtime("owner bundle layout misses an item's x");
            S.addr0 = xt16 | (xr << 13);
          }
          for (int q = 0; q < c.ncols; ++q) {
            const int j = c.cols[q];
            const u32 zl = layout_z(L, static_cast<u32>(e.zl[j]));
            OwnerKey& K = S.key[q];
            const u64 zt = host_compress(e.z[j], d.free_mask);
            if (zt >> 30) fail_runtime("owner sweep supports at most 42 qubits");
            K.w = static_cast<u32>(zt) | (((zl >> 9) & 1U) << 30) | (((zl >> 10) & 1U) << 31);
            K.zlt = zl & 511U;
            K.info = (((zl >> 11) & 1U) << 2) | (static_cast<u32>(j) << 4) | (1U << 15);
          }
        }
        bundles.push_back(bd);
        ++pl.owner_count;
      });
    });
    groups += pl.owner_count;
    pl.item_groups = std::max<u32>(1, groups);
    if (pl.owner_count) {
      const u64 ntiles = 1ULL << (d.n - d.k);
      u64 tb = std::max<u64>(ntiles / kOwnerTilesPerBlock,
                             (kOwnerMinBlocks + pl.item_groups - 1) / pl.item_groups);
      u64 pow2 = 8;
      while (pow2 < tb) pow2 <<= 1;
      while (pow2 > 8 && pow2 * pl.num_items > kOwnerMaxPassRows) pow2 >>= 1;
      pl.tile_blocks = static_cast<u32>(std::min<u64>(pow2, std::max<u64>(8, ntiles)));
    }
  }
}

// ------------------------------------------------------------ plan

namespace {

// Sigma key records, pass by pass (pass 0 from item 0: it also carries the
// diagonal items), each pass's list padded to a whole number of chunks with
// zero records. Sets each pass's [sigma_begin, sigma_end).
void build_sigma_records(const std::vector<ExpItem>& items, std::vector<PassLaunch>& passes,
                         std::vector<SigmaKey>& sigma_keys) {
  for (std::size_t p = 0; p < passes.size(); ++p) {
    PassLaunch& pl = passes[p];
    const u32 first = p == 0 ? 0 : pl.first_item;
    pl.sigma_begin = static_cast<u32>(sigma_keys.size());
    for (u32 i = first; i < pl.first_item + pl.num_items; ++i) {
      const ExpItem& it = items[i];
      for (int j = 0; j < it.nz; ++j) {
        SigmaKey k{};
        k.z = it.z[j];
        k.zl = static_cast<u32>(it.zl[j]);
        k.xl = static_cast<u32>(it.xl) | (j + 1 == it.nz ? kSigmaLast : 0U);
        for (u32 r = 1; r < 4; ++r) {
          k.sh[r - 1] = 0x3FF00000U | (host_parity(k.zl & (r << 10)) ? 0x80000000U : 0U);
        }
        k.key = it.key[j];
        sigma_keys.push_back(k);
      }
    }
    SigmaKey pad{};
    pad.key = kSigmaPadKey;
    while ((sigma_keys.size() - pl.sigma_begin) % kSigmaChunk != 0) sigma_keys.push_back(pad);
    pl.sigma_end = static_cast<u32>(sigma_keys.size());
  }
}

}  // namespace

ExpectationPlan::~ExpectationPlan() {
  int prev = -1;
  const hipError_t r = hipGetDevice(&prev);
  (void)hipSetDevice(device);
  device_free(d_items);
  device_free(d_partial_offset);
  device_free(d_partial_count);
  device_free(d_term_key);
  device_free(d_group_x);
  device_free(d_group_begin);
  device_free(d_key_z);
  device_free(d_key_y);
  device_free(d_key_c);
  device_free(d_key_term_begin);
  device_free(d_passes);
  device_free(d_flat_key);
  device_free(d_key_terms);
  device_free(d_bundles);
  device_free(d_owner_table);
  device_free(d_owner_entry);
  device_free(d_sigma_keys);
  if (r == hipSuccess && prev >= 0) (void)hipSetDevice(prev);
}

namespace {

std::shared_ptr<ExpectationPlan> build_plan(const wg_pauli_sum* op, hipStream_t stream) {
  auto plan = std::make_shared<ExpectationPlan>();
  plan->n = op->num_qubits;
  plan->T = op->num_terms;
  plan->device = op->device;
  const std::size_t T = op->num_terms, W = op->words;
  std::vector<u64> hx(T * W), hz(T * W);
  std::vector<double2> hc(T);
  if (T * W) {
    check_hip(hipMemcpyAsync(hx.data(), op->x, T * W * sizeof(u64), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    check_hip(hipMemcpyAsync(hz.data(), op->z, T * W * sizeof(u64), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
  }
  if (T) {
    check_hip(hipMemcpyAsync(hc.data(), op->coeffs, T * sizeof(double2), hipMemcpyDeviceToHost,
                             stream),
              "copy device buffer to host");
  }
  sync(stream);

  // Distinct keys in first-appearance order; groups of keys by x, in
  // first-appearance order of x.
  struct PairHash {
    std::size_t operator()(const std::pair<u64, u64>& k) const noexcept {
      return (k.first * 0x9E3779B97F4A7C15ULL) ^ (k.second + 0x632BE59BD9B4E019ULL + (k.first << 7));
    }
  };
  std::unordered_map<std::pair<u64, u64>, u32, PairHash> key_index;
  std::vector<u64> kx, kz;
  std::vector<double2> kc;
  std::vector<u32> term_key(T);
  for (std::size_t t = 0; t < T; ++t) {
    const u64 xm = W ? hx[t * W] : 0, zm = W ? hz[t * W] : 0;
    auto [itr, inserted] = key_index.try_emplace({xm, zm}, static_cast<u32>(kx.size()));
    if (inserted) {
      kx.push_back(xm);
      kz.push_back(zm);
      kc.push_back(hc[t]);
    } else {
      kc[itr->second].x += hc[t].x;
      kc[itr->second].y += hc[t].y;
    }
    term_key[t] = itr->second;
  }
  const std::size_t K = kx.size();
  std::unordered_map<u64, u32> group_index;
  std::vector<u64> gx;
  std::vector<std::vector<u32>> gkeys;
  for (std::size_t k = 0; k < K; ++k) {
    auto [itr, inserted] = group_index.try_emplace(kx[k], static_cast<u32>(gx.size()));
    if (inserted) {
      gx.push_back(kx[k]);
      gkeys.emplace_back();
    }
    gkeys[itr->second].push_back(static_cast<u32>(k));
  }
  std::vector<ExpItem> items;
  std::vector<u64> flat_z;
  std::vector<int> flat_y;
  std::vector<double2> flat_c;
  std::vector<u32> group_begin;
  for (std::size_t g = 0; g < gx.size(); ++g) {
    group_begin.push_back(static_cast<u32>(flat_z.size()));
    const auto& ks = gkeys[g];
    for (std::size_t s = 0; s < ks.size(); s += kZ) {
      ExpItem it{};
      it.x = gx[g];
      it.nz = static_cast<int>(std::min<std::size_t>(kZ, ks.size() - s));
      for (int j = 0; j < it.nz; ++j) {
        it.z[j] = kz[ks[s + j]];
        it.key[j] = ks[s + j];
        it.y[j] = __builtin_popcountll(gx[g] & kz[ks[s + j]]) & 3;
      }
      items.push_back(it);
    }
    for (u32 k : ks) {
      flat_z.push_back(kz[k]);
      flat_y.push_back(__builtin_popcountll(kx[k] & kz[k]) & 3);
      flat_c.push_back(kc[k]);
    }
  }
  group_begin.push_back(static_cast<u32>(flat_z.size()));
  // Terms of each key (keys in group order, terms in term order).
  std::vector<u32> flat_pos(K);
  {
    u32 pos = 0;
    for (std::size_t g = 0; g < gx.size(); ++g) {
      for (u32 k : gkeys[g]) flat_pos[k] = pos++;
    }
  }
  std::vector<u32> flat_key(K);
  for (std::size_t k = 0; k < K; ++k) flat_key[flat_pos[k]] = static_cast<u32>(k);
  std::vector<u32> key_term_begin(K + 1, 0), key_terms(T);
  for (std::size_t t = 0; t < T; ++t) key_term_begin[flat_pos[term_key[t]] + 1]++;
  for (std::size_t k = 0; k < K; ++k) key_term_begin[k + 1] += key_term_begin[k];
  {
    std::vector<u32> fill(key_term_begin.begin(), key_term_begin.end() - 1);
    for (std::size_t t = 0; t < T; ++t) key_terms[fill[flat_pos[term_key[t]]]++] = static_cast<u32>(t);
  }

  plan->num_keys = K;
  plan->num_groups = gx.size();
  plan->distinct_x = gx.size();
  plan->direct_items = items;
  plan->h_group_x = gx;
  plan->h_group_begin = group_begin;
  plan->h_key_z = flat_z;
  plan->h_key_y = flat_y;
  plan->h_key_c = flat_c;
  plan->h_term_key = term_key;
  plan->h_key_term_begin = key_term_begin;
  plan->h_key_terms = key_terms;
  plan->tiled = plan->n >= kTiledMinQubits && !items.empty();
  std::vector<u32> offset, count;
  std::vector<SigmaKey> sigma_keys;
  if (plan->tiled) {
    std::vector<ExpItem> tiled_items = items;
    std::vector<PassLaunch> passes;
    std::size_t diag_keys = 0;
    for (const ExpItem& it : items) diag_keys += it.x == 0 ? static_cast<std::size_t>(it.nz) : 0;
    // The transform costs about as much as a few items per tile, so the
    // diagonal keys take it once they fill more than kDiagWhtMinItems items.
    const bool diag_separate = diag_keys > static_cast<std::size_t>(kDiagWhtMinItems) * kZ;
    u32 diag_front = 0;
    const std::size_t max_passes = max_tiled_passes(plan->n, gx.size());
    plan_tiled_passes(plan->n, tiled_items, passes, diag_separate, &diag_front, max_passes);
    if (passes.size() > max_passes) {
      plan->tiled = false;  // launch-bound: one launch per pass would dominate
    } else {
      for (const PassLaunch& p : passes) {
        if (p.desc.k != kTileMaxLog) fail_runtime("tiled expectation plan has a short tile");
      }
      if (kOwnerSweep) plan_owner_bundles(tiled_items, passes, plan->bundles);
      items = std::move(tiled_items);
      plan->passes = std::move(passes);
      plan->diag_items = diag_front;
      plan->diag_keys = diag_front ? static_cast<u32>(diag_keys) : 0;
      const u64 ntiles = u64{1} << (plan->n - kTileMaxLog);
      plan->diag_blocks = static_cast<u32>(std::min<u64>(ntiles, kDiagWhtBlocks));
    }
  }
  if (plan->tiled) {
    offset.resize(items.size());
    count.resize(items.size());
    std::size_t slot = 0;
    for (u32 i = 0; i < plan->diag_items; ++i) {
      offset[i] = static_cast<u32>(slot);
      count[i] = plan->diag_blocks;
      slot += plan->diag_blocks;
    }
    for (PassLaunch& p : plan->passes) {
      p.partial_base = static_cast<u32>(slot);
      for (u32 i = 0; i < p.num_items; ++i) {
        offset[p.first_item + i] = static_cast<u32>(slot + static_cast<std::size_t>(i) * p.tile_blocks);
        count[p.first_item + i] = p.tile_blocks;
      }
      slot += static_cast<std::size_t>(p.num_items) * p.tile_blocks;
    }
    if (slot > 0xFFFFFFFFULL) fail_runtime("tiled expectation plan has too many partial rows");
    plan->partial_slots = slot;
    build_sigma_records(items, plan->passes, sigma_keys);
    plan->sigma_keys = sigma_keys.size();
    int cus = 0;
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, op->device),
              "query compute units");
    const u64 ntiles = u64{1} << (plan->n - kTileMaxLog);
    u32 lb = 0;
    while ((u64{2} << lb) <= std::min<u64>(ntiles, static_cast<u64>(std::max(cus, 1)))) ++lb;
    plan->sigma_log_blocks = lb;
  } else {
    const u64 range = plan->n == 0 ? 1 : (u64{1} << (plan->n - 1));
    u64 bpi = (range + kBlock * 64 - 1) / (kBlock * 64);
    bpi = std::max<u64>(1, std::min<u64>(bpi, 1024));
    plan->blocks_per_item = static_cast<u32>(bpi);
    offset.resize(items.size());
    count.resize(items.size());
    for (std::size_t i = 0; i < items.size(); ++i) {
      offset[i] = static_cast<u32>(i * bpi);
      count[i] = static_cast<u32>(bpi);
    }
    plan->partial_slots = items.size() * bpi;
  }
  plan->num_items = items.size();

  DeviceScope scope(op->device);
  auto upload = [&](auto*& dst, const auto& src) {
    using T2 = std::remove_reference_t<decltype(*dst)>;
    dst = device_alloc<T2>(src.size());
    if (!src.empty()) {
      check_hip(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T2), hipMemcpyHostToDevice,
                               stream),
                "copy host buffer to device");
    }
  };
  upload(plan->d_items, items);
  upload(plan->d_partial_offset, offset);
  upload(plan->d_partial_count, count);
  upload(plan->d_term_key, term_key);
  upload(plan->d_group_x, gx);
  upload(plan->d_group_begin, group_begin);
  upload(plan->d_key_z, flat_z);
  upload(plan->d_key_y, flat_y);
  upload(plan->d_key_c, flat_c);
  upload(plan->d_key_term_begin, key_term_begin);
  upload(plan->d_flat_key, flat_key);
  {
    std::vector<PassDesc> descs;
    for (const PassLaunch& p : plan->passes) descs.push_back(p.desc);
    upload(plan->d_passes, descs);
  }
  upload(plan->d_key_terms, key_terms);
  upload(plan->d_bundles, plan->bundles);
  {
    // One owner launch for all passes: each pass's blocks are contiguous.
    std::vector<u32> entry;
    u64 blocks = 0;
    for (std::size_t p = 0; p < plan->passes.size(); ++p) {
      const PassLaunch& pl = plan->passes[p];
      if (!pl.owner_count) continue;
      const u64 n_blocks = static_cast<u64>(pl.owner_count) * pl.tile_blocks;
      if (blocks + n_blocks > 0x7fffffffULL) fail_runtime("owner sweep launch exceeds the grid limit");
      plan->owner_table.push_back({static_cast<u32>(p), pl.owner_first, pl.owner_count, pl.tile_blocks,
                                   pl.partial_base, static_cast<u32>(blocks)});
      entry.insert(entry.end(), n_blocks / 8, static_cast<u32>(plan->owner_table.size() - 1));
      blocks += n_blocks;
    }
    plan->owner_blocks = blocks;
    upload(plan->d_owner_table, plan->owner_table);
    upload(plan->d_owner_entry, entry);
  }
  upload(plan->d_sigma_keys, sigma_keys);
  sync(stream);
  return plan;
}

}  // namespace

std::shared_ptr<ExpectationPlan> get_plan(wg_pauli_sum* op, hipStream_t stream) {
  std::lock_guard<std::mutex> lk(op->plan_mutex);
  if (!op->plan) op->plan = build_plan(op, stream);
  return op->plan;
}

namespace {

// Runs the forward path on a device-resident complex128 psi: <P> per distinct
// key, E into out_dev, per-term values optionally.
void run_forward(const double2* coeffs, const ExpectationPlan& plan, const double2* psi,
                 double2* term_values, double2* out_dev, Scratch& scratch, hipStream_t stream) {
  double* key_vals = scratch.take<double>(std::max<std::size_t>(plan.num_keys, 1));
  double2* partials =
      plan.num_items
          ? scratch.take<double2>(checked_mul(plan.partial_slots, kZ, "expectation partials"))
          : nullptr;
  forward_into(coeffs, plan, psi, term_values, out_dev, key_vals, partials, stream);
}

std::size_t validate_psi(const wg_pauli_sum* op, const void* psi, std::size_t length);

}  // namespace

std::size_t forward_scratch_bytes(const ExpectationPlan& plan) {
  const std::size_t keys = (std::max<std::size_t>(plan.num_keys, 1) * sizeof(double) + 255) & ~std::size_t{255};
  return keys + (plan.num_items ? checked_mul(checked_mul(plan.partial_slots, kZ, "expectation partials"),
                                              sizeof(double2), "expectation partials")
                                : 0);
}

// Every partial row and key value the reduce and finalize kernels read is
// written earlier in the same sequence, so the buffers need no clearing.
void forward_into(const double2* coeffs, const ExpectationPlan& plan, const double2* psi,
                  double2* term_values, double2* out_dev, double* key_vals, double2* partials,
                  hipStream_t stream) {
  if (plan.num_items) {
    if (plan.tiled) {
      for (u32 k0 = 0; k0 < plan.diag_keys; k0 += kWhtMaxKeys) {
        ProfileScope prof("expect_diag_wht", stream);
        expect_diag_wht_kernel<<<plan.diag_blocks, kTileBlock, 0, stream>>>(
            psi, static_cast<u32>(plan.n - kTileMaxLog), plan.d_items, k0,
            std::min<u32>(kWhtMaxKeys, plan.diag_keys - k0), 0, partials);
        check_launch("expect_diag_wht");
      }
      for (const PassLaunch& p : plan.passes) {
        for (const SubLaunch& sub : p.subs) {
          const unsigned grid = sub.groups * p.tile_blocks;
          ProfileScope prof("expect_tiled", stream);
          kTiledKernels[sub.variant]<<<grid, kTileBlock, 0, stream>>>(
              psi, plan.d_passes + (&p - plan.passes.data()), plan.d_items + p.first_item + sub.first,
              sub.count, sub.groups, p.tile_blocks, p.partial_base + sub.first * p.tile_blocks,
              partials);
          check_launch("expect_tiled");
        }
      }
      if (plan.owner_blocks) {
        ProfileScope prof("expect_owner", stream);
        expect_owner_kernel<<<static_cast<unsigned>(plan.owner_blocks), kOwnerBlock, 0, stream>>>(
            psi, plan.d_passes, plan.d_bundles, plan.d_owner_table, plan.d_owner_entry, partials);
        check_launch("expect_owner");
      }
    } else {
      const u64 len = u64{1} << plan.n;
      ProfileScope prof("expect_direct", stream);
      expect_direct_kernel<<<grid_for(plan.num_items * plan.blocks_per_item, 1), kBlock, 0,
                             stream>>>(psi, len, plan.d_items, plan.blocks_per_item, partials);
      check_launch("expect_direct");
    }
    expect_reduce_kernel<<<static_cast<unsigned>(plan.num_items), kReduceBlock, 0, stream>>>(
        plan.d_items, plan.d_partial_offset, plan.d_partial_count, partials, key_vals);
    check_launch("expect_reduce");
  }
  expect_finalize_kernel<<<1, 1024, 0, stream>>>(coeffs, plan.d_term_key, plan.T, key_vals,
                                                 term_values, out_dev);
  check_launch("expect_finalize");
}

namespace {

std::size_t validate_psi(const wg_pauli_sum* op, const void* psi, std::size_t length) {
  if (op->num_qubits > 63) {
    fail_invalid("HIP expectation_statevector requires num_qubits <= 63");
  }
  const std::size_t expected = std::size_t{1} << op->num_qubits;
  if (length != expected) {
    fail_invalid("expectation_statevector requires len(psi) == 2 ** num_qubits");
  }
  if (psi == nullptr) fail_invalid("HIP statevector data pointer must be non-null");
  return expected;
}

}  // namespace
}  // namespace wg

using namespace wg;

extern "C" {

int wg_expectation(const wg_pauli_sum* op_c, const void* psi, int dtype, size_t length,
                   int psi_on_device, void* stream_ptr, double* out2) {
  return guard([&] {
    require_op(op_c);
    auto* op = const_cast<wg_pauli_sum*>(op_c);
    if (dtype != WG_DTYPE_COMPLEX64 && dtype != WG_DTYPE_COMPLEX128) {
      fail_type("psi dtype must be complex64 or complex128");
    }
    const std::size_t len = validate_psi(op, psi, length);
    out2[0] = 0.0;
    out2[1] = 0.0;
    if (op->num_terms == 0) return;
    DeviceScope scope(op->device);
    hipStream_t stream = as_stream(stream_ptr);
    auto plan = get_plan(op, stream);
    Scratch scratch(op->device);
    const double2* dpsi;
    if (dtype == WG_DTYPE_COMPLEX128 && psi_on_device) {
      dpsi = static_cast<const double2*>(psi);
    } else {
      double2* buf = scratch.take<double2>(len);
      if (dtype == WG_DTYPE_COMPLEX128) {
        check_hip(hipMemcpyAsync(buf, psi, len * sizeof(double2), hipMemcpyHostToDevice, stream),
                  "copy host buffer to device");
      } else {
        const float2* src = static_cast<const float2*>(psi);
        if (!psi_on_device) {
          float2* staged = scratch.take<float2>(len);
          check_hip(hipMemcpyAsync(staged, psi, len * sizeof(float2), hipMemcpyHostToDevice,
                                   stream),
                    "copy host buffer to device");
          src = staged;
        }
        widen_c64_kernel<<<stride_grid(len, 256), 256, 0, stream>>>(src, len, buf);
        check_launch("widen_c64");
      }
      dpsi = buf;
    }
    double2* dout = scratch.take<double2>(1);
    run_forward(op->coeffs, *plan, dpsi, nullptr, dout, scratch, stream);
    double2 h;
    check_hip(hipMemcpyAsync(&h, dout, sizeof(double2), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
    out2[0] = h.x;
    out2[1] = h.y;
  });
}

}  // extern "C"

namespace wg {

namespace {

__global__ __launch_bounds__(256) void clear_c128_kernel(double2* __restrict__ out, std::size_t n) {
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    out[i] = make_double2(0.0, 0.0);
  }
}

std::size_t align256(std::size_t b) { return (b + 255) & ~std::size_t{255}; }

}  // namespace

std::size_t tiled_adjoint_bytes(const ExpectationPlan& plan) {
  const std::size_t K = std::max<std::size_t>(plan.num_keys, 1);
  return forward_scratch_bytes(plan) + 3 * align256(K * sizeof(double2)) +
         align256(std::max<std::size_t>(plan.sigma_keys, 1) * sizeof(SigmaKey));
}

void tiled_adjoint_enqueue(const ExpectationPlan& plan, const double2* op_coeffs, const double2* psi,
                           const double2* coeffs, double2* term_values, double2* sigma,
                           double2* sigma_dagger, double2* dout, void* buffer, hipStream_t stream) {
  const std::size_t K = std::max<std::size_t>(plan.num_keys, 1);
  auto* base = static_cast<unsigned char*>(buffer);
  const std::size_t key_bytes = align256(K * sizeof(double));
  auto* key_vals = reinterpret_cast<double*>(base);
  auto* partials = reinterpret_cast<double2*>(base + key_bytes);
  base += forward_scratch_bytes(plan);
  auto* kc = reinterpret_cast<double2*>(base);
  auto* cy = reinterpret_cast<double2*>(base + align256(K * sizeof(double2)));
  auto* cy_dag = reinterpret_cast<double2*>(base + 2 * align256(K * sizeof(double2)));
  auto* skeys = reinterpret_cast<SigmaKey*>(base + 3 * align256(K * sizeof(double2)));
  const double2* c = coeffs ? coeffs : op_coeffs;
  const std::size_t len = std::size_t{1} << plan.n;
  // Forward sweeps for E and the per-key values (only when either is wanted),
  // then sigma over the same passes (one read of psi and one sigma update per
  // pass).
  if (dout || term_values) {
    forward_into(c, plan, psi, term_values, dout, key_vals, plan.num_items ? partials : nullptr, stream);
  }
  if (!sigma && !sigma_dagger) return;
  const double2* key_c = plan.d_key_c;
  if (coeffs) {
    key_coeff_kernel<<<grid_for(plan.num_keys, 256), 256, 0, stream>>>(
        c, plan.d_key_term_begin, plan.d_key_terms, plan.num_keys, kc);
    check_launch("key_coeff");
    key_c = kc;
  }
  key_cy_kernel<<<grid_for(plan.num_keys, 256), 256, 0, stream>>>(key_c, plan.d_key_y, plan.d_flat_key,
                                                                   plan.num_keys, cy, cy_dag);
  check_launch("key_cy");
  // Per-call key records, then one sweep per pass: sigma and sigma^dagger
  // together when both are wanted. The first launched pass writes sigma
  // without reading it.
  sigma_keys_kernel<<<grid_for(plan.sigma_keys, 256), 256, 0, stream>>>(
      plan.d_sigma_keys, plan.sigma_keys, cy, cy_dag, sigma ? 0 : 1, skeys);
  check_launch("sigma_keys");
  double2* out = sigma ? sigma : sigma_dagger;
  double2* out_dag = sigma && sigma_dagger ? sigma_dagger : nullptr;
  const unsigned grid = 1U << plan.sigma_log_blocks;
  u32 init = 1;
  for (std::size_t p = 0; p < plan.passes.size(); ++p) {
    const PassLaunch& pl = plan.passes[p];
    if (pl.sigma_end == pl.sigma_begin) continue;
    ProfileScope prof("sigma_tiled", stream);
    if (out_dag) {
      sigma_tiled_kernel<true><<<grid, kSigmaBlock, 0, stream>>>(
          psi, plan.d_passes + p, skeys, pl.sigma_begin, pl.sigma_end, plan.sigma_log_blocks, init, out,
          out_dag);
    } else {
      sigma_tiled_kernel<false><<<grid, kSigmaBlock, 0, stream>>>(
          psi, plan.d_passes + p, skeys, pl.sigma_begin, pl.sigma_end, plan.sigma_log_blocks, init, out,
          nullptr);
    }
    check_launch("sigma_tiled");
    init = 0;
  }
  if (init) {
    // No key records at all: sigma is zero (a fill kernel, so the sequence
    // can be captured into a graph).
    if (sigma) clear_c128_kernel<<<stride_grid(len, 256), 256, 0, stream>>>(sigma, len);
    if (sigma_dagger) clear_c128_kernel<<<stride_grid(len, 256), 256, 0, stream>>>(sigma_dagger, len);
    check_launch("clear_sigma");
  }
}

namespace {

// Slices of x-groups for the direct sigma sweep: enough blocks to fill the GPU
// (about 2,048), each slice's partial rows bounded to 2^22 entries per output.
void sigma_direct_enqueue(const ExpectationPlan& plan, const double2* psi, u64 len,
                          const double2* key_c, double2* sigma, double2* sigma_dag,
                          Scratch& scratch, hipStream_t stream) {
  const u32 G = static_cast<u32>(plan.num_groups);
  const u64 chunks = (len + kSigChunk - 1) / kSigChunk;
  if (chunks > 0x7fffffffULL) fail_invalid("request exceeds HIP launch grid limits");
  u64 slices = std::max<u64>(1, 2048 / chunks);
  slices = std::min<u64>(slices, std::max<u64>(1, (u64{1} << 22) / len));
  slices = std::min<u64>(slices, std::max<u32>(G, 1));
  const u32 per = static_cast<u32>((G + slices - 1) / slices);
  const u32 S = static_cast<u32>((G + per - 1) / std::max<u32>(per, 1));
  double2* rs = sigma;
  double2* rd = sigma_dag;
  if (S > 1) {
    rs = sigma ? scratch.take<double2>(checked_mul(S, len, "sigma partial rows")) : nullptr;
    rd = sigma_dag ? scratch.take<double2>(checked_mul(S, len, "sigma partial rows")) : nullptr;
  }
  {
    ProfileScope prof("sigma_direct", stream);
    sigma_direct_kernel<<<dim3(static_cast<unsigned>(chunks), std::max<u32>(S, 1)), kSigBlock, 0,
                          stream>>>(psi, len, plan.d_group_x, plan.d_group_begin, G,
                                    std::max<u32>(per, 1), plan.d_key_z, plan.d_key_y, key_c, rs, rd);
    check_launch("sigma_direct");
  }
  if (S > 1) {
    if (sigma) {
      sigma_rows_sum_kernel<<<stride_grid(len, 256), 256, 0, stream>>>(rs, S, len, sigma);
      check_launch("sigma_rows_sum");
    }
    if (sigma_dag) {
      sigma_rows_sum_kernel<<<stride_grid(len, 256), 256, 0, stream>>>(rd, S, len, sigma_dag);
      check_launch("sigma_rows_sum");
    }
  }
}

}  // namespace

void adjoint_enqueue(wg_pauli_sum* op, const double* psi, std::size_t len, const double* coeffs,
                     double* term_values, double* sigma, double* sigma_dagger, hipStream_t stream,
                     Scratch& scratch, double2* dout) {
  auto plan = get_plan(op, stream);
  const double2* dpsi = reinterpret_cast<const double2*>(psi);
  const double2* c = coeffs ? reinterpret_cast<const double2*>(coeffs) : op->coeffs;
  if ((sigma || sigma_dagger) && plan->tiled) {
    void* buf = scratch.take_bytes(tiled_adjoint_bytes(*plan));
    tiled_adjoint_enqueue(*plan, op->coeffs, dpsi, reinterpret_cast<const double2*>(coeffs),
                          reinterpret_cast<double2*>(term_values), reinterpret_cast<double2*>(sigma),
                          reinterpret_cast<double2*>(sigma_dagger), dout, buf, stream);
  } else if (sigma || sigma_dagger) {
    // E and the per-term values from the forward sweep; sigma and sigma^dagger
    // from the direct sigma sweep over output chunks x x-group slices.
    if (term_values || dout) {
      run_forward(c, *plan, dpsi, reinterpret_cast<double2*>(term_values), dout, scratch, stream);
    }
    const double2* key_c = plan->d_key_c;
    if (coeffs) {
      double2* kc = scratch.take<double2>(plan->num_keys);
      key_coeff_kernel<<<grid_for(plan->num_keys, 256), 256, 0, stream>>>(
          c, plan->d_key_term_begin, plan->d_key_terms, plan->num_keys, kc);
      check_launch("key_coeff");
      key_c = kc;
    }
    sigma_direct_enqueue(*plan, dpsi, len, key_c, reinterpret_cast<double2*>(sigma),
                         reinterpret_cast<double2*>(sigma_dagger), scratch, stream);
  } else {
    run_forward(c, *plan, dpsi, reinterpret_cast<double2*>(term_values), dout, scratch, stream);
  }
}

}  // namespace wg

extern "C" {

int wg_expectation_adjoint(const wg_pauli_sum* op_c, const double* psi, size_t length,
                           const double* coeffs, double* term_values, double* sigma,
                           double* sigma_dagger, void* stream_ptr, double* out2) {
  return guard([&] {
    require_op(op_c);
    auto* op = const_cast<wg_pauli_sum*>(op_c);
    const std::size_t len = validate_psi(op, psi, length);
    DeviceScope scope(op->device);
    hipStream_t stream = as_stream(stream_ptr);
    if (out2) {
      out2[0] = 0.0;
      out2[1] = 0.0;
    }
    if (op->num_terms == 0) {
      if (sigma) check_hip(hipMemsetAsync(sigma, 0, len * 16, stream), "clear sigma");
      if (sigma_dagger) check_hip(hipMemsetAsync(sigma_dagger, 0, len * 16, stream), "clear sigma");
      sync(stream);
      return;
    }
    Scratch scratch(op->device);
    double2* dout = out2 ? scratch.take<double2>(1) : nullptr;
    adjoint_enqueue(op, psi, len, coeffs, term_values, sigma, sigma_dagger, stream, scratch, dout);
    if (out2) {
      double2 h;
      check_hip(hipMemcpyAsync(&h, dout, sizeof(double2), hipMemcpyDeviceToHost, stream),
                "copy device buffer to host");
      sync(stream);
      out2[0] = h.x;
      out2[1] = h.y;
    } else {
      sync(stream);
    }
  });
}

int wg_expectation_plan_info(const wg_pauli_sum* op_c, size_t* distinct_keys, size_t* x_groups,
                             size_t* sweeps, int* tiled, void* stream_ptr) {
  return guard([&] {
    require_op(op_c);
    auto* op = const_cast<wg_pauli_sum*>(op_c);
    DeviceScope scope(op->device);
    auto plan = get_plan(op, as_stream(stream_ptr));
    if (distinct_keys) *distinct_keys = plan->num_keys;
    if (x_groups) *x_groups = plan->distinct_x;
    if (sweeps) *sweeps = plan->tiled ? plan->passes.size() : plan->num_items;
    if (tiled) *tiled = plan->tiled ? 1 : 0;
  });
}

// Host-only: the tiled pass plan of a W = 1 operator, without replay. Writes
// up to max_passes rows of {k, items, item_groups, tile_blocks, diagonal
// items, off-diagonal items, owner bundles, owner slots} and the pass count.
constexpr int kDebugPassColumns = 8;
}  // extern "C"

namespace wg {
namespace {

// Work items of a W = 1 operator given as host masks, as the device plan
// forms them: distinct keys numbered in first-appearance order, grouped by x
// in first-appearance order of x, kZ keys per item. *num_keys gets the count.
std::vector<ExpItem> debug_items(std::size_t num_terms, const uint64_t* x, const uint64_t* z,
                                 std::size_t* num_keys) {
  std::unordered_map<u64, std::vector<std::pair<u64, u32>>> zs_of;
  std::vector<u64> order_x;
  std::unordered_map<u64, std::unordered_map<u64, u32>> seen;
  u32 keys = 0;
  for (std::size_t t = 0; t < num_terms; ++t) {
    if (seen[x[t]].emplace(z[t], keys).second) {
      if (zs_of.find(x[t]) == zs_of.end()) order_x.push_back(x[t]);
      zs_of[x[t]].push_back({z[t], keys++});
    }
  }
  std::vector<ExpItem> items;
  for (u64 gx : order_x) {
    const auto& zs = zs_of[gx];
    for (std::size_t s = 0; s < zs.size(); s += kZ) {
      ExpItem it{};
      it.x = gx;
      it.nz = static_cast<int>(std::min<std::size_t>(kZ, zs.size() - s));
      for (int j = 0; j < it.nz; ++j) {
        it.z[j] = zs[s + j].first;
        it.key[j] = zs[s + j].second;
        it.y[j] = __builtin_popcountll(gx & zs[s + j].first) & 3;
      }
      items.push_back(it);
    }
  }
  if (num_keys) *num_keys = keys;
  return items;
}

}  // namespace
}  // namespace wg

extern "C" {

int wg_debug_plan_passes(size_t num_qubits, size_t num_terms, const uint64_t* x, const uint64_t* z,
                         int* rows, size_t max_passes, size_t* num_passes) {
  return guard([&] {
    std::vector<ExpItem> items = debug_items(num_terms, x, z, nullptr);
    std::vector<PassLaunch> passes;
    std::size_t diag_keys = 0;
    for (const ExpItem& it : items) diag_keys += it.x == 0 ? static_cast<std::size_t>(it.nz) : 0;
    plan_tiled_passes(num_qubits, items, passes,
                      diag_keys > static_cast<std::size_t>(kDiagWhtMinItems) * kZ, nullptr,
                      SIZE_MAX);
    std::vector<OwnerBundle> bundles;
    if (kOwnerSweep) plan_owner_bundles(items, passes, bundles);
    *num_passes = passes.size();
    for (std::size_t p = 0; p < passes.size() && p < max_passes; ++p) {
      int diag = 0;
      for (u32 i = 0; i < passes[p].num_items; ++i) diag += items[passes[p].first_item + i].x == 0;
      int slots = 0;
      for (u32 b = passes[p].owner_first; b < passes[p].owner_first + passes[p].owner_count; ++b) {
        slots += kOwnerShapeSlots[bundles[b].kind >> 2];
      }
      int* row = rows + kDebugPassColumns * p;
      row[0] = passes[p].desc.k;
      row[1] = static_cast<int>(passes[p].num_items);
      row[2] = static_cast<int>(passes[p].item_groups);
      row[3] = static_cast<int>(passes[p].tile_blocks);
      row[4] = diag;
      row[5] = static_cast<int>(passes[p].num_items) - diag;
      row[6] = static_cast<int>(passes[p].owner_count);
      row[7] = slots;
    }
  });
}

// Host-only: where each distinct key (first-appearance order) of a W = 1
// operator is evaluated by the tiled plan: key_pass[k] = the pass holding it
// (-1: the diagonal Walsh-Hadamard sweep), key_kind[k] = the kind word of the
// first owner bundle holding it (bit 0 odd |x & z|, bit 1 lane pivot, bits
// 2.. shape; -1 when no owner bundle holds it). Both arrays hold num_terms
// entries; *num_keys gets the key count.
int wg_debug_key_sweeps(size_t num_qubits, size_t num_terms, const uint64_t* x, const uint64_t* z,
                        int* key_pass, int* key_kind, size_t* num_keys) {
  return guard([&] {
    std::size_t K = 0;
    std::vector<ExpItem> items = debug_items(num_terms, x, z, &K);
    *num_keys = K;
    for (std::size_t k = 0; k < K; ++k) {
      key_pass[k] = -1;
      key_kind[k] = -1;
    }
    std::vector<PassLaunch> passes;
    std::size_t diag_keys = 0;
    for (const ExpItem& it : items) diag_keys += it.x == 0 ? static_cast<std::size_t>(it.nz) : 0;
    u32 diag_front = 0;
    plan_tiled_passes(num_qubits, items, passes,
                      diag_keys > static_cast<std::size_t>(kDiagWhtMinItems) * kZ, &diag_front,
                      SIZE_MAX);
    std::vector<OwnerBundle> bundles;
    if (kOwnerSweep) plan_owner_bundles(items, passes, bundles);
    for (std::size_t p = 0; p < passes.size(); ++p) {
      const PassLaunch& pl = passes[p];
      for (u32 i = 0; i < pl.num_items; ++i) {
        const ExpItem& it = items[pl.first_item + i];
        for (int j = 0; j < it.nz; ++j) key_pass[it.key[j]] = static_cast<int>(p);
      }
      for (u32 b = pl.owner_first; b < pl.owner_first + pl.owner_count; ++b) {
        for (int sl = 0; sl < kOwnerMaxSlots; ++sl) {
          const OwnerSlot& slot = bundles[b].slot[sl];
          bool used = false;
          for (int kk = 0; kk < kOwnerMaxKeys; ++kk) used |= (slot.key[kk].info & (1U << 15)) != 0;
          if (!used || slot.item >= pl.num_items) continue;
          const ExpItem& it = items[pl.first_item + slot.item];
          for (int j = 0; j < it.nz; ++j) {
            if (key_kind[it.key[j]] < 0) key_kind[it.key[j]] = static_cast<int>(bundles[b].kind);
          }
        }
      }
    }
  });
}

// Host-only replay of the tiled sigma sweep's records (test hook): builds the
// items, passes and sigma records of a W = 1 operator as the device plan does
// and evaluates them point by point on the CPU with the kernel's sign
// decomposition (uniform, lane and register parts; the constant sign
// (-1)^{parity(zl & xl)} folded into the per-call coefficients), accumulating
// sigma = H psi and H^dagger psi.
int wg_debug_sigma_plan(size_t num_qubits, size_t num_terms, const uint64_t* x, const uint64_t* z,
                        const double* coeffs, const double* psi, double* sigma, double* sigma_dag,
                        size_t* num_records) {
  return guard([&] {
    const std::size_t n = num_qubits;
    if (n < kTiledMinQubits || n > 24) fail_invalid("debug sigma plan supports 18 <= n <= 24");
    std::size_t K = 0;
    std::vector<ExpItem> items = debug_items(num_terms, x, z, &K);
    // Key coefficients summed in term order (first-appearance key ids).
    std::vector<double2> kc(K, make_double2(0.0, 0.0));
    std::vector<u64> kxm(K), kzm(K);
    std::vector<char> seen(K, 0);
    {
      std::unordered_map<u64, std::unordered_map<u64, u32>> idx;
      u32 next = 0;
      for (std::size_t t = 0; t < num_terms; ++t) {
        auto [itr, ins] = idx[x[t]].emplace(z[t], next);
        if (ins) ++next;
        const u32 k = itr->second;
        const double2 c = make_double2(coeffs[2 * t], coeffs[2 * t + 1]);
        if (!seen[k]) {
          kc[k] = c;
          seen[k] = 1;
          kxm[k] = x[t];
          kzm[k] = z[t];
        } else {
          kc[k].x += c.x;
          kc[k].y += c.y;
        }
      }
    }
    std::vector<PassLaunch> passes;
    std::size_t diag_keys = 0;
    for (const ExpItem& it : items) diag_keys += it.x == 0 ? static_cast<std::size_t>(it.nz) : 0;
    u32 diag_front = 0;
    plan_tiled_passes(n, items, passes, diag_keys > static_cast<std::size_t>(kDiagWhtMinItems) * kZ,
                      &diag_front, SIZE_MAX);
    std::vector<SigmaKey> recs;
    build_sigma_records(items, passes, recs);
    *num_records = recs.size();
    auto times_iy = [](double2 v, int y) {
      switch (y & 3) {
        case 0: return v;
        case 1: return make_double2(-v.y, v.x);
        case 2: return make_double2(-v.x, -v.y);
        default: return make_double2(v.y, -v.x);
      }
    };
    for (SigmaKey& r : recs) {
      if (r.key == kSigmaPadKey) continue;
      const int y = __builtin_popcountll(kxm[r.key] & kzm[r.key]) & 3;
      const bool neg = host_parity(r.zl & r.xl & ~kSigmaLast) != 0;
      const double2 a = times_iy(kc[r.key], y), b = times_iy(make_double2(kc[r.key].x, -kc[r.key].y), y);
      r.c = neg ? make_double2(-a.x, -a.y) : a;
      r.cd = neg ? make_double2(-b.x, -b.y) : b;
    }
    const std::size_t len = std::size_t{1} << n;
    const auto* amp = reinterpret_cast<const double2*>(psi);
    auto* out = reinterpret_cast<double2*>(sigma);
    auto* out_d = reinterpret_cast<double2*>(sigma_dag);
    std::fill(out, out + len, make_double2(0.0, 0.0));
    std::fill(out_d, out_d + len, make_double2(0.0, 0.0));
    std::vector<double2> tile(kTilePoints);
    std::vector<u64> gidx(kTilePoints);
    for (const PassLaunch& pl : passes) {
      const PassDesc& d = pl.desc;
      if (d.k != kTileMaxLog) fail_runtime("sigma replay expects full tiles");
      const u64 ntiles = u64{1} << (n - d.k);
      for (u64 t = 0; t < ntiles; ++t) {
        u64 rep = 0;
        for (int j = 0; j < d.nfree; ++j) rep |= ((t >> j) & 1ULL) << d.free_pos[j];
        for (u32 u = 0; u < kTilePoints; ++u) {
          gidx[u] = (rep ^ d.seg_off[u >> d.l]) | (u & ((1U << d.l) - 1));
          tile[u] = amp[gidx[u]];
        }
        for (u32 u = 0; u < kTilePoints; ++u) {
          // Point u of the tile is point I of thread tid (u = tid + 1024 I).
          const u32 tid = u & 1023U, I = u >> 10;
          double2 c = make_double2(0, 0), e = make_double2(0, 0);
          for (u32 r = pl.sigma_begin; r < pl.sigma_end; ++r) {
            const SigmaKey& k = recs[r];
            const u32 xl = k.xl & ~kSigmaLast;
            const bool sgn = (host_parity(k.zl & tid) ^ host_parity(k.z & rep) ^
                              host_parity(k.zl & (I << 10))) != 0;
            c.x += sgn ? -k.c.x : k.c.x;
            c.y += sgn ? -k.c.y : k.c.y;
            e.x += sgn ? -k.cd.x : k.cd.x;
            e.y += sgn ? -k.cd.y : k.cd.y;
            if (k.xl & kSigmaLast) {
              const double2 g = tile[u ^ xl];
              double2& o = out[gidx[u]];
              o.x += c.x * g.x - c.y * g.y;
              o.y += c.x * g.y + c.y * g.x;
              double2& od = out_d[gidx[u]];
              od.x += e.x * g.x - e.y * g.y;
              od.y += e.x * g.y + e.y * g.x;
              c = e = make_double2(0, 0);
            }
          }
        }
      }
    }
  });
}

// Host-only check of the tiled pass plan for an operator given as host masks
// (W = 1): builds the items and passes exactly as the device plan does, then
// replays the tile decomposition on the CPU over `psi` (length 2^n) with the
// same local-coordinate arithmetic as expect_tiled_kernel and writes <P> per
// distinct key (first-appearance order) to key_vals. Returns the number of
// passes in *num_passes. Test-only: no product entry point calls it.
int wg_debug_tiled_plan(size_t num_qubits, size_t num_terms, const uint64_t* x, const uint64_t* z,
                        const double* psi, double* key_vals, size_t* num_keys, size_t* num_passes) {
  return guard([&] {
    const std::size_t n = num_qubits;
    if (n < 9 || n > 26) fail_invalid("debug tiled plan supports 9 <= n <= 26");
    // Distinct keys and groups, as build_plan.
    std::vector<u64> kx, kz;
    std::unordered_map<u64, std::vector<u32>> by_x;
    std::vector<u64> order_x;
    {
      std::unordered_map<u64, std::unordered_map<u64, u32>> idx;
      for (std::size_t t = 0; t < num_terms; ++t) {
        auto& zs = idx[x[t]];
        if (zs.emplace(z[t], static_cast<u32>(kx.size())).second) {
          if (by_x.find(x[t]) == by_x.end()) order_x.push_back(x[t]);
          by_x[x[t]].push_back(static_cast<u32>(kx.size()));
          kx.push_back(x[t]);
          kz.push_back(z[t]);
        }
      }
    }
    std::vector<ExpItem> items;
    for (u64 gx : order_x) {
      const auto& ks = by_x[gx];
      for (std::size_t s = 0; s < ks.size(); s += kZ) {
        ExpItem it{};
        it.x = gx;
        it.nz = static_cast<int>(std::min<std::size_t>(kZ, ks.size() - s));
        for (int j = 0; j < it.nz; ++j) {
          it.z[j] = kz[ks[s + j]];
          it.key[j] = ks[s + j];
          it.y[j] = __builtin_popcountll(gx & kz[ks[s + j]]) & 3;
        }
        items.push_back(it);
      }
    }
    std::vector<PassLaunch> passes;
    std::size_t diag_keys = 0;
    for (const ExpItem& it : items) diag_keys += it.x == 0 ? static_cast<std::size_t>(it.nz) : 0;
    const bool diag_separate =
        n >= kTileMaxLog + 3 && diag_keys > static_cast<std::size_t>(kDiagWhtMinItems) * kZ;
    u32 diag_front = 0;
    plan_tiled_passes(n, items, passes, diag_separate, &diag_front, SIZE_MAX);
    std::vector<OwnerBundle> bundles;
    if (kOwnerSweep) plan_owner_bundles(items, passes, bundles);
    *num_passes = passes.size();
    *num_keys = kx.size();
    const auto* amp = reinterpret_cast<const double2*>(psi);
    // Diagonal items swept by the Walsh-Hadamard kernel: same tile split
    // (consecutive runs of kTilePoints) and key lookup.
    if (diag_front) {
      std::vector<double> P(kTilePoints);
      std::vector<double> acc(static_cast<std::size_t>(diag_front) * kZ, 0.0);
      for (u64 t = 0; t < (u64{1} << (n - kTileMaxLog)); ++t) {
        for (u32 u = 0; u < kTilePoints; ++u) {
          const double2 a = amp[(t << kTileMaxLog) + u];
          P[u] = a.x * a.x + a.y * a.y;
        }
        for (u32 h = 1; h < kTilePoints; h <<= 1) {
          for (u32 u = 0; u < kTilePoints; ++u) {
            if (!(u & h)) {
              const double a = P[u], c = P[u | h];
              P[u] = a + c;
              P[u | h] = a - c;
            }
          }
        }
        for (u32 i = 0; i < diag_front; ++i) {
          const ExpItem& it = items[i];
          if (it.x != 0) fail_runtime("diagonal sweep holds an off-diagonal item");
          for (int j = 0; j < it.nz; ++j) {
            const double v = P[it.z[j] & (kTilePoints - 1)];
            acc[i * kZ + j] += (__builtin_popcountll((it.z[j] >> kTileMaxLog) & t) & 1) ? -v : v;
          }
        }
      }
      for (u32 i = 0; i < diag_front; ++i) {
        for (int j = 0; j < items[i].nz; ++j) key_vals[items[i].key[j]] = acc[i * kZ + j];
      }
    }
    std::vector<double2> tile;
    std::vector<char> seen(std::size_t{1} << n);
    for (const PassLaunch& p : passes) {
      const PassDesc& d = p.desc;
      const u32 tile_size = 1U << d.k;
      const u64 ntiles = 1ULL << (n - d.k);
      std::fill(seen.begin(), seen.end(), 0);
      tile.assign(tile_size, make_double2(0, 0));
      std::vector<double2> acc(static_cast<std::size_t>(p.num_items) * kZ, make_double2(0, 0));
      std::vector<double2> owner_acc(acc.size(), make_double2(0, 0));
      for (u64 t = 0; t < ntiles; ++t) {
        u64 rep = 0;
        for (int j = 0; j < d.nfree; ++j) rep |= ((t >> j) & 1ULL) << d.free_pos[j];
        for (u32 u = 0; u < tile_size; ++u) {
          const u64 b = (rep ^ d.seg_off[u >> d.l]) | (u & ((1U << d.l) - 1));
          if (b >= (u64{1} << n) || seen[b]) fail_runtime("tiles do not partition the index space");
          seen[b] = 1;
          tile[u] = amp[b];
        }
        // Owner bundles: the kernel's representatives, partners and signs,
        // with the points addressed through each bundle's own layout.
        for (u32 bi = p.owner_first; bi < p.owner_first + p.owner_count; ++bi) {
          const OwnerBundle& bd = bundles[bi];
          const int M = kOwnerShapeSlots[bd.kind >> 2], NZ = kOwnerShapeKeys[bd.kind >> 2];
          const bool lanex = (bd.kind >> 1) & 1, odd = bd.kind & 1;
          auto point = [&](u32 u) {
            u64 off = 0;
            for (int i = 0; i < kSegBits; ++i) {
              if ((u >> (kTileLowLog + i)) & 1U) off ^= bd.seg_basis[i];
            }
            return (rep ^ off) | (u & ((1U << kTileLowLog) - 1U));
          };
          if (t == 0) {
            std::vector<u64> pts(kTilePoints), ref(kTilePoints);
            for (u32 u = 0; u < kTilePoints; ++u) {
              pts[u] = point(u);
              ref[u] = (rep ^ d.seg_off[u >> d.l]) | (u & ((1U << d.l) - 1U));
            }
            std::sort(pts.begin(), pts.end());
            std::sort(ref.begin(), ref.end());
            if (pts != ref) fail_runtime("owner bundle layout does not span the pass's tile");
          }
          for (int m = 0; m < M; ++m) {
            const OwnerSlot& S = bd.slot[m];
            for (u32 tid = 0; tid < static_cast<u32>(kOwnerBlock); ++tid) {
              const u32 h = lanex ? (tid >> S.hbit) & 1U : 0U;
              for (u32 j = 0; j < 4; ++j) {
                const u32 i_own = lanex ? j + 4 * h : j;
                const u32 par = lanex ? (((tid << 4) ^ S.addr0) | (h << 15)) + (j << 13)
                                      : (tid << 4) ^ S.addr0 ^ (j << 13);
                if (par >= kTilePoints * 16U) fail_runtime("owner partner outside the tile");
                const double2 A = amp[point(tid + kOwnerBlock * i_own)];
                const double2 P = amp[point(par >> 4)];
                const double v = odd ? P.x * A.y - P.y * A.x : P.x * A.x + P.y * A.y;
                for (int q = 0; q < NZ; ++q) {
                  const OwnerKey& K = S.key[q];
                  if (!((K.info >> 15) & 1U)) continue;
                  u32 sb = (__builtin_popcount(K.w & 0x3FFFFFFFU & static_cast<u32>(t)) & 1U) ^
                           (__builtin_popcount(K.zlt & tid) & 1U) ^ ((K.w >> 30) & j & 1U) ^
                           ((K.w >> 31) & (j >> 1) & 1U);
                  if (lanex) sb ^= (K.info >> 2) & h & 1U;
                  double2& a2 = owner_acc[static_cast<std::size_t>(S.item) * kZ + ((K.info >> 4) & 15U)];
                  a2.x += sb ? -v : v;
                  a2.y += sb ? -v : v;
                }
              }
            }
          }
        }
        for (u32 i = 0; i < p.num_items; ++i) {
          const ExpItem& it = items[p.first_item + i];
          if (kOwnerSweep && it.x != 0) continue;
          const u32 xl = static_cast<u32>(it.xl), pbit = static_cast<u32>(it.pbit);
          auto ins0 = [&](u32 v) {
            const u32 low = pbit - 1;
            return ((v & ~low) << 1) | (v & low);
          };
          const u32 span = xl ? tile_size / 2 : tile_size;
          const u32 steps = span > 64 ? span / 64 : 1;
          if (xl && it.cls_log != kNoClasses && steps == kOffSteps) {
            // The kernel's class decomposition: steps enumerated through
            // ebasis, class = e >> (5 - D), class sums folded with lam.
            const int D = it.cls_log;
            for (u32 lane = 0; lane < 64; ++lane) {
              const u32 f = ins0(lane);
              double qr[1 << kMaxClassLog] = {}, qi[1 << kMaxClassLog] = {};
              for (int e = 0; e < kOffSteps; ++e) {
                u32 st = 0;
                for (int l = 0; l < 5; ++l) {
                  if ((e >> l) & 1) st ^= it.ebasis[l];
                }
                const u32 u = f ^ ins0(st << 6);
                const double2 A = tile[u], B = tile[u ^ xl];
                qr[e >> (5 - D)] += B.x * A.x + B.y * A.y;
                qi[e >> (5 - D)] += B.x * A.y - B.y * A.x;
              }
              for (int j = 0; j < it.nz; ++j) {
                const u32 lane_par = __builtin_popcount(static_cast<u32>(it.zl[j]) & f) & 1;
                const u32 srep = __builtin_popcountll(it.z[j] & rep) & 1;
                double2& acc2 = acc[static_cast<std::size_t>(i) * kZ + j];
                for (int c = 0; c < (1 << D); ++c) {
                  const u32 neg = lane_par ^ srep ^ (__builtin_popcount(it.lam[j] & c) & 1);
                  acc2.x += neg ? -qr[c] : qr[c];
                  acc2.y += neg ? -qi[c] : qi[c];
                }
              }
            }
            continue;
          }
          for (int j = 0; j < it.nz; ++j) {
            const u32 zlj = static_cast<u32>(it.zl[j]);
            const u32 srep = __builtin_popcountll(it.z[j] & rep) & 1;
            double2& acc2 = acc[static_cast<std::size_t>(i) * kZ + j];
            // Same decomposition as the kernel: lane part + uniform step part.
            for (u32 lane = 0; lane < 64 && lane < span; ++lane) {
              const u32 f = xl ? ins0(lane) : lane;
              const u32 lane_par = __builtin_popcount(zlj & f) & 1;
              double lr = 0.0, li = 0.0;
              for (u32 st = 0; st < steps; ++st) {
                const u32 g = xl ? ins0(st << 6) : (st << 6);
                const u32 u = f ^ g;
                const u32 sbit = (__builtin_popcount(zlj & g) & 1) ^ srep;
                const double sg = sbit ? -1.0 : 1.0;
                if (xl) {
                  const double2 A = tile[u], B = tile[u ^ xl];
                  lr += sg * (B.x * A.x + B.y * A.y);
                  li += sg * (B.x * A.y - B.y * A.x);
                } else {
                  lr += sg * (tile[u].x * tile[u].x + tile[u].y * tile[u].y);
                }
              }
              acc2.x += lane_par ? -lr : lr;
              acc2.y += lane_par ? -li : li;
            }
          }
        }
      }
      for (char c : seen) {
        if (!c) fail_runtime("tiles do not cover the index space");
      }
      for (u32 i = 0; i < p.num_items; ++i) {
        const ExpItem& it = items[p.first_item + i];
        for (int j = 0; j < it.nz; ++j) {
          const double2 S = (kOwnerSweep && it.x != 0 ? owner_acc : acc)[static_cast<std::size_t>(i) * kZ + j];
          double v;
          if (it.x == 0) {
            v = S.x;
          } else {
            switch (it.y[j] & 3) {
              case 0: v = 2 * S.x; break;
              case 1: v = -2 * S.y; break;
              case 2: v = -2 * S.x; break;
              default: v = 2 * S.y; break;
            }
          }
          key_vals[it.key[j]] = v;
        }
      }
    }
  });
}

}  // extern "C"

// Greedy commuting groups on the device (reference PauliSum::group_commuting,
// src/grouping.cpp:197-290, strategy "largest_first").
//
// The visiting order (descending weight |x | z|, then canonical key order,
// then index; reference src/grouping.cpp:36-62) is built on the host, and the
// masks are gathered into that order on the device, so a term is named by its
// visiting position s. Each term takes the first group that accepts it given
// every earlier placement:
//   * full: a group accepts s when s commutes with every member; group g keeps
//     a conflict bit row over positions (the OR of its members'
//     anticommutation rows, from the packed commutation matrix of the
//     reordered operator);
//   * qubit-wise: group g keeps Xb, Zb (ORs of its members' masks) and accepts
//     s iff (x_s & Zb) ^ (z_s & Xb) = 0 on every word (reference
//     src/grouping.cpp:64-85). Members of a group never disagree on a qubit,
//     so this equals "compatible with every member".
// In both modes acceptance is "no member conflicts with s", and adding
// members only removes acceptances.
//
// The positions are resolved in batches of kBatch, with the same result as one
// position at a time:
//   1. group_candidates_kernel (one block per batch position, whole GPU): the
//      first kCands groups that accept s as the groups stood before the batch,
//      the next accepting group after them (where to resume a scan), and the
//      conflict bits of s against the earlier positions of its batch;
//   2. group_resolve_kernel (one block): walks the batch in order. Position i
//      stamps the groups of its conflicting earlier batch positions (the only
//      members the candidate kernel did not see), then takes the first
//      candidate without its stamp; if every candidate is stamped, it scans
//      the remaining pre-batch groups from the resume point; then the groups
//      opened earlier in this batch (no pre-batch members: accepted unless
//      stamped); else it opens a new group;
//   3. group_commit_*_kernel: folds the batch's placements into the group
//      state for the next batch.
// A group rejects s exactly when a pre-batch member conflicts (step 1) or a
// member placed earlier in the batch conflicts (the stamps), so each position
// gets the group the one-at-a-time walk would give it.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <memory>
#include <numeric>
#include <thread>
#include <type_traits>
#include <vector>

#include "wg_internal.hpp"

namespace wg {
namespace {

constexpr u32 kBatch = 1024;         // visiting positions resolved per batch
constexpr int kCandBlock = 256;      // candidate kernel block (one batch position)
constexpr int kResolveBlock = 256;   // resolve kernel: 4 waves, 4 batch positions per lane
constexpr int kCands = 15;           // pre-batch candidate groups kept per position
constexpr int kCandWords = 16;       // the candidates and the resume group
constexpr u32 kNone = 0xFFFFFFFFu;
constexpr u32 kLdsGroups = 32768;    // stamps live in LDS when every group id is below this

static_assert(kBatch == 4 * kResolveBlock, "four batch positions per resolve thread");

enum GroupMode : int { kQwc = 0, kFull = 1 };

struct GroupState {
  const u64* px;        // masks in visiting order [T x W]
  const u64* pz;
  u64 W;
  const u64* commute;   // full: packed commutation rows in visiting order [T x row_words]
  u64 row_words;
  u64* conflict;        // full: per group, OR of the members' anticommutation rows
  u64* xb;              // qubit-wise: per group, OR of the members' masks [T x W]
  u64* zb;
};

// Does group g (its members so far) accept position s?
template <int kMode>
__device__ inline bool group_accepts(const GroupState& st, u32 g, u64 s) {
  if constexpr (kMode == kFull) {
    return ((st.conflict[static_cast<u64>(g) * st.row_words + (s >> 6)] >> (s & 63)) & 1) == 0;
  } else {
    const u64* xs = st.px + s * st.W;
    const u64* zs = st.pz + s * st.W;
    const u64* xg = st.xb + static_cast<u64>(g) * st.W;
    const u64* zg = st.zb + static_cast<u64>(g) * st.W;
    for (u64 k = 0; k < st.W; ++k) {
      if ((xs[k] & zg[k]) ^ (zs[k] & xg[k])) return false;
    }
    return true;
  }
}

// May positions a and b not share a group?
template <int kMode>
__device__ inline bool terms_conflict(const GroupState& st, u64 a, u64 b) {
  if constexpr (kMode == kFull) {
    return ((st.commute[a * st.row_words + (b >> 6)] >> (b & 63)) & 1) == 0;
  } else {
    const u64* xa = st.px + a * st.W;
    const u64* za = st.pz + a * st.W;
    const u64* xc = st.px + b * st.W;
    const u64* zc = st.pz + b * st.W;
    for (u64 k = 0; k < st.W; ++k) {
      if ((xa[k] & zc[k]) ^ (za[k] & xc[k])) return true;
    }
    return false;
  }
}

__global__ __launch_bounds__(256) void gather_masks_kernel(const u32* __restrict__ order,
                                                           const u64* __restrict__ x,
                                                           const u64* __restrict__ z, u64 T, u64 W,
                                                           u64* __restrict__ px,
                                                           u64* __restrict__ pz) {
  const u64 n = T * W;
  for (u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<u64>(gridDim.x) * blockDim.x) {
    const u64 s = i / W, k = i % W;
    const u64 src = static_cast<u64>(order[s]) * W + k;
    px[i] = x[src];
    pz[i] = z[src];
  }
}

// Block i: position s = s0 + i of the batch.
//   cmask[i][t], bit m: s conflicts with batch position 256 (t >> 6) + 64 m + (t & 63),
//     for earlier positions only (the resolve kernel's register layout);
//   cand[i][0..kCands): the first kCands groups (index order) that accept s given
//     the groups before this batch, kNone-padded; cand[i][kCands]: the next
//     accepting group after them, kNone when there is none.
template <int kMode>
__global__ __launch_bounds__(kCandBlock) void group_candidates_kernel(
    GroupState st, u64 s0, const u32* __restrict__ num_groups, u32* __restrict__ cand,
    std::uint8_t* __restrict__ cmask) {
  __shared__ u64 ballots[4][kCandBlock / 64];
  __shared__ u32 list[kCandWords];
  __shared__ u32 found_s, resume_s;
  const u32 i = blockIdx.x;
  const u64 s = s0 + i;
  const u32 t = threadIdx.x, wave = t >> 6, lane = t & 63;

  u32 mask = 0;
#pragma unroll
  for (u32 m = 0; m < 4; ++m) {
    const u32 j = 256 * wave + 64 * m + lane;
    if (j < i && terms_conflict<kMode>(st, s, s0 + j)) mask |= 1U << m;
  }
  cmask[static_cast<u64>(i) * kResolveBlock + t] = static_cast<std::uint8_t>(mask);

  const u32 G = *num_groups;
  u32 found = 0, resume = kNone;
  for (u32 base = 0; base < G && resume == kNone; base += 4 * kCandBlock) {
    bool acc[4];
#pragma unroll
    for (u32 q = 0; q < 4; ++q) {
      const u32 g = base + q * kCandBlock + t;
      acc[q] = g < G && group_accepts<kMode>(st, g < G ? g : 0, s);
    }
#pragma unroll
    for (u32 q = 0; q < 4; ++q) {
      const u64 b = __ballot(acc[q]);
      if (lane == 0) ballots[q][wave] = b;
    }
    __syncthreads();
    if (t == 0) {
      u32 f = found, r = kNone;
      for (u32 q = 0; q < 4 && r == kNone; ++q) {
        for (u32 w = 0; w < kCandBlock / 64 && r == kNone; ++w) {
          u64 b = ballots[q][w];
          while (b != 0) {
            const u32 g = base + q * kCandBlock + w * 64 + static_cast<u32>(__builtin_ctzll(b));
            if (f < kCands) {
              list[f++] = g;
            } else {
              r = g;
              break;
            }
            b &= b - 1;
          }
        }
      }
      found_s = f;
      resume_s = r;
    }
    __syncthreads();
    found = found_s;
    resume = resume_s;
  }
  if (t < kCandWords) {
    cand[static_cast<u64>(i) * kCandWords + t] = t < kCands ? (t < found ? list[t] : kNone) : resume;
  }
}

// First pre-batch group at or after `from` that accepts s and is not stamped.
template <int kMode, typename Stamp>
__device__ u32 scan_old_groups(const GroupState& st, u64 s, u32 from, u32 G_old,
                               const Stamp* cur, Stamp stamp, u32 lane) {
  for (u32 base = from; base < G_old; base += 64) {
    const u32 g = base + lane;
    const u32 gs = g < G_old ? g : 0;
    const bool ok = g < G_old && group_accepts<kMode>(st, gs, s) && cur[gs] != stamp;
    const u64 b = __ballot(ok);
    if (b != 0) return base + static_cast<u32>(__builtin_ctzll(b));
  }
  return kNone;
}

// One block walks the batch. Thread t holds the groups of batch positions
// 256 (t >> 6) + 64 q + (t & 63), q < 4. Stamps alternate between two arrays
// (position i uses array i & 1), so one barrier per position suffices: a
// wave stamps for position i + 2 only after every wave has passed position
// i + 1's barrier, i.e. finished reading position i's stamps. With kLds the
// stamps are 16-bit batch-local (i + 1) and zeroed per launch; otherwise
// 32-bit position stamps (s + 1) in global memory, zeroed once.
// The candidate lists and conflict masks are staged through LDS kChunk
// positions at a time: the next chunk's are loaded into registers at the start
// of a chunk and written to LDS at its end, so their latency is hidden.
constexpr u32 kChunk = 32;

template <int kMode, bool kLds>
__global__ __launch_bounds__(kResolveBlock) void group_resolve_kernel(
    GroupState st, u64 s0, u32 nb, const u32* __restrict__ cand,
    const std::uint8_t* __restrict__ cmask, u32* __restrict__ num_groups, u32* __restrict__ gpos,
    u32* __restrict__ gstamps, u64 gstamp_len) {
  using Stamp = std::conditional_t<kLds, std::uint16_t, u32>;
  constexpr u32 kLdsLen = kLdsGroups + kResolveBlock;
  __shared__ Stamp lds_stamps[kLds ? 2 * kLdsLen : 1];
  __shared__ uint2 s_cand[kChunk * kCandWords / 2];
  __shared__ uint4 s_cmask[kChunk * kResolveBlock / 16];
  __shared__ u32 s_gpos[kBatch];
  static_assert(kChunk * kCandWords / 2 == kResolveBlock && kChunk * kResolveBlock / 16 == 2 * kResolveBlock,
                "one uint2 and two uint4 of a chunk per thread");
  const u32 t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const u32 G_old = *num_groups;
  u32 G_cur = G_old;
  Stamp* st0;
  Stamp* st1;
  u32 dummy;  // a private slot for threads with nothing to stamp
  if constexpr (kLds) {
    st0 = lds_stamps;
    st1 = lds_stamps + kLdsLen;
    dummy = kLdsGroups + t;
    for (u32 k = t; k < G_old + nb; k += kResolveBlock) {
      st0[k] = 0;
      st1[k] = 0;
    }
  } else {
    st0 = gstamps;
    st1 = gstamps + gstamp_len;
    dummy = static_cast<u32>(gstamp_len) - kResolveBlock + t;
  }
  // The buffers hold kBatch positions, so chunk loads past nb stay in bounds.
  const uint2* cand2 = reinterpret_cast<const uint2*>(cand);
  const uint4* cmask4 = reinterpret_cast<const uint4*>(cmask);
  uint2 nc = cand2[t];
  uint4 nm0 = cmask4[t], nm1 = cmask4[kResolveBlock + t];
  s_cand[t] = nc;
  s_cmask[t] = nm0;
  s_cmask[kResolveBlock + t] = nm1;
  __syncthreads();
  const std::uint8_t* s_mask8 = reinterpret_cast<const std::uint8_t*>(s_cmask);
  const u32* s_cand32 = reinterpret_cast<const u32*>(s_cand);
  const u32 chunks = (nb + kChunk - 1) / kChunk;

  u32 pl[4] = {kNone, kNone, kNone, kNone};
  for (u32 ch = 0; ch < chunks; ++ch) {
    {
      const u32 next = min(ch + 1, kBatch / kChunk - 1);
      nc = cand2[next * kResolveBlock + t];
      nm0 = cmask4[next * 2 * kResolveBlock + t];
      nm1 = cmask4[next * 2 * kResolveBlock + kResolveBlock + t];
    }
    const u32 n_here = min(kChunk, nb - ch * kChunk);
    for (u32 ii = 0; ii < n_here; ++ii) {
      const u32 i = ch * kChunk + ii;
      const u32 c = s_cand32[ii * kCandWords + (lane & 15)];
      const u32 m = s_mask8[ii * kResolveBlock + t];
      Stamp* cur = (i & 1) ? st1 : st0;
      const Stamp stamp = kLds ? static_cast<Stamp>(i + 1) : static_cast<Stamp>(s0 + i + 1);
#pragma unroll
      for (u32 q = 0; q < 4; ++q) cur[((m >> q) & 1) ? pl[q] : dummy] = stamp;
      if constexpr (kLds) {
        // LDS stamps: the barrier needs only the LDS writes done. A full
        // __syncthreads() would also drain the next chunk's global loads.
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
      } else {
        __syncthreads();
      }

      // Stamp reads are unconditional (invalid lanes read their dummy slot),
      // so they are issued together rather than as masked branches.
      const u32 cc = lane < kCands ? c : kNone;
      const Stamp cs = cur[cc != kNone ? cc : dummy];
      const bool ok = cc != kNone && cs != stamp;
      const u64 b = __ballot(ok);
      u32 g = kNone;
      if (b != 0) {
        g = __builtin_amdgcn_readlane(cc, static_cast<int>(__builtin_ctzll(b)));
      } else {
        const u32 resume = __builtin_amdgcn_readlane(c, kCands);
        if (resume != kNone) g = scan_old_groups<kMode, Stamp>(st, s0 + i, resume, G_old, cur, stamp, lane);
        // Groups opened earlier in this batch, 256 stamps per round trip.
        for (u32 base = G_old; g == kNone && base < G_cur; base += 256) {
          Stamp v[4];
#pragma unroll
          for (u32 q = 0; q < 4; ++q) {
            const u32 gg = base + 64 * q + lane;
            v[q] = cur[gg < G_cur ? gg : dummy];
          }
          bool open[4];
#pragma unroll
          for (u32 q = 0; q < 4; ++q) open[q] = base + 64 * q + lane < G_cur && v[q] != stamp;
#pragma unroll
          for (u32 q = 0; q < 4; ++q) {
            const u64 fb = __ballot(open[q]);
            if (g == kNone && fb != 0) g = base + 64 * q + static_cast<u32>(__builtin_ctzll(fb));
          }
        }
        if (g == kNone) g = G_cur++;
      }
      if (wave == (i >> 8) && lane == (i & 63)) {
        const u32 q = (i >> 6) & 3;
#pragma unroll
        for (u32 k = 0; k < 4; ++k) pl[k] = k == q ? g : pl[k];
      }
      if (t == 0) s_gpos[i] = g;
    }
    // Every wave is done with this chunk's staged inputs: replace them.
    __syncthreads();
    s_cand[t] = nc;
    s_cmask[t] = nm0;
    s_cmask[kResolveBlock + t] = nm1;
    __syncthreads();
  }
  for (u32 k = t; k < nb; k += kResolveBlock) gpos[s0 + k] = s_gpos[k];
  if (t == 0) *num_groups = G_cur;
}

// Full mode: conflict[g] |= anticommutation row of each batch position placed
// in g, over the words from w0 on (later positions only). Block (x, i).
__global__ __launch_bounds__(256) void group_commit_full_kernel(GroupState st, u64 s0, u64 w0,
                                                                const u32* __restrict__ gpos) {
  const u64 s = s0 + blockIdx.y;
  const u64 k = w0 + static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= st.row_words) return;
  const u64 v = ~st.commute[s * st.row_words + k];
  if (v != 0) {
    atomicOr(reinterpret_cast<unsigned long long*>(st.conflict + static_cast<u64>(gpos[s]) * st.row_words + k),
             static_cast<unsigned long long>(v));
  }
}

// Qubit-wise mode: Xb[g] |= x_s, Zb[g] |= z_s for each batch position s in g.
__global__ __launch_bounds__(256) void group_commit_qwc_kernel(GroupState st, u64 s0, u32 nb,
                                                               const u32* __restrict__ gpos) {
  const u64 idx = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<u64>(nb) * st.W) return;
  const u64 s = s0 + idx / st.W, k = idx % st.W;
  const u64 dst = static_cast<u64>(gpos[s]) * st.W + k;
  const u64 xv = st.px[s * st.W + k], zv = st.pz[s * st.W + k];
  if (xv != 0) atomicOr(reinterpret_cast<unsigned long long*>(st.xb + dst), static_cast<unsigned long long>(xv));
  if (zv != 0) atomicOr(reinterpret_cast<unsigned long long*>(st.zb + dst), static_cast<unsigned long long>(zv));
}

template <int kMode>
void run_batches(const GroupState& st, u64 T, u32* count, u32* cand, std::uint8_t* cmask,
                 u32* gpos, u32* gstamps, u64 gstamp_len, hipStream_t stream) {
  const bool lds = T <= kLdsGroups;
  for (u64 s0 = 0; s0 < T; s0 += kBatch) {
    const u32 nb = static_cast<u32>(std::min<u64>(kBatch, T - s0));
    group_candidates_kernel<kMode><<<nb, kCandBlock, 0, stream>>>(st, s0, count, cand, cmask);
    check_launch("group_candidates");
    if (lds) {
      group_resolve_kernel<kMode, true><<<1, kResolveBlock, 0, stream>>>(st, s0, nb, cand, cmask, count,
                                                                         gpos, gstamps, gstamp_len);
    } else {
      group_resolve_kernel<kMode, false><<<1, kResolveBlock, 0, stream>>>(st, s0, nb, cand, cmask, count,
                                                                          gpos, gstamps, gstamp_len);
    }
    check_launch("group_resolve");
    if (s0 + nb >= T) break;
    if constexpr (kMode == kFull) {
      const u64 w0 = (s0 + nb) >> 6;
      const u64 words = st.row_words - w0;
      const dim3 grid(static_cast<unsigned>((words + 255) / 256), nb);
      group_commit_full_kernel<<<grid, 256, 0, stream>>>(st, s0, w0, gpos);
      check_launch("group_commit_full");
    } else if (st.W != 0) {
      group_commit_qwc_kernel<<<grid_for(static_cast<u64>(nb) * st.W, 256), 256, 0, stream>>>(st, s0, nb, gpos);
      check_launch("group_commit_qwc");
    }
  }
}

// Visiting orders of at least this many terms are sorted on four threads.
constexpr std::size_t kParallelSortTerms = 8192;

// Largest term count for the full mode (the packed matrix and the conflict
// rows are each T^2 / 8 bytes).
constexpr std::size_t kMaxFullGroupingTerms = std::size_t{1} << 18;

}  // namespace
}  // namespace wg

using namespace wg;

extern "C" {

int wg_group_commuting(const wg_pauli_sum* op, int full, void* stream_ptr, size_t* num_groups,
                       uint32_t* sizes, uint32_t* members) {
  return guard([&] {
    require_op(op);
    *num_groups = 0;
    const std::size_t T = op->num_terms, W = op->words;
    if (T == 0) return;
    if (T > 0xFFFFFFFEULL) fail_invalid("group_commuting supports at most 2^32 - 2 terms");
    if (full && T > kMaxFullGroupingTerms) {
      fail_invalid("group_commuting(mode='full') on the device supports at most 262144 terms");
    }
    DeviceScope scope(op->device);
    hipStream_t stream = as_stream(stream_ptr);
    // Visiting order (reference src/grouping.cpp:36-62).
    std::vector<u64> hx(T * W), hz(T * W);
    if (W) {
      check_hip(hipMemcpyAsync(hx.data(), op->x, T * W * sizeof(u64), hipMemcpyDeviceToHost, stream),
                "copy device buffer to host");
      check_hip(hipMemcpyAsync(hz.data(), op->z, T * W * sizeof(u64), hipMemcpyDeviceToHost, stream),
                "copy device buffer to host");
      sync(stream);
    }
    std::vector<u32> order(T), weight(T, 0);
    for (std::size_t t = 0; t < T; ++t) {
      order[t] = static_cast<u32>(t);
      for (std::size_t w = 0; w < W; ++w) weight[t] += __builtin_popcountll(hx[t * W + w] | hz[t * W + w]);
    }
    const auto before = [&](u32 a, u32 b) {
      if (weight[a] != weight[b]) return weight[a] > weight[b];
      for (std::size_t w = 0; w < W; ++w) {
        if (hx[a * W + w] != hx[b * W + w]) return hx[a * W + w] < hx[b * W + w];
        if (hz[a * W + w] != hz[b * W + w]) return hz[a * W + w] < hz[b * W + w];
      }
      return a < b;
    };
    // The order is total (ties end on the index), so sorting quarters on
    // their own threads and merging gives the same permutation.
    if (T >= kParallelSortTerms) {
      const std::size_t q[5] = {0, T / 4, T / 2, 3 * T / 4, T};
      std::vector<std::thread> workers;
      for (int k = 1; k < 4; ++k) {
        workers.emplace_back([&, k] { std::sort(order.begin() + q[k], order.begin() + q[k + 1], before); });
      }
      std::sort(order.begin(), order.begin() + q[1], before);
      for (auto& th : workers) th.join();
      std::inplace_merge(order.begin(), order.begin() + q[1], order.begin() + q[2], before);
      std::inplace_merge(order.begin() + q[2], order.begin() + q[3], order.end(), before);
      std::inplace_merge(order.begin(), order.begin() + q[2], order.end(), before);
    } else {
      std::sort(order.begin(), order.end(), before);
    }
    Scratch scratch(op->device);
    u32* d_order = scratch.take<u32>(T);
    u32* gpos = scratch.take<u32>(T);
    u32* d_count = scratch.take<u32>(1);
    u32* cand = scratch.take<u32>(std::size_t{kBatch} * kCandWords);
    auto* cmask = scratch.take<std::uint8_t>(std::size_t{kBatch} * kResolveBlock);
    check_hip(hipMemcpyAsync(d_order, order.data(), T * sizeof(u32), hipMemcpyHostToDevice, stream),
              "copy host buffer to device");
    check_hip(hipMemsetAsync(d_count, 0, sizeof(u32), stream), "memset");
    GroupState st{};
    st.W = W;
    u64* px = scratch.take<u64>(std::max<std::size_t>(1, T * W));
    u64* pz = scratch.take<u64>(std::max<std::size_t>(1, T * W));
    st.px = px;
    st.pz = pz;
    if (W) {
      gather_masks_kernel<<<grid_for(std::min<std::size_t>(T * W, std::size_t{1} << 22), 256), 256, 0,
                            stream>>>(d_order, op->x, op->z, T, W, px, pz);
      check_launch("gather_masks");
    }
    // Stamps of the resolve kernel when some group id may not fit its LDS.
    u32* gstamps = nullptr;
    const std::size_t gstamp_len = T + kResolveBlock;
    if (T > kLdsGroups) {
      gstamps = scratch.take<u32>(2 * gstamp_len);
      check_hip(hipMemsetAsync(gstamps, 0, 2 * gstamp_len * sizeof(u32), stream), "memset");
    }
    if (full) {
      const std::size_t row_words = (T + 63) / 64;
      const std::size_t words = checked_mul(T, row_words, "group conflict rows");
      u64* commute = scratch.take<u64>(words);
      u64* conflict = scratch.take<u64>(words);
      launch_commute_raw(px, pz, T, px, pz, T, W, commute, row_words, stream);
      check_hip(hipMemsetAsync(conflict, 0, words * sizeof(u64), stream), "memset");
      st.commute = commute;
      st.row_words = row_words;
      st.conflict = conflict;
      ProfileScope prof("group_full", stream);
      run_batches<kFull>(st, T, d_count, cand, cmask, gpos, gstamps, gstamp_len, stream);
    } else {
      u64* xb = scratch.take<u64>(std::max<std::size_t>(1, T * W));
      u64* zb = scratch.take<u64>(std::max<std::size_t>(1, T * W));
      if (W) {
        check_hip(hipMemsetAsync(xb, 0, T * W * sizeof(u64), stream), "memset");
        check_hip(hipMemsetAsync(zb, 0, T * W * sizeof(u64), stream), "memset");
      }
      st.xb = xb;
      st.zb = zb;
      ProfileScope prof("group_qwc", stream);
      run_batches<kQwc>(st, T, d_count, cand, cmask, gpos, gstamps, gstamp_len, stream);
    }
    std::vector<u32> group_of(T);
    u32 G = 0;
    std::vector<u32> gof_pos(T);
    check_hip(hipMemcpyAsync(gof_pos.data(), gpos, T * sizeof(u32), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    check_hip(hipMemcpyAsync(&G, d_count, sizeof(u32), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
    for (std::size_t p = 0; p < T; ++p) group_of[order[p]] = gof_pos[p];
    // Groups in creation order, members in placement (visiting) order.
    std::vector<u32> count(G, 0), start(G + 1, 0);
    for (std::size_t t = 0; t < T; ++t) {
      if (group_of[t] >= G) fail_runtime("group_commuting produced an invalid group id");
      ++count[group_of[t]];
    }
    for (u32 g = 0; g < G; ++g) start[g + 1] = start[g] + count[g];
    std::vector<u32> fill(start.begin(), start.end() - 1);
    for (u32 t : order) members[fill[group_of[t]]++] = t;
    for (u32 g = 0; g < G; ++g) sizes[g] = count[g];
    *num_groups = G;
  });
}

}  // extern "C"

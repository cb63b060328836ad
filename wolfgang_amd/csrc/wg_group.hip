// Greedy commuting groups on the device (reference PauliSum::group_commuting,
// src/grouping.cpp:197-290, strategy "largest_first").
//
// The visiting order (descending weight |x | z|, then canonical key order,
// then index; reference src/grouping.cpp:36-62) is built on the host. The
// placement itself is sequential by definition (each term takes the first
// group that accepts it, given every earlier placement), so one 1024-thread
// block walks the order and, per term, tests all open groups in parallel:
//   * full: group g keeps a conflict bit row over all terms (OR of the
//     anticommutation rows of its members, taken from the packed device
//     commutation matrix); term t is accepted iff bit t of the row is clear,
//     and joining ORs t's anticommutation row into it;
//   * qubit-wise: group g keeps Xb, Zb (ORs of its members' masks); t is
//     accepted iff (x_t & Zb) ^ (z_t & Xb) = 0 on every word (reference
//     src/grouping.cpp:64-85).
// The first accepting group is a block-wide minimum. The result equals the
// reference's graph and streaming paths, which differ only in how they test
// acceptance.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <memory>
#include <numeric>
#include <vector>

#include "wg_internal.hpp"

namespace wg {
namespace {

constexpr int kGroupBlock = 1024;

// Block-wide minimum of v; every thread gets the result.
__device__ u32 block_min(u32 v, u32* scratch) {
  for (int d = 32; d > 0; d >>= 1) v = min(v, static_cast<u32>(__shfl_xor(static_cast<int>(v), d, 64)));
  const u32 wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  u32 r = UINT_MAX;
  for (u32 w = 0; w < kGroupBlock / 64; ++w) r = min(r, scratch[w]);
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kGroupBlock) void greedy_full_kernel(
    const u32* __restrict__ order, u32 T, const u64* __restrict__ commute_bits, u64 row_words,
    u64* __restrict__ conflict, u32* __restrict__ group_of, u32* __restrict__ num_groups) {
  __shared__ u32 scratch[kGroupBlock / 64];
  u32 G = 0;  // identical in every thread
  for (u32 s = 0; s < T; ++s) {
    const u32 t = order[s];
    const u64 w = t >> 6;
    const u64 bit = 1ULL << (t & 63);
    u32 mine = UINT_MAX;
    for (u32 g = threadIdx.x; g < G; g += kGroupBlock) {
      if (!(conflict[g * row_words + w] & bit)) {
        mine = g;
        break;
      }
    }
    u32 g = block_min(mine, scratch);
    if (g == UINT_MAX) g = G++;
    const u64* crow = commute_bits + static_cast<u64>(t) * row_words;
    u64* row = conflict + static_cast<u64>(g) * row_words;
    for (u64 k = threadIdx.x; k < row_words; k += kGroupBlock) row[k] |= ~crow[k];
    if (threadIdx.x == 0) group_of[t] = g;
    __syncthreads();
  }
  if (threadIdx.x == 0) *num_groups = G;
}

__global__ __launch_bounds__(kGroupBlock) void greedy_qwc_kernel(
    const u32* __restrict__ order, u32 T, const u64* __restrict__ x, const u64* __restrict__ z,
    u64 W, u64* __restrict__ xb, u64* __restrict__ zb, u32* __restrict__ group_of,
    u32* __restrict__ num_groups) {
  __shared__ u32 scratch[kGroupBlock / 64];
  u32 G = 0;
  for (u32 s = 0; s < T; ++s) {
    const u32 t = order[s];
    const u64* xt = x + static_cast<u64>(t) * W;
    const u64* zt = z + static_cast<u64>(t) * W;
    u32 mine = UINT_MAX;
    for (u32 g = threadIdx.x; g < G; g += kGroupBlock) {
      bool ok = true;
      for (u64 k = 0; k < W && ok; ++k) {
        ok = ((xt[k] & zb[g * W + k]) ^ (zt[k] & xb[g * W + k])) == 0;
      }
      if (ok) {
        mine = g;
        break;
      }
    }
    u32 g = block_min(mine, scratch);
    if (g == UINT_MAX) g = G++;
    for (u64 k = threadIdx.x; k < W; k += kGroupBlock) {
      xb[g * W + k] |= xt[k];
      zb[g * W + k] |= zt[k];
    }
    if (threadIdx.x == 0) group_of[t] = g;
    __syncthreads();
  }
  if (threadIdx.x == 0) *num_groups = G;
}

// Largest term count for the full mode (packed matrix and conflict rows are
// each T^2 / 8 bytes).
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
    std::sort(order.begin(), order.end(), [&](u32 a, u32 b) {
      if (weight[a] != weight[b]) return weight[a] > weight[b];
      for (std::size_t w = 0; w < W; ++w) {
        if (hx[a * W + w] != hx[b * W + w]) return hx[a * W + w] < hx[b * W + w];
        if (hz[a * W + w] != hz[b * W + w]) return hz[a * W + w] < hz[b * W + w];
      }
      return a < b;
    });
    // The packed commutation matrix first (its entry points take the
    // device's scratch arena themselves, so it must not be held here yet).
    wg_comm_matrix* matrix = nullptr;
    std::unique_ptr<wg_comm_matrix, void (*)(wg_comm_matrix*)> guard_matrix(nullptr,
                                                                            wg_comm_matrix_destroy);
    const uint64_t* bits = nullptr;
    size_t rows = 0, cols = 0, row_words = 0;
    if (full) {
      if (wg_comm_matrix_create(T, T, op->device, &matrix) != WG_OK) fail_runtime(wg_last_error());
      guard_matrix.reset(matrix);
      if (wg_commutes_with_device(op, op, matrix, T * T, stream_ptr) != WG_OK) {
        fail_runtime(wg_last_error());
      }
      int dev = 0;
      if (wg_comm_matrix_device_bits(matrix, &bits) != WG_OK ||
          wg_comm_matrix_info(matrix, &rows, &cols, &row_words, &dev) != WG_OK) {
        fail_runtime(wg_last_error());
      }
    }
    Scratch scratch(op->device);
    u32* d_order = scratch.take<u32>(T);
    u32* d_group = scratch.take<u32>(T);
    u32* d_count = scratch.take<u32>(1);
    check_hip(hipMemcpyAsync(d_order, order.data(), T * sizeof(u32), hipMemcpyHostToDevice, stream),
              "copy host buffer to device");
    if (full) {
      u64* conflict = scratch.take<u64>(checked_mul(T, row_words, "group conflict rows"));
      check_hip(hipMemsetAsync(conflict, 0, T * row_words * sizeof(u64), stream), "memset");
      ProfileScope prof("group_full", stream);
      greedy_full_kernel<<<1, kGroupBlock, 0, stream>>>(
          d_order, static_cast<u32>(T), reinterpret_cast<const u64*>(bits), row_words, conflict, d_group,
          d_count);
      check_launch("greedy_full");
      sync(stream);
    } else {
      u64* xb = scratch.take<u64>(std::max<std::size_t>(1, T * W));
      u64* zb = scratch.take<u64>(std::max<std::size_t>(1, T * W));
      if (W) {
        check_hip(hipMemsetAsync(xb, 0, T * W * sizeof(u64), stream), "memset");
        check_hip(hipMemsetAsync(zb, 0, T * W * sizeof(u64), stream), "memset");
      }
      ProfileScope prof("group_qwc", stream);
      greedy_qwc_kernel<<<1, kGroupBlock, 0, stream>>>(d_order, static_cast<u32>(T), op->x, op->z, W,
                                                       xb, zb, d_group, d_count);
      check_launch("greedy_qwc");
    }
    std::vector<u32> group_of(T);
    u32 G = 0;
    check_hip(hipMemcpyAsync(group_of.data(), d_group, T * sizeof(u32), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    check_hip(hipMemcpyAsync(&G, d_count, sizeof(u32), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
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

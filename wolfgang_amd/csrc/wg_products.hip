// Batched Hamiltonian products: many (lhs @ rhs) pairs of uneven term and
// qubit counts, every term-pair product of every pair served from one device
// work queue by one persistent kernel, then each pair's sort-reduce simplify.
//
// A unit of work is (pair, lhs term, block of 1024 rhs terms); its products
// land at the pair's slot of a shared intermediate buffer in the same
// lhs-outer, rhs-inner order as the single-pair path, with the same phase and
// coefficient arithmetic (reference src/multiply.cpp:267-402), so every
// pair's simplified result is bit-identical to its own matmul. Units are
// ordered by pair (largest first) and claimed with one atomic per block.
//
// Pairs are processed in chunks whose intermediate products fit a byte
// budget; each chunk's queue reset and persistent launch are captured once
// into a HIP graph (the counter is cleared by a one-thread kernel, no memset
// node) and replayed by every run. Simplify has data-dependent sizes and runs
// after the replay, pair by pair, on the same stream.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <numeric>
#include <vector>

#include "wg_internal.hpp"
#include "wg_product_term.hpp"
#include "wg_simplify.hpp"

namespace wg {
namespace {

constexpr int kBlock = 256;
constexpr int kPerThread = 4;
constexpr u32 kChunk = kBlock * kPerThread;

struct ProductPair {
  const u64* lx;
  const u64* lz;
  const double2* lc;
  const u64* rx;
  const u64* rz;
  const double2* rc;
  u64* ox;
  u64* oz;
  double2* oc;
  u64 R;
  u64 W;
  u64 unit_begin;  // first unit of this pair in the chunk
  u64 chunks;      // rhs blocks per lhs term
};

__global__ void reset_counter_kernel(unsigned long long* counter) { *counter = 0; }

// Persistent: each block claims units until the queue is drained. Every block
// reaches the exit test after each claim, so the grid drains.
__global__ __launch_bounds__(kBlock) void product_batch_kernel(
    const ProductPair* __restrict__ pairs, u32 num_pairs, u64 num_units,
    unsigned long long* __restrict__ counter) {
  __shared__ u64 unit_s;
  for (;;) {
    if (threadIdx.x == 0) unit_s = atomicAdd(counter, 1ULL);
    __syncthreads();
    const u64 unit = unit_s;
    __syncthreads();
    if (unit >= num_units) return;
    // Pair of this unit: last pair whose first unit is <= unit.
    u32 lo = 0, hi = num_pairs - 1;
    while (lo < hi) {
      const u32 mid = (lo + hi + 1) / 2;
      if (pairs[mid].unit_begin <= unit) lo = mid; else hi = mid - 1;
    }
    const ProductPair& p = pairs[lo];
    const u64 local = unit - p.unit_begin;
    const u64 i = local / p.chunks;
    const u64 j0 = (local % p.chunks) * kChunk;
    const u64 W = p.W;
    int ly = 0;
    for (u64 w = 0; w < W; ++w) ly += __popcll(p.lx[i * W + w] & p.lz[i * W + w]);
    const double2 a = p.lc[i];
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
      const u64 j = j0 + static_cast<u64>(q) * kBlock + threadIdx.x;
      if (j >= p.R) break;
      const u64 o = i * p.R + j;
      int k = -ly;
      for (u64 w = 0; w < W; ++w) {
        const u64 lxw = p.lx[i * W + w], lzw = p.lz[i * W + w];
        const u64 rxw = p.rx[j * W + w], rzw = p.rz[j * W + w];
        const u64 oxw = lxw ^ rxw, ozw = lzw ^ rzw;
        p.ox[o * W + w] = oxw;
        p.oz[o * W + w] = ozw;
        k += __popcll(oxw & ozw) - __popcll(rxw & rzw) + 2 * __popcll(lxw & rzw);
      }
      const double2 b = p.rc[j];
      p.oc[o] = apply_phase(complex_product(a, b), k);
    }
  }
}

}  // namespace
}  // namespace wg

using namespace wg;

struct wg_product_batch {
  struct Chunk {
    std::vector<std::size_t> pairs;  // pair indices, in launch order
    std::vector<u64> product_offset; // per pair of the chunk: first product in the buffer
    std::size_t products = 0;
    u64 units = 0;
    ProductPair* d_pairs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  int device = 0;
  std::vector<const wg_pauli_sum*> lhs, rhs;
  std::vector<std::size_t> products;  // per pair
  std::vector<int> hash_seed;         // per pair: simplify seed rule of matmul
  std::vector<Chunk> chunks;
  u64* d_x = nullptr;
  u64* d_z = nullptr;
  double2* d_c = nullptr;
  std::size_t cap_words = 0, cap_products = 0;
  unsigned long long* d_counter = nullptr;
  unsigned grid = 0;

  ~wg_product_batch() {
    int prev = -1;
    const hipError_t r = hipGetDevice(&prev);
    (void)hipSetDevice(device);
    for (Chunk& c : chunks) {
      if (c.exec) (void)hipGraphExecDestroy(c.exec);
      if (c.graph) (void)hipGraphDestroy(c.graph);
      device_free(c.d_pairs);
    }
    device_free(d_x);
    device_free(d_z);
    device_free(d_c);
    device_free(d_counter);
    if (r == hipSuccess && prev >= 0) (void)hipSetDevice(prev);
  }
};

namespace {

std::size_t pair_products(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, std::size_t max_terms) {
  require_op(lhs);
  require_op(rhs);
  require_same_device("HIP matmul", lhs, rhs);
  if (lhs->num_qubits != rhs->num_qubits) fail_invalid("PauliSum matmul requires the same num_qubits");
  const std::size_t L = lhs->num_terms, R = rhs->num_terms;
  if (L != 0 && R > SIZE_MAX / L) fail_invalid("matmul intermediate term count overflows size_t");
  if (L * R > max_terms) fail_invalid("matmul intermediate term count exceeds max_intermediate_terms");
  return L * R;
}

}  // namespace

extern "C" {

int wg_product_batch_create(size_t count, const wg_pauli_sum* const* lhs,
                            const wg_pauli_sum* const* rhs, size_t max_intermediate_terms,
                            size_t budget_bytes, void* stream_ptr, wg_product_batch** out) {
  return guard([&] {
    *out = nullptr;
    if (count == 0) fail_invalid("product batch must contain at least one pair");
    if (lhs == nullptr || rhs == nullptr) fail_invalid("product batch operand arrays must be non-null");
    std::unique_ptr<wg_product_batch> b(new wg_product_batch());
    for (std::size_t i = 0; i < count; ++i) {
      b->products.push_back(pair_products(lhs[i], rhs[i], max_intermediate_terms));
      if (lhs[i]->device != lhs[0]->device) fail_invalid("batched products must share one device");
      b->lhs.push_back(lhs[i]);
      b->rhs.push_back(rhs[i]);
    }
    b->device = lhs[0]->device;
    DeviceScope scope(b->device);
    hipStream_t stream = as_stream(stream_ptr);
    if (budget_bytes == 0) budget_bytes = std::size_t{16} << 30;
    // matmul's simplify seed rule, per pair (reference src/multiply.cpp:107-261).
    for (std::size_t i = 0; i < count; ++i) {
      const std::size_t W = lhs[i]->words;
      const bool hash = (W == 1 || W == 2) && b->products[i] != 0 &&
                        (operand_duplicate_heavy(lhs[i], stream) ||
                         operand_duplicate_heavy(rhs[i], stream));
      b->hash_seed.push_back(hash ? 1 : -1);
    }
    // Chunks: pairs by decreasing size, packed greedily under the budget.
    std::vector<std::size_t> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](std::size_t a, std::size_t c) {
      return b->products[a] > b->products[c];
    });
    auto bytes_of = [&](std::size_t i) {
      return b->products[i] * (2 * 8 * lhs[i]->words + sizeof(double2));
    };
    wg_product_batch::Chunk cur;
    std::size_t cur_bytes = 0;
    for (std::size_t i : order) {
      if (b->products[i] == 0) continue;
      if (!cur.pairs.empty() && cur_bytes + bytes_of(i) > budget_bytes) {
        b->chunks.push_back(std::move(cur));
        cur = wg_product_batch::Chunk();
        cur_bytes = 0;
      }
      cur.pairs.push_back(i);
      cur_bytes += bytes_of(i);
    }
    if (!cur.pairs.empty()) b->chunks.push_back(std::move(cur));
    // Buffer capacity: the largest chunk's words and products.
    for (auto& c : b->chunks) {
      std::size_t words = 0, prods = 0;
      for (std::size_t i : c.pairs) {
        c.product_offset.push_back(prods);
        words += b->products[i] * lhs[i]->words;
        prods += b->products[i];
      }
      c.products = prods;
      b->cap_words = std::max(b->cap_words, words);
      b->cap_products = std::max(b->cap_products, prods);
    }
    b->d_x = device_alloc<u64>(std::max<std::size_t>(b->cap_words, 1));
    b->d_z = device_alloc<u64>(std::max<std::size_t>(b->cap_words, 1));
    b->d_c = device_alloc<double2>(std::max<std::size_t>(b->cap_products, 1));
    b->d_counter = device_alloc<unsigned long long>(1);
    int cus = 0;
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device),
              "query device attribute");
    b->grid = static_cast<unsigned>(std::max(1, cus) * 8);
    for (auto& c : b->chunks) {
      std::vector<ProductPair> pd;
      std::size_t word_off = 0;
      u64 unit = 0;
      for (std::size_t k = 0; k < c.pairs.size(); ++k) {
        const std::size_t i = c.pairs[k];
        const wg_pauli_sum* L = lhs[i];
        const wg_pauli_sum* R = rhs[i];
        ProductPair p{};
        p.lx = L->x;
        p.lz = L->z;
        p.lc = L->coeffs;
        p.rx = R->x;
        p.rz = R->z;
        p.rc = R->coeffs;
        p.ox = b->d_x + word_off;
        p.oz = b->d_z + word_off;
        p.oc = b->d_c + c.product_offset[k];
        p.R = R->num_terms;
        p.W = L->words;
        p.chunks = (R->num_terms + kChunk - 1) / kChunk;
        p.unit_begin = unit;
        unit += static_cast<u64>(L->num_terms) * p.chunks;
        word_off += b->products[i] * L->words;
        pd.push_back(p);
      }
      c.units = unit;
      c.d_pairs = device_alloc<ProductPair>(pd.size());
      check_hip(hipMemcpyAsync(c.d_pairs, pd.data(), pd.size() * sizeof(ProductPair),
                               hipMemcpyHostToDevice, stream),
                "copy host buffer to device");
      sync(stream);
      hipStream_t cs;
      check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "create capture stream");
      check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
      reset_counter_kernel<<<1, 1, 0, cs>>>(b->d_counter);
      product_batch_kernel<<<b->grid, kBlock, 0, cs>>>(c.d_pairs, static_cast<u32>(pd.size()),
                                                        c.units, b->d_counter);
      hipGraph_t g = nullptr;
      const hipError_t r = hipStreamEndCapture(cs, &g);
      (void)hipStreamDestroy(cs);
      check_hip(r, "end capture");
      c.graph = g;
      check_hip(hipGraphInstantiate(&c.exec, g, nullptr, nullptr, 0), "instantiate graph");
    }
    *out = b.release();
  });
}

int wg_product_batch_run(wg_product_batch* b, int simplify_output, void* stream_ptr,
                         wg_pauli_sum** outs) {
  return guard([&] {
    if (b == nullptr) fail_invalid("product batch handle is null");
    if (outs == nullptr) fail_invalid("product batch output array must be non-null");
    DeviceScope scope(b->device);
    hipStream_t stream = as_stream(stream_ptr);
    const std::size_t count = b->lhs.size();
    std::vector<std::unique_ptr<wg_pauli_sum>> results(count);
    for (std::size_t i = 0; i < count; ++i) {
      if (b->products[i] == 0) results[i].reset(new_pauli_sum(b->lhs[i]->num_qubits, 0, b->device));
    }
    Scratch scratch(b->device);
    for (auto& c : b->chunks) {
      check_hip(hipGraphLaunch(c.exec, stream), "launch graph");
      check_launch("product_batch");
      std::size_t word_off = 0;
      for (std::size_t k = 0; k < c.pairs.size(); ++k) {
        const std::size_t i = c.pairs[k];
        const wg_pauli_sum* L = b->lhs[i];
        const std::size_t P = b->products[i], W = L->words;
        const u64* px = b->d_x + word_off;
        const u64* pz = b->d_z + word_off;
        const double2* pc = b->d_c + c.product_offset[k];
        if (simplify_output) {
          // simplify_device synchronizes before it returns: its scratch is
          // free again for the next pair.
          const std::size_t mark = scratch.mark();
          results[i].reset(simplify_device(px, pz, pc, L->num_qubits, P, b->device, 1.0e-12, 0.0,
                                           b->hash_seed[i], scratch, stream));
          scratch.release(mark);
        } else {
          std::unique_ptr<wg_pauli_sum> r(new_pauli_sum(L->num_qubits, P, b->device));
          if (W) {
            check_hip(hipMemcpyAsync(r->x, px, P * W * sizeof(u64), hipMemcpyDeviceToDevice, stream),
                      "copy device buffer");
            check_hip(hipMemcpyAsync(r->z, pz, P * W * sizeof(u64), hipMemcpyDeviceToDevice, stream),
                      "copy device buffer");
          }
          check_hip(hipMemcpyAsync(r->coeffs, pc, P * sizeof(double2), hipMemcpyDeviceToDevice,
                                   stream),
                    "copy device buffer");
          results[i] = std::move(r);
        }
        word_off += P * W;
      }
      sync(stream);
    }
    for (std::size_t i = 0; i < count; ++i) outs[i] = results[i].release();
  });
}

int wg_product_batch_info(const wg_product_batch* b, size_t* pairs, size_t* chunks,
                          size_t* total_products) {
  return guard([&] {
    if (b == nullptr) fail_invalid("product batch handle is null");
    if (pairs) *pairs = b->lhs.size();
    if (chunks) *chunks = b->chunks.size();
    if (total_products) {
      *total_products = std::accumulate(b->products.begin(), b->products.end(), std::size_t{0});
    }
  });
}

void wg_product_batch_destroy(wg_product_batch* b) { delete b; }

}  // extern "C"

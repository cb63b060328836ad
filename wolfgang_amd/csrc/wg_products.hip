// Batched Hamiltonian products: many (lhs @ rhs) pairs of uneven term and
// qubit counts, every unit of every pair served from one device work queue by
// one persistent kernel, then each pair's sort-reduce simplify.
//
// A unit of work is (pair, lhs term, block of rhs terms: 2,048 when the unit
// writes sort keys, 1,024 when it writes products). With simplify
// (the matmul path) a unit writes its products' word-0 sort keys, indices and
// key statistics, as matmul's key pass does, and the products themselves are
// never materialized: each pair's sort, segment heads and segment sums read
// them through the pair's operand records, which the graph rebuilds on every
// replay (so later changes to the operands are seen). Without simplify a unit
// writes its products at the pair's slot of a shared buffer. Either way the
// order (lhs-outer, rhs-inner) and the phase and coefficient arithmetic are
// the single-pair path's (reference src/multiply.cpp:267-402), so every pair's
// result is bit-identical to its own matmul. Units are ordered by pair
// (largest first) and claimed with one atomic per block.
//
// Pairs are processed in chunks whose intermediate products fit a byte
// budget; each chunk's queue reset and persistent launch are captured once
// into a HIP graph (the counter is cleared by a one-thread kernel, no memset
// node) and replayed by every run (the materializing graphs are captured on
// the first run without simplify). Simplify has data-dependent sizes and runs
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

// Units claimed per atomic: one counter serializes its atomics, so claiming
// single 1,024-product units would cost more than the units themselves.
constexpr u64 kClaim = 32;

// Persistent: each block claims kClaim units at a time until the queue is
// drained. Every block reaches the exit test after each claim, so the grid
// drains.
__global__ __launch_bounds__(kBlock) void product_batch_kernel(
    const ProductPair* __restrict__ pairs, u32 num_pairs, u64 num_units,
    unsigned long long* __restrict__ counter) {
  __shared__ u64 unit_s;
  for (;;) {
    if (threadIdx.x == 0) unit_s = atomicAdd(counter, kClaim);
    __syncthreads();
    const u64 first = unit_s;
    __syncthreads();
    if (first >= num_units) return;
    const u64 last = std::min<u64>(first + kClaim, num_units);
    for (u64 unit = first; unit < last; ++unit) {
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
}

// A pair of the key pass: its operand records, where its keys, indices and
// per-unit statistics go, and its first unit in the chunk's queue. shape is
// the mask word count W (1..4), or 0 for one-word keys x << 32 | z.
struct KeyPair {
  ProductTerms P;
  u64* keys;
  u32* idx;
  u64* stats;
  u64 unit_begin;
  u32 chunks;
  u32 units;
  u32 shape;
};

// Persistent key pass: each block claims kClaim units at a time until the
// queue is drained; every block reaches the exit test after each claim, so
// the grid drains.
__global__ __launch_bounds__(kProductKeyBlock) void product_keys_batch_kernel(
    const KeyPair* __restrict__ pairs, u32 num_pairs, u64 num_units,
    unsigned long long* __restrict__ counter) {
  __shared__ u64 unit_s;
  __shared__ u64 lds[kProductKeyBlock / 64][kProductMaxStats];
  for (;;) {
    if (threadIdx.x == 0) unit_s = atomicAdd(counter, kClaim);
    __syncthreads();
    const u64 first = unit_s;
    __syncthreads();
    if (first >= num_units) return;
    const u64 last = std::min<u64>(first + kClaim, num_units);
    for (u64 unit = first; unit < last; ++unit) {
      u32 lo = 0, hi = num_pairs - 1;
      while (lo < hi) {
        const u32 mid = (lo + hi + 1) / 2;
        if (pairs[mid].unit_begin <= unit) lo = mid; else hi = mid - 1;
      }
      const KeyPair& p = pairs[lo];
      const u32 local = static_cast<u32>(unit - p.unit_begin);
      const u32 i = local / p.chunks;
      const u32 j0 = (local - i * p.chunks) * kProductChunk;
      switch (p.shape) {
        case 0:
          product_keys_unit<1, true, false>(p.P, i, j0, p.keys, p.idx, p.stats, p.units, local, lds);
          break;
        case 1:
          product_keys_unit<1, false, false>(p.P, i, j0, p.keys, p.idx, p.stats, p.units, local, lds);
          break;
        case 2:
          product_keys_unit<2, false, false>(p.P, i, j0, p.keys, p.idx, p.stats, p.units, local, lds);
          break;
        case 3:
          product_keys_unit<3, false, false>(p.P, i, j0, p.keys, p.idx, p.stats, p.units, local, lds);
          break;
        default:
          product_keys_unit<4, false, false>(p.P, i, j0, p.keys, p.idx, p.stats, p.units, local, lds);
          break;
      }
    }
  }
}

}  // namespace
}  // namespace wg

using namespace wg;

struct wg_product_batch {
  // Pairs packed under the byte budget: for runs with simplify the key pass
  // (keys, indices, statistics), for runs without it the materialized products.
  struct Chunk {
    std::vector<std::size_t> pairs;  // pair indices, in launch order
    std::vector<u64> offset;         // per pair: first product in the buffers
    std::vector<u64> stat_offset;    // per pair (key chunks): first statistics word
    u64 units = 0;
    void* d_desc = nullptr;          // KeyPair[] or ProductPair[]
    hipGraphExec_t exec = nullptr;
  };
  int device = 0;
  std::vector<const wg_pauli_sum*> lhs, rhs;
  std::vector<std::size_t> products;  // per pair
  std::vector<int> hash_seed;         // per pair: simplify seed rule of matmul
  std::vector<bool> key_pass;         // per pair: 1 to 4 mask words (products not materialized)
  std::vector<u64*> lrec, rrec;       // per key-pass pair: operand records
  std::vector<std::size_t> units;     // per key-pass pair: key units
  std::vector<Chunk> key_chunks;      // key-pass pairs
  std::vector<Chunk> mat_chunks;      // every pair; graphs captured on first use
  u64* d_x = nullptr;
  u64* d_z = nullptr;
  double2* d_c = nullptr;
  u64* d_keys = nullptr;
  u32* d_idx = nullptr;
  u64* d_stats = nullptr;
  std::size_t cap_words = 0, cap_products = 0;
  unsigned long long* d_counter = nullptr;
  unsigned grid = 0;

  ~wg_product_batch() {
    int prev = -1;
    const hipError_t r = hipGetDevice(&prev);
    (void)hipSetDevice(device);
    for (auto* list : {&key_chunks, &mat_chunks}) {
      for (Chunk& c : *list) {
        if (c.exec) (void)hipGraphExecDestroy(c.exec);
        device_free(c.d_desc);
      }
    }
    for (u64* p : lrec) device_free(p);
    for (u64* p : rrec) device_free(p);
    device_free(d_x);
    device_free(d_z);
    device_free(d_c);
    device_free(d_keys);
    device_free(d_idx);
    device_free(d_stats);
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

// Greedy packing of `pairs` (largest first) into chunks of at most `budget`
// bytes (a larger pair gets a chunk of its own).
template <class Bytes>
std::vector<wg_product_batch::Chunk> pack_chunks(std::vector<std::size_t> pairs,
                                                 const std::vector<std::size_t>& products,
                                                 std::size_t budget, Bytes&& bytes_of) {
  std::stable_sort(pairs.begin(), pairs.end(),
                   [&](std::size_t a, std::size_t c) { return products[a] > products[c]; });
  std::vector<wg_product_batch::Chunk> chunks;
  wg_product_batch::Chunk cur;
  std::size_t cur_bytes = 0;
  for (std::size_t i : pairs) {
    if (!cur.pairs.empty() && cur_bytes + bytes_of(i) > budget) {
      chunks.push_back(std::move(cur));
      cur = wg_product_batch::Chunk();
      cur_bytes = 0;
    }
    cur.pairs.push_back(i);
    cur_bytes += bytes_of(i);
  }
  if (!cur.pairs.empty()) chunks.push_back(std::move(cur));
  return chunks;
}

// Captures `enqueue(stream)` into a graph and instantiates it.
template <class Fn>
hipGraphExec_t capture(Fn&& enqueue) {
  hipStream_t cs;
  check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "create capture stream");
  check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
  enqueue(cs);
  hipGraph_t g = nullptr;
  const hipError_t r = hipStreamEndCapture(cs, &g);
  (void)hipStreamDestroy(cs);
  check_hip(r, "end capture");
  hipGraphExec_t exec = nullptr;
  const hipError_t ri = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  check_hip(ri, "instantiate graph");
  return exec;
}

template <class T>
T* upload(const std::vector<T>& host, hipStream_t stream) {
  T* d = device_alloc<T>(host.size());
  check_hip(hipMemcpyAsync(d, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, stream),
            "copy host buffer to device");
  sync(stream);
  return d;
}

// The materializing graphs (runs without simplify): every product of every
// pair at the pair's slot of (d_x, d_z, d_c).
void build_materializing_graphs(wg_product_batch* b, hipStream_t stream) {
  b->d_x = device_alloc<u64>(std::max<std::size_t>(b->cap_words, 1));
  b->d_z = device_alloc<u64>(std::max<std::size_t>(b->cap_words, 1));
  b->d_c = device_alloc<double2>(std::max<std::size_t>(b->cap_products, 1));
  for (auto& c : b->mat_chunks) {
    std::vector<ProductPair> pd;
    std::size_t word_off = 0;
    u64 unit = 0;
    for (std::size_t k = 0; k < c.pairs.size(); ++k) {
      const std::size_t i = c.pairs[k];
      const wg_pauli_sum* L = b->lhs[i];
      const wg_pauli_sum* R = b->rhs[i];
      ProductPair p{};
      p.lx = L->x;
      p.lz = L->z;
      p.lc = L->coeffs;
      p.rx = R->x;
      p.rz = R->z;
      p.rc = R->coeffs;
      p.ox = b->d_x + word_off;
      p.oz = b->d_z + word_off;
      p.oc = b->d_c + c.offset[k];
      p.R = R->num_terms;
      p.W = L->words;
      p.chunks = (R->num_terms + kChunk - 1) / kChunk;
      p.unit_begin = unit;
      unit += static_cast<u64>(L->num_terms) * p.chunks;
      word_off += b->products[i] * L->words;
      pd.push_back(p);
    }
    c.units = unit;
    ProductPair* d = upload(pd, stream);
    c.d_desc = d;
    c.exec = capture([&](hipStream_t cs) {
      reset_counter_kernel<<<1, 1, 0, cs>>>(b->d_counter);
      product_batch_kernel<<<b->grid, kBlock, 0, cs>>>(d, static_cast<u32>(pd.size()), c.units,
                                                        b->d_counter);
    });
  }
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
    std::vector<std::size_t> key_pairs, all_pairs;
    b->units.assign(count, 0);
    for (std::size_t i = 0; i < count; ++i) {
      // matmul's simplify seed rule and path, per pair (reference
      // src/multiply.cpp:107-261; wg_matmul). A pair past simplify's term
      // limit takes the materialized path, whose simplify reports the limit.
      b->hash_seed.push_back(b->products[i] ? matmul_seed_mode(lhs[i], rhs[i], stream) : -1);
      const std::size_t W = lhs[i]->words;
      const bool kp = W >= 1 && W <= 4 && b->products[i] != 0 && b->products[i] < 0xFFFFFFFFULL;
      b->key_pass.push_back(kp);
      if (kp) {
        key_pairs.push_back(i);
        b->units[i] = lhs[i]->num_terms * ((rhs[i]->num_terms + kProductChunk - 1) / kProductChunk);
      }
      if (b->products[i]) all_pairs.push_back(i);
    }
    auto key_bytes = [&](std::size_t i) {
      return b->products[i] * (sizeof(u64) + sizeof(u32)) +
             b->units[i] * (4 * lhs[i]->words + 1) * sizeof(u64);
    };
    auto mat_bytes = [&](std::size_t i) {
      return b->products[i] * (2 * 8 * lhs[i]->words + sizeof(double2));
    };
    b->key_chunks = pack_chunks(key_pairs, b->products, budget_bytes, key_bytes);
    b->mat_chunks = pack_chunks(all_pairs, b->products, budget_bytes, mat_bytes);
    std::size_t cap_keys = 0, cap_stats = 0;
    for (auto& c : b->key_chunks) {
      std::size_t prods = 0, stats = 0;
      for (std::size_t i : c.pairs) {
        c.offset.push_back(prods);
        c.stat_offset.push_back(stats);
        prods += b->products[i];
        stats += b->units[i] * (4 * lhs[i]->words + 1);
      }
      cap_keys = std::max(cap_keys, prods);
      cap_stats = std::max(cap_stats, stats);
    }
    for (auto& c : b->mat_chunks) {
      std::size_t words = 0, prods = 0;
      for (std::size_t i : c.pairs) {
        c.offset.push_back(prods);
        words += b->products[i] * lhs[i]->words;
        prods += b->products[i];
      }
      b->cap_words = std::max(b->cap_words, words);
      b->cap_products = std::max(b->cap_products, prods);
    }
    b->d_keys = device_alloc<u64>(std::max<std::size_t>(cap_keys, 1));
    b->d_idx = device_alloc<u32>(std::max<std::size_t>(cap_keys, 1));
    b->d_stats = device_alloc<u64>(std::max<std::size_t>(cap_stats, 1));
    b->d_counter = device_alloc<unsigned long long>(1);
    b->lrec.assign(count, nullptr);
    b->rrec.assign(count, nullptr);
    for (std::size_t i : key_pairs) {
      const std::size_t RW = product_record_words(static_cast<u32>(lhs[i]->words));
      b->lrec[i] = device_alloc<u64>(checked_mul(lhs[i]->num_terms, RW, "operand records"));
      b->rrec[i] = device_alloc<u64>(checked_mul(rhs[i]->num_terms, RW, "operand records"));
    }
    int cus = 0;
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device),
              "query device attribute");
    b->grid = static_cast<unsigned>(std::max(1, cus) * 8);
    wg_product_batch* bp = b.get();
    for (auto& c : b->key_chunks) {
      std::vector<KeyPair> kp;
      u64 unit = 0;
      for (std::size_t k = 0; k < c.pairs.size(); ++k) {
        const std::size_t i = c.pairs[k];
        KeyPair p{};
        p.P = ProductTerms{b->lrec[i], b->rrec[i], static_cast<u32>(rhs[i]->num_terms),
                           static_cast<u32>(lhs[i]->words)};
        p.keys = b->d_keys + c.offset[k];
        p.idx = b->d_idx + c.offset[k];
        p.stats = b->d_stats + c.stat_offset[k];
        p.unit_begin = unit;
        p.chunks = static_cast<u32>((rhs[i]->num_terms + kProductChunk - 1) / kProductChunk);
        p.units = static_cast<u32>(b->units[i]);
        p.shape = (lhs[i]->words == 1 && lhs[i]->num_qubits <= 32) ? 0U
                                                                     : static_cast<u32>(lhs[i]->words);
        unit += b->units[i];
        kp.push_back(p);
      }
      c.units = unit;
      KeyPair* d = upload(kp, stream);
      c.d_desc = d;
      c.exec = capture([&](hipStream_t cs) {
        // Records are rebuilt on every replay, so the run reads the operands'
        // current contents.
        for (std::size_t i : c.pairs) {
          build_operand_records(bp->lhs[i], bp->lrec[i], cs);
          build_operand_records(bp->rhs[i], bp->rrec[i], cs);
        }
        reset_counter_kernel<<<1, 1, 0, cs>>>(bp->d_counter);
        product_keys_batch_kernel<<<bp->grid, kProductKeyBlock, 0, cs>>>(
            d, static_cast<u32>(kp.size()), c.units, bp->d_counter);
      });
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
    if (simplify_output) {
      for (auto& c : b->key_chunks) {
        check_hip(hipGraphLaunch(c.exec, stream), "launch graph");
        check_launch("product_keys_batch");
        for (std::size_t k = 0; k < c.pairs.size(); ++k) {
          const std::size_t i = c.pairs[k];
          const wg_pauli_sum* L = b->lhs[i];
          const ProductTerms PT{b->lrec[i], b->rrec[i], static_cast<u32>(b->rhs[i]->num_terms),
                                static_cast<u32>(L->words)};
          // Each simplify synchronizes before it returns: its scratch is free
          // again for the next pair.
          const std::size_t mark = scratch.mark();
          results[i].reset(simplify_product_keys(
              PT, L->num_qubits, b->products[i], b->device, b->d_keys + c.offset[k],
              b->d_idx + c.offset[k], b->d_stats + c.stat_offset[k], b->units[i], 1.0e-12, 0.0,
              b->hash_seed[i], scratch, stream));
          scratch.release(mark);
        }
      }
      // Pairs without a key pass (no mask words, more than four, or past the
      // term limit): matmul's materialized path.
      for (std::size_t i = 0; i < count; ++i) {
        if (b->products[i] == 0 || b->key_pass[i]) continue;
        const std::size_t mark = scratch.mark();
        results[i].reset(matmul_device(b->lhs[i], b->rhs[i], b->hash_seed[i], scratch, stream));
        scratch.release(mark);
      }
    } else {
      if (!b->mat_chunks.empty() && b->mat_chunks[0].exec == nullptr) {
        build_materializing_graphs(b, stream);
      }
      for (auto& c : b->mat_chunks) {
        check_hip(hipGraphLaunch(c.exec, stream), "launch graph");
        check_launch("product_batch");
        std::size_t word_off = 0;
        for (std::size_t k = 0; k < c.pairs.size(); ++k) {
          const std::size_t i = c.pairs[k];
          const wg_pauli_sum* L = b->lhs[i];
          const std::size_t P = b->products[i], W = L->words;
          std::unique_ptr<wg_pauli_sum> r(new_pauli_sum(L->num_qubits, P, b->device));
          if (W) {
            check_hip(hipMemcpyAsync(r->x, b->d_x + word_off, P * W * sizeof(u64),
                                     hipMemcpyDeviceToDevice, stream),
                      "copy device buffer");
            check_hip(hipMemcpyAsync(r->z, b->d_z + word_off, P * W * sizeof(u64),
                                     hipMemcpyDeviceToDevice, stream),
                      "copy device buffer");
          }
          check_hip(hipMemcpyAsync(r->coeffs, b->d_c + c.offset[k], P * sizeof(double2),
                                   hipMemcpyDeviceToDevice, stream),
                    "copy device buffer");
          results[i] = std::move(r);
          word_off += P * W;
        }
        sync(stream);
      }
    }
    for (std::size_t i = 0; i < count; ++i) outs[i] = results[i].release();
  });
}

int wg_product_batch_info(const wg_product_batch* b, size_t* pairs, size_t* chunks,
                          size_t* total_products) {
  return guard([&] {
    if (b == nullptr) fail_invalid("product batch handle is null");
    if (pairs) *pairs = b->lhs.size();
    if (chunks) *chunks = b->key_chunks.size() + (b->key_chunks.empty() ? b->mat_chunks.size() : 0);
    if (total_products) {
      *total_products = std::accumulate(b->products.begin(), b->products.end(), std::size_t{0});
    }
  });
}

void wg_product_batch_destroy(wg_product_batch* b) { delete b; }

}  // extern "C"

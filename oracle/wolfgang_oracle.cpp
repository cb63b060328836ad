// CPU oracle for the packed symplectic Pauli hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP product path
// is compared against; it is never linked into, loaded by, or called from the
// product library (wolfgang_amd/). Only tests/, the smoke entry point and the
// benchmark's CPU-baseline leg may load liboracle.
//
// It restates the reference's scalar C++ algorithms (and its AVX-512 / threaded
// commutation baselines) in plain C++20 over an extern "C" interface. Build with
// -ffp-contract=off and no -march flags so that the x86-64 arithmetic matches
// the reference build (no FMA contraction of (ac - bd, ad + bc)).
//
// Layout (reference include/wolfgang/pauli_sum.hpp:125-138):
//   x[t*W + w], z[t*W + w] : uint64 words, qubit q is bit q%64 of word q/64
//   c[2*t], c[2*t+1]       : complex128 coefficient (re, im)
//
// What pins each function (golden vectors, closed forms, the compiled
// reference) is stated in the comments below and exercised by tests/.

#include <algorithm>
#include <atomic>
#include <bit>
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include <immintrin.h>

namespace {

using u64 = std::uint64_t;
using cplx = std::complex<double>;

thread_local std::string g_error;

struct Op {
  std::size_t n = 0;
  std::size_t W = 0;
  std::size_t T = 0;
  std::vector<u64> x;
  std::vector<u64> z;
  std::vector<cplx> c;
};

Op make_op(std::size_t n, std::size_t T, const u64* x, const u64* z, const double* c) {
  Op op;
  op.n = n;
  op.W = n == 0 ? 0 : (n + 63) / 64;
  op.T = T;
  op.x.assign(x, x + T * op.W);
  op.z.assign(z, z + T * op.W);
  op.c.resize(T);
  for (std::size_t t = 0; t < T; ++t) op.c[t] = cplx(c[2 * t], c[2 * t + 1]);
  return op;
}

// i^k applied to a value by exact swap/negate
// (reference src/detail/phase.hpp:24-36).
cplx apply_phase(cplx v, std::int64_t k) {
  switch (static_cast<u64>(k) & 3U) {
    case 0: return v;
    case 1: return {-v.imag(), v.real()};
    case 2: return {-v.real(), -v.imag()};
    default: return {v.imag(), -v.real()};
  }
}

// i^k as a complex constant (reference src/detail/phase.hpp:8-21).
cplx phase_value(std::int64_t k) {
  switch (static_cast<u64>(k) & 3U) {
    case 0: return {1.0, 0.0};
    case 1: return {0.0, 1.0};
    case 2: return {-1.0, 0.0};
    default: return {0.0, -1.0};
  }
}

// ---------------------------------------------------------------- simplify
// Reference src/simplify.cpp:67-143 (sampled duplicate-rate dispatch) and
// :147-430 (hash accumulation / stable key sort + construction-order sums).

int key_cmp(const Op& op, std::size_t a, std::size_t b) {
  // Lexicographic over (x0, z0, x1, z1, ...) as unsigned words
  // (reference src/detail/packed_key.hpp:86-133).
  for (std::size_t w = 0; w < op.W; ++w) {
    u64 xa = op.x[a * op.W + w], xb = op.x[b * op.W + w];
    if (xa != xb) return xa < xb ? -1 : 1;
    u64 za = op.z[a * op.W + w], zb = op.z[b * op.W + w];
    if (za != zb) return za < zb ? -1 : 1;
  }
  return 0;
}

struct KeyView {
  const Op* op;
  std::size_t t;
};

struct KeyHash {
  std::size_t operator()(const KeyView& k) const noexcept {
    // Any hash works for the oracle: only membership matters, never iteration
    // order (survivors are sorted before output).
    std::size_t h = 1469598103934665603ULL;
    for (std::size_t w = 0; w < k.op->W; ++w) {
      h ^= k.op->x[k.t * k.op->W + w] * 0x9E3779B97F4A7C15ULL;
      h = (h << 7) | (h >> 57);
      h ^= k.op->z[k.t * k.op->W + w] * 0xC2B2AE3D27D4EB4FULL;
      h = (h << 11) | (h >> 53);
    }
    return h;
  }
};

struct KeyEq {
  bool operator()(const KeyView& a, const KeyView& b) const noexcept {
    return key_cmp(*a.op, a.t, b.t) == 0;
  }
};

double sampled_duplicate_rate(const Op& op, std::size_t sample) {
  if (sample == 0) return 0.0;
  std::unordered_set<KeyView, KeyHash, KeyEq> seen;
  seen.reserve(sample);
  for (std::size_t t = 0; t < sample; ++t) seen.insert({&op, t});
  return 1.0 - static_cast<double>(seen.size()) / static_cast<double>(sample);
}

// Reference src/simplify.cpp:107-143: only W in {1, 2} with T >= 4096 takes the
// hash path; first sample 512, then 4096 when the first rate is in (0.05, 0.25).
bool simplify_uses_hash_path(const Op& op) {
  if (op.W != 1 && op.W != 2) return false;
  if (op.T < 4096) return false;
  const double first = sampled_duplicate_rate(op, std::min<std::size_t>(op.T, 512));
  if (first >= 0.25) return true;
  if (first <= 0.05) return false;
  return sampled_duplicate_rate(op, std::min<std::size_t>(op.T, 4096)) >= 0.25;
}

template <typename Fn>
void parallel_for(std::size_t count, int threads, Fn fn) {
  if (threads <= 1 || count <= 1) {
    for (std::size_t i = 0; i < count; ++i) fn(i);
    return;
  }
  std::atomic<std::size_t> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&]() {
      for (std::size_t i = next.fetch_add(1); i < count; i = next.fetch_add(1)) fn(i);
    });
  }
  for (auto& th : pool) th.join();
}

// Worker threads for the oracle's sorts and products (orc_set_threads;
// default 1). Any thread count gives the same result: the canonical order is
// a total order (key, then construction index).
int g_threads = 1;

template <int W>
struct KeyRec {
  u64 k[2 * W];  // x0, z0, x1, z1, ...
  u64 idx;
  bool operator<(const KeyRec& o) const {
    for (int i = 0; i < 2 * W; ++i) {
      if (k[i] != o.k[i]) return k[i] < o.k[i];
    }
    return idx < o.idx;
  }
};

// Sorts [0, n) of v with `threads` chunk sorts and pairwise merges.
template <typename V>
void parallel_sort(std::vector<V>& v, int threads) {
  const std::size_t n = v.size();
  const std::size_t parts = std::max<std::size_t>(1, std::min<std::size_t>(threads, n / 4096 + 1));
  std::vector<std::size_t> bounds(parts + 1);
  for (std::size_t i = 0; i <= parts; ++i) bounds[i] = n * i / parts;
  {
    std::vector<std::thread> pool;
    for (std::size_t i = 0; i < parts; ++i) {
      pool.emplace_back([&, i] { std::sort(v.begin() + bounds[i], v.begin() + bounds[i + 1]); });
    }
    for (auto& t : pool) t.join();
  }
  std::vector<V> tmp(n);
  for (std::size_t width = 1; width < parts; width *= 2) {
    std::vector<std::thread> pool;
    for (std::size_t i = 0; i < parts; i += 2 * width) {
      const std::size_t lo = bounds[i], mid = bounds[std::min(parts, i + width)],
                        hi = bounds[std::min(parts, i + 2 * width)];
      pool.emplace_back([&, lo, mid, hi] {
        std::merge(v.begin() + lo, v.begin() + mid, v.begin() + mid, v.begin() + hi, tmp.begin() + lo);
      });
    }
    for (auto& t : pool) t.join();
    v.swap(tmp);
  }
}

// Term indices in canonical key order, equal keys in construction order (a
// stable sort by key), and run_start[i] = 1 where order[i] starts a new key.
struct Canonical {
  std::vector<std::size_t> order;
  std::vector<char> run_start;
};

template <int W>
Canonical canonical_order_w(const Op& op) {
  const std::size_t T = op.T;
  std::vector<KeyRec<W>> recs(T);
  const std::size_t blocks = (T + 65535) / 65536;
  parallel_for(blocks, g_threads, [&](std::size_t b) {
    for (std::size_t t = b * 65536; t < std::min(T, (b + 1) * 65536); ++t) {
      for (int w = 0; w < W; ++w) {
        recs[t].k[2 * w] = op.x[t * W + w];
        recs[t].k[2 * w + 1] = op.z[t * W + w];
      }
      recs[t].idx = t;
    }
  });
  parallel_sort(recs, g_threads);
  Canonical out;
  out.order.resize(T);
  out.run_start.resize(T);
  parallel_for(blocks, g_threads, [&](std::size_t b) {
    for (std::size_t i = b * 65536; i < std::min(T, (b + 1) * 65536); ++i) {
      out.order[i] = recs[i].idx;
      bool start = i == 0;
      for (int k = 0; k < 2 * W && !start; ++k) start = recs[i].k[k] != recs[i - 1].k[k];
      out.run_start[i] = start ? 1 : 0;
    }
  });
  return out;
}

Canonical canonical_order(const Op& op) {
  switch (op.W) {
    case 1: return canonical_order_w<1>(op);
    case 2: return canonical_order_w<2>(op);
    case 3: return canonical_order_w<3>(op);
    case 4: return canonical_order_w<4>(op);
    default: break;
  }
  Canonical out;
  out.order.resize(op.T);
  std::iota(out.order.begin(), out.order.end(), std::size_t{0});
  std::stable_sort(out.order.begin(), out.order.end(), [&](std::size_t a, std::size_t b) {
    return key_cmp(op, a, b) < 0;
  });
  out.run_start.resize(op.T);
  for (std::size_t i = 0; i < op.T; ++i) {
    out.run_start[i] = (i == 0 || key_cmp(op, out.order[i - 1], out.order[i]) != 0) ? 1 : 0;
  }
  return out;
}

// Sums each run of equal keys of `src` (canonical order `can`) in
// construction order, the accumulator seeded with the run's first
// coefficient (seed_first) or with 0.0, and keeps the sums with
// |sum| > threshold. Runs are split over threads at run starts and the pieces
// concatenated in order, so the output does not depend on the thread count.
Op reduce_runs(const Op& src, const Canonical& can, bool seed_first, double threshold) {
  const std::size_t T = src.T, W = src.W;
  const std::size_t parts = std::max<std::size_t>(1, std::min<std::size_t>(g_threads * 4, T / 65536 + 1));
  std::vector<std::size_t> cut(parts + 1);
  for (std::size_t i = 0; i <= parts; ++i) {
    std::size_t c = T * i / parts;
    while (c < T && c > 0 && !can.run_start[c]) ++c;
    cut[i] = c;
  }
  cut[parts] = T;
  std::vector<Op> piece(parts);
  parallel_for(parts, g_threads, [&](std::size_t part) {
    Op& out = piece[part];
    std::size_t pos = cut[part];
    while (pos < cut[part + 1]) {
      std::size_t end = pos + 1;
      while (end < T && !can.run_start[end]) ++end;
      cplx acc = seed_first ? src.c[can.order[pos]] : cplx(0.0, 0.0);
      for (std::size_t i = seed_first ? pos + 1 : pos; i < end; ++i) acc += src.c[can.order[i]];
      if (std::abs(acc) > threshold) {
        const std::size_t t = can.order[pos];
        for (std::size_t w = 0; w < W; ++w) {
          out.x.push_back(src.x[t * W + w]);
          out.z.push_back(src.z[t * W + w]);
        }
        out.c.push_back(acc);
      }
      pos = end;
    }
  });
  Op out;
  out.n = src.n;
  out.W = W;
  for (const Op& pc : piece) {
    out.x.insert(out.x.end(), pc.x.begin(), pc.x.end());
    out.z.insert(out.z.end(), pc.z.begin(), pc.z.end());
    out.c.insert(out.c.end(), pc.c.begin(), pc.c.end());
  }
  out.T = out.c.size();
  return out;
}

Op simplify_impl(const Op& op, double atol, double rtol) {
  if (atol < 0.0 || rtol < 0.0 || !std::isfinite(atol) || !std::isfinite(rtol)) {
    throw std::invalid_argument("simplify tolerances must be non-negative finite values");
  }
  Op out;
  out.n = op.n;
  out.W = op.W;
  if (op.T == 0) return out;

  double max_abs = 0.0;
  for (const cplx& v : op.c) max_abs = std::max(max_abs, std::abs(v));
  const double threshold = atol + rtol * max_abs;

  // Distinct keys in canonical order, each with the construction-ordered list
  // of its terms.
  const Canonical can = canonical_order(op);
  // Hash path: the accumulator is seeded with the first coefficient
  // (try_emplace) and later ones are added. Sort path: seeded with 0.0.
  const bool hash_path = simplify_uses_hash_path(op);
  return reduce_runs(op, can, hash_path, threshold);
}

// ---------------------------------------------------------------- matmul
// Reference src/multiply.cpp:73-402.

bool operand_duplicate_heavy(const Op& op) {
  // Reference src/multiply.cpp:29-69: first min(T, 512) terms, >= 128 terms,
  // duplicate rate >= 0.25.
  if (op.T < 128) return false;
  const std::size_t sample = std::min<std::size_t>(op.T, 512);
  if (sample < 128) return false;
  return sampled_duplicate_rate(op, sample) >= 0.25;
}

std::int64_t y_parity(const Op& op, std::size_t t) {
  std::int64_t y = 0;
  for (std::size_t w = 0; w < op.W; ++w) {
    y += std::popcount(op.x[t * op.W + w] & op.z[t * op.W + w]);
  }
  return y;
}

Op matmul_impl(const Op& lhs, const Op& rhs, bool simplify_output, std::size_t max_terms) {
  if (lhs.n != rhs.n) throw std::invalid_argument("PauliSum matmul requires the same num_qubits");
  if (lhs.T != 0 && rhs.T > std::numeric_limits<std::size_t>::max() / lhs.T) {
    throw std::invalid_argument("matmul intermediate term count overflows size_t");
  }
  const std::size_t P = lhs.T * rhs.T;
  if (P > max_terms) {
    throw std::invalid_argument("matmul intermediate term count exceeds max_intermediate_terms");
  }
  Op prod;
  prod.n = lhs.n;
  prod.W = lhs.W;
  if (P == 0) return prod;

  const std::size_t W = lhs.W;
  std::vector<std::int64_t> ly(lhs.T), ry(rhs.T);
  for (std::size_t i = 0; i < lhs.T; ++i) ly[i] = y_parity(lhs, i);
  for (std::size_t j = 0; j < rhs.T; ++j) ry[j] = y_parity(rhs, j);

  prod.T = P;
  prod.x.resize(P * W);
  prod.z.resize(P * W);
  prod.c.resize(P);
  parallel_for(lhs.T, g_threads, [&](std::size_t i) {
    std::size_t p = i * rhs.T;
    for (std::size_t j = 0; j < rhs.T; ++j, ++p) {
      // k = |out_x & out_z| - |lx & lz| - |rx & rz| + 2 |lx & rz|
      // (reference src/multiply.cpp:279-283, :329-334, :368-382).
      std::int64_t k = -ly[i] - ry[j];
      for (std::size_t w = 0; w < W; ++w) {
        const u64 lx = lhs.x[i * W + w], lz = lhs.z[i * W + w];
        const u64 rx = rhs.x[j * W + w], rz = rhs.z[j * W + w];
        const u64 ox = lx ^ rx, oz = lz ^ rz;
        prod.x[p * W + w] = ox;
        prod.z[p * W + w] = oz;
        k += std::popcount(ox & oz);
        k += 2 * static_cast<std::int64_t>(std::popcount(lx & rz));
      }
      prod.c[p] = apply_phase(lhs.c[i] * rhs.c[j], k);
    }
  });
  if (!simplify_output) return prod;

  // Reference src/multiply.cpp:107-261: for W in {1,2} with a duplicate-heavy
  // operand the products go through a hash accumulator (seeded by the first
  // product) with a fixed 1e-12 drop; otherwise simplify() of the materialized
  // products. Both accumulate in lhs-outer / rhs-inner order, so they differ only
  // in how an all-negative-zero component is seeded.
  const bool hash_path = (W == 1 || W == 2) &&
                         (operand_duplicate_heavy(lhs) || operand_duplicate_heavy(rhs));
  if (!hash_path) return simplify_impl(prod, 1.0e-12, 0.0);

  return reduce_runs(prod, canonical_order(prod), true, 1.0e-12);
}

// ---------------------------------------------------------------- commutation
// Reference src/simd/commute_kernels_scalar.cpp:29-90 and
// src/detail/commutation.hpp:10-46: entry (i, j) is 1 iff the parity of
// popcount((lx_i & rz_j) ^ (lz_i & rx_j)) summed over words is even.

void check_commutation_entries(std::size_t L, std::size_t R, std::size_t max_entries) {
  if (L != 0 && R > std::numeric_limits<std::size_t>::max() / L) {
    throw std::invalid_argument("commutation matrix entry count overflows size_t");
  }
  if (L * R > max_entries) {
    throw std::invalid_argument(
        "commutation matrix entry count exceeds max_commutation_matrix_entries");
  }
}

void commute_rows_scalar(const Op& l, const Op& r, std::uint8_t* out, std::size_t row_begin,
                         std::size_t row_end) {
  const std::size_t W = l.W;
  for (std::size_t i = row_begin; i < row_end; ++i) {
    for (std::size_t j = 0; j < r.T; ++j) {
      unsigned parity = 0;
      for (std::size_t w = 0; w < W; ++w) {
        const u64 conflicts = (l.x[i * W + w] & r.z[j * W + w]) ^ (l.z[i * W + w] & r.x[j * W + w]);
        parity ^= static_cast<unsigned>(std::popcount(conflicts)) & 1U;
      }
      out[i * r.T + j] = parity == 0 ? 1 : 0;
    }
  }
}

// AVX-512 VPOPCNTDQ baseline for W in {1, 2}: 8 rhs terms per vector
// (reference src/simd/commute_kernels_avx512.cpp:44-149).
__attribute__((target("avx512f,avx512bw,avx512vl,avx512vpopcntdq"))) void commute_rows_avx512(
    const Op& l, const Op& r, std::uint8_t* out, std::size_t row_begin, std::size_t row_end) {
  const std::size_t W = l.W;
  const std::size_t R = r.T;
  // Strided gathers of word w of 8 consecutive rhs terms.
  const long long sw = static_cast<long long>(W);
  const __m512i stride_idx = _mm512_set_epi64(7 * sw, 6 * sw, 5 * sw, 4 * sw, 3 * sw, 2 * sw, sw, 0);
  for (std::size_t i = row_begin; i < row_end; ++i) {
    std::size_t j = 0;
    for (; j + 8 <= R; j += 8) {
      __m512i counts = _mm512_setzero_si512();
      for (std::size_t w = 0; w < W; ++w) {
        const __m512i lx = _mm512_set1_epi64(static_cast<long long>(l.x[i * W + w]));
        const __m512i lz = _mm512_set1_epi64(static_cast<long long>(l.z[i * W + w]));
        const __m512i idx = _mm512_add_epi64(
            stride_idx, _mm512_set1_epi64(static_cast<long long>(j * W + w)));
        const __m512i rz = _mm512_i64gather_epi64(idx, r.z.data(), 8);
        const __m512i rx = _mm512_i64gather_epi64(idx, r.x.data(), 8);
        const __m512i conflicts =
            _mm512_xor_si512(_mm512_and_si512(lx, rz), _mm512_and_si512(lz, rx));
        counts = _mm512_add_epi64(counts, _mm512_popcnt_epi64(conflicts));
      }
      const __mmask8 odd = _mm512_test_epi64_mask(counts, _mm512_set1_epi64(1));
      for (int b = 0; b < 8; ++b) out[i * R + j + b] = ((odd >> b) & 1U) ? 0 : 1;
    }
    for (; j < R; ++j) {
      unsigned parity = 0;
      for (std::size_t w = 0; w < W; ++w) {
        const u64 conflicts = (l.x[i * W + w] & r.z[j * W + w]) ^ (l.z[i * W + w] & r.x[j * W + w]);
        parity ^= static_cast<unsigned>(std::popcount(conflicts)) & 1U;
      }
      out[i * R + j] = parity == 0 ? 1 : 0;
    }
  }
}

bool cpu_has_avx512_popcnt() {
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
         __builtin_cpu_supports("avx512vl") && __builtin_cpu_supports("avx512vpopcntdq");
}

// ---------------------------------------------------------------- expectation
// Reference src/expectation.cpp:37-155, restated term by term, in the same
// order and with the same roundings (no FMA).

cplx expectation_reference_order(const Op& op, const cplx* psi, std::size_t len) {
  if (op.n > 63) throw std::invalid_argument("expectation_statevector requires num_qubits <= 63");
  if (len != (std::size_t{1} << op.n)) {
    throw std::invalid_argument("expectation_statevector requires len(psi) == 2 ** num_qubits");
  }
  if (op.T == 0) return {0.0, 0.0};
  const std::size_t N = len;
  auto xw = [&](std::size_t t) -> u64 { return op.W == 0 ? 0 : op.x[t * op.W]; };
  auto zw = [&](std::size_t t) -> u64 { return op.W == 0 ? 0 : op.z[t * op.W]; };

  bool all_diag = true;
  for (std::size_t t = 0; t < op.T && all_diag; ++t) all_diag = xw(t) == 0;

  if (all_diag) {
    std::vector<double> prob(N);
    for (std::size_t b = 0; b < N; ++b) prob[b] = std::norm(psi[b]);
    auto zsum = [&](u64 zmask) {
      double s = 0.0;
      for (std::size_t b = 0; b < N; ++b) {
        const bool odd = (std::popcount(zmask & static_cast<u64>(b)) & 1U) != 0;
        s += odd ? -prob[b] : prob[b];
      }
      return s;
    };
    cplx result{0.0, 0.0};
    if (op.T >= 64) {
      // Coefficients pre-summed per unique z in first-appearance order
      // (reference src/expectation.cpp:66-99).
      std::unordered_map<u64, std::size_t> index;
      std::vector<u64> masks;
      std::vector<cplx> sums;
      for (std::size_t t = 0; t < op.T; ++t) {
        auto [it, inserted] = index.try_emplace(zw(t), masks.size());
        if (inserted) {
          masks.push_back(zw(t));
          sums.push_back(op.c[t]);
        } else {
          sums[it->second] += op.c[t];
        }
      }
      for (std::size_t m = 0; m < masks.size(); ++m) result += sums[m] * zsum(masks[m]);
      return result;
    }
    for (std::size_t t = 0; t < op.T; ++t) result += op.c[t] * zsum(zw(t));
    return result;
  }

  cplx result{0.0, 0.0};
  for (std::size_t t = 0; t < op.T; ++t) {
    const u64 xm = xw(t), zm = zw(t);
    const cplx yz = phase_value(std::popcount(xm & zm));
    double te = 0.0;
    if (xm == 0) {
      for (std::size_t b = 0; b < N; ++b) {
        const bool odd = (std::popcount(zm & static_cast<u64>(b)) & 1U) != 0;
        const double p = std::norm(psi[b]);
        te += odd ? -p : p;
      }
    } else {
      const u64 pivot = xm & (0 - xm);
      for (std::size_t b = 0; b < N; ++b) {
        if ((static_cast<u64>(b) & pivot) != 0) continue;
        const std::size_t tgt = b ^ static_cast<std::size_t>(xm);
        const bool odd = (std::popcount(zm & static_cast<u64>(b)) & 1U) != 0;
        const cplx ph = odd ? -yz : yz;
        const cplx amp = std::conj(psi[tgt]) * psi[b];
        te += 2.0 * (ph * amp).real();
      }
    }
    result += op.c[t] * te;
  }
  return result;
}

// Compensated (double-double) single-term value <psi|P|psi>, real for the
// Hermitian P = i^{|x&z|} X^x Z^z. Used to state errors at sizes where the
// plain sequential sum carries ~sqrt(N) eps of its own rounding.
struct DD {
  double hi = 0.0, lo = 0.0;
  void add(double v) {
    const double s = hi + v;
    const double bp = s - hi;
    const double err = (hi - (s - bp)) + (v - bp);
    hi = s;
    lo += err;
  }
  double value() const { return hi + lo; }
};

double term_value_compensated(u64 xm, u64 zm, const cplx* psi, std::size_t N) {
  DD acc;
  if (xm == 0) {
    for (std::size_t b = 0; b < N; ++b) {
      const double re = psi[b].real(), im = psi[b].imag();
      // Exact squares as two-products: re^2 = p + e (FMA-free Dekker split).
      const double p1 = re * re, p2 = im * im;
      const bool odd = (std::popcount(zm & static_cast<u64>(b)) & 1U) != 0;
      acc.add(odd ? -p1 : p1);
      acc.add(odd ? -p2 : p2);
    }
    return acc.value();
  }
  const std::int64_t y = std::popcount(xm & zm);
  const u64 pivot = xm & (0 - xm);
  for (std::size_t b = 0; b < N; ++b) {
    if ((static_cast<u64>(b) & pivot) != 0) continue;
    const cplx a = psi[b], t = psi[b ^ static_cast<std::size_t>(xm)];
    // conj(t) * a = (tr ar + ti ai) + i (tr ai - ti ar)
    const double re1 = t.real() * a.real(), re2 = t.imag() * a.imag();
    const double im1 = t.real() * a.imag(), im2 = -(t.imag() * a.real());
    const bool odd = (std::popcount(zm & static_cast<u64>(b)) & 1U) != 0;
    // Re(i^y * s * amp) picks +-re or +-im of amp.
    double sgn = odd ? -2.0 : 2.0;
    switch (y & 3) {
      case 0: acc.add(sgn * re1); acc.add(sgn * re2); break;
      case 1: acc.add(-sgn * im1); acc.add(-sgn * im2); break;
      case 2: acc.add(-sgn * re1); acc.add(-sgn * re2); break;
      default: acc.add(sgn * im1); acc.add(sgn * im2); break;
    }
  }
  return acc.value();
}


template <typename Fn>
int guarded(Fn fn) {
  try {
    fn();
    return 0;
  } catch (const std::invalid_argument& e) {
    g_error = e.what();
    return 1;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 2;
  }
}

struct Result {
  Op op;
};

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }

// Result handles for variable-size outputs.
std::size_t orc_result_terms(const void* h) { return static_cast<const Result*>(h)->op.T; }
std::size_t orc_result_qubits(const void* h) { return static_cast<const Result*>(h)->op.n; }
void orc_result_copy(const void* h, u64* x, u64* z, double* c) {
  const Op& op = static_cast<const Result*>(h)->op;
  if (!op.x.empty()) std::memcpy(x, op.x.data(), op.x.size() * sizeof(u64));
  if (!op.z.empty()) std::memcpy(z, op.z.data(), op.z.size() * sizeof(u64));
  for (std::size_t t = 0; t < op.T; ++t) {
    c[2 * t] = op.c[t].real();
    c[2 * t + 1] = op.c[t].imag();
  }
}
void orc_result_free(void* h) { delete static_cast<Result*>(h); }

// PauliSum::simplify (reference src/simplify.cpp:147-430).
int orc_simplify(std::size_t n, std::size_t T, const u64* x, const u64* z, const double* c,
                 double atol, double rtol, void** out) {
  return guarded([&] {
    auto* r = new Result{simplify_impl(make_op(n, T, x, z, c), atol, rtol)};
    *out = r;
  });
}

// Which accumulation seed simplify uses (1 = hash path seeded by the first
// coefficient, 0 = sort path seeded by 0.0); reference src/simplify.cpp:107-143.
int orc_simplify_hash_path(std::size_t n, std::size_t T, const u64* x, const u64* z) {
  std::vector<double> zeros(2 * T, 0.0);
  return simplify_uses_hash_path(make_op(n, T, x, z, zeros.data())) ? 1 : 0;
}

// PauliSum::matmul (reference src/multiply.cpp:73-402).
int orc_matmul(std::size_t n, std::size_t TL, const u64* lx, const u64* lz, const double* lc,
               std::size_t TR, const u64* rx, const u64* rz, const double* rc, int simplify_output,
               std::size_t max_terms, void** out) {
  return guarded([&] {
    const Op l = make_op(n, TL, lx, lz, lc);
    const Op r = make_op(n, TR, rx, rz, rc);
    *out = new Result{matmul_impl(l, r, simplify_output != 0, max_terms)};
  });
}

// PauliSum::commutes_with (reference src/commute.cpp:27-93). variant: 0 scalar,
// 1 AVX-512 VPOPCNTDQ (W <= 2), 2 row-partitioned threads (reference
// src/parallel/commute_kernels_tbb.cpp:15-71 partitions lhs rows the same way).
int orc_commutes_with(std::size_t n, std::size_t TL, const u64* lx, const u64* lz, std::size_t TR,
                      const u64* rx, const u64* rz, std::size_t max_entries, int variant,
                      int threads, std::uint8_t* out) {
  return guarded([&] {
    std::vector<double> zl(2 * TL, 0.0), zr(2 * TR, 0.0);
    const Op l = make_op(n, TL, lx, lz, zl.data());
    const Op r = make_op(n, TR, rx, rz, zr.data());
    check_commutation_entries(TL, TR, max_entries);
    if (variant == 1) {
      if (l.W != 1 && l.W != 2) {
        throw std::runtime_error(
            "WOLFGANG_CPU_BACKEND=avx512 supports commutation kernels only for packed widths of "
            "1 or 2 uint64 words; got " + std::to_string(l.W) + " words");
      }
      if (!cpu_has_avx512_popcnt()) throw std::runtime_error("avx512 backend is hardware_unavailable");
      commute_rows_avx512(l, r, out, 0, TL);
    } else if (variant == 2) {
      const std::size_t chunk = 16;
      const std::size_t blocks = (TL + chunk - 1) / chunk;
      const bool simd = (l.W == 1 || l.W == 2) && cpu_has_avx512_popcnt();
      parallel_for(blocks, threads, [&](std::size_t b) {
        const std::size_t lo = b * chunk, hi = std::min(TL, lo + chunk);
        if (simd) {
          commute_rows_avx512(l, r, out, lo, hi);
        } else {
          commute_rows_scalar(l, r, out, lo, hi);
        }
      });
    } else {
      commute_rows_scalar(l, r, out, 0, TL);
    }
  });
}

// Worker threads for simplify's sort and matmul's products (results do not
// depend on it).
void orc_set_threads(int threads) { g_threads = threads < 1 ? 1 : threads; }

int orc_has_avx512_popcnt(void) { return cpu_has_avx512_popcnt() ? 1 : 0; }

// expectation_statevector_complex128 (reference src/expectation.cpp:37-155,209-212).
int orc_expectation(std::size_t n, std::size_t T, const u64* x, const u64* z, const double* c,
                    const double* psi, std::size_t len, double* out2) {
  return guarded([&] {
    const Op op = make_op(n, T, x, z, c);
    const cplx r = expectation_reference_order(op, reinterpret_cast<const cplx*>(psi), len);
    out2[0] = r.real();
    out2[1] = r.imag();
  });
}

// expectation_statevector_complex64 (reference src/expectation.cpp:214-217):
// promote each amplitude to complex128, then the same arithmetic.
int orc_expectation_c64(std::size_t n, std::size_t T, const u64* x, const u64* z, const double* c,
                        const float* psi, std::size_t len, double* out2) {
  return guarded([&] {
    const Op op = make_op(n, T, x, z, c);
    std::vector<cplx> wide(len);
    for (std::size_t b = 0; b < len; ++b) wide[b] = cplx(psi[2 * b], psi[2 * b + 1]);
    const cplx r = expectation_reference_order(op, wide.data(), len);
    out2[0] = r.real();
    out2[1] = r.imag();
  });
}

// expectation_z_counts (reference src/expectation.cpp:157-247): diagonal
// terms only; counts validated (non-negative, finite, positive total) before
// the bitstrings are packed (character i is qubit n - 1 - i); then, term by
// term, the signed counts summed in row order, divided by the total and
// scaled by the coefficient, accumulated in term order. Row r's bitstring is
// chars[offsets[r] .. offsets[r + 1]).
int orc_expectation_z_counts(std::size_t n, std::size_t T, const u64* x, const u64* z,
                             const double* c, const char* chars, const std::size_t* offsets,
                             const double* counts, std::size_t R, double* out2) {
  return guarded([&] {
    const std::size_t W = n == 0 ? 0 : (n + 63) / 64;
    for (std::size_t i = 0; i < T * W; ++i) {
      if (x[i] != 0) throw std::invalid_argument("expectation_z_counts requires diagonal Pauli terms");
    }
    double total = 0.0;
    for (std::size_t r = 0; r < R; ++r) {
      if (counts[r] < 0.0) throw std::invalid_argument("Z-count values must be non-negative");
      if (!std::isfinite(counts[r])) throw std::invalid_argument("Z-count values must be finite");
      total += counts[r];
    }
    if (total <= 0.0) throw std::invalid_argument("Z-count total count must be positive");
    std::vector<u64> packed(R * W, 0);
    for (std::size_t r = 0; r < R; ++r) {
      const std::size_t len = offsets[r + 1] - offsets[r];
      if (len != n) throw std::invalid_argument("Z-count bitstring length must equal num_qubits");
      for (std::size_t i = 0; i < len; ++i) {
        const char bit = chars[offsets[r] + i];
        if (bit != '0' && bit != '1') {
          throw std::invalid_argument("Z-count bitstrings may contain only 0 or 1");
        }
        if (bit == '1') {
          const std::size_t q = n - 1 - i;
          packed[r * W + q / 64] |= u64{1} << (q % 64);
        }
      }
    }
    cplx result{0.0, 0.0};
    for (std::size_t t = 0; t < T; ++t) {
      double weighted = 0.0;
      for (std::size_t r = 0; r < R; ++r) {
        bool odd = false;
        for (std::size_t w = 0; w < W; ++w) odd ^= (std::popcount(z[t * W + w] & packed[r * W + w]) & 1) != 0;
        weighted += counts[r] * (odd ? -1.0 : 1.0);
      }
      result += cplx(c[2 * t], c[2 * t + 1]) * (weighted / total);
    }
    out2[0] = result.real();
    out2[1] = result.imag();
  });
}

// Per-term <psi|P_k|psi> (real) with compensated summation, terms spread over
// threads. The reference computes the same quantity as `term_expectation`
// (src/expectation.cpp:117-153); the adjoint needs it per term (dE/dc_k).
int orc_term_values(std::size_t n, std::size_t T, const u64* x, const u64* z, const double* psi,
                    std::size_t len, int threads, double* out) {
  return guarded([&] {
    if (n > 63) throw std::invalid_argument("expectation_statevector requires num_qubits <= 63");
    if (len != (std::size_t{1} << n)) {
      throw std::invalid_argument("expectation_statevector requires len(psi) == 2 ** num_qubits");
    }
    const std::size_t W = n == 0 ? 0 : (n + 63) / 64;
    const cplx* p = reinterpret_cast<const cplx*>(psi);
    // Distinct (x, z) keys are evaluated once.
    std::unordered_map<std::size_t, std::size_t> first;
    std::vector<std::size_t> uniq;
    std::vector<std::size_t> owner(T);
    struct PairHash {
      std::size_t operator()(const std::pair<u64, u64>& k) const noexcept {
        return k.first * 0x9E3779B97F4A7C15ULL ^ (k.second + 0x632BE59BD9B4E019ULL);
      }
    };
    std::unordered_map<std::pair<u64, u64>, std::size_t, PairHash> idx;
    for (std::size_t t = 0; t < T; ++t) {
      const u64 xm = W ? x[t * W] : 0, zm = W ? z[t * W] : 0;
      auto [it, ins] = idx.try_emplace({xm, zm}, uniq.size());
      if (ins) uniq.push_back(t);
      owner[t] = it->second;
    }
    std::vector<double> vals(uniq.size());
    parallel_for(uniq.size(), threads, [&](std::size_t u) {
      const std::size_t t = uniq[u];
      vals[u] = term_value_compensated(W ? x[t * W] : 0, W ? z[t * W] : 0, p, len);
    });
    for (std::size_t t = 0; t < T; ++t) out[t] = vals[owner[t]];
  });
}

// E = sum_k c_k <P_k> from compensated per-term values, summed in term order
// with a compensated complex accumulator.
int orc_expectation_compensated(std::size_t n, std::size_t T, const u64* x, const u64* z,
                                const double* c, const double* psi, std::size_t len, int threads,
                                double* out2) {
  std::vector<double> vals(T);
  const int rc = orc_term_values(n, T, x, z, psi, len, threads, vals.data());
  if (rc != 0) return rc;
  DD re, im;
  for (std::size_t t = 0; t < T; ++t) {
    re.add(c[2 * t] * vals[t]);
    im.add(c[2 * t + 1] * vals[t]);
  }
  out2[0] = re.value();
  out2[1] = im.value();
  return 0;
}

// sigma = H psi, with (P psi)[b ^ x] = i^{|x&z|} (-1)^{|z&b|} psi[b]. The
// gradient dE/dpsi* of E = <psi|H|psi>. Rows of sigma are spread over threads;
// each output amplitude sums its terms in term order.
int orc_apply(std::size_t n, std::size_t T, const u64* x, const u64* z, const double* c,
              const double* psi, std::size_t len, int threads, double* sigma) {
  return guarded([&] {
    if (n > 63) throw std::invalid_argument("expectation_statevector requires num_qubits <= 63");
    if (len != (std::size_t{1} << n)) {
      throw std::invalid_argument("expectation_statevector requires len(psi) == 2 ** num_qubits");
    }
    const std::size_t W = n == 0 ? 0 : (n + 63) / 64;
    const cplx* p = reinterpret_cast<const cplx*>(psi);
    cplx* s = reinterpret_cast<cplx*>(sigma);
    const std::size_t chunk = 1024;
    const std::size_t blocks = (len + chunk - 1) / chunk;
    parallel_for(blocks, threads, [&](std::size_t blk) {
      for (std::size_t out_b = blk * chunk; out_b < std::min(len, (blk + 1) * chunk); ++out_b) {
        cplx acc{0.0, 0.0};
        for (std::size_t t = 0; t < T; ++t) {
          const u64 xm = W ? x[t * W] : 0, zm = W ? z[t * W] : 0;
          const std::size_t b = out_b ^ static_cast<std::size_t>(xm);
          const bool odd = (std::popcount(zm & static_cast<u64>(b)) & 1U) != 0;
          cplx v = apply_phase(p[b], std::popcount(xm & zm));
          if (odd) v = -v;
          acc += cplx(c[2 * t], c[2 * t + 1]) * v;
        }
        s[out_b] = acc;
      }
    });
  });
}

// PauliSum::group_commuting (reference src/grouping.cpp:36-62 ordering,
// :64-85 qubit-wise acceptance, :197-290 greedy placement): terms are visited
// by descending weight |x | z|, then canonical key order, then index; each
// goes to the first existing group that accepts it, else opens a new group.
// mode 0 ("qwc"): accepted when (x_t & Zb) ^ (z_t & Xb) = 0 on every word,
// with Xb/Zb the OR of the group's masks; mode 1 ("full"): accepted when the
// term commutes with every member. Output: group sizes (in group order) and
// the member indices, groups concatenated, members in placement order.
int orc_group_commuting(std::size_t n, std::size_t T, const u64* x, const u64* z, int mode,
                        std::size_t* num_groups, unsigned* sizes, unsigned* members) {
  return guarded([&] {
    const std::size_t W = n == 0 ? 0 : (n + 63) / 64;
    std::vector<std::size_t> order(T);
    std::vector<std::size_t> weight(T, 0);
    for (std::size_t t = 0; t < T; ++t) {
      order[t] = t;
      for (std::size_t w = 0; w < W; ++w) weight[t] += std::popcount(x[t * W + w] | z[t * W + w]);
    }
    auto key_less = [&](std::size_t a, std::size_t b) {
      for (std::size_t w = 0; w < W; ++w) {
        if (x[a * W + w] != x[b * W + w]) return x[a * W + w] < x[b * W + w] ? -1 : 1;
        if (z[a * W + w] != z[b * W + w]) return z[a * W + w] < z[b * W + w] ? -1 : 1;
      }
      return 0;
    };
    std::sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) {
      if (weight[a] != weight[b]) return weight[a] > weight[b];
      const int k = key_less(a, b);
      if (k != 0) return k < 0;
      return a < b;
    });
    auto commute = [&](std::size_t a, std::size_t b) {
      unsigned parity = 0;
      for (std::size_t w = 0; w < W; ++w) {
        parity ^= static_cast<unsigned>(
                      std::popcount((x[a * W + w] & z[b * W + w]) ^ (z[a * W + w] & x[b * W + w]))) & 1U;
      }
      return parity == 0;
    };
    std::vector<std::vector<std::size_t>> groups;
    std::vector<std::vector<u64>> xb, zb;
    for (std::size_t t : order) {
      bool placed = false;
      for (std::size_t g = 0; g < groups.size() && !placed; ++g) {
        bool ok = true;
        if (mode == 0) {
          for (std::size_t w = 0; w < W && ok; ++w) {
            ok = ((x[t * W + w] & zb[g][w]) ^ (z[t * W + w] & xb[g][w])) == 0;
          }
        } else {
          for (std::size_t m : groups[g]) {
            if (!commute(t, m)) {
              ok = false;
              break;
            }
          }
        }
        if (ok) {
          groups[g].push_back(t);
          if (mode == 0) {
            for (std::size_t w = 0; w < W; ++w) {
              xb[g][w] |= x[t * W + w];
              zb[g][w] |= z[t * W + w];
            }
          }
          placed = true;
        }
      }
      if (!placed) {
        groups.push_back({t});
        xb.emplace_back(W, 0);
        zb.emplace_back(W, 0);
        for (std::size_t w = 0; w < W; ++w) {
          xb.back()[w] = x[t * W + w];
          zb.back()[w] = z[t * W + w];
        }
      }
    }
    *num_groups = groups.size();
    std::size_t k = 0;
    for (std::size_t g = 0; g < groups.size(); ++g) {
      sizes[g] = static_cast<unsigned>(groups[g].size());
      for (std::size_t m : groups[g]) members[k++] = static_cast<unsigned>(m);
    }
  });
}

}  // extern "C"

namespace {

// One x-group's pairs inside a chunk pair (see orc_key_values_accumulate):
// representatives only (pivot bit clear), signs applied by flipping the sign
// bit (no data-dependent branches), flushed every 256 representatives.
template <typename Flush>
__attribute__((target("popcnt"))) void sweep_group(const cplx* own, const cplx* other, u64 xl,
                                                   u64 pivot, u64 mask, const u64* zl,
                                                   const int* odd, std::size_t nk, double* blk,
                                                   Flush& flush) {
  const u64 low = pivot <= mask ? pivot - 1 : mask;
  const u64 reps = pivot <= mask ? (mask + 1) >> 1 : mask + 1;
  for (u64 r0 = 0; r0 < reps; r0 += 256) {
    const u64 r1 = std::min<u64>(reps, r0 + 256);
    for (u64 r = r0; r < r1; ++r) {
      const u64 u = pivot <= mask ? (((r & ~low) << 1) | (r & low)) : r;
      const cplx av = own[u], tv = other[u ^ xl];
      const double re = tv.real() * av.real() + tv.imag() * av.imag();
      const double im = tv.real() * av.imag() - tv.imag() * av.real();
      for (std::size_t j = 0; j < nk; ++j) {
        double v = odd[j] ? im : re;
        u64 bits;
        std::memcpy(&bits, &v, 8);
        bits ^= static_cast<u64>(__builtin_popcountll(zl[j] & u) & 1) << 63;
        std::memcpy(&v, &bits, 8);
        blk[j] += v;
      }
    }
    flush();
  }
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- checks at full size
// Packed commutation bits (row-major, row_words = ceil(TR/64), bit j%64 of
// word j/64 = entry (i, j)) from the row-threaded kernel above, so that a
// 65,536^2 matrix is checked in 512 MiB of host memory instead of 4 GiB.
int orc_commutes_packed(std::size_t n, std::size_t TL, const u64* lx, const u64* lz,
                        std::size_t TR, const u64* rx, const u64* rz, int threads, u64* out) {
  return guarded([&] {
    std::vector<double> zl(2 * TL, 0.0), zr(2 * TR, 0.0);
    const Op l = make_op(n, TL, lx, lz, zl.data());
    const Op r = make_op(n, TR, rx, rz, zr.data());
    const std::size_t row_words = (TR + 63) / 64;
    const std::size_t chunk = 16;
    const std::size_t blocks = (TL + chunk - 1) / chunk;
    const bool simd = (l.W == 1 || l.W == 2) && cpu_has_avx512_popcnt();
    parallel_for(blocks, threads, [&](std::size_t b) {
      const std::size_t lo = b * chunk, hi = std::min(TL, lo + chunk);
      Op sub;
      sub.n = l.n;
      sub.W = l.W;
      sub.T = hi - lo;
      sub.x.assign(l.x.begin() + lo * l.W, l.x.begin() + hi * l.W);
      sub.z.assign(l.z.begin() + lo * l.W, l.z.begin() + hi * l.W);
      std::vector<std::uint8_t> rows(sub.T * TR);
      if (simd) {
        commute_rows_avx512(sub, r, rows.data(), 0, sub.T);
      } else {
        commute_rows_scalar(sub, r, rows.data(), 0, sub.T);
      }
      for (std::size_t i = 0; i < sub.T; ++i) {
        u64* dst = out + (lo + i) * row_words;
        std::fill(dst, dst + row_words, 0ULL);
        for (std::size_t j = 0; j < TR; ++j) dst[j / 64] |= static_cast<u64>(rows[i * TR + j]) << (j % 64);
      }
    });
  });
}

// sigma = H psi on a subset of output rows only (as orc_apply, same term order
// per row): out[i] = sigma[rows[i]].
int orc_apply_rows(std::size_t n, std::size_t T, const u64* x, const u64* z, const double* c,
                   const double* psi, std::size_t len, const u64* rows, std::size_t nrows,
                   int threads, double* out) {
  return guarded([&] {
    if (n > 63) throw std::invalid_argument("expectation_statevector requires num_qubits <= 63");
    if (len != (std::size_t{1} << n)) {
      throw std::invalid_argument("expectation_statevector requires len(psi) == 2 ** num_qubits");
    }
    const std::size_t W = n == 0 ? 0 : (n + 63) / 64;
    const cplx* p = reinterpret_cast<const cplx*>(psi);
    cplx* s = reinterpret_cast<cplx*>(out);
    parallel_for(nrows, threads, [&](std::size_t i) {
      const std::size_t out_b = rows[i];
      if (out_b >= len) throw std::invalid_argument("row index out of range");
      cplx acc{0.0, 0.0};
      for (std::size_t t = 0; t < T; ++t) {
        const u64 xm = W ? x[t * W] : 0, zm = W ? z[t * W] : 0;
        const std::size_t b = out_b ^ static_cast<std::size_t>(xm);
        const bool odd = (std::popcount(zm & static_cast<u64>(b)) & 1U) != 0;
        cplx v = apply_phase(p[b], std::popcount(xm & zm));
        if (odd) v = -v;
        acc += cplx(c[2 * t], c[2 * t + 1]) * v;
      }
      s[i] = acc;
    });
  });
}

// Distinct (x, z) keys of a one-word operator in first-appearance order:
// key_of_term[t] = key id; returns the key count in *num_keys.
int orc_distinct_keys(std::size_t n, std::size_t T, const u64* x, const u64* z,
                      unsigned* key_of_term, std::size_t* num_keys) {
  return guarded([&] {
    if (n > 64) throw std::invalid_argument("distinct keys are formed for one-word operators");
    const bool w = n != 0;
    struct PairHash {
      std::size_t operator()(const std::pair<u64, u64>& k) const noexcept {
        return k.first * 0x9E3779B97F4A7C15ULL ^ (k.second + 0x632BE59BD9B4E019ULL);
      }
    };
    std::unordered_map<std::pair<u64, u64>, unsigned, PairHash> idx;
    for (std::size_t t = 0; t < T; ++t) {
      auto [it, ins] = idx.try_emplace({w ? x[t] : 0, w ? z[t] : 0}, static_cast<unsigned>(idx.size()));
      key_of_term[t] = it->second;
    }
    *num_keys = idx.size();
  });
}

// Per-key values <P_k> = <psi|P_k|psi> of a one-word operator, accumulated
// over the chunks [chunk_begin, chunk_end) of 2^C amplitudes each
// (C = min(n, 16)) into caller-held double-double sums (key_hi, key_lo; keys
// as numbered by orc_distinct_keys), so a caller can report progress between
// calls. The evaluation is organised differently from the term-by-term
// oracle above, so the two check each other:
//  * diagonal keys (x = 0): per chunk beta, p(u) = |psi[beta 2^C + u]|^2 and its
//    Walsh-Hadamard transform P(v) (pairwise butterflies); key z gains
//    (-1)^{|(z >> C) & beta|} P(z mod 2^C);
//  * off-diagonal keys: grouped by x, and the groups by h = x >> C; per chunk
//    beta and h, one sweep of the pairs (b, b ^ x) whose representative b (pivot
//    bit of x clear) lies in beta, reading the partner chunk beta ^ h, serves every
//    key of every group with that h; per key the signed products
//    (-1)^{|z & b|} conj(psi[b ^ x]) psi[b] are summed in blocks of 256 in double
//    and the block sums enter the double-double accumulator
//    (value = 2 Re(i^{|x & z|} S), reference src/expectation.cpp:137-152).
__attribute__((target("popcnt"))) int orc_key_values_accumulate(
    std::size_t n, std::size_t T, const u64* x, const u64* z, const double* psi, std::size_t len,
    std::size_t chunk_begin, std::size_t chunk_end, int threads, double* key_hi, double* key_lo) {
  return guarded([&] {
    if (n > 40) throw std::invalid_argument("full-size key values support num_qubits <= 40");
    if (len != (std::size_t{1} << n)) {
      throw std::invalid_argument("expectation_statevector requires len(psi) == 2 ** num_qubits");
    }
    const bool w = n != 0;
    const int C = static_cast<int>(std::min<std::size_t>(n, 16));
    const u64 mask = (u64{1} << C) - 1;
    const std::size_t chunks = len >> C;
    if (chunk_end > chunks || chunk_begin > chunk_end) throw std::invalid_argument("chunk range out of bounds");
    const cplx* p = reinterpret_cast<const cplx*>(psi);
    // Keys in first-appearance order (as orc_distinct_keys numbers them).
    struct PairHash {
      std::size_t operator()(const std::pair<u64, u64>& k) const noexcept {
        return k.first * 0x9E3779B97F4A7C15ULL ^ (k.second + 0x632BE59BD9B4E019ULL);
      }
    };
    std::unordered_map<std::pair<u64, u64>, unsigned, PairHash> idx;
    std::vector<u64> kx, kz;
    for (std::size_t t = 0; t < T; ++t) {
      const u64 xm = w ? x[t] : 0, zm = w ? z[t] : 0;
      if (idx.try_emplace({xm, zm}, static_cast<unsigned>(kx.size())).second) {
        kx.push_back(xm);
        kz.push_back(zm);
      }
    }
    const std::size_t K = kx.size();
    std::vector<unsigned> diag;
    std::unordered_map<u64, std::vector<unsigned>> by_x;
    std::vector<u64> xs;
    for (unsigned k = 0; k < K; ++k) {
      if (kx[k] == 0) {
        diag.push_back(k);
      } else {
        auto [it, ins] = by_x.try_emplace(kx[k]);
        if (ins) xs.push_back(kx[k]);
        it->second.push_back(k);
      }
    }
    // Off-diagonal x-groups bucketed by their high part h, with the per-key
    // constants of the sweep precomputed.
    struct Group {
      u64 xl, pivot;
      std::vector<unsigned> keys;
      std::vector<u64> zl, zh;
      std::vector<int> odd;
      std::vector<double> f;  // 2 Re(i^y S) = f * S_re (y even) or f * S_im (y odd)
    };
    std::unordered_map<u64, std::size_t> bucket_of;
    std::vector<u64> hs;
    std::vector<std::vector<Group>> buckets;
    for (u64 xm : xs) {
      auto [it, ins] = bucket_of.try_emplace(xm >> C, buckets.size());
      if (ins) {
        hs.push_back(xm >> C);
        buckets.emplace_back();
      }
      Group g;
      g.xl = xm & mask;
      g.pivot = xm & (0 - xm);
      for (unsigned k : by_x[xm]) {
        const int y = std::popcount(xm & kz[k]) & 3;
        g.keys.push_back(k);
        g.zl.push_back(kz[k] & mask);
        g.zh.push_back(kz[k] >> C);
        g.odd.push_back(y & 1);
        g.f.push_back((y == 0 || y == 3) ? 2.0 : -2.0);
      }
      buckets[it->second].push_back(std::move(g));
    }
    const int nt = std::max(1, threads);
    std::vector<std::vector<DD>> acc(nt, std::vector<DD>(K));
    std::atomic<std::size_t> next{chunk_begin};
    auto worker = [&](int tid) {
      std::vector<DD>& a = acc[tid];
      std::vector<double> pw(std::size_t{1} << C);
      std::vector<double> blk;
      for (std::size_t beta = next.fetch_add(1); beta < chunk_end; beta = next.fetch_add(1)) {
        const cplx* own = p + (beta << C);
        if (!diag.empty()) {
          for (std::size_t u = 0; u <= mask; ++u) {
            pw[u] = own[u].real() * own[u].real() + own[u].imag() * own[u].imag();
          }
          for (std::size_t h = 1; h <= mask; h <<= 1) {
            for (std::size_t i = 0; i <= mask; i += 2 * h) {
              for (std::size_t j = i; j < i + h; ++j) {
                const double a0 = pw[j], a1 = pw[j + h];
                pw[j] = a0 + a1;
                pw[j + h] = a0 - a1;
              }
            }
          }
          for (unsigned k : diag) {
            const double v = pw[kz[k] & mask];
            a[k].add((std::popcount((kz[k] >> C) & beta) & 1) ? -v : v);
          }
        }
        for (std::size_t hb = 0; hb < hs.size(); ++hb) {
          const cplx* other = p + ((beta ^ hs[hb]) << C);
          for (const Group& g : buckets[hb]) {
            const std::size_t nk = g.keys.size();
            // The pivot bit lies in the chunk index: this chunk holds every
            // representative of the group's pairs or none.
            if (g.pivot > mask && ((static_cast<u64>(beta) << C) & g.pivot)) continue;
            blk.assign(nk, 0.0);
            auto flush = [&]() {
              for (std::size_t j = 0; j < nk; ++j) {
                const double cs = (std::popcount(g.zh[j] & beta) & 1) ? -g.f[j] : g.f[j];
                a[g.keys[j]].add(cs * blk[j]);
                blk[j] = 0.0;
              }
            };
            sweep_group(own, other, g.xl, g.pivot, mask, g.zl.data(), g.odd.data(), nk, blk.data(),
                        flush);
          }
        }
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto& th : pool) th.join();
    for (std::size_t k = 0; k < K; ++k) {
      DD tot;
      tot.hi = key_hi[k];
      tot.lo = key_lo[k];
      for (int t = 0; t < nt; ++t) {
        tot.add(acc[t][k].hi);
        tot.add(acc[t][k].lo);
      }
      key_hi[k] = tot.hi;
      key_lo[k] = tot.lo;
    }
  });
}

}  // extern "C"


// extern "C" shim over the reference's own scalar C++ core, compiled from the
// reference sources in place (see oracle/Makefile, target `ref`). It exists only
// to validate the oracle restatement (wolfgang_oracle.cpp) and to time the
// reference's CPU path as a reported baseline. TEST INFRASTRUCTURE ONLY: never
// linked into the product library.
//
// Operators cross the boundary as packed masks, rebuilt with
// PauliSum::from_sparse_list (reference src/parse.cpp:56-97) so that only public
// reference API is used.

#include "wolfgang/pauli_sum.hpp"

#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <exception>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

wolfgang::PauliSum build(std::size_t n, std::size_t T, const std::uint64_t* x,
                         const std::uint64_t* z, const double* c) {
  const std::size_t W = n == 0 ? 0 : (n + 63) / 64;
  std::vector<wolfgang::PauliSum::SparseTerm> terms(T);
  for (std::size_t t = 0; t < T; ++t) {
    auto& term = terms[t];
    for (std::size_t q = 0; q < n; ++q) {
      const bool xb = (x[t * W + q / 64] >> (q % 64)) & 1U;
      const bool zb = (z[t * W + q / 64] >> (q % 64)) & 1U;
      if (!xb && !zb) continue;
      term.local_pauli_string.push_back(xb && zb ? 'Y' : (xb ? 'X' : 'Z'));
      term.qubit_indices.push_back(q);
    }
    term.coefficient = {c ? c[2 * t] : 1.0, c ? c[2 * t + 1] : 0.0};
  }
  return wolfgang::PauliSum::from_sparse_list(terms, n);
}

struct Result {
  wolfgang::PauliSum op;
};

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

}  // namespace

extern "C" {

const char* wgref_last_error(void) { return g_error.c_str(); }

std::size_t wgref_result_terms(const void* h) { return static_cast<const Result*>(h)->op.num_terms(); }

void wgref_result_copy(const void* h, std::uint64_t* x, std::uint64_t* z, double* c) {
  const auto& op = static_cast<const Result*>(h)->op;
  const auto& xs = op.x_words();
  const auto& zs = op.z_words();
  const auto& cs = op.coeffs();
  if (!xs.empty()) std::memcpy(x, xs.data(), xs.size() * sizeof(std::uint64_t));
  if (!zs.empty()) std::memcpy(z, zs.data(), zs.size() * sizeof(std::uint64_t));
  for (std::size_t t = 0; t < cs.size(); ++t) {
    c[2 * t] = cs[t].real();
    c[2 * t + 1] = cs[t].imag();
  }
}

void wgref_result_free(void* h) { delete static_cast<Result*>(h); }

int wgref_simplify(std::size_t n, std::size_t T, const std::uint64_t* x, const std::uint64_t* z,
                   const double* c, double atol, double rtol, void** out) {
  return guarded([&] { *out = new Result{build(n, T, x, z, c).simplify(atol, rtol)}; });
}

int wgref_matmul(std::size_t n, std::size_t TL, const std::uint64_t* lx, const std::uint64_t* lz,
                 const double* lc, std::size_t TR, const std::uint64_t* rx,
                 const std::uint64_t* rz, const double* rc, int simplify_output,
                 std::size_t max_terms, void** out) {
  return guarded([&] {
    const auto l = build(n, TL, lx, lz, lc);
    const auto r = build(n, TR, rx, rz, rc);
    *out = new Result{l.matmul(r, simplify_output != 0, max_terms)};
  });
}

// Runs whatever WOLFGANG_CPU_BACKEND selects (reference src/commute.cpp:27-93).
int wgref_commutes_with(std::size_t n, std::size_t TL, const std::uint64_t* lx,
                        const std::uint64_t* lz, std::size_t TR, const std::uint64_t* rx,
                        const std::uint64_t* rz, std::size_t max_entries, std::uint8_t* out) {
  return guarded([&] {
    const auto l = build(n, TL, lx, lz, nullptr);
    const auto r = build(n, TR, rx, rz, nullptr);
    const auto flags = l.commutes_with(r, max_entries);
    if (!flags.empty()) std::memcpy(out, flags.data(), flags.size());
  });
}

int wgref_expectation(std::size_t n, std::size_t T, const std::uint64_t* x,
                      const std::uint64_t* z, const double* c, const double* psi,
                      std::size_t len, double* out2) {
  return guarded([&] {
    const auto op = build(n, T, x, z, c);
    const auto r = op.expectation_statevector_complex128(
        std::span<const std::complex<double>>(reinterpret_cast<const std::complex<double>*>(psi), len));
    out2[0] = r.real();
    out2[1] = r.imag();
  });
}

int wgref_expectation_z_counts(std::size_t n, std::size_t T, const std::uint64_t* x,
                               const std::uint64_t* z, const double* c, const char* chars,
                               const std::size_t* offsets, const double* counts, std::size_t R,
                               double* out2) {
  return guarded([&] {
    const auto op = build(n, T, x, z, c);
    std::vector<std::string> bitstrings;
    std::vector<double> values(counts, counts + R);
    for (std::size_t r = 0; r < R; ++r) {
      bitstrings.emplace_back(chars + offsets[r], offsets[r + 1] - offsets[r]);
    }
    const auto e = op.expectation_z_counts(bitstrings, values);
    out2[0] = e.real();
    out2[1] = e.imag();
  });
}

int wgref_expectation_c64(std::size_t n, std::size_t T, const std::uint64_t* x,
                          const std::uint64_t* z, const double* c, const float* psi,
                          std::size_t len, double* out2) {
  return guarded([&] {
    const auto op = build(n, T, x, z, c);
    const auto r = op.expectation_statevector_complex64(
        std::span<const std::complex<float>>(reinterpret_cast<const std::complex<float>*>(psi), len));
    out2[0] = r.real();
    out2[1] = r.imag();
  });
}

// PauliSum::group_commuting (reference src/grouping.cpp:197-290): the groups
// concatenated into (x, z, c) arrays of num_terms rows, with the group sizes.
int wgref_group_commuting(std::size_t n, std::size_t T, const std::uint64_t* x,
                          const std::uint64_t* z, const double* c, const char* mode,
                          std::size_t max_terms_for_graph, std::size_t* num_groups,
                          std::size_t* sizes, std::uint64_t* gx, std::uint64_t* gz, double* gc) {
  return guarded([&] {
    const auto op = build(n, T, x, z, c);
    const auto groups = op.group_commuting(mode, "largest_first", max_terms_for_graph);
    *num_groups = groups.size();
    std::size_t row = 0;
    const std::size_t W = op.words();
    for (std::size_t g = 0; g < groups.size(); ++g) {
      const auto& grp = groups[g];
      sizes[g] = grp.num_terms();
      if (W) {
        std::memcpy(gx + row * W, grp.x_words().data(), grp.num_terms() * W * sizeof(std::uint64_t));
        std::memcpy(gz + row * W, grp.z_words().data(), grp.num_terms() * W * sizeof(std::uint64_t));
      }
      for (std::size_t t = 0; t < grp.num_terms(); ++t) {
        gc[2 * (row + t)] = grp.coeffs()[t].real();
        gc[2 * (row + t) + 1] = grp.coeffs()[t].imag();
      }
      row += grp.num_terms();
    }
  });
}

}  // extern "C"

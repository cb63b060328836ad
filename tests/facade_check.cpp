// Drop-in check of the C++ facade (wolfgang_amd/facade/wolfgang_hip_amd.cpp):
// the reference's own C++ API (include/wolfgang/*.hpp) driven exactly as its
// users drive it, with the device members answered by libwolfgang_amd.so and
// every result compared with the reference's own CPU PauliSum members (its
// scalar sources compiled in place into oracle/_ref/libwolfgang_ref.so), which
// act as the checker here.
//
//   facade_check nodevice   no GPU visible: the status members and the errors
//   facade_check device     a GPU visible: every member against the CPU path
//
// Built by `make -C oracle facade` into oracle/_ref/facade_check. Prints one
// line per check group and "facade_check: OK" at the end; exits 1 on the first
// mismatch.

#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "wolfgang/device_commutation_matrix.hpp"
#include "wolfgang/device_pauli_sum.hpp"
#include "wolfgang/pauli_sum.hpp"

using wolfgang::AcceleratorBackend;
using wolfgang::DeviceCommutationMatrix;
using wolfgang::DevicePauliSum;
using wolfgang::PauliSum;

namespace {

int g_checks = 0;

void expect(bool ok, const std::string& what) {
  ++g_checks;
  if (!ok) {
    std::fprintf(stderr, "facade_check: FAILED: %s\n", what.c_str());
    std::exit(1);
  }
}

template <class Exc, class F>
void expect_throw(F&& f, const std::string& fragment, const std::string& what) {
  bool thrown = false, right_type = false;
  std::string msg;
  try {
    f();
  } catch (const std::exception& e) {
    thrown = true;
    if constexpr (std::is_same_v<Exc, std::exception>) {
      right_type = true;
    } else {
      right_type = dynamic_cast<const Exc*>(&e) != nullptr;
    }
    msg = e.what();
  }
  expect(thrown, what + ": no exception");
  expect(right_type, what + ": wrong exception type: " + msg);
  expect(fragment.empty() || msg.find(fragment) != std::string::npos,
         what + ": message '" + msg + "' lacks '" + fragment + "'");
}

// Random dense labels (right-most character = qubit 0), each qubit I with
// probability 1 - density, and complex coefficients.
PauliSum random_sum(std::size_t n, std::size_t terms, double density, std::uint64_t seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  std::vector<std::string> labels(terms, std::string(n, 'I'));
  std::vector<std::complex<double>> coeffs(terms);
  const char paulis[3] = {'X', 'Y', 'Z'};
  for (std::size_t t = 0; t < terms; ++t) {
    for (std::size_t q = 0; q < n; ++q) {
      if (u(rng) < density) labels[t][q] = paulis[rng() % 3];
    }
    coeffs[t] = {u(rng) * 2.0 - 1.0, u(rng) * 2.0 - 1.0};
  }
  return PauliSum::from_labels(labels, coeffs);
}

bool same_bits(const PauliSum& a, const PauliSum& b) {
  if (a.num_qubits() != b.num_qubits() || a.num_terms() != b.num_terms()) return false;
  if (a.x_words() != b.x_words() || a.z_words() != b.z_words()) return false;
  const auto& ca = a.coeffs();
  const auto& cb = b.coeffs();
  return ca.empty() || std::memcmp(ca.data(), cb.data(), ca.size() * sizeof(ca[0])) == 0;
}

double sum_abs(const PauliSum& p) {
  double s = 0.0;
  for (const auto& c : p.coeffs()) s += std::abs(c);
  return s;
}

void check_nodevice() {
  expect(!DevicePauliSum::hip_available(), "hip_available() is false without a device");
  const auto status = DevicePauliSum::hip_status();
  expect(status.built && !status.runtime_available && status.device_count == 0 &&
             !status.skip_reason.empty(),
         "hip_status() reports a built backend, no device and a skip reason");
  expect(!DevicePauliSum::cuda_available() && !DevicePauliSum::cuda_status().built,
         "CUDA is not built");
  expect(!DevicePauliSum::metal_available() && !DevicePauliSum::metal_status().built,
         "Metal is not built");
  const PauliSum h = random_sum(8, 10, 0.5, 1);
  expect_throw<std::runtime_error>([&] { (void)DevicePauliSum::from_host(h, 0); }, "",
                                   "from_host without a device");
  expect_throw<std::runtime_error>(
      [&] { (void)DevicePauliSum::from_host(h, AcceleratorBackend::Cuda, 0); },
      "built without CUDA support", "from_host(Cuda)");
  DevicePauliSum empty;
  expect(empty.num_terms() == 0 && empty.num_qubits() == 0 && empty.backend() == "hip",
         "default-constructed accessors");
  expect_throw<std::runtime_error>([&] { (void)empty.to_host(); },
                                   "DevicePauliSum is empty or moved-from", "to_host moved-from");
  std::printf("nodevice: status members and errors ok\n");
}

void check_algebra() {
  // 70 qubits: two packed words per mask.
  const PauliSum a = random_sum(70, 300, 0.3, 11);
  const PauliSum b = random_sum(70, 200, 0.3, 12);
  DevicePauliSum da = DevicePauliSum::from_host(a, 0);
  DevicePauliSum db = DevicePauliSum::from_host(b, AcceleratorBackend::Hip, 0);
  expect(da.num_qubits() == 70 && da.num_terms() == 300 && da.words() == 2 && da.device() == 0,
         "accessors");
  expect(same_bits(da.to_host(), a), "to_host round trip");

  expect(same_bits(da.matmul(db).to_host(), a.matmul(b)), "matmul + simplify, bit-exact");
  expect(same_bits(da.matmul(db, false).to_host(), a.matmul(b, false)),
         "matmul without simplify, bit-exact");
  const PauliSum dup = a.add(b).add(a.scalar_multiply({0.5, -0.25}));
  DevicePauliSum ddup = DevicePauliSum::from_host(dup, 0);
  expect(same_bits(ddup.simplify().to_host(), dup.simplify()), "simplify, bit-exact");
  expect(same_bits(ddup.simplify(1e-3, 0.1).to_host(), dup.simplify(1e-3, 0.1)),
         "simplify with atol/rtol, bit-exact");
  std::printf("algebra: to_host, matmul, simplify ok\n");

  const std::vector<std::uint8_t> want = a.commutes_with(b);
  expect(da.commutes_with(db) == want, "commutes_with bytes");
  std::vector<std::uint8_t> into(want.size(), 7);
  da.commutes_with_into(db, into);
  expect(into == want, "commutes_with_into bytes");
  DeviceCommutationMatrix m = da.commutes_with_device(db);
  expect(m.rows() == 300 && m.cols() == 200 && m.num_entries() == 60000 && m.device() == 0 &&
             m.backend() == "hip" && m.dlpack_device_type() == 10,
         "matrix accessors");
  expect(m.to_host() == want, "commutes_with_device to_host bytes");
  std::uint64_t total = 0;
  std::vector<std::uint64_t> rows(300, 0), cols(200, 0);
  for (std::size_t i = 0; i < 300; ++i) {
    for (std::size_t j = 0; j < 200; ++j) {
      const std::uint64_t v = want[i * 200 + j];
      total += v;
      rows[i] += v;
      cols[j] += v;
    }
  }
  expect(m.count_commuting() == total, "count_commuting");
  expect(m.count_commuting_rows() == rows, "count_commuting_rows");
  expect(m.count_commuting_cols() == cols, "count_commuting_cols");
  wolfgang::copy_device_commutation_matrix_from_host_for_testing(
      m, std::vector<std::uint8_t>(60000, 1));
  expect(m.count_commuting() == 60000, "testing copy fills the matrix");
  da.commutes_with_device_into(db, m);
  expect(m.to_host() == want, "commutes_with_device_into overwrites a poisoned matrix");
  expect_throw<std::runtime_error>([&] { (void)m.data_pointer_for_dlpack(); }, "packed bits",
                                   "byte data pointer of a packed matrix");
  std::printf("commutation: bytes, device matrix, counts ok\n");

  const PauliSum other = random_sum(71, 5, 0.3, 13);
  DevicePauliSum dother = DevicePauliSum::from_host(other, 0);
  expect_throw<std::invalid_argument>([&] { (void)da.matmul(dother); }, "", "matmul qubit mismatch");
  expect_throw<std::invalid_argument>([&] { (void)da.commutes_with(dother); }, "",
                                      "commutation qubit mismatch");
  expect_throw<std::exception>([&] { (void)da.matmul(db, true, 100); }, "max_intermediate_terms",
                               "matmul max_intermediate_terms");
  expect_throw<std::exception>([&] { (void)da.commutes_with(db, 100); },
                               "max_commutation_matrix_entries",
                               "commutes_with max_commutation_matrix_entries");
  DevicePauliSum moved = std::move(da);
  expect_throw<std::runtime_error>([&] { (void)da.matmul(db); },
                                   "DevicePauliSum is empty or moved-from", "moved-from matmul");
  expect(moved.num_terms() == 300, "moved-to object keeps the operator");
  std::printf("errors: qubit mismatch, limits, moved-from ok\n");
}

void check_expectation() {
  for (std::size_t n : {6, 14, 19}) {
    const PauliSum h = random_sum(n, 800, 0.3, 100 + n);
    DevicePauliSum dh = DevicePauliSum::from_host(h, 0);
    std::mt19937_64 rng(200 + n);
    std::normal_distribution<double> g(0.0, 1.0);
    std::vector<std::complex<double>> psi(std::size_t{1} << n);
    double norm = 0.0;
    for (auto& v : psi) {
      v = {g(rng), g(rng)};
      norm += std::norm(v);
    }
    for (auto& v : psi) v /= std::sqrt(norm);
    const double tol = 1e-12 * sum_abs(h);
    const std::complex<double> want = h.expectation_statevector_complex128(psi);
    expect(std::abs(dh.expectation_statevector_complex128(psi) - want) <= tol,
           "expectation complex128, n=" + std::to_string(n));

    std::vector<std::complex<float>> psi32(psi.size());
    for (std::size_t i = 0; i < psi.size(); ++i) {
      psi32[i] = {static_cast<float>(psi[i].real()), static_cast<float>(psi[i].imag())};
    }
    const std::complex<double> want32 = h.expectation_statevector_complex64(psi32);
    expect(std::abs(dh.expectation_statevector_complex64(psi32) - want32) <= tol,
           "expectation complex64, n=" + std::to_string(n));

    void* dev = nullptr;
    expect(hipMalloc(&dev, psi.size() * sizeof(psi[0])) == hipSuccess, "hipMalloc");
    expect(hipMemcpy(dev, psi.data(), psi.size() * sizeof(psi[0]), hipMemcpyHostToDevice) ==
               hipSuccess,
           "hipMemcpy");
    const std::complex<double> got_dev = dh.expectation_statevector_device_pointer(
        reinterpret_cast<std::uintptr_t>(dev), wolfgang::DeviceStatevectorDtype::Complex128,
        psi.size());
    expect(std::abs(got_dev - want) <= tol, "expectation device pointer, n=" + std::to_string(n));
    expect(hipFree(dev) == hipSuccess, "hipFree");
  }
  const PauliSum h = random_sum(6, 10, 0.3, 5);
  DevicePauliSum dh = DevicePauliSum::from_host(h, 0);
  std::vector<std::complex<double>> wrong(63);
  expect_throw<std::invalid_argument>(
      [&] { (void)dh.expectation_statevector_complex128(wrong); }, "", "statevector length");
  std::printf("expectation: complex128, complex64, device pointer ok\n");
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "";
  if (mode == "nodevice") {
    check_nodevice();
  } else if (mode == "device") {
    expect(DevicePauliSum::hip_available(), "a HIP device is visible");
    const auto status = DevicePauliSum::hip_status();
    expect(status.device_count >= 1 && !status.devices.empty() &&
               status.devices[0].gfx_target.rfind("gfx", 0) == 0,
           "hip_status() lists the device");
    check_algebra();
    check_expectation();
  } else {
    std::fprintf(stderr, "usage: facade_check nodevice|device\n");
    return 2;
  }
  std::printf("facade_check: OK (%d checks)\n", g_checks);
  return 0;
}

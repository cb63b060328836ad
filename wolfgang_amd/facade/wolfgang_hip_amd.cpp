// The reference's DevicePauliSum / DeviceCommutationMatrix members, forwarded
// to libwolfgang_amd.so's C ABI.
//
// A maintainer of the reference tree compiles this one translation unit in
// place of the reference's HIP backend sources (src/hip/*.hip.cpp and the
// backend-less stubs src/device_pauli_sum_stub.cpp and
// src/device_commutation_matrix_stub.cpp) and links libwolfgang_amd.so; the
// public headers (include/wolfgang/device_pauli_sum.hpp,
// include/wolfgang/device_commutation_matrix.hpp) and the Python binding stay
// as they are. Every member the two headers declare is defined here, with the
// reference's exception types: WG_ERR_INVALID -> std::invalid_argument,
// WG_ERR_OVERFLOW -> std::overflow_error, anything else -> std::runtime_error,
// carrying the library's message (the reference's texts).
//
// Build (from the reference tree's root):
//   g++ -std=c++20 -fPIC -c -Iinclude -I<wolfgang_amd>/include wolfgang_hip_amd.cpp
//
// Deliberate differences from the reference's HIP backend: the commutation
// matrix is stored as packed bits, so the two raw data-pointer accessors (which
// hand out one byte per entry) raise instead of returning an address, and
// complex64 statevectors are widened to complex128 on the device before the
// sweep.

#include <complex>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "wolfgang/accelerator_status.hpp"
#include "wolfgang/device_commutation_matrix.hpp"
#include "wolfgang/device_pauli_sum.hpp"
#include "wolfgang/pauli_sum.hpp"
#include "wolfgang_amd.h"

namespace wolfgang {

namespace {

void check(int rc) {
  if (rc == WG_OK) return;
  std::string msg = wg_last_error();
  if (rc == WG_ERR_INVALID || rc == WG_ERR_TYPE) throw std::invalid_argument(msg);
  if (rc == WG_ERR_OVERFLOW) throw std::overflow_error(msg);
  throw std::runtime_error(msg);
}

std::string version_string(int v) {
  // HIP encodes major * 10^7 + minor * 10^5 + patch.
  if (v <= 0) return {};
  return std::to_string(v / 10000000) + "." + std::to_string((v / 100000) % 100) + "." +
         std::to_string(v % 100000);
}

// The live implementation of a handle object, or the reference's moved-from
// error. (A template, so the private Impl types are named only inside members.)
template <class Impl>
const Impl* require(const std::unique_ptr<Impl>& impl, const char* what) {
  if (!impl) throw std::runtime_error(std::string(what) + " is empty or moved-from");
  return impl.get();
}

// Only the HIP backend exists behind this facade.
void require_hip(AcceleratorBackend requested, bool hip_ready) {
  const AcceleratorBackend selected =
      select_accelerator_backend(requested, false, false, true, hip_ready, false, false);
  if (selected != AcceleratorBackend::Hip) {
    throw std::runtime_error(accelerator_not_built_message(selected));
  }
}

}  // namespace

// ------------------------------------------------------------ DevicePauliSum

struct DevicePauliSum::Impl {
  wg_pauli_sum* h = nullptr;
  std::size_t num_qubits = 0;
  std::size_t num_terms = 0;
  std::size_t words = 0;
  int device = -1;

  explicit Impl(wg_pauli_sum* handle) : h(handle) {
    const int rc = wg_pauli_sum_info(h, &num_qubits, &num_terms, &words, &device);
    if (rc != WG_OK) {
      wg_pauli_sum_destroy(h);  // the destructor does not run when construction throws
      check(rc);
    }
  }
  ~Impl() { wg_pauli_sum_destroy(h); }
  Impl(const Impl&) = delete;
  Impl& operator=(const Impl&) = delete;
};

DevicePauliSum::DevicePauliSum() noexcept = default;
DevicePauliSum::DevicePauliSum(std::unique_ptr<Impl> impl) noexcept : impl_(std::move(impl)) {}
DevicePauliSum::~DevicePauliSum() = default;
DevicePauliSum::DevicePauliSum(DevicePauliSum&& other) noexcept = default;
DevicePauliSum& DevicePauliSum::operator=(DevicePauliSum&& other) noexcept = default;

DevicePauliSum DevicePauliSum::from_host(const PauliSum& host, int device) {
  return from_host(host, AcceleratorBackend::None, device);
}

DevicePauliSum DevicePauliSum::from_host(const PauliSum& host, AcceleratorBackend backend,
                                         int device) {
  require_hip(backend, hip_available());
  wg_pauli_sum* h = nullptr;
  check(wg_pauli_sum_create(host.num_qubits(), host.num_terms(), host.x_words().data(),
                            host.z_words().data(),
                            reinterpret_cast<const double*>(host.coeffs().data()), device, &h));
  return DevicePauliSum(std::make_unique<Impl>(h));
}

CudaStatus DevicePauliSum::cuda_status() {
  CudaStatus status;
  status.skip_reason = accelerator_not_built_message(AcceleratorBackend::Cuda);
  return status;
}

bool DevicePauliSum::cuda_available() { return false; }

HipStatus DevicePauliSum::hip_status() {
  HipStatus status;
  status.built = true;
  int runtime = 0, driver = 0;
  if (wg_runtime_versions(&runtime, &driver) == WG_OK) {
    status.runtime_version = version_string(runtime);
    status.driver_version = version_string(driver);
    status.toolkit_version = status.runtime_version;
  }
  int count = 0;
  if (wg_device_count(&count) != WG_OK || count == 0) {
    status.skip_reason = wg_last_error();
    if (status.skip_reason.empty()) status.skip_reason = "no HIP device is available";
    return status;
  }
  status.device_count = count;
  for (int ordinal = 0; ordinal < count; ++ordinal) {
    char name[256] = {}, gfx[64] = {};
    std::size_t bytes = 0;
    int cus = 0;
    if (wg_device_info(ordinal, name, sizeof name, gfx, sizeof gfx, &bytes, &cus) != WG_OK) {
      status.skip_reason = wg_last_error();
      status.devices.clear();
      return status;
    }
    HipDeviceInfo info;
    info.ordinal = ordinal;
    info.name = name;
    info.gfx_target = gfx;
    info.total_memory_bytes = bytes;
    status.devices.push_back(std::move(info));
  }
  status.runtime_available = true;
  return status;
}

bool DevicePauliSum::hip_available() { return hip_status().runtime_available; }

MetalStatus DevicePauliSum::metal_status() {
  MetalStatus status;
  status.skip_reason = accelerator_not_built_message(AcceleratorBackend::Metal);
  return status;
}

bool DevicePauliSum::metal_available() { return false; }

PauliSum DevicePauliSum::to_host() const {
  const Impl* p = require(impl_, "DevicePauliSum");
  std::vector<std::uint64_t> x(p->num_terms * p->words), z(p->num_terms * p->words);
  std::vector<std::complex<double>> c(p->num_terms);
  check(wg_pauli_sum_to_host(p->h, x.data(), z.data(), reinterpret_cast<double*>(c.data())));
  return PauliSum(p->num_qubits, p->words, p->num_terms, std::move(x), std::move(z),
                  std::move(c));
}

DevicePauliSum DevicePauliSum::simplify(double atol, double rtol) const {
  const Impl* p = require(impl_, "DevicePauliSum");
  wg_pauli_sum* out = nullptr;
  check(wg_simplify(p->h, atol, rtol, nullptr, &out));
  return DevicePauliSum(std::make_unique<Impl>(out));
}

std::complex<double> DevicePauliSum::expectation_statevector_complex128(
    std::span<const std::complex<double>> psi) const {
  const Impl* p = require(impl_, "DevicePauliSum");
  double e[2] = {0.0, 0.0};
  check(wg_expectation(p->h, psi.data(), WG_DTYPE_COMPLEX128, psi.size(), 0, nullptr, e));
  return {e[0], e[1]};
}

std::complex<double> DevicePauliSum::expectation_statevector_complex64(
    std::span<const std::complex<float>> psi) const {
  const Impl* p = require(impl_, "DevicePauliSum");
  double e[2] = {0.0, 0.0};
  check(wg_expectation(p->h, psi.data(), WG_DTYPE_COMPLEX64, psi.size(), 0, nullptr, e));
  return {e[0], e[1]};
}

std::complex<double> DevicePauliSum::expectation_statevector_device_pointer(
    std::uintptr_t device_pointer, DeviceStatevectorDtype dtype, std::size_t length) const {
  return expectation_statevector_device_pointer_impl(device_pointer, dtype, length, true);
}

std::complex<double> DevicePauliSum::expectation_statevector_device_pointer_impl(
    std::uintptr_t device_pointer, DeviceStatevectorDtype dtype, std::size_t length,
    bool /*synchronize_input: every call orders after prior work on the null stream*/) const {
  const Impl* p = require(impl_, "DevicePauliSum");
  double e[2] = {0.0, 0.0};
  const int wg_dtype =
      dtype == DeviceStatevectorDtype::Complex64 ? WG_DTYPE_COMPLEX64 : WG_DTYPE_COMPLEX128;
  check(wg_expectation(p->h, reinterpret_cast<const void*>(device_pointer), wg_dtype, length, 1,
                       nullptr, e));
  return {e[0], e[1]};
}

std::vector<std::uint8_t> DevicePauliSum::commutes_with(
    const DevicePauliSum& rhs, std::size_t max_commutation_matrix_entries) const {
  const Impl* p = require(impl_, "DevicePauliSum");
  const Impl* q = require(rhs.impl_, "DevicePauliSum");
  // The entry count is validated before the host buffer exists.
  const std::size_t entries = PauliSum::checked_commutation_matrix_entries_for_testing(
      p->num_terms, q->num_terms, max_commutation_matrix_entries);
  std::vector<std::uint8_t> out(entries);
  check(wg_commutes_with_host(p->h, q->h, out.data(), out.size(), max_commutation_matrix_entries,
                              nullptr));
  return out;
}

void DevicePauliSum::commutes_with_into(const DevicePauliSum& rhs, std::span<std::uint8_t> output,
                                        std::size_t max_commutation_matrix_entries) const {
  const Impl* p = require(impl_, "DevicePauliSum");
  const Impl* q = require(rhs.impl_, "DevicePauliSum");
  check(wg_commutes_with_host(p->h, q->h, output.data(), output.size(),
                              max_commutation_matrix_entries, nullptr));
}

DeviceCommutationMatrix DevicePauliSum::commutes_with_device(
    const DevicePauliSum& rhs, std::size_t max_commutation_matrix_entries) const {
  const Impl* p = require(impl_, "DevicePauliSum");
  const Impl* q = require(rhs.impl_, "DevicePauliSum");
  (void)PauliSum::checked_commutation_matrix_entries_for_testing(p->num_terms, q->num_terms,
                                                                 max_commutation_matrix_entries);
  DeviceCommutationMatrix out =
      DeviceCommutationMatrix::empty(p->num_terms, q->num_terms, AcceleratorBackend::Hip, p->device);
  commutes_with_device_into(rhs, out, max_commutation_matrix_entries);
  return out;
}

// ------------------------------------------------------ DeviceCommutationMatrix

struct DeviceCommutationMatrix::Impl {
  wg_comm_matrix* m = nullptr;
  std::size_t rows = 0;
  std::size_t cols = 0;
  std::size_t row_words = 0;
  int device = -1;

  explicit Impl(wg_comm_matrix* handle) : m(handle) {
    const int rc = wg_comm_matrix_info(m, &rows, &cols, &row_words, &device);
    if (rc != WG_OK) {
      wg_comm_matrix_destroy(m);
      check(rc);
    }
  }
  ~Impl() { wg_comm_matrix_destroy(m); }
  Impl(const Impl&) = delete;
  Impl& operator=(const Impl&) = delete;
};

void DevicePauliSum::commutes_with_device_into(const DevicePauliSum& rhs,
                                               DeviceCommutationMatrix& output,
                                               std::size_t max_commutation_matrix_entries) const {
  const Impl* p = require(impl_, "DevicePauliSum");
  const Impl* q = require(rhs.impl_, "DevicePauliSum");
  const DeviceCommutationMatrix::Impl* o = require(output.impl_, "DeviceCommutationMatrix");
  check(wg_commutes_with_device(p->h, q->h, o->m, max_commutation_matrix_entries, nullptr));
}

DevicePauliSum DevicePauliSum::matmul(const DevicePauliSum& rhs, bool simplify_output,
                                      std::size_t max_intermediate_terms) const {
  const Impl* p = require(impl_, "DevicePauliSum");
  const Impl* q = require(rhs.impl_, "DevicePauliSum");
  wg_pauli_sum* out = nullptr;
  check(wg_matmul(p->h, q->h, simplify_output ? 1 : 0, max_intermediate_terms, nullptr, &out));
  return DevicePauliSum(std::make_unique<Impl>(out));
}

std::size_t DevicePauliSum::num_qubits() const noexcept { return impl_ ? impl_->num_qubits : 0; }
std::size_t DevicePauliSum::num_terms() const noexcept { return impl_ ? impl_->num_terms : 0; }
std::size_t DevicePauliSum::words() const noexcept { return impl_ ? impl_->words : 0; }
int DevicePauliSum::device() const noexcept { return impl_ ? impl_->device : -1; }
std::string DevicePauliSum::backend() const { return "hip"; }

DeviceCommutationMatrix::DeviceCommutationMatrix() noexcept = default;
DeviceCommutationMatrix::DeviceCommutationMatrix(std::unique_ptr<Impl> impl) noexcept
    : impl_(std::move(impl)) {}
DeviceCommutationMatrix::~DeviceCommutationMatrix() = default;
DeviceCommutationMatrix::DeviceCommutationMatrix(DeviceCommutationMatrix&& other) noexcept =
    default;
DeviceCommutationMatrix& DeviceCommutationMatrix::operator=(
    DeviceCommutationMatrix&& other) noexcept = default;

DeviceCommutationMatrix DeviceCommutationMatrix::empty(std::size_t rows, std::size_t cols,
                                                       int device) {
  return empty(rows, cols, AcceleratorBackend::None, device);
}

DeviceCommutationMatrix DeviceCommutationMatrix::empty(std::size_t rows, std::size_t cols,
                                                       AcceleratorBackend backend, int device) {
  require_hip(backend, DevicePauliSum::hip_available());
  wg_comm_matrix* m = nullptr;
  check(wg_comm_matrix_create(rows, cols, device, &m));
  return DeviceCommutationMatrix(std::make_unique<Impl>(m));
}

std::uint64_t DeviceCommutationMatrix::count_commuting() const {
  const Impl* p = require(impl_, "DeviceCommutationMatrix");
  std::uint64_t total = 0;
  check(wg_comm_matrix_count(p->m, -1, 0, &total, 1, nullptr));
  return total;
}

std::vector<std::uint64_t> DeviceCommutationMatrix::count_commuting_rows() const {
  const Impl* p = require(impl_, "DeviceCommutationMatrix");
  std::vector<std::uint64_t> out(p->rows);
  check(wg_comm_matrix_count(p->m, 1, 0, out.data(), out.size(), nullptr));
  return out;
}

std::vector<std::uint64_t> DeviceCommutationMatrix::count_commuting_cols() const {
  const Impl* p = require(impl_, "DeviceCommutationMatrix");
  std::vector<std::uint64_t> out(p->cols);
  check(wg_comm_matrix_count(p->m, 0, 0, out.data(), out.size(), nullptr));
  return out;
}

std::vector<std::uint8_t> DeviceCommutationMatrix::to_host() const {
  const Impl* p = require(impl_, "DeviceCommutationMatrix");
  std::vector<std::uint8_t> out(p->rows * p->cols);
  check(wg_comm_matrix_to_host_u8(p->m, out.data(), out.size(), nullptr));
  return out;
}

std::size_t DeviceCommutationMatrix::rows() const noexcept { return impl_ ? impl_->rows : 0; }
std::size_t DeviceCommutationMatrix::cols() const noexcept { return impl_ ? impl_->cols : 0; }
std::size_t DeviceCommutationMatrix::num_entries() const noexcept {
  return impl_ ? impl_->rows * impl_->cols : 0;
}
int DeviceCommutationMatrix::device() const noexcept { return impl_ ? impl_->device : -1; }
std::string DeviceCommutationMatrix::backend() const { return "hip"; }

std::uintptr_t DeviceCommutationMatrix::data_pointer_for_cuda_array_interface() const {
  (void)require(impl_, "DeviceCommutationMatrix");
  throw std::runtime_error(
      "DeviceCommutationMatrix stores packed bits on this backend (row i: ceil(cols/64) uint64 "
      "words, bit j%64 of word j/64 = entry (i, j)); wg_comm_matrix_device_bits gives their "
      "device address");
}

std::uintptr_t DeviceCommutationMatrix::data_pointer_for_dlpack() const {
  return data_pointer_for_cuda_array_interface();
}

int DeviceCommutationMatrix::dlpack_device_type() const {
  (void)require(impl_, "DeviceCommutationMatrix");
  return 10;  // kDLROCM
}

std::uint8_t* DeviceCommutationMatrix::mutable_data_for_device_write() {
  (void)data_pointer_for_cuda_array_interface();
  return nullptr;
}

void copy_device_commutation_matrix_from_host_for_testing(DeviceCommutationMatrix& matrix,
                                                          std::span<const std::uint8_t> values) {
  const DeviceCommutationMatrix::Impl* p = require(matrix.impl_, "DeviceCommutationMatrix");
  check(wg_comm_matrix_from_host_u8(p->m, values.data(), values.size(), nullptr));
}

}  // namespace wolfgang

// Allocation-free statevector expectation for callers that cannot allocate
// device memory: code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc.
//
// wg_expectation_noalloc runs the same kernels as wg_expectation (same plan,
// same sweeps and fixed-order reductions, so the same bits) but takes its
// partial sums and per-key values from `scratch`, a file-scope device array of
// kNoAllocScratchBytes, instead of the library's growable arena; E goes to a
// caller-owned device location and nothing is synchronized. Each device that
// runs it holds its own copy of the array until the process ends (this
// translation unit's code object is loaded on a device only when the entry
// point is first called there).
//
// Host code never uses the array's name as an address: at the entry point's
// first call on a device, with that device current, hipGetSymbolAddress gives
// the device copy's address, kept per device, and only that pointer is passed
// to kernels.
//
// Every call on one device shares that device's array and leaves the last
// call's contents in it. Calls that overlap on one device (on different
// streams, or from several host threads) race on it. Nothing in the array
// needs clearing: each partial row and key value the reduction reads is
// written earlier in the same call.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "wg_expect.hpp"
#include "wg_internal.hpp"

namespace wg {
namespace {

constexpr std::size_t kNoAllocScratchBytes = std::size_t{256} << 20;

}  // namespace
}  // namespace wg

static __device__ __align__(256) unsigned char scratch[wg::kNoAllocScratchBytes];

namespace wg {
namespace {

__global__ void zero_energy_kernel(double2* out) { *out = make_double2(0.0, 0.0); }

std::mutex g_scratch_mutex;
std::vector<void*> g_scratch_ptr;  // device address of `scratch`, per device

// The device copy of `scratch` on `device`, which must be current.
void* scratch_on(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (g_scratch_ptr.size() <= static_cast<std::size_t>(device)) {
    g_scratch_ptr.resize(static_cast<std::size_t>(device) + 1, nullptr);
  }
  void*& p = g_scratch_ptr[static_cast<std::size_t>(device)];
  if (p == nullptr) {
    check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "get scratch symbol address");
    if (p == nullptr) fail_runtime("HIP get scratch symbol address returned null");
  }
  return p;
}

}  // namespace
}  // namespace wg

using namespace wg;

extern "C" {

int wg_expectation_noalloc(const wg_pauli_sum* op_c, const double* psi, size_t length,
                           double* out2_device, void* stream_ptr) {
  return guard([&] {
    require_op(op_c);
    auto* op = const_cast<wg_pauli_sum*>(op_c);
    if (op->num_qubits > 63) {
      fail_invalid("HIP expectation_statevector requires num_qubits <= 63");
    }
    if (length != (std::size_t{1} << op->num_qubits)) {
      fail_invalid("expectation_statevector requires len(psi) == 2 ** num_qubits");
    }
    if (psi == nullptr) fail_invalid("HIP statevector data pointer must be non-null");
    if (out2_device == nullptr) fail_invalid("expectation output pointer must be non-null");
    if (reinterpret_cast<std::uintptr_t>(out2_device) % 16 != 0) {
      fail_invalid("expectation output pointer must be 16-byte aligned");
    }
    if (op->num_terms == 0) {
      DeviceScope scope(op->device);
      zero_energy_kernel<<<1, 1, 0, as_stream(stream_ptr)>>>(reinterpret_cast<double2*>(out2_device));
      check_launch("zero_energy");
      return;
    }
    std::shared_ptr<ExpectationPlan> plan;
    {
      std::lock_guard<std::mutex> lk(op->plan_mutex);
      plan = op->plan;
    }
    if (!plan) {
      fail_invalid(
          "wg_expectation_noalloc needs the operator's expectation plan: build it first with "
          "wg_expectation_plan_info or any expectation call");
    }
    const std::size_t need = forward_scratch_bytes(*plan);
    if (need > kNoAllocScratchBytes) {
      fail_invalid("wg_expectation_noalloc: the plan needs " + std::to_string(need) +
                   " bytes of scratch, more than the " + std::to_string(kNoAllocScratchBytes) +
                   "-byte maximum");
    }
    DeviceScope scope(op->device);
    auto* base = static_cast<unsigned char*>(scratch_on(op->device));
    const std::size_t key_bytes =
        (std::max<std::size_t>(plan->num_keys, 1) * sizeof(double) + 255) & ~std::size_t{255};
    forward_into(op->coeffs, *plan, reinterpret_cast<const double2*>(psi), nullptr,
                 reinterpret_cast<double2*>(out2_device), reinterpret_cast<double*>(base),
                 plan->num_items ? reinterpret_cast<double2*>(base + key_bytes) : nullptr,
                 as_stream(stream_ptr));
  });
}

}  // extern "C"

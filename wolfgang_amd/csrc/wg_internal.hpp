// Internal runtime pieces shared by the wolfgang_amd HIP translation units:
// handle layouts, error plumbing behind the C ABI, device scoping and a
// per-device scratch arena.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/wolfgang_amd.h"

namespace wg {

using u64 = std::uint64_t;
using u32 = std::uint32_t;

// ------------------------------------------------------------------ errors
// Exceptions carry the C ABI category; the guard() wrapper turns them into a
// status code plus a thread-local message (reference error convention:
// std::invalid_argument for validation, std::runtime_error for runtime/HIP).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail_invalid(const std::string& m) { throw Error(WG_ERR_INVALID, m); }
[[noreturn]] inline void fail_runtime(const std::string& m) { throw Error(WG_ERR_RUNTIME, m); }
[[noreturn]] inline void fail_type(const std::string& m) { throw Error(WG_ERR_TYPE, m); }
[[noreturn]] inline void fail_overflow(const std::string& m) { throw Error(WG_ERR_OVERFLOW, m); }

void set_last_error(const std::string& m);
void clear_last_error();

template <typename Fn>
int guard(Fn&& fn) noexcept {
  try {
    clear_last_error();
    fn();
    return WG_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return WG_ERR_RUNTIME;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return WG_ERR_RUNTIME;
  } catch (...) {
    set_last_error("unknown error");
    return WG_ERR_RUNTIME;
  }
}

// "HIP <action> failed: <hipGetErrorString>"
// (reference src/hip/device_pauli_sum.hip.hpp:40-45).
inline void check_hip(hipError_t r, const char* action) {
  if (r != hipSuccess) {
    (void)hipGetLastError();
    fail_runtime(std::string("HIP ") + action + " failed: " + hipGetErrorString(r));
  }
}

inline void check_launch(const char* kernel) {
  const hipError_t r = hipGetLastError();
  if (r != hipSuccess) {
    fail_runtime(std::string("HIP launch ") + kernel + " failed: " + hipGetErrorString(r));
  }
}

// Makes `device` current for the scope and restores the previous device
// (reference src/hip/device_pauli_sum.hip.hpp:56-76).
class DeviceScope {
 public:
  explicit DeviceScope(int device) {
    check_hip(hipGetDevice(&prev_), "get current device");
    if (prev_ != device) check_hip(hipSetDevice(device), "set device");
  }
  DeviceScope(const DeviceScope&) = delete;
  DeviceScope& operator=(const DeviceScope&) = delete;
  ~DeviceScope() {
    if (prev_ >= 0) (void)hipSetDevice(prev_);
  }

 private:
  int prev_ = -1;
};

// Validates a device ordinal before any allocation
// (reference src/hip/device_pauli_sum.hip.cpp:70-82).
void validate_device(int device);

// Checked size arithmetic (reference src/detail/checked_arithmetic.hpp:20-32).
inline std::size_t checked_mul(std::size_t a, std::size_t b, const char* what) {
  if (a != 0 && b > SIZE_MAX / a) fail_invalid(std::string(what) + " size overflows size_t");
  return a * b;
}

inline std::size_t words_for(std::size_t n) { return n == 0 ? 0 : (n + 63) / 64; }

// ------------------------------------------------------------------ memory
// Device memory goes through a per-device block cache: a freed block is kept
// (up to a byte limit per device) and handed to the next request of the same
// size class, so steady-state loops (products, batches, operators rebuilt per
// step) make no hipMalloc / hipFree calls; mapping and unmapping large buffers
// costs milliseconds each. A block is returned to the cache only after the
// device has finished all queued work (as hipFree would), so reuse never races
// a kernel still reading it. `action` names the allocation in the error.
void* device_alloc_bytes(std::size_t bytes, const char* action = "allocate device buffer");
void device_free(void* p);

template <typename T>
T* device_alloc(std::size_t count) {
  if (count == 0) return nullptr;
  return static_cast<T*>(device_alloc_bytes(checked_mul(count, sizeof(T), "device buffer")));
}

// Cache control: give every cached block of `device` back to the runtime; set
// the per-device byte limit (default: a quarter of the device's memory; returns
// the previous override, SIZE_MAX if none); cached bytes, cached blocks and
// bytes handed out and not yet freed on `device`.
void memory_cache_release(int device);
std::size_t memory_cache_set_limit(std::size_t bytes);
void memory_cache_info(int device, std::size_t* cached, std::size_t* blocks, std::size_t* live);

// Per-device growable scratch arena. A call takes the device lock for its whole
// duration (all entry points synchronize before returning), carves aligned
// sub-buffers and releases them on scope exit.
class Scratch {
 public:
  explicit Scratch(int device);
  ~Scratch();
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;

  // Reserves `bytes` (rounded to 256) and returns a device pointer.
  template <typename T>
  T* take(std::size_t count) {
    return static_cast<T*>(take_bytes(checked_mul(count, sizeof(T), "scratch buffer")));
  }
  void* take_bytes(std::size_t bytes);

  // mark() / release(m): buffers taken after the mark are given back, so a
  // loop of calls reuses the same memory. The work that used them must have
  // completed (the caller has synchronized the stream).
  std::size_t mark() const { return used_; }
  void release(std::size_t m);

  struct Arena;

 private:
  Arena* arena_;
  std::unique_lock<std::mutex> lock_;
  std::size_t used_ = 0;
  std::size_t high_ = 0;    // high-water mark of used_
  bool overflowed_ = false;  // some take did not fit the arena
  std::vector<void*> overflow_;
  std::vector<std::size_t> overflow_at_;  // used_ before each overflow buffer
};

// ------------------------------------------------------------------ handles
struct ExpectationPlan;  // wg_expect.hip

}  // namespace wg

struct wg_pauli_sum {
  std::size_t num_qubits = 0;
  std::size_t words = 0;
  std::size_t num_terms = 0;
  wg::u64* x = nullptr;       // [num_terms * words]
  wg::u64* z = nullptr;       // [num_terms * words]
  double2* coeffs = nullptr;  // [num_terms]
  int device = 0;
  // Host-side expectation plan, built on first use and reused by later calls.
  std::shared_ptr<wg::ExpectationPlan> plan;
  std::mutex plan_mutex;

  ~wg_pauli_sum();
};

struct wg_comm_matrix {
  std::size_t rows = 0;
  std::size_t cols = 0;
  std::size_t row_words = 0;
  wg::u64* bits = nullptr;  // [rows * row_words]
  int device = 0;

  ~wg_comm_matrix();
};

namespace wg {

// Allocates an owning handle with room for `terms` terms on `device`.
wg_pauli_sum* new_pauli_sum(std::size_t n, std::size_t terms, int device);

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

inline void sync(hipStream_t s) { check_hip(hipStreamSynchronize(s), "synchronize stream"); }

inline unsigned grid_for(std::size_t work, unsigned block) {
  const std::size_t g = (work + block - 1) / block;
  if (g == 0) return 1;
  if (g > 0x7fffffffULL) fail_invalid("request exceeds HIP launch grid limits");
  return static_cast<unsigned>(g);
}

// The moved-from / empty-handle message of the reference
// (reference src/hip/matmul_hip.hip.cpp:95-97).
inline void require_op(const wg_pauli_sum* op) {
  if (op == nullptr) fail_runtime("DevicePauliSum is empty or moved-from");
}

// Same backend + same device check (reference src/accelerator_status.cpp:126-143).
void require_same_device(const char* operation, const wg_pauli_sum* a, const wg_pauli_sum* b);

// Event-pair timing of named kernel launches (enabled by wg_profile_enable).
class ProfileScope {
 public:
  ProfileScope(const char* name, hipStream_t stream);
  ~ProfileScope();
  ProfileScope(const ProfileScope&) = delete;
  ProfileScope& operator=(const ProfileScope&) = delete;

 private:
  const char* name_;
  hipStream_t stream_;
  hipEvent_t start_ = nullptr;
};

// Device-wide helpers (wg_runtime.hip).
void exclusive_scan_u32(const u32* in, u32* out, std::size_t n, u32* total_dev, Scratch& scratch,
                        hipStream_t stream);

// Packed commutation bits (wg_commute.hip): bit c of row r of out
// [L x row_words] is 1 when term r of (lx, lz) commutes with term c of (rx, rz);
// both operands hold W mask words per term.
void launch_commute_raw(const u64* lx, const u64* lz, std::size_t L, const u64* rx, const u64* rz,
                        std::size_t R, std::size_t W, u64* out, std::size_t row_words,
                        hipStream_t stream);

}  // namespace wg

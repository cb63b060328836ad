// Runtime: device discovery, operator handles, host<->device transfer, the
// scratch arena and a device-wide exclusive scan used by simplify.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "wg_internal.hpp"

namespace wg {
namespace {

thread_local std::string t_last_error;

std::mutex g_arena_map_mutex;

constexpr std::size_t kAlign = 256;

}  // namespace

void set_last_error(const std::string& m) { t_last_error = m; }
void clear_last_error() { t_last_error.clear(); }

void validate_device(int device) {
  int count = 0;
  const hipError_t r = hipGetDeviceCount(&count);
  if (r != hipSuccess) {
    (void)hipGetLastError();
    if (r == hipErrorNoDevice) fail_runtime("no HIP device is available");
    fail_runtime(std::string("HIP runtime library is unavailable or failed during device "
                             "discovery: ") + hipGetErrorString(r));
  }
  if (count == 0) fail_runtime("no HIP device is available");
  if (device < 0 || device >= count) fail_invalid("HIP device ordinal is out of range");
}

void require_same_device(const char* operation, const wg_pauli_sum* a, const wg_pauli_sum* b) {
  if (a->device == b->device) return;
  fail_invalid(std::string(operation) +
               " requires operands on the same accelerator backend and same device "
               "(operation=" + operation + ", left_backend=hip, left_device=" +
               std::to_string(a->device) + ", right_backend=hip, right_device=" +
               std::to_string(b->device) + ", failure_stage=pre_allocation)");
}

// ------------------------------------------------------------ block cache

namespace {

struct BlockCache {
  std::mutex mutex;
  struct Block {
    int device;
    std::size_t size;  // size class actually allocated
  };
  std::map<void*, Block> live;                                     // every block we own
  std::map<std::pair<int, std::size_t>, std::vector<void*>> free;  // cached, by (device, class)
  std::map<int, std::size_t> cached_bytes;
  std::map<int, std::size_t> limit;  // per device; set on first use
  std::size_t limit_override = SIZE_MAX;  // wg_memory_cache_limit
};

BlockCache& block_cache() {
  static auto* c = new BlockCache();  // lives until process exit
  return *c;
}

// Size classes: 256-byte steps up to 4 KiB, then 8 steps per power of two
// (at most 12.5 % rounding).
std::size_t size_class(std::size_t b) {
  if (b <= 4096) return (b + 255) & ~std::size_t{255};
  const int hi = 64 - __builtin_clzll(b - 1);  // 2^(hi-1) < b <= 2^hi
  const std::size_t step = std::size_t{1} << (hi - 3);
  return (b + step - 1) & ~(step - 1);
}

std::size_t default_limit(int device) {
  std::size_t free_b = 0, total_b = 0;
  DeviceScope scope(device);
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    return std::size_t{8} << 30;
  }
  return total_b / 4;
}

// Frees cached blocks of `device` (largest first) until at most `keep` bytes
// stay cached. Caller holds the cache mutex.
void trim_locked(BlockCache& c, int device, std::size_t keep) {
  auto& cached = c.cached_bytes[device];
  auto it = c.free.lower_bound({device + 1, 0});
  while (cached > keep && it != c.free.begin()) {
    --it;
    if (it->first.first != device) break;
    auto& list = it->second;
    while (cached > keep && !list.empty()) {
      void* p = list.back();
      list.pop_back();
      cached -= it->first.second;
      c.live.erase(p);
      (void)hipFree(p);
    }
    if (list.empty()) it = c.free.erase(it);
  }
  (void)hipGetLastError();
}

}  // namespace

void* device_alloc_bytes(std::size_t bytes, const char* action) {
  if (bytes == 0) return nullptr;
  int device = 0;
  check_hip(hipGetDevice(&device), "get current device");
  const std::size_t cls = size_class(bytes);
  BlockCache& c = block_cache();
  {
    std::lock_guard<std::mutex> lk(c.mutex);
    auto it = c.free.find({device, cls});
    if (it != c.free.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      c.cached_bytes[device] -= cls;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t r = hipMalloc(&p, cls);
  if (r == hipErrorOutOfMemory) {
    // Give the cached blocks of this device back to the runtime and retry once.
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> lk(c.mutex);
      trim_locked(c, device, 0);
    }
    r = hipMalloc(&p, cls);
  }
  check_hip(r, action);
  std::lock_guard<std::mutex> lk(c.mutex);
  c.live[p] = BlockCache::Block{device, cls};
  return p;
}

void device_free(void* p) {
  if (p == nullptr) return;
  BlockCache& c = block_cache();
  BlockCache::Block b{};
  {
    std::lock_guard<std::mutex> lk(c.mutex);
    auto it = c.live.find(p);
    if (it == c.live.end()) {
      (void)hipFree(p);  // not from this cache
      (void)hipGetLastError();
      return;
    }
    b = it->second;
  }
  {
    // Work queued on the device may still use the block: wait for it, as
    // hipFree does, before the block can be handed out again.
    int prev = -1;
    const bool have_prev = hipGetDevice(&prev) == hipSuccess;
    if (!have_prev || prev != b.device) (void)hipSetDevice(b.device);
    (void)hipDeviceSynchronize();
    if (have_prev && prev != b.device) (void)hipSetDevice(prev);
    (void)hipGetLastError();
  }
  std::lock_guard<std::mutex> lk(c.mutex);
  auto lim = c.limit.find(b.device);
  if (lim == c.limit.end()) {
    lim = c.limit.emplace(b.device, default_limit(b.device)).first;
  }
  const std::size_t limit = std::min(lim->second, c.limit_override);
  if (b.size > limit) {
    c.live.erase(p);
    (void)hipFree(p);
    (void)hipGetLastError();
    return;
  }
  c.free[{b.device, b.size}].push_back(p);
  c.cached_bytes[b.device] += b.size;
  if (c.cached_bytes[b.device] > limit) trim_locked(c, b.device, limit);
}

void memory_cache_release(int device) {
  BlockCache& c = block_cache();
  std::lock_guard<std::mutex> lk(c.mutex);
  trim_locked(c, device, 0);
}

std::size_t memory_cache_set_limit(std::size_t bytes) {
  BlockCache& c = block_cache();
  std::lock_guard<std::mutex> lk(c.mutex);
  const std::size_t prev = c.limit_override;
  c.limit_override = bytes;
  for (auto& [device, cached] : c.cached_bytes) {
    const auto lim = c.limit.find(device);
    const std::size_t l = std::min(lim == c.limit.end() ? SIZE_MAX : lim->second, bytes);
    if (cached > l) trim_locked(c, device, l);
  }
  return prev;
}

void memory_cache_info(int device, std::size_t* cached, std::size_t* blocks, std::size_t* live) {
  BlockCache& c = block_cache();
  std::lock_guard<std::mutex> lk(c.mutex);
  std::size_t owned = 0, free_blocks = 0;
  for (const auto& [key, list] : c.free) {
    if (key.first == device) free_blocks += list.size();
  }
  for (const auto& [ptr, blk] : c.live) {
    if (blk.device == device) owned += blk.size;
  }
  const auto it = c.cached_bytes.find(device);
  *cached = it == c.cached_bytes.end() ? 0 : it->second;
  *blocks = free_blocks;
  *live = owned - *cached;  // bytes in blocks handed out and not yet freed
}

struct Scratch::Arena {
  std::mutex mutex;
  void* base = nullptr;
  std::size_t capacity = 0;
};

namespace {
// Arenas live until process exit (one per device that was used).
Scratch::Arena* arena_for(int device) {
  static auto* arenas = new std::map<int, Scratch::Arena*>();
  std::lock_guard<std::mutex> lk(g_arena_map_mutex);
  auto it = arenas->find(device);
  if (it != arenas->end()) return it->second;
  auto* a = new Scratch::Arena();
  (*arenas)[device] = a;
  return a;
}
}  // namespace

Scratch::Scratch(int device) : arena_(arena_for(device)), lock_(arena_->mutex) {}

Scratch::~Scratch() {
  if (overflowed_) {
    // Work using the overflow buffers has completed (entry points synchronize
    // before returning); grow the arena to this call's high-water mark.
    for (void* p : overflow_) device_free(p);
    device_free(arena_->base);
    arena_->base = nullptr;
    arena_->capacity = 0;
    const std::size_t top = std::max(high_, used_);
    const std::size_t want = top + top / 4;
    try {
      arena_->base = device_alloc_bytes(want, "allocate scratch buffer");
      arena_->capacity = want;
    } catch (...) {
      // The arena stays empty; the next call overflows again.
    }
  }
}

void* Scratch::take_bytes(std::size_t bytes) {
  const std::size_t rounded = (bytes + kAlign - 1) / kAlign * kAlign;
  if (rounded == 0) return nullptr;
  const std::size_t offset = used_;
  used_ += rounded;
  if (overflow_.empty() && used_ <= arena_->capacity) {
    return static_cast<char*>(arena_->base) + offset;
  }
  void* p = device_alloc_bytes(rounded, "allocate scratch buffer");
  overflow_.push_back(p);
  overflow_at_.push_back(offset);
  overflowed_ = true;
  return p;
}

void Scratch::release(std::size_t m) {
  while (!overflow_.empty() && overflow_at_.back() >= m) {
    device_free(overflow_.back());
    overflow_.pop_back();
    overflow_at_.pop_back();
  }
  high_ = std::max(high_, used_);
  used_ = std::min(used_, m);
}

wg_pauli_sum* new_pauli_sum(std::size_t n, std::size_t terms, int device) {
  auto* op = new wg_pauli_sum();
  op->num_qubits = n;
  op->words = words_for(n);
  op->num_terms = terms;
  op->device = device;
  try {
    const std::size_t packed = checked_mul(terms, op->words, "HIP device packed words");
    op->x = device_alloc<u64>(packed);
    op->z = device_alloc<u64>(packed);
    op->coeffs = device_alloc<double2>(terms);
  } catch (...) {
    delete op;
    throw;
  }
  return op;
}

// ------------------------------------------------------------------ scan

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr std::size_t kScanTile = kScanBlock * kScanItems;

__device__ inline u32 wave_inclusive_scan(u32 v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const u32 o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Exclusive block scan of one value per thread; returns the block total.
__device__ inline u32 block_exclusive_scan(u32 v, u32* lds_waves, u32* total) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const u32 inc = wave_inclusive_scan(v);
  if (lane == 63) lds_waves[wave] = inc;
  __syncthreads();
  u32 before = 0, sum = 0;
  for (int w = 0; w < kScanBlock / 64; ++w) {
    const u32 s = lds_waves[w];
    if (w < wave) before += s;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return before + inc - v;
}

__global__ __launch_bounds__(kScanBlock) void scan_reduce_kernel(const u32* in, std::size_t n,
                                                                 u32* block_sums) {
  __shared__ u32 lds[kScanBlock / 64];
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kScanTile;
  u32 s = 0;
  for (int i = 0; i < kScanItems; ++i) {
    const std::size_t idx = base + static_cast<std::size_t>(i) * kScanBlock + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  u32 total;
  (void)block_exclusive_scan(s, lds, &total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void scan_apply_kernel(const u32* in, std::size_t n,
                                                                const u32* block_offsets,
                                                                u32* out, u32* total_out) {
  __shared__ u32 lds[kScanBlock / 64];
  // Thread t owns the contiguous run [base + t*kScanItems, +kScanItems).
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kScanTile +
                           static_cast<std::size_t>(threadIdx.x) * kScanItems;
  // Full runs of 16-byte aligned buffers move as 16-byte vectors (base is a
  // multiple of kScanItems).
  const bool full = base + kScanItems <= n &&
                    ((reinterpret_cast<std::uintptr_t>(in) | reinterpret_cast<std::uintptr_t>(out)) &
                     15U) == 0;
  u32 vals[kScanItems];
  u32 s = 0;
  if (full) {
#pragma unroll
    for (int i = 0; i < kScanItems; i += 4) {
      const uint4 v = *reinterpret_cast<const uint4*>(in + base + i);
      vals[i] = v.x;
      vals[i + 1] = v.y;
      vals[i + 2] = v.z;
      vals[i + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const std::size_t idx = base + i;
      vals[i] = idx < n ? in[idx] : 0;
    }
  }
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) s += vals[i];
  u32 total;
  u32 run = block_exclusive_scan(s, lds, &total);
  const u32 off = block_offsets ? block_offsets[blockIdx.x] : 0;
  run += off;
  if (full) {
#pragma unroll
    for (int i = 0; i < kScanItems; i += 4) {
      uint4 v;
      v.x = run;
      run += vals[i];
      v.y = run;
      run += vals[i + 1];
      v.z = run;
      run += vals[i + 2];
      v.w = run;
      run += vals[i + 3];
      *reinterpret_cast<uint4*>(out + base + i) = v;
    }
  } else {
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const std::size_t idx = base + i;
      if (idx < n) out[idx] = run;
      run += vals[i];
    }
  }
  if (total_out != nullptr && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    *total_out = off + total;
  }
}

}  // namespace

void exclusive_scan_u32(const u32* in, u32* out, std::size_t n, u32* total_dev, Scratch& scratch,
                        hipStream_t stream) {
  if (n == 0) {
    if (total_dev) check_hip(hipMemsetAsync(total_dev, 0, sizeof(u32), stream), "memset");
    return;
  }
  const std::size_t blocks = (n + kScanTile - 1) / kScanTile;
  if (blocks == 1) {
    scan_apply_kernel<<<1, kScanBlock, 0, stream>>>(in, n, nullptr, out, total_dev);
    check_launch("scan_apply");
    return;
  }
  u32* sums = scratch.take<u32>(blocks);
  u32* offsets = scratch.take<u32>(blocks);
  scan_reduce_kernel<<<grid_for(blocks, 1), kScanBlock, 0, stream>>>(in, n, sums);
  check_launch("scan_reduce");
  exclusive_scan_u32(sums, offsets, blocks, nullptr, scratch, stream);
  scan_apply_kernel<<<grid_for(blocks, 1), kScanBlock, 0, stream>>>(in, n, offsets, out,
                                                                    total_dev);
  check_launch("scan_apply");
}

}  // namespace wg

wg_pauli_sum::~wg_pauli_sum() {
  if (x == nullptr && z == nullptr && coeffs == nullptr) return;
  int prev = -1;
  const hipError_t r = hipGetDevice(&prev);
  (void)hipSetDevice(device);
  wg::device_free(x);
  wg::device_free(z);
  wg::device_free(coeffs);
  if (r == hipSuccess && prev >= 0 && prev != device) (void)hipSetDevice(prev);
}

wg_comm_matrix::~wg_comm_matrix() {
  if (bits == nullptr) return;
  int prev = -1;
  const hipError_t r = hipGetDevice(&prev);
  (void)hipSetDevice(device);
  wg::device_free(bits);
  if (r == hipSuccess && prev >= 0 && prev != device) (void)hipSetDevice(prev);
}

namespace wg {
const char* last_error_cstr() { return t_last_error.c_str(); }

namespace {
struct PendingTiming {
  std::string name;
  hipEvent_t start;
  hipEvent_t stop;
};
struct ProfileTotals {
  double ms = 0.0;
  std::uint64_t launches = 0;
};
std::mutex g_profile_mutex;
bool g_profile_enabled = false;
std::vector<PendingTiming>& pending_timings() {
  static auto* v = new std::vector<PendingTiming>();
  return *v;
}
std::map<std::string, ProfileTotals>& profile_totals() {
  static auto* m = new std::map<std::string, ProfileTotals>();
  return *m;
}
void resolve_pending_locked() {
  for (auto& p : pending_timings()) {
    float ms = 0.0f;
    if (hipEventSynchronize(p.stop) == hipSuccess && hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      auto& t = profile_totals()[p.name];
      t.ms += ms;
      t.launches += 1;
    }
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  (void)hipGetLastError();
  pending_timings().clear();
}
}  // namespace

ProfileScope::ProfileScope(const char* name, hipStream_t stream) : name_(name), stream_(stream) {
  std::lock_guard<std::mutex> lk(g_profile_mutex);
  if (!g_profile_enabled) return;
  // Launches being captured into a graph are not timed here.
  hipStreamCaptureStatus capturing = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream_, &capturing) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  if (capturing != hipStreamCaptureStatusNone) return;
  if (hipEventCreate(&start_) != hipSuccess || hipEventRecord(start_, stream_) != hipSuccess) {
    (void)hipGetLastError();
    start_ = nullptr;
  }
}

ProfileScope::~ProfileScope() {
  if (start_ == nullptr) return;
  hipEvent_t stop = nullptr;
  if (hipEventCreate(&stop) != hipSuccess || hipEventRecord(stop, stream_) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipEventDestroy(start_);
    return;
  }
  std::lock_guard<std::mutex> lk(g_profile_mutex);
  pending_timings().push_back({name_, start_, stop});
}
}  // namespace wg

using namespace wg;

extern "C" {

const char* wg_last_error(void) { return last_error_cstr(); }

int wg_abi_version(void) { return 1; }

int wg_device_count(int* count) {
  return guard([&] {
    *count = 0;
    int c = 0;
    const hipError_t r = hipGetDeviceCount(&c);
    if (r != hipSuccess) {
      (void)hipGetLastError();
      // Absence is reported, not raised (reference src/hip/device_pauli_sum.hip.cpp:36-48).
      set_last_error(r == hipErrorNoDevice
                         ? std::string("no HIP device is available")
                         : std::string("HIP runtime library is unavailable or failed during "
                                       "device discovery: ") + hipGetErrorString(r));
      return;
    }
    *count = c;
    if (c == 0) set_last_error("no HIP device is available");
  });
}

int wg_device_info(int ordinal, char* name, size_t name_len, char* gfx_target, size_t gfx_len,
                   size_t* total_memory_bytes, int* compute_units) {
  return guard([&] {
    validate_device(ordinal);
    hipDeviceProp_t prop{};
    check_hip(hipGetDeviceProperties(&prop, ordinal), "get device properties");
    if (name && name_len) {
      std::strncpy(name, prop.name, name_len - 1);
      name[name_len - 1] = '\0';
    }
    if (gfx_target && gfx_len) {
      std::strncpy(gfx_target, prop.gcnArchName, gfx_len - 1);
      gfx_target[gfx_len - 1] = '\0';
    }
    if (total_memory_bytes) *total_memory_bytes = prop.totalGlobalMem;
    if (compute_units) *compute_units = prop.multiProcessorCount;
  });
}

int wg_memory_cache_limit(size_t bytes, size_t* previous) {
  return guard([&] {
    const std::size_t prev = memory_cache_set_limit(bytes);
    if (previous) *previous = prev;
  });
}

int wg_memory_cache_release(int device) {
  return guard([&] {
    validate_device(device);
    memory_cache_release(device);
  });
}

int wg_memory_cache_info(int device, size_t* cached_bytes, size_t* cached_blocks,
                         size_t* live_bytes) {
  return guard([&] {
    std::size_t c = 0, b = 0, l = 0;
    memory_cache_info(device, &c, &b, &l);
    if (cached_bytes) *cached_bytes = c;
    if (cached_blocks) *cached_blocks = b;
    if (live_bytes) *live_bytes = l;
  });
}

int wg_runtime_versions(int* runtime_version, int* driver_version) {
  return guard([&] {
    *runtime_version = 0;
    *driver_version = 0;
    (void)hipRuntimeGetVersion(runtime_version);
    (void)hipDriverGetVersion(driver_version);
    (void)hipGetLastError();
  });
}

int wg_pauli_sum_create(size_t num_qubits, size_t num_terms, const uint64_t* x, const uint64_t* z,
                        const double* coeffs, int device, wg_pauli_sum** out) {
  return guard([&] {
    *out = nullptr;
    validate_device(device);
    const std::size_t W = words_for(num_qubits);
    const std::size_t packed = checked_mul(num_terms, W, "HIP device packed words");
    if (packed != 0 && (x == nullptr || z == nullptr)) fail_invalid("packed mask pointers must be non-null");
    if (num_terms != 0 && coeffs == nullptr) fail_invalid("coefficient pointer must be non-null");
    DeviceScope scope(device);
    std::unique_ptr<wg_pauli_sum> op(new_pauli_sum(num_qubits, num_terms, device));
    if (packed) {
      check_hip(hipMemcpy(op->x, x, packed * sizeof(u64), hipMemcpyHostToDevice),
                "copy host buffer to device");
      check_hip(hipMemcpy(op->z, z, packed * sizeof(u64), hipMemcpyHostToDevice),
                "copy host buffer to device");
    }
    if (num_terms) {
      check_hip(hipMemcpy(op->coeffs, coeffs, num_terms * sizeof(double2), hipMemcpyHostToDevice),
                "copy host buffer to device");
    }
    *out = op.release();
  });
}

int wg_pauli_sum_to_host(const wg_pauli_sum* op, uint64_t* x, uint64_t* z, double* coeffs) {
  return guard([&] {
    require_op(op);
    DeviceScope scope(op->device);
    const std::size_t packed = op->num_terms * op->words;
    if (packed) {
      check_hip(hipMemcpy(x, op->x, packed * sizeof(u64), hipMemcpyDeviceToHost),
                "copy device buffer to host");
      check_hip(hipMemcpy(z, op->z, packed * sizeof(u64), hipMemcpyDeviceToHost),
                "copy device buffer to host");
    }
    if (op->num_terms) {
      check_hip(hipMemcpy(coeffs, op->coeffs, op->num_terms * sizeof(double2),
                          hipMemcpyDeviceToHost),
                "copy device buffer to host");
    }
  });
}

int wg_pauli_sum_info(const wg_pauli_sum* op, size_t* num_qubits, size_t* num_terms, size_t* words,
                      int* device) {
  return guard([&] {
    require_op(op);
    if (num_qubits) *num_qubits = op->num_qubits;
    if (num_terms) *num_terms = op->num_terms;
    if (words) *words = op->words;
    if (device) *device = op->device;
  });
}

int wg_pauli_sum_device_buffers(const wg_pauli_sum* op, const uint64_t** x, const uint64_t** z,
                                const double** coeffs) {
  return guard([&] {
    require_op(op);
    if (x) *x = op->x;
    if (z) *z = op->z;
    if (coeffs) *coeffs = reinterpret_cast<const double*>(op->coeffs);
  });
}

void wg_pauli_sum_destroy(wg_pauli_sum* op) { delete op; }

int wg_profile_enable(int enabled) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(g_profile_mutex);
    g_profile_enabled = enabled != 0;
  });
}

int wg_profile_reset(void) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(g_profile_mutex);
    resolve_pending_locked();
    profile_totals().clear();
  });
}

int wg_profile_get(const char* kernel, double* total_ms, uint64_t* launches) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(g_profile_mutex);
    resolve_pending_locked();
    auto it = profile_totals().find(kernel ? kernel : "");
    *total_ms = it == profile_totals().end() ? 0.0 : it->second.ms;
    *launches = it == profile_totals().end() ? 0 : it->second.launches;
  });
}

}  // extern "C"

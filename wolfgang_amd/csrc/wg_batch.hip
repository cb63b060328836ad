// Batched forward + adjoint over many (psi, H) instances.
//
// All instances' work is cut into units and served from one device work queue
// by a persistent kernel:
//   * forward units (instance, item, chunk of its sweep range) produce partial
//     sums S_z for one x-group slice, as the direct sweep does;
//   * sigma units (instance, range of output amplitudes) produce sigma = H psi
//     (and H^dagger psi) for those outputs.
// Units are ordered by decreasing cost, so the queue behaves as a
// longest-processing-time-first schedule across instances of very different
// sizes. A reduce kernel and a per-instance finalize kernel follow. The whole
// sequence (queue reset, persistent kernel, reduce, finalize) is captured once
// into a HIP graph and replayed; the queue counter is reset by a one-thread
// kernel inside the graph (no memset node).
//
// Instances whose expectation plan is tiled (large states with many x-groups)
// stay out of the queue: the direct sweeps of a unit read psi once per x-group,
// the tiled passes once per pass. They run the per-instance tiled path on the
// same stream after the graph.
//
// Tunables (picked by a timing search at creation unless given): forward chunk
// (amplitude pairs per unit, 2^fwd_log), sigma chunk (outputs per unit) and the
// persistent grid size.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "wg_expect.hpp"
#include "wg_internal.hpp"

namespace wg {
namespace {

constexpr int kBlock = 256;

struct BatchInst {
  const double2* psi;
  double2* sigma;
  double2* sigma_dag;
  double2* term_values;
  const double2* coeffs;
  u64 len;
  u32 group_base;
  u32 num_groups;
  u32 term_base;
  u32 num_terms;
  u32 key_base;
  u32 slot;  // the instance's position in the batch (its energy entry)
};

struct BatchUnit {
  u32 kind;   // 0 forward, 1 sigma
  u32 inst;
  u32 item;   // forward: global item index
  u32 slot;   // forward: partial row
  u64 begin;
  u64 end;
};

__device__ inline u64 insert_zero(u64 v, u64 pivot) {
  const u64 low = pivot - 1;
  return ((v & ~low) << 1) | (v & low);
}

__device__ inline double2 times_i_pow(double2 v, int k) {
  switch (k & 3) {
    case 0: return v;
    case 1: return make_double2(-v.y, v.x);
    case 2: return make_double2(-v.x, -v.y);
    default: return make_double2(v.y, -v.x);
  }
}

__global__ void reset_counter_kernel(u32* counter) { *counter = 0; }

__global__ __launch_bounds__(kBlock) void batch_persistent_kernel(
    const BatchUnit* __restrict__ units, u32 num_units, u32* __restrict__ counter,
    const BatchInst* __restrict__ insts, const ExpItem* __restrict__ items,
    const u64* __restrict__ group_x, const u32* __restrict__ group_begin,
    const u64* __restrict__ key_z, const int* __restrict__ key_y,
    const double2* __restrict__ key_c, double2* __restrict__ partials) {
  __shared__ u32 next;
  __shared__ double2 lds[kBlock / 64][kZ];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1U);
    __syncthreads();
    const u32 uid = next;
    __syncthreads();
    if (uid >= num_units) break;
    const BatchUnit un = units[uid];
    const BatchInst in = insts[un.inst];
    if (un.kind == 0) {
      const ExpItem& it = items[un.item];
      const int nz = it.nz;
      double ar[kZ], ai[kZ];
#pragma unroll
      for (int j = 0; j < kZ; ++j) {
        ar[j] = 0.0;
        ai[j] = 0.0;
      }
      const double2* psi = in.psi;
      if (it.x) {
        const u64 pivot = it.x & (0 - it.x);
        for (u64 v = un.begin + threadIdx.x; v < un.end; v += kBlock) {
          const u64 b = insert_zero(v, pivot);
          const double2 a = psi[b];
          const double2 t = psi[b ^ it.x];
          const double re = fma(t.x, a.x, t.y * a.y);
          const double im = fma(t.x, a.y, -(t.y * a.x));
#pragma unroll
          for (int j = 0; j < kZ; ++j) {
            if (j < nz) {
              const bool odd = __popcll(it.z[j] & b) & 1;
              ar[j] += odd ? -re : re;
              ai[j] += odd ? -im : im;
            }
          }
        }
      } else {
        for (u64 b = un.begin + threadIdx.x; b < un.end; b += kBlock) {
          const double2 a = psi[b];
          const double p = fma(a.x, a.x, a.y * a.y);
#pragma unroll
          for (int j = 0; j < kZ; ++j) {
            if (j < nz) {
              const bool odd = __popcll(it.z[j] & b) & 1;
              ar[j] += odd ? -p : p;
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < kZ; ++j) {
        if (j < nz) {
          double r = ar[j], i = ai[j];
          for (int d = 32; d > 0; d >>= 1) {
            r += __shfl_xor(r, d, 64);
            i += __shfl_xor(i, d, 64);
          }
          if (lane == 0) lds[wave][j] = make_double2(r, i);
        }
      }
      __syncthreads();
      if (threadIdx.x < static_cast<unsigned>(nz)) {
        double r = 0.0, i = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) {
          r += lds[w][threadIdx.x].x;
          i += lds[w][threadIdx.x].y;
        }
        partials[static_cast<std::size_t>(un.slot) * kZ + threadIdx.x] = make_double2(r, i);
      }
      // lds is reused by the next unit only after the barrier at the loop head.
    } else {
      for (u64 o = un.begin + threadIdx.x; o < un.end; o += kBlock) {
        double sr = 0.0, si = 0.0, dr = 0.0, di = 0.0;
        for (u32 g = 0; g < in.num_groups; ++g) {
          const u32 gg = in.group_base + g;
          const u64 src = o ^ group_x[gg];
          const double2 t = in.psi[src];
          for (u32 k = group_begin[gg]; k < group_begin[gg + 1]; ++k) {
            double2 w = times_i_pow(t, key_y[k]);
            if (__popcll(key_z[k] & src) & 1) w = make_double2(-w.x, -w.y);
            const double2 c = key_c[k];
            sr = fma(c.x, w.x, fma(-c.y, w.y, sr));
            si = fma(c.x, w.y, fma(c.y, w.x, si));
            dr = fma(c.x, w.x, fma(c.y, w.y, dr));
            di = fma(c.x, w.y, fma(-c.y, w.x, di));
          }
        }
        if (in.sigma) in.sigma[o] = make_double2(sr, si);
        if (in.sigma_dag) in.sigma_dag[o] = make_double2(dr, di);
      }
    }
  }
}

// Ordered per-(item, slot) reduction into per-key values (global key ids).
__global__ __launch_bounds__(256) void batch_reduce_kernel(const ExpItem* __restrict__ items,
                                                           u32 num_items,
                                                           const u32* __restrict__ slot_begin,
                                                           const double2* __restrict__ partials,
                                                           double* __restrict__ key_vals) {
  const std::size_t gid = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const std::size_t item_id = gid / kZ;
  const int j = static_cast<int>(gid % kZ);
  if (item_id >= num_items) return;
  const ExpItem& it = items[item_id];
  if (j >= it.nz) return;
  double r = 0.0, i = 0.0;
  for (u32 s = slot_begin[item_id]; s < slot_begin[item_id + 1]; ++s) {
    const double2 p = partials[static_cast<std::size_t>(s) * kZ + j];
    r += p.x;
    i += p.y;
  }
  double val;
  if (it.x == 0) {
    val = r;
  } else {
    switch (it.y[j] & 3) {
      case 0: val = 2.0 * r; break;
      case 1: val = -2.0 * i; break;
      case 2: val = -2.0 * r; break;
      default: val = 2.0 * i; break;
    }
  }
  key_vals[it.key[j]] = val;
}

// One block per instance: E_i = sum_t c_t <P_key(t)> and per-term values.
__global__ __launch_bounds__(256) void batch_finalize_kernel(const BatchInst* __restrict__ insts,
                                                             const u32* __restrict__ term_key,
                                                             const double* __restrict__ key_vals,
                                                             double2* __restrict__ energies) {
  __shared__ double2 lds[4];
  const BatchInst in = insts[blockIdx.x];
  double r = 0.0, i = 0.0;
  for (u32 t = threadIdx.x; t < in.num_terms; t += blockDim.x) {
    const double v = key_vals[in.key_base + term_key[in.term_base + t]];
    const double2 c = in.coeffs[t];
    r = fma(c.x, v, r);
    i = fma(c.y, v, i);
    if (in.term_values) in.term_values[t] = make_double2(v, 0.0);
  }
  for (int d = 32; d > 0; d >>= 1) {
    r += __shfl_xor(r, d, 64);
    i += __shfl_xor(i, d, 64);
  }
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = make_double2(r, i);
  __syncthreads();
  if (threadIdx.x == 0) {
    double sr = 0.0, si = 0.0;
    for (int w = 0; w < 4; ++w) {
      sr += lds[w].x;
      si += lds[w].y;
    }
    energies[in.slot] = make_double2(sr, si);
  }
}

template <typename T>
T* upload(const std::vector<T>& v, hipStream_t stream) {
  T* d = device_alloc<T>(std::max<std::size_t>(v.size(), 1));
  if (!v.empty()) {
    check_hip(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream),
              "copy host buffer to device");
  }
  return d;
}

}  // namespace
}  // namespace wg

using namespace wg;

struct wg_batch {
  int device = 0;
  std::vector<std::shared_ptr<ExpectationPlan>> plans;  // keeps instance plans alive
  std::vector<BatchInst> insts;
  std::vector<ExpItem> items;         // direct items, keys offset to global ids
  std::vector<u64> group_x;
  std::vector<u32> group_begin;       // global CSR over keys (flat, group order)
  std::vector<u64> key_z;
  std::vector<int> key_y;
  std::vector<double2> key_c;
  std::vector<u32> term_key;          // per instance, local key ids
  std::size_t total_keys = 0;
  // Device copies.
  BatchInst* d_insts = nullptr;
  ExpItem* d_items = nullptr;
  u64* d_group_x = nullptr;
  u32* d_group_begin = nullptr;
  u64* d_key_z = nullptr;
  int* d_key_y = nullptr;
  double2* d_key_c = nullptr;
  u32* d_term_key = nullptr;
  double* d_key_vals = nullptr;
  double2* d_energies = nullptr;      // caller-owned
  u32* d_counter = nullptr;
  // Schedule for the current tunables.
  int fwd_log = 16;
  u32 sigma_chunk = 4096;
  u32 grid = 0;
  std::size_t num_units = 0;
  BatchUnit* d_units = nullptr;
  u32* d_slot_begin = nullptr;
  double2* d_partials = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void free_schedule() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    device_free(d_units);
    device_free(d_slot_begin);
    device_free(d_partials);
    d_units = nullptr;
    d_slot_begin = nullptr;
    d_partials = nullptr;
  }

  ~wg_batch() {
    int prev = -1;
    const hipError_t r = hipGetDevice(&prev);
    (void)hipSetDevice(device);
    free_schedule();
    device_free(d_insts);
    device_free(d_items);
    device_free(d_group_x);
    device_free(d_group_begin);
    device_free(d_key_z);
    device_free(d_key_y);
    device_free(d_key_c);
    device_free(d_term_key);
    device_free(d_key_vals);
    device_free(d_counter);
    if (r == hipSuccess && prev >= 0) (void)hipSetDevice(prev);
  }

  // Builds units for (fwd_log, sigma_chunk) and uploads them.
  void schedule(int fwd_log_, u32 sigma_chunk_, u32 grid_, hipStream_t stream) {
    free_schedule();
    fwd_log = fwd_log_;
    sigma_chunk = sigma_chunk_;
    grid = grid_;
    struct Costed {
      double cost;
      BatchUnit u;
    };
    std::vector<Costed> work;
    std::vector<u32> slot_begin(items.size() + 1, 0);
    u32 slot = 0;
    const u64 fchunk = u64{1} << fwd_log;
    for (std::size_t i = 0, inst = 0; i < items.size(); ++i) {
      const ExpItem& it = items[i];
      // Items are stored instance by instance.
      while (inst + 1 < insts.size() && item_first[inst + 1] <= i) ++inst;
      const u64 range = it.x ? (insts[inst].len >> 1) : insts[inst].len;
      slot_begin[i] = slot;
      for (u64 b = 0; b < range; b += fchunk) {
        BatchUnit u{};
        u.kind = 0;
        u.inst = static_cast<u32>(inst);
        u.item = static_cast<u32>(i);
        u.slot = slot++;
        u.begin = b;
        u.end = std::min(range, b + fchunk);
        work.push_back({static_cast<double>(u.end - u.begin) * (4.0 + it.nz), u});
      }
    }
    slot_begin[items.size()] = slot;
    for (std::size_t inst = 0; inst < insts.size(); ++inst) {
      const BatchInst& in = insts[inst];
      if (!in.sigma && !in.sigma_dag) continue;
      const u32 keys = group_begin[in.group_base + in.num_groups] - group_begin[in.group_base];
      // Outputs per unit scaled down for operators with many keys.
      u64 chunk = sigma_chunk;
      while (chunk > 256 && chunk * (keys + 1) > (u64{1} << 22)) chunk >>= 1;
      for (u64 o = 0; o < in.len; o += chunk) {
        BatchUnit u{};
        u.kind = 1;
        u.inst = static_cast<u32>(inst);
        u.begin = o;
        u.end = std::min<u64>(in.len, o + chunk);
        work.push_back({static_cast<double>(u.end - u.begin) * (2.0 + 3.0 * keys), u});
      }
    }
    std::stable_sort(work.begin(), work.end(),
                     [](const Costed& a, const Costed& b) { return a.cost > b.cost; });
    std::vector<BatchUnit> units;
    units.reserve(work.size());
    for (const auto& w : work) units.push_back(w.u);
    num_units = units.size();
    d_units = upload(units, stream);
    d_slot_begin = upload(slot_begin, stream);
    d_partials = device_alloc<double2>(std::max<std::size_t>(1, static_cast<std::size_t>(slot) * kZ));
    sync(stream);
  }

  void launch(hipStream_t stream) {
    if (insts.empty()) return;
    reset_counter_kernel<<<1, 1, 0, stream>>>(d_counter);
    batch_persistent_kernel<<<grid, kBlock, 0, stream>>>(
        d_units, static_cast<u32>(num_units), d_counter, d_insts, d_items, d_group_x,
        d_group_begin, d_key_z, d_key_y, d_key_c, d_partials);
    if (!items.empty()) {
      batch_reduce_kernel<<<grid_for(items.size() * kZ, 256), 256, 0, stream>>>(
          d_items, static_cast<u32>(items.size()), d_slot_begin, d_partials, d_key_vals);
    }
    batch_finalize_kernel<<<static_cast<unsigned>(insts.size()), 256, 0, stream>>>(
        d_insts, d_term_key, d_key_vals, d_energies);
  }

  std::vector<std::size_t> item_first;  // first item of each instance

  // Instances whose plan is tiled (the largest ones) run the tiled sweeps of
  // the per-instance path after the graph: the queue's direct sweeps would
  // read psi once per x-group.
  struct TiledInst {
    wg_pauli_sum* op;
    const double* psi;
    const double* coeffs;
    double* term_values;
    double* sigma;
    double* sigma_dag;
    u32 slot;
  };
  std::vector<TiledInst> tiled;
};

namespace {

// One timed launch of the batch (after a warm-up launch when `warm`).
float time_launch(wg_batch* b, hipStream_t stream, bool warm) {
  hipEvent_t e0, e1;
  check_hip(hipEventCreate(&e0), "create event");
  check_hip(hipEventCreate(&e1), "create event");
  if (warm) b->launch(stream);
  check_hip(hipEventRecord(e0, stream), "record event");
  b->launch(stream);
  check_hip(hipEventRecord(e1, stream), "record event");
  check_hip(hipEventSynchronize(e1), "synchronize event");
  float ms = 0.0f;
  check_hip(hipEventElapsedTime(&ms, e0, e1), "event elapsed time");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  check_launch("batch");
  return ms;
}

void capture(wg_batch* b, hipStream_t stream) {
  hipStream_t cs;
  check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "create capture stream");
  check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
  b->launch(cs);
  hipGraph_t g = nullptr;
  const hipError_t r = hipStreamEndCapture(cs, &g);
  (void)hipStreamDestroy(cs);
  check_hip(r, "end capture");
  b->graph = g;
  check_hip(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "instantiate graph");
  (void)stream;
}

}  // namespace

extern "C" {

int wg_batch_create(size_t count, const wg_pauli_sum* const* ops, const double* const* psi,
                    const double* const* coeffs, double* const* term_values,
                    double* const* sigma, double* const* sigma_dagger, double* energies,
                    const int* tunables, void* stream_ptr, wg_batch** out) {
  return guard([&] {
    *out = nullptr;
    if (count == 0) fail_invalid("batch must contain at least one instance");
    if (ops == nullptr || psi == nullptr || energies == nullptr) {
      fail_invalid("batch operator, statevector and energy pointers must be non-null");
    }
    for (std::size_t i = 0; i < count; ++i) {
      require_op(ops[i]);
      if (ops[i]->device != ops[0]->device) {
        fail_invalid("batched instances must share one device");
      }
      if (ops[i]->num_qubits > 40) fail_invalid("batched expectation supports num_qubits <= 40");
      if (psi[i] == nullptr) fail_invalid("HIP statevector data pointer must be non-null");
    }
    const int device = ops[0]->device;
    DeviceScope scope(device);
    hipStream_t stream = as_stream(stream_ptr);
    std::unique_ptr<wg_batch> b(new wg_batch());
    b->device = device;
    b->d_energies = reinterpret_cast<double2*>(energies);
    for (std::size_t i = 0; i < count; ++i) {
      auto* op = const_cast<wg_pauli_sum*>(ops[i]);
      auto plan = get_plan(op, stream);
      b->plans.push_back(plan);
      if (plan->tiled && op->num_terms > 0) {
        b->tiled.push_back({op, psi[i], coeffs ? coeffs[i] : nullptr,
                            term_values ? term_values[i] : nullptr, sigma ? sigma[i] : nullptr,
                            sigma_dagger ? sigma_dagger[i] : nullptr, static_cast<u32>(i)});
        continue;
      }
      BatchInst in{};
      in.slot = static_cast<u32>(i);
      in.psi = reinterpret_cast<const double2*>(psi[i]);
      in.len = u64{1} << op->num_qubits;
      in.coeffs = coeffs && coeffs[i] ? reinterpret_cast<const double2*>(coeffs[i]) : op->coeffs;
      in.term_values = term_values ? reinterpret_cast<double2*>(term_values[i]) : nullptr;
      in.sigma = sigma ? reinterpret_cast<double2*>(sigma[i]) : nullptr;
      in.sigma_dag = sigma_dagger ? reinterpret_cast<double2*>(sigma_dagger[i]) : nullptr;
      in.group_base = static_cast<u32>(b->group_x.size());
      in.num_groups = static_cast<u32>(plan->h_group_x.size());
      in.term_base = static_cast<u32>(b->term_key.size());
      in.num_terms = static_cast<u32>(op->num_terms);
      in.key_base = static_cast<u32>(b->total_keys);
      const u32 flat_base = static_cast<u32>(b->key_z.size());
      b->item_first.push_back(b->items.size());
      for (ExpItem it : plan->direct_items) {
        for (int j = 0; j < it.nz; ++j) it.key[j] += in.key_base;
        b->items.push_back(it);
      }
      for (std::size_t g = 0; g < plan->h_group_x.size(); ++g) {
        b->group_x.push_back(plan->h_group_x[g]);
        b->group_begin.push_back(flat_base + plan->h_group_begin[g]);
      }
      b->key_z.insert(b->key_z.end(), plan->h_key_z.begin(), plan->h_key_z.end());
      b->key_y.insert(b->key_y.end(), plan->h_key_y.begin(), plan->h_key_y.end());
      if (coeffs && coeffs[i]) {
        // Per-key sums of the given coefficients, keys in group order.
        std::vector<double2> hc(op->num_terms);
        check_hip(hipMemcpy(hc.data(), coeffs[i], op->num_terms * sizeof(double2),
                            hipMemcpyDeviceToHost),
                  "copy device buffer to host");
        std::vector<double2> kc(plan->num_keys, make_double2(0, 0));
        std::vector<char> seen(plan->num_keys, 0);
        for (std::size_t t = 0; t < op->num_terms; ++t) {
          const u32 k = plan->h_term_key[t];
          if (!seen[k]) {
            kc[k] = hc[t];
            seen[k] = 1;
          } else {
            kc[k].x += hc[t].x;
            kc[k].y += hc[t].y;
          }
        }
        // Map first-appearance key ids to group order through the items.
        std::vector<double2> flat(plan->h_key_z.size());
        std::size_t pos = 0;
        for (const ExpItem& it : plan->direct_items) {
          for (int j = 0; j < it.nz; ++j) flat[pos++] = kc[it.key[j]];
        }
        b->key_c.insert(b->key_c.end(), flat.begin(), flat.end());
      } else {
        b->key_c.insert(b->key_c.end(), plan->h_key_c.begin(), plan->h_key_c.end());
      }
      b->term_key.insert(b->term_key.end(), plan->h_term_key.begin(), plan->h_term_key.end());
      b->total_keys += plan->num_keys;
      b->insts.push_back(in);
    }
    b->group_begin.push_back(static_cast<u32>(b->key_z.size()));
    b->d_insts = upload(b->insts, stream);
    b->d_items = upload(b->items, stream);
    b->d_group_x = upload(b->group_x, stream);
    b->d_group_begin = upload(b->group_begin, stream);
    b->d_key_z = upload(b->key_z, stream);
    b->d_key_y = upload(b->key_y, stream);
    b->d_key_c = upload(b->key_c, stream);
    b->d_term_key = upload(b->term_key, stream);
    b->d_key_vals = device_alloc<double>(std::max<std::size_t>(1, b->total_keys));
    b->d_counter = device_alloc<u32>(1);
    sync(stream);

    hipDeviceProp_t prop{};
    check_hip(hipGetDeviceProperties(&prop, device), "get device properties");
    const u32 cus = static_cast<u32>(prop.multiProcessorCount);
    if (b->insts.empty()) {
      // Every instance is tiled: nothing for the queue.
    } else if (tunables) {
      b->schedule(tunables[0], static_cast<u32>(tunables[1]), static_cast<u32>(tunables[2]), stream);
    } else {
      // Timing search on the real batch, one tunable at a time (forward
      // chunk, then sigma chunk, then grid), each from the best so far.
      int best_f = 16;
      u32 best_s = 4096, best_g = cus * 8;
      b->schedule(best_f, best_s, best_g, stream);
      float best = time_launch(b.get(), stream, true);
      auto trial = [&](int f, u32 s, u32 g) {
        if (f == best_f && s == best_s && g == best_g) return;
        b->schedule(f, s, g, stream);
        const float ms = time_launch(b.get(), stream, best < 50.0f);
        if (ms < best) {
          best = ms;
          best_f = f;
          best_s = s;
          best_g = g;
        }
      };
      for (int f : {14, 18}) trial(f, best_s, best_g);
      const int f_pick = best_f;
      for (u32 s : {1024u, 16384u}) trial(f_pick, s, best_g);
      const u32 s_pick = best_s;
      for (u32 g : {cus * 4, cus * 16}) trial(f_pick, s_pick, g);
      b->schedule(best_f, best_s, best_g, stream);
    }
    if (!b->insts.empty()) capture(b.get(), stream);
    *out = b.release();
  });
}

int wg_batch_run(wg_batch* b, void* stream_ptr) {
  return guard([&] {
    if (b == nullptr) fail_runtime("batch is empty or destroyed");
    DeviceScope scope(b->device);
    hipStream_t stream = as_stream(stream_ptr);
    if (b->exec) {
      ProfileScope prof("batch_graph", stream);
      check_hip(hipGraphLaunch(b->exec, stream), "launch graph");
    }
    // Tiled instances behind the graph on the same stream, one at a time
    // (each synchronizes before its scratch buffers are released).
    for (const auto& t : b->tiled) {
      Scratch scratch(b->device);
      adjoint_enqueue(t.op, t.psi, std::size_t{1} << t.op->num_qubits, t.coeffs, t.term_values,
                      t.sigma, t.sigma_dag, stream, scratch, b->d_energies + t.slot);
      sync(stream);
    }
  });
}

int wg_batch_sync(wg_batch* b, void* stream_ptr) {
  return guard([&] {
    if (b == nullptr) fail_runtime("batch is empty or destroyed");
    DeviceScope scope(b->device);
    sync(as_stream(stream_ptr));
  });
}

int wg_batch_info(const wg_batch* b, int* tunables, size_t* units, size_t* items) {
  return guard([&] {
    if (b == nullptr) fail_runtime("batch is empty or destroyed");
    if (tunables) {
      tunables[0] = b->fwd_log;
      tunables[1] = static_cast<int>(b->sigma_chunk);
      tunables[2] = static_cast<int>(b->grid);
    }
    if (units) *units = b->num_units;
    if (items) *items = b->items.size();
  });
}

void wg_batch_destroy(wg_batch* b) { delete b; }

}  // extern "C"

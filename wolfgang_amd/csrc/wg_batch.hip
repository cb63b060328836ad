// Batched forward + adjoint over many (psi, H) instances, one graph replay.
//
// Instances whose expectation plan is direct (small states, or few x-groups)
// are cut into units served from one device work queue by a persistent
// kernel:
//   * forward units (instance, item, chunk of its sweep range) produce partial
//     sums S_z for one x-group slice, as the direct sweep does;
//   * sigma units (instance, range of output amplitudes) produce sigma = H psi
//     (and H^dagger psi) for those outputs.
// Units are ordered by decreasing cost, so the queue behaves as a
// longest-processing-time-first schedule across instances of very different
// sizes. The per-key coefficients of these instances are recomputed from the
// callers' coefficient buffers at the start of every replay, then the queue
// counter is reset by a one-thread kernel (no memset node), and a reduce and a
// per-instance finalize kernel follow the persistent kernel.
//
// Instances whose plan is tiled (18 qubits and up with many x-groups) run the
// tiled sweeps (forward owner sweep, then one sigma launch per pass) with
// their temporary buffers carved from one allocation made at creation. Their
// launch sequences are spread over `streams` worker streams,
// longest-estimated-first onto the least loaded stream, forked from and
// joined back to the caller's stream by events, so instances too small to
// fill the GPU on their own (an 18-qubit pass has 64 tiles) run side by side.
// That fork/join is captured into the same HIP graph as the queue (one
// hipGraphLaunch per run, no host synchronization), or, when timing at
// creation shows it faster, enqueued on the worker streams at each run.
//
// Tunables (picked by timing the actual batch at creation unless given):
// forward chunk (amplitude pairs per unit, 2^fwd_log), sigma chunk (outputs per
// unit), persistent grid, worker streams for the tiled instances, and whether
// the tiled fork/join is inside the graph.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <thread>
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

// Per-key coefficient sums of the queued instances, recomputed from the
// instances' coefficient buffers on every replay (keys in the flat group
// order of the batch, terms of a key in term order: the same additions as the
// per-instance plan, so the same bits).
__global__ __launch_bounds__(256) void batch_key_coeff_kernel(const BatchInst* __restrict__ insts,
                                                              const u32* __restrict__ key_inst,
                                                              const u32* __restrict__ key_term_begin,
                                                              const u32* __restrict__ key_terms,
                                                              u32 num_keys, double2* __restrict__ key_c) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= num_keys) return;
  const double2* c = insts[key_inst[k]].coeffs;
  const u32 b = key_term_begin[k], e = key_term_begin[k + 1];
  double2 acc = c[key_terms[b]];
  for (u32 i = b + 1; i < e; ++i) {
    acc.x += c[key_terms[i]].x;
    acc.y += c[key_terms[i]].y;
  }
  key_c[k] = acc;
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

constexpr int kMaxTiledStreams = 4;

struct wg_batch {
  int device = 0;
  std::vector<std::shared_ptr<ExpectationPlan>> plans;  // keeps instance plans alive
  // ---- instances with a direct plan: the persistent queue
  std::vector<BatchInst> insts;
  std::vector<std::size_t> item_first;  // first item of each queued instance
  std::vector<ExpItem> items;           // direct items, keys offset to global ids
  std::vector<u64> group_x;
  std::vector<u32> group_begin;         // global CSR over keys (flat, group order)
  std::vector<u64> key_z;
  std::vector<int> key_y;
  std::vector<u32> key_inst;            // queued instance of each flat key
  std::vector<u32> key_term_begin;      // global CSR: flat key -> its terms
  std::vector<u32> key_terms;           // term indices local to the instance
  std::vector<u32> term_key;            // per instance, local key ids
  std::size_t total_keys = 0;
  BatchInst* d_insts = nullptr;
  ExpItem* d_items = nullptr;
  u64* d_group_x = nullptr;
  u32* d_group_begin = nullptr;
  u64* d_key_z = nullptr;
  int* d_key_y = nullptr;
  double2* d_key_c = nullptr;           // written by batch_key_coeff_kernel every replay
  u32* d_key_inst = nullptr;
  u32* d_key_term_begin = nullptr;
  u32* d_key_terms = nullptr;
  u32* d_term_key = nullptr;
  double* d_key_vals = nullptr;
  double2* d_energies = nullptr;        // caller-owned
  u32* d_counter = nullptr;
  int fwd_log = 16;
  u32 sigma_chunk = 4096;
  u32 grid = 0;
  std::size_t num_units = 0;
  BatchUnit* d_units = nullptr;
  u32* d_slot_begin = nullptr;
  double2* d_partials = nullptr;
  // ---- instances with a tiled plan
  struct TiledInst {
    const ExpectationPlan* plan;
    const double2* op_coeffs;
    const double2* psi;
    const double2* coeffs;  // caller override or null
    double2* term_values;
    double2* sigma;
    double2* sigma_dag;
    u32 slot;
    std::size_t buf_offset;
    double cost;
  };
  std::vector<TiledInst> tiled;  // longest-estimated first
  void* d_tiled_buf = nullptr;
  int streams = 1;
  bool tiled_in_graph = true;
  hipStream_t worker[kMaxTiledStreams] = {};
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join[kMaxTiledStreams] = {};
  // ---- the graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void free_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }

  void free_schedule() {
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
    free_graph();
    free_schedule();
    for (void* p : {static_cast<void*>(d_insts), static_cast<void*>(d_items),
                    static_cast<void*>(d_group_x), static_cast<void*>(d_group_begin),
                    static_cast<void*>(d_key_z), static_cast<void*>(d_key_y),
                    static_cast<void*>(d_key_c), static_cast<void*>(d_key_inst),
                    static_cast<void*>(d_key_term_begin), static_cast<void*>(d_key_terms),
                    static_cast<void*>(d_term_key), static_cast<void*>(d_key_vals),
                    static_cast<void*>(d_counter), d_tiled_buf}) {
      device_free(p);
    }
    for (hipStream_t& s : worker) {
      if (s) (void)hipStreamDestroy(s);
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    for (hipEvent_t& e : join) {
      if (e) (void)hipEventDestroy(e);
    }
    if (r == hipSuccess && prev >= 0) (void)hipSetDevice(prev);
  }

  // Builds the queue's units for (fwd_log, sigma_chunk) and uploads them.
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

  // The queued instances: key coefficients, queue reset, persistent kernel,
  // reduce, finalize.
  void launch_queue(hipStream_t stream) {
    if (insts.empty()) return;
    if (total_keys) {
      batch_key_coeff_kernel<<<grid_for(key_z.size(), 256), 256, 0, stream>>>(
          d_insts, d_key_inst, d_key_term_begin, d_key_terms, static_cast<u32>(key_z.size()), d_key_c);
      check_launch("batch_key_coeff");
    }
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
    check_launch("batch");
  }

  // The tiled instances run on `nstreams` worker streams (longest first onto
  // the least loaded one), forked from `stream` before the queue's launches
  // and joined back after them, so the queue and the tiled sweeps overlap.
  void fork(hipStream_t stream, int nstreams) {
    check_hip(hipEventRecord(fork_ev, stream), "record fork event");
    for (int w = 0; w < nstreams; ++w) {
      check_hip(hipStreamWaitEvent(worker[w], fork_ev, 0), "wait fork event");
    }
  }

  void enqueue_tiled(int nstreams) {
    std::vector<double> load(nstreams, 0.0);
    for (const TiledInst& t : tiled) {
      const int w = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
      load[w] += t.cost;
      tiled_adjoint_enqueue(*t.plan, t.op_coeffs, t.psi, t.coeffs, t.term_values, t.sigma,
                            t.sigma_dag, d_energies + t.slot,
                            static_cast<unsigned char*>(d_tiled_buf) + t.buf_offset, worker[w]);
    }
  }

  void join_back(hipStream_t stream, int nstreams) {
    for (int w = 0; w < nstreams; ++w) {
      check_hip(hipEventRecord(join[w], worker[w]), "record join event");
      check_hip(hipStreamWaitEvent(stream, join[w], 0), "wait join event");
    }
  }

  // Queue launches on `stream` (when `queue`) and, when `with_tiled`, the
  // tiled instances on the worker streams, all ordered after the work already
  // on `stream` and before what follows it.
  void launch(hipStream_t stream, bool queue, bool with_tiled, int nstreams) {
    const bool t = with_tiled && !tiled.empty();
    if (t) fork(stream, nstreams);
    if (queue) launch_queue(stream);
    if (t) {
      enqueue_tiled(nstreams);
      join_back(stream, nstreams);
    }
  }

  void capture(bool with_tiled) {
    free_graph();
    hipStream_t cs;
    check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "create capture stream");
    check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
    launch(cs, true, with_tiled, streams);
    hipGraph_t g = nullptr;
    const hipError_t r = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    check_hip(r, "end capture");
    graph = g;
    check_hip(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0), "instantiate graph");
  }

  void run(hipStream_t stream) {
    if (tiled_in_graph || tiled.empty()) {
      if (exec) check_hip(hipGraphLaunch(exec, stream), "launch graph");
      return;
    }
    fork(stream, streams);
    if (exec) check_hip(hipGraphLaunch(exec, stream), "launch graph");
    enqueue_tiled(streams);
    join_back(stream, streams);
  }
};

namespace {

// Milliseconds of one call of fn(stream) after `warm` untimed calls.
template <typename Fn>
float time_ms(hipStream_t stream, int warm, Fn&& fn) {
  hipEvent_t e0, e1;
  check_hip(hipEventCreate(&e0), "create event");
  check_hip(hipEventCreate(&e1), "create event");
  for (int i = 0; i < warm; ++i) fn(stream);
  check_hip(hipEventRecord(e0, stream), "record event");
  fn(stream);
  check_hip(hipEventRecord(e1, stream), "record event");
  check_hip(hipEventSynchronize(e1), "synchronize event");
  float ms = 0.0f;
  check_hip(hipEventElapsedTime(&ms, e0, e1), "event elapsed time");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  check_launch("batch");
  return ms;
}

}  // namespace

namespace wg {
namespace {

// The instances' expectation plans are independent host work (planning, then
// uploads): the ones not built yet are built on up to 16 host threads, each
// with a stream of its own. Errors are rethrown on the calling thread.
void build_plans_parallel(const wg_pauli_sum* const* ops, std::size_t count, int device) {
  std::vector<wg_pauli_sum*> todo;
  for (std::size_t i = 0; i < count; ++i) {
    auto* op = const_cast<wg_pauli_sum*>(ops[i]);
    std::lock_guard<std::mutex> lk(op->plan_mutex);
    if (!op->plan && op->num_terms > 0) todo.push_back(op);
  }
  std::sort(todo.begin(), todo.end());
  todo.erase(std::unique(todo.begin(), todo.end()), todo.end());
  const std::size_t hw = std::max(1U, std::thread::hardware_concurrency());
  const std::size_t nthreads = std::min<std::size_t>({todo.size(), hw, 16});
  if (nthreads < 2) return;  // the caller's loop builds them
  std::atomic<std::size_t> next{0};
  std::vector<std::exception_ptr> errors(nthreads);
  std::vector<std::thread> pool;
  pool.reserve(nthreads);
  for (std::size_t t = 0; t < nthreads; ++t) {
    pool.emplace_back([&, t] {
      try {
        DeviceScope scope(device);
        hipStream_t st = nullptr;
        check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "create stream");
        struct StreamGuard {
          hipStream_t s;
          ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard_stream{st};
        for (;;) {
          const std::size_t k = next.fetch_add(1);
          if (k >= todo.size()) break;
          (void)get_plan(todo[k], st);
        }
      } catch (...) {
        errors[t] = std::current_exception();
      }
    });
  }
  for (std::thread& th : pool) th.join();
  for (const std::exception_ptr& e : errors) {
    if (e) std::rethrow_exception(e);
  }
}

}  // namespace
}  // namespace wg

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
    if (tunables && (tunables[0] < 8 || tunables[0] > 30 || tunables[1] < 1 || tunables[2] < 1 ||
                     tunables[3] < 1 || tunables[3] > kMaxTiledStreams)) {
      fail_invalid("batch tunables out of range");
    }
    const int device = ops[0]->device;
    DeviceScope scope(device);
    hipStream_t stream = as_stream(stream_ptr);
    std::unique_ptr<wg_batch> b(new wg_batch());
    b->device = device;
    b->d_energies = reinterpret_cast<double2*>(energies);
    build_plans_parallel(ops, count, device);
    std::size_t tiled_bytes = 0;
    for (std::size_t i = 0; i < count; ++i) {
      auto* op = const_cast<wg_pauli_sum*>(ops[i]);
      auto plan = get_plan(op, stream);
      b->plans.push_back(plan);
      const double2* ci = coeffs && coeffs[i] ? reinterpret_cast<const double2*>(coeffs[i]) : nullptr;
      if (plan->tiled && op->num_terms > 0) {
        wg_batch::TiledInst t{};
        t.plan = plan.get();
        t.op_coeffs = op->coeffs;
        t.psi = reinterpret_cast<const double2*>(psi[i]);
        t.coeffs = ci;
        t.term_values = term_values ? reinterpret_cast<double2*>(term_values[i]) : nullptr;
        t.sigma = sigma ? reinterpret_cast<double2*>(sigma[i]) : nullptr;
        t.sigma_dag = sigma_dagger ? reinterpret_cast<double2*>(sigma_dagger[i]) : nullptr;
        t.slot = static_cast<u32>(i);
        t.buf_offset = tiled_bytes;
        tiled_bytes += (tiled_adjoint_bytes(*plan) + 255) & ~std::size_t{255};
        // Work estimate: psi (and sigma) streamed once per pass.
        t.cost = static_cast<double>(plan->passes.size()) * static_cast<double>(u64{1} << op->num_qubits);
        b->tiled.push_back(t);
        continue;
      }
      BatchInst in{};
      in.slot = static_cast<u32>(i);
      in.psi = reinterpret_cast<const double2*>(psi[i]);
      in.len = u64{1} << op->num_qubits;
      in.coeffs = ci ? ci : op->coeffs;
      in.term_values = term_values ? reinterpret_cast<double2*>(term_values[i]) : nullptr;
      in.sigma = sigma ? reinterpret_cast<double2*>(sigma[i]) : nullptr;
      in.sigma_dag = sigma_dagger ? reinterpret_cast<double2*>(sigma_dagger[i]) : nullptr;
      in.group_base = static_cast<u32>(b->group_x.size());
      in.num_groups = static_cast<u32>(plan->h_group_x.size());
      in.term_base = static_cast<u32>(b->term_key.size());
      in.num_terms = static_cast<u32>(op->num_terms);
      in.key_base = static_cast<u32>(b->total_keys);
      const u32 flat_base = static_cast<u32>(b->key_z.size());
      const u32 inst = static_cast<u32>(b->insts.size());
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
      // Flat keys of this instance: the terms to sum per replay.
      const u32 term_off = static_cast<u32>(b->key_terms.size());
      for (std::size_t k = 0; k < plan->h_key_z.size(); ++k) {
        b->key_inst.push_back(inst);
        b->key_term_begin.push_back(term_off + plan->h_key_term_begin[k]);
      }
      b->key_terms.insert(b->key_terms.end(), plan->h_key_terms.begin(), plan->h_key_terms.end());
      b->term_key.insert(b->term_key.end(), plan->h_term_key.begin(), plan->h_term_key.end());
      b->total_keys += plan->num_keys;
      b->insts.push_back(in);
    }
    b->group_begin.push_back(static_cast<u32>(b->key_z.size()));
    b->key_term_begin.push_back(static_cast<u32>(b->key_terms.size()));
    // Longest-estimated tiled instances first.
    std::stable_sort(b->tiled.begin(), b->tiled.end(),
                     [](const wg_batch::TiledInst& x, const wg_batch::TiledInst& y) { return x.cost > y.cost; });
    b->d_insts = upload(b->insts, stream);
    b->d_items = upload(b->items, stream);
    b->d_group_x = upload(b->group_x, stream);
    b->d_group_begin = upload(b->group_begin, stream);
    b->d_key_z = upload(b->key_z, stream);
    b->d_key_y = upload(b->key_y, stream);
    b->d_key_inst = upload(b->key_inst, stream);
    b->d_key_term_begin = upload(b->key_term_begin, stream);
    b->d_key_terms = upload(b->key_terms, stream);
    b->d_term_key = upload(b->term_key, stream);
    b->d_key_c = device_alloc<double2>(std::max<std::size_t>(1, b->key_z.size()));
    b->d_key_vals = device_alloc<double>(std::max<std::size_t>(1, b->total_keys));
    b->d_counter = device_alloc<u32>(1);
    if (tiled_bytes) b->d_tiled_buf = device_alloc<unsigned char>(tiled_bytes);
    for (int w = 0; w < kMaxTiledStreams; ++w) {
      check_hip(hipStreamCreateWithFlags(&b->worker[w], hipStreamNonBlocking), "create worker stream");
      check_hip(hipEventCreateWithFlags(&b->join[w], hipEventDisableTiming), "create event");
    }
    check_hip(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming), "create event");
    sync(stream);

    hipDeviceProp_t prop{};
    check_hip(hipGetDeviceProperties(&prop, device), "get device properties");
    const u32 cus = static_cast<u32>(prop.multiProcessorCount);
    if (tunables) {
      b->streams = tunables[3];
      b->tiled_in_graph = tunables[4] != 0;
      if (!b->insts.empty()) {
        b->schedule(tunables[0], static_cast<u32>(tunables[1]), static_cast<u32>(tunables[2]), stream);
      }
    } else {
      if (!b->insts.empty()) {
        // Timing search on the queued part, one tunable at a time (forward
        // chunk, then sigma chunk, then grid), each from the best so far.
        int best_f = 16;
        u32 best_s = 4096, best_g = cus * 8;
        b->schedule(best_f, best_s, best_g, stream);
        auto queue = [&](hipStream_t s) { b->launch(s, true, false, 1); };
        float best = time_ms(stream, 1, queue);
        auto trial = [&](int f, u32 s, u32 g) {
          if (f == best_f && s == best_s && g == best_g) return;
          b->schedule(f, s, g, stream);
          const float ms = time_ms(stream, best < 50.0f ? 1 : 0, queue);
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
      if (!b->tiled.empty()) {
        // Worker streams for the tiled instances: 1, 2 or 4, timed directly
        // (one warm-up run first).
        float best = 0.0f;
        for (int k : {1, 2, 4}) {
          const float ms =
              time_ms(stream, k == 1 ? 1 : 0, [&](hipStream_t s) { b->launch(s, false, true, k); });
          if (k == 1 || ms < best) {
            best = ms;
            b->streams = k;
          }
        }
        // Inside the graph or enqueued per run: time whole runs both ways.
        b->tiled_in_graph = true;
        b->capture(true);
        const float in_graph = time_ms(stream, 0, [&](hipStream_t s) { b->run(s); });
        b->tiled_in_graph = false;
        b->capture(false);
        const float outside = time_ms(stream, 0, [&](hipStream_t s) { b->run(s); });
        b->tiled_in_graph = in_graph <= outside;
      }
    }
    if (!b->insts.empty() || b->tiled_in_graph) b->capture(b->tiled_in_graph);
    sync(stream);
    *out = b.release();
  });
}

int wg_batch_run(wg_batch* b, void* stream_ptr) {
  return guard([&] {
    if (b == nullptr) fail_runtime("batch is empty or destroyed");
    DeviceScope scope(b->device);
    hipStream_t stream = as_stream(stream_ptr);
    ProfileScope prof("batch_graph", stream);
    b->run(stream);
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
      tunables[3] = b->streams;
      tunables[4] = b->tiled_in_graph ? 1 : 0;
    }
    if (units) *units = b->num_units;
    if (items) *items = b->items.size();
  });
}

int wg_batch_tiled_count(const wg_batch* b, size_t* tiled) {
  return guard([&] {
    if (b == nullptr) fail_runtime("batch is empty or destroyed");
    *tiled = b->tiled.size();
  });
}

void wg_batch_destroy(wg_batch* b) { delete b; }

}  // extern "C"

// Statevector expectation E = <psi|H|psi> and its adjoint.
//
// Reference semantics (src/expectation.cpp:37-155): for P = i^{|x&z|} X^x Z^z,
//   <P> = sum_{b : b & pivot = 0} 2 Re( i^{|x&z|} (-1)^{|z&b|} conj(psi[b^x]) psi[b] )
// with pivot any set bit of x (the pair {b, b^x} contributes a value and its
// conjugate), and <P> = sum_b (-1)^{|z&b|} |psi[b]|^2 for x = 0.
//
// Device plan (built once per operator on the host, cached on the handle):
//   * distinct (x, z) keys, with every term mapped to its key;
//   * keys grouped by x-mask; a group's keys share the amplitude products
//     a_b = conj(psi[b^x]) psi[b], so one streaming pass over psi per x-group
//     serves all of its z-masks ("items" hold up to kZ keys of one group);
//   * per item, per block, partial sums S_z = sum_b (-1)^{|z&b|} a_b written to
//     a partial buffer; a second kernel reduces them in a fixed order and a
//     third forms E = sum_t c_t <P_key(t)> in a fixed tree order, so results
//     are bit-identical from call to call.
//
// Adjoint: sigma = H psi = dE/dpsi* with (P psi)[b ^ x] = i^{|x&z|}
// (-1)^{|z&b|} psi[b]; each output amplitude is owned by one thread which
// gathers psi[o ^ x] once per x-group and applies all of the group's keys.
// dE/dc_k = <P_k> comes from the forward plan.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "wg_expect.hpp"
#include "wg_internal.hpp"

namespace wg {

namespace {

constexpr int kBlock = 256;

__device__ inline u64 insert_zero(u64 v, u64 pivot) {
  const u64 low = pivot - 1;
  return ((v & ~low) << 1) | (v & low);
}

// ------------------------------------------------------------ forward pass

template <int ZC>
__global__ __launch_bounds__(kBlock) void expect_direct_kernel(const double2* __restrict__ psi,
                                                               u64 len,
                                                               const ExpItem* __restrict__ items,
                                                               u32 blocks_per_item,
                                                               double2* __restrict__ partials) {
  __shared__ double2 lds[kBlock / 64][ZC];
  const u32 item_id = blockIdx.x / blocks_per_item;
  const u32 chunk = blockIdx.x % blocks_per_item;
  const ExpItem& it = items[item_id];
  const u64 x = it.x;
  const int nz = it.nz;
  const u64 range = x ? (len >> 1) : len;
  const u64 per = (range + blocks_per_item - 1) / blocks_per_item;
  const u64 begin = static_cast<u64>(chunk) * per;
  const u64 end = begin + per < range ? begin + per : range;

  double ar[ZC], ai[ZC];
#pragma unroll
  for (int j = 0; j < ZC; ++j) {
    ar[j] = 0.0;
    ai[j] = 0.0;
  }
  if (x) {
    const u64 pivot = x & (0 - x);
    for (u64 v = begin + threadIdx.x; v < end; v += kBlock) {
      const u64 b = insert_zero(v, pivot);
      const double2 a = psi[b];
      const double2 t = psi[b ^ x];
      // conj(t) * a
      const double re = fma(t.x, a.x, t.y * a.y);
      const double im = fma(t.x, a.y, -(t.y * a.x));
#pragma unroll
      for (int j = 0; j < ZC; ++j) {
        if (j < nz) {
          const bool odd = __popcll(it.z[j] & b) & 1;
          ar[j] += odd ? -re : re;
          ai[j] += odd ? -im : im;
        }
      }
    }
  } else {
    for (u64 b = begin + threadIdx.x; b < end; b += kBlock) {
      const double2 a = psi[b];
      const double p = fma(a.x, a.x, a.y * a.y);
#pragma unroll
      for (int j = 0; j < ZC; ++j) {
        if (j < nz) {
          const bool odd = __popcll(it.z[j] & b) & 1;
          ar[j] += odd ? -p : p;
        }
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < ZC; ++j) {
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
  if (threadIdx.x < nz) {
    double r = 0.0, i = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) {
      r += lds[w][threadIdx.x].x;
      i += lds[w][threadIdx.x].y;
    }
    partials[(static_cast<std::size_t>(item_id) * blocks_per_item + chunk) * ZC + threadIdx.x] =
        make_double2(r, i);
  }
}

// One thread per (item, slot): ordered sum over the item's blocks, then the
// key value <P> = 2 Re(i^y S) (x != 0) or Re S (x == 0).
__global__ __launch_bounds__(256) void expect_reduce_kernel(const ExpItem* __restrict__ items,
                                                            u32 num_items, u32 blocks_per_item,
                                                            const double2* __restrict__ partials,
                                                            int zc, double* __restrict__ key_vals) {
  const std::size_t gid = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const std::size_t item_id = gid / zc;
  const int j = static_cast<int>(gid % zc);
  if (item_id >= num_items) return;
  const ExpItem& it = items[item_id];
  if (j >= it.nz) return;
  double r = 0.0, i = 0.0;
  for (u32 c = 0; c < blocks_per_item; ++c) {
    const double2 p = partials[(item_id * blocks_per_item + c) * zc + j];
    r += p.x;
    i += p.y;
  }
  double val;
  if (it.x == 0) {
    val = r;
  } else {
    switch (__popcll(it.x & it.z[j]) & 3) {
      case 0: val = 2.0 * r; break;
      case 1: val = -2.0 * i; break;
      case 2: val = -2.0 * r; break;
      default: val = 2.0 * i; break;
    }
  }
  key_vals[it.key[j]] = val;
}

// E = sum_t c_t <P_key(t)>: one block, strided per-thread sums in term order,
// then a fixed-shape tree. Optionally writes per-term values (complex, imag 0).
__global__ __launch_bounds__(1024) void expect_finalize_kernel(const double2* __restrict__ coeffs,
                                                               const u32* __restrict__ term_key,
                                                               std::size_t T,
                                                               const double* __restrict__ key_vals,
                                                               double2* __restrict__ term_values,
                                                               double2* __restrict__ out) {
  __shared__ double2 lds[16];
  double r = 0.0, i = 0.0;
  for (std::size_t t = threadIdx.x; t < T; t += blockDim.x) {
    const double v = key_vals[term_key[t]];
    const double2 c = coeffs[t];
    r = fma(c.x, v, r);
    i = fma(c.y, v, i);
    if (term_values) term_values[t] = make_double2(v, 0.0);
  }
  for (int d = 32; d > 0; d >>= 1) {
    r += __shfl_xor(r, d, 64);
    i += __shfl_xor(i, d, 64);
  }
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = make_double2(r, i);
  __syncthreads();
  if (threadIdx.x == 0) {
    double sr = 0.0, si = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x / 64); ++w) {
      sr += lds[w].x;
      si += lds[w].y;
    }
    *out = make_double2(sr, si);
  }
}

__global__ __launch_bounds__(256) void widen_c64_kernel(const float2* __restrict__ in,
                                                        std::size_t len,
                                                        double2* __restrict__ out) {
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < len;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    out[i] = make_double2(static_cast<double>(in[i].x), static_cast<double>(in[i].y));
  }
}

// ------------------------------------------------------------ sigma = H psi

__device__ inline double2 times_i_pow(double2 v, int k) {
  switch (k & 3) {
    case 0: return v;
    case 1: return make_double2(-v.y, v.x);
    case 2: return make_double2(-v.x, -v.y);
    default: return make_double2(v.y, -v.x);
  }
}

__global__ __launch_bounds__(kBlock) void sigma_direct_kernel(
    const double2* __restrict__ psi, u64 len, const u64* __restrict__ group_x,
    const u32* __restrict__ group_begin, u32 num_groups, const u64* __restrict__ key_z,
    const int* __restrict__ key_y, const double2* __restrict__ key_c, double2* __restrict__ sigma,
    double2* __restrict__ sigma_dag) {
  for (u64 o = static_cast<u64>(blockIdx.x) * kBlock + threadIdx.x; o < len;
       o += static_cast<u64>(gridDim.x) * kBlock) {
    double sr = 0.0, si = 0.0, dr = 0.0, di = 0.0;
    for (u32 g = 0; g < num_groups; ++g) {
      const u64 src = o ^ group_x[g];
      const double2 t = psi[src];
      const u32 kb = group_begin[g], ke = group_begin[g + 1];
      for (u32 k = kb; k < ke; ++k) {
        double2 w = times_i_pow(t, key_y[k]);
        if (__popcll(key_z[k] & src) & 1) w = make_double2(-w.x, -w.y);
        const double2 c = key_c[k];
        // sigma += c w ; sigma_dag += conj(c) w
        sr = fma(c.x, w.x, fma(-c.y, w.y, sr));
        si = fma(c.x, w.y, fma(c.y, w.x, si));
        dr = fma(c.x, w.x, fma(c.y, w.y, dr));
        di = fma(c.x, w.y, fma(-c.y, w.x, di));
      }
    }
    if (sigma) sigma[o] = make_double2(sr, si);
    if (sigma_dag) sigma_dag[o] = make_double2(dr, di);
  }
}

unsigned stride_grid(std::size_t work, unsigned block) {
  const std::size_t g = (work + block - 1) / block;
  return static_cast<unsigned>(std::max<std::size_t>(1, std::min<std::size_t>(g, 1u << 20)));
}

}  // namespace

// ------------------------------------------------------------ plan

ExpectationPlan::~ExpectationPlan() {
  if (!d_items && !d_term_key && !d_group_x) return;
  int prev = -1;
  const hipError_t r = hipGetDevice(&prev);
  (void)hipSetDevice(device);
  device_free(d_items);
  device_free(d_term_key);
  device_free(d_group_x);
  device_free(d_group_begin);
  device_free(d_key_z);
  device_free(d_key_y);
  device_free(d_key_c);
  if (r == hipSuccess && prev >= 0) (void)hipSetDevice(prev);
}

std::shared_ptr<ExpectationPlan> build_plan(const wg_pauli_sum* op, hipStream_t stream) {
  auto plan = std::make_shared<ExpectationPlan>();
  plan->n = op->num_qubits;
  plan->T = op->num_terms;
  plan->device = op->device;
  const std::size_t T = op->num_terms, W = op->words;
  std::vector<u64> hx(T * W), hz(T * W);
  std::vector<double2> hc(T);
  if (T * W) {
    check_hip(hipMemcpyAsync(hx.data(), op->x, T * W * sizeof(u64), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    check_hip(hipMemcpyAsync(hz.data(), op->z, T * W * sizeof(u64), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
  }
  if (T) {
    check_hip(hipMemcpyAsync(hc.data(), op->coeffs, T * sizeof(double2), hipMemcpyDeviceToHost,
                             stream),
              "copy device buffer to host");
  }
  sync(stream);

  // Distinct keys in first-appearance order; groups of keys by x, in
  // first-appearance order of x.
  struct PairHash {
    std::size_t operator()(const std::pair<u64, u64>& k) const noexcept {
      return (k.first * 0x9E3779B97F4A7C15ULL) ^ (k.second + 0x632BE59BD9B4E019ULL + (k.first << 7));
    }
  };
  std::unordered_map<std::pair<u64, u64>, u32, PairHash> key_index;
  std::vector<u64> kx, kz;
  std::vector<double2> kc;
  std::vector<u32> term_key(T);
  for (std::size_t t = 0; t < T; ++t) {
    const u64 xm = W ? hx[t * W] : 0, zm = W ? hz[t * W] : 0;
    auto [itr, inserted] = key_index.try_emplace({xm, zm}, static_cast<u32>(kx.size()));
    if (inserted) {
      kx.push_back(xm);
      kz.push_back(zm);
      kc.push_back(hc[t]);
    } else {
      kc[itr->second].x += hc[t].x;
      kc[itr->second].y += hc[t].y;
    }
    term_key[t] = itr->second;
  }
  const std::size_t K = kx.size();
  std::unordered_map<u64, u32> group_index;
  std::vector<u64> gx;
  std::vector<std::vector<u32>> gkeys;
  for (std::size_t k = 0; k < K; ++k) {
    auto [itr, inserted] = group_index.try_emplace(kx[k], static_cast<u32>(gx.size()));
    if (inserted) {
      gx.push_back(kx[k]);
      gkeys.emplace_back();
    }
    gkeys[itr->second].push_back(static_cast<u32>(k));
  }
  // Items of up to kZ keys.
  std::vector<ExpItem> items;
  std::vector<u64> flat_z;
  std::vector<int> flat_y;
  std::vector<double2> flat_c;
  std::vector<u32> group_begin;
  std::vector<u32> key_order;  // key ids in group order
  for (std::size_t g = 0; g < gx.size(); ++g) {
    group_begin.push_back(static_cast<u32>(flat_z.size()));
    const auto& ks = gkeys[g];
    for (std::size_t s = 0; s < ks.size(); s += kZ) {
      ExpItem it{};
      it.x = gx[g];
      it.nz = static_cast<int>(std::min<std::size_t>(kZ, ks.size() - s));
      for (int j = 0; j < it.nz; ++j) {
        it.z[j] = kz[ks[s + j]];
        it.key[j] = ks[s + j];
      }
      items.push_back(it);
    }
    for (u32 k : ks) {
      flat_z.push_back(kz[k]);
      flat_y.push_back(__builtin_popcountll(kx[k] & kz[k]) & 3);
      flat_c.push_back(kc[k]);
    }
  }
  group_begin.push_back(static_cast<u32>(flat_z.size()));

  plan->num_keys = K;
  plan->num_items = items.size();
  plan->num_groups = gx.size();
  const u64 half = plan->n == 0 ? 1 : (u64{1} << (plan->n - 1));
  plan->half_range = plan->n == 0 ? 0 : half;
  // Blocks per item: enough to fill the device, each thread doing a few
  // hundred amplitude pairs at large n.
  const u64 range = plan->n == 0 ? 1 : half;
  u64 bpi = (range + kBlock * 64 - 1) / (kBlock * 64);
  bpi = std::max<u64>(1, std::min<u64>(bpi, 1024));
  plan->blocks_per_item = static_cast<u32>(bpi);

  DeviceScope scope(op->device);
  auto upload = [&](auto*& dst, const auto& src) {
    using T2 = std::remove_reference_t<decltype(*dst)>;
    dst = device_alloc<T2>(src.size());
    if (!src.empty()) {
      check_hip(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T2), hipMemcpyHostToDevice,
                               stream),
                "copy host buffer to device");
    }
  };
  upload(plan->d_items, items);
  upload(plan->d_term_key, term_key);
  upload(plan->d_group_x, gx);
  upload(plan->d_group_begin, group_begin);
  upload(plan->d_key_z, flat_z);
  upload(plan->d_key_y, flat_y);
  upload(plan->d_key_c, flat_c);
  sync(stream);
  return plan;
}

std::shared_ptr<ExpectationPlan> get_plan(wg_pauli_sum* op, hipStream_t stream) {
  std::lock_guard<std::mutex> lk(op->plan_mutex);
  if (!op->plan) op->plan = build_plan(op, stream);
  return op->plan;
}

// Runs the forward path on a device-resident complex128 psi. key_vals receives
// <P> per distinct key; out (device) receives E; term_values optional.
void run_forward(const wg_pauli_sum* op, const ExpectationPlan& plan, const double2* psi,
                 double2* term_values, double2* out_dev, Scratch& scratch, hipStream_t stream) {
  double* key_vals = scratch.take<double>(std::max<std::size_t>(plan.num_keys, 1));
  if (plan.num_items) {
    const std::size_t nblocks = plan.num_items * plan.blocks_per_item;
    double2* partials = scratch.take<double2>(checked_mul(nblocks, kZ, "expectation partials"));
    const u64 len = u64{1} << plan.n;
    ProfileScope prof("expect_direct", stream);
    expect_direct_kernel<kZ><<<grid_for(nblocks, 1), kBlock, 0, stream>>>(
        psi, len, plan.d_items, plan.blocks_per_item, partials);
    check_launch("expect_direct");
    expect_reduce_kernel<<<grid_for(plan.num_items * kZ, 256), 256, 0, stream>>>(
        plan.d_items, static_cast<u32>(plan.num_items), plan.blocks_per_item, partials, kZ,
        key_vals);
    check_launch("expect_reduce");
  }
  expect_finalize_kernel<<<1, 1024, 0, stream>>>(op->coeffs, plan.d_term_key, plan.T, key_vals,
                                                 term_values, out_dev);
  check_launch("expect_finalize");
}

void run_sigma(const ExpectationPlan& plan, const double2* psi, std::size_t len, double2* sigma,
               double2* sigma_dag, hipStream_t stream) {
  if (plan.num_groups == 0) {
    if (sigma) check_hip(hipMemsetAsync(sigma, 0, len * sizeof(double2), stream), "clear sigma");
    if (sigma_dag) {
      check_hip(hipMemsetAsync(sigma_dag, 0, len * sizeof(double2), stream), "clear sigma");
    }
    return;
  }
  ProfileScope prof("sigma_direct", stream);
  sigma_direct_kernel<<<stride_grid(len, kBlock), kBlock, 0, stream>>>(
      psi, len, plan.d_group_x, plan.d_group_begin, static_cast<u32>(plan.num_groups),
      plan.d_key_z, plan.d_key_y, plan.d_key_c, sigma, sigma_dag);
  check_launch("sigma_direct");
}

std::size_t validate_psi(const wg_pauli_sum* op, const void* psi, std::size_t length) {
  if (op->num_qubits > 63) {
    fail_invalid("HIP expectation_statevector requires num_qubits <= 63");
  }
  const std::size_t expected = std::size_t{1} << op->num_qubits;
  if (length != expected) {
    fail_invalid("expectation_statevector requires len(psi) == 2 ** num_qubits");
  }
  if (psi == nullptr) fail_invalid("HIP statevector data pointer must be non-null");
  return expected;
}

}  // namespace wg

using namespace wg;

extern "C" {

int wg_expectation(const wg_pauli_sum* op_c, const void* psi, int dtype, size_t length,
                   int psi_on_device, void* stream_ptr, double* out2) {
  return guard([&] {
    require_op(op_c);
    auto* op = const_cast<wg_pauli_sum*>(op_c);
    if (dtype != WG_DTYPE_COMPLEX64 && dtype != WG_DTYPE_COMPLEX128) {
      fail_type("psi dtype must be complex64 or complex128");
    }
    const std::size_t len = validate_psi(op, psi, length);
    out2[0] = 0.0;
    out2[1] = 0.0;
    if (op->num_terms == 0) return;
    DeviceScope scope(op->device);
    hipStream_t stream = as_stream(stream_ptr);
    auto plan = get_plan(op, stream);
    Scratch scratch(op->device);
    const double2* dpsi;
    if (dtype == WG_DTYPE_COMPLEX128 && psi_on_device) {
      dpsi = static_cast<const double2*>(psi);
    } else {
      double2* buf = scratch.take<double2>(len);
      if (dtype == WG_DTYPE_COMPLEX128) {
        check_hip(hipMemcpyAsync(buf, psi, len * sizeof(double2), hipMemcpyHostToDevice, stream),
                  "copy host buffer to device");
      } else {
        const float2* src = static_cast<const float2*>(psi);
        if (!psi_on_device) {
          float2* staged = scratch.take<float2>(len);
          check_hip(hipMemcpyAsync(staged, psi, len * sizeof(float2), hipMemcpyHostToDevice,
                                   stream),
                    "copy host buffer to device");
          src = staged;
        }
        widen_c64_kernel<<<stride_grid(len, 256), 256, 0, stream>>>(src, len, buf);
        check_launch("widen_c64");
      }
      dpsi = buf;
    }
    double2* dout = scratch.take<double2>(1);
    run_forward(op, *plan, dpsi, nullptr, dout, scratch, stream);
    double2 h;
    check_hip(hipMemcpyAsync(&h, dout, sizeof(double2), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
    out2[0] = h.x;
    out2[1] = h.y;
  });
}

int wg_expectation_adjoint(const wg_pauli_sum* op_c, const double* psi, size_t length,
                           double* term_values, double* sigma, double* sigma_dagger,
                           void* stream_ptr, double* out2) {
  return guard([&] {
    require_op(op_c);
    auto* op = const_cast<wg_pauli_sum*>(op_c);
    const std::size_t len = validate_psi(op, psi, length);
    DeviceScope scope(op->device);
    hipStream_t stream = as_stream(stream_ptr);
    out2[0] = 0.0;
    out2[1] = 0.0;
    if (op->num_terms == 0) {
      if (sigma) check_hip(hipMemsetAsync(sigma, 0, len * 16, stream), "clear sigma");
      if (sigma_dagger) check_hip(hipMemsetAsync(sigma_dagger, 0, len * 16, stream), "clear sigma");
      sync(stream);
      return;
    }
    auto plan = get_plan(op, stream);
    Scratch scratch(op->device);
    const double2* dpsi = reinterpret_cast<const double2*>(psi);
    double2* dout = scratch.take<double2>(1);
    run_forward(op, *plan, dpsi, reinterpret_cast<double2*>(term_values), dout, scratch, stream);
    if (sigma || sigma_dagger) {
      run_sigma(*plan, dpsi, len, reinterpret_cast<double2*>(sigma),
                reinterpret_cast<double2*>(sigma_dagger), stream);
    }
    double2 h;
    check_hip(hipMemcpyAsync(&h, dout, sizeof(double2), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
    out2[0] = h.x;
    out2[1] = h.y;
  });
}

}  // extern "C"

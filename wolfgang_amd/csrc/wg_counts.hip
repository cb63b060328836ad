// Expectation over computational-basis counts on the device (reference
// PauliSum::expectation_z_counts, src/expectation.cpp:157-247; Python binding
// bindings/python/pauli_sum_py.cpp:261-300,1614-1631).
//
// For diagonal terms (x = 0):
//   E = sum_t c_t * (sum_r n_r (-1)^{|z_t & b_r|}) / sum_r n_r,
// b_r the packed bitstring of row r (character i is qubit n - 1 - i, i.e. the
// right-most character is qubit 0). Validation follows the reference's order
// and messages: diagonal terms, then the counts (non-negative, finite,
// positive total), then each bitstring (length, characters).
//
// Work layout: one thread per term, blockIdx.y a chunk of rows. The row words
// and counts are the same for every thread of a block (uniform addresses, read
// through the scalar cache); the term's z words stay in registers for up to 4
// words. Partial sums per (chunk, term) are reduced in chunk order, then one
// block combines c_t * (sum_t / total) in term order with a fixed tree, so the
// result does not depend on scheduling.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "wg_internal.hpp"

namespace wg {
namespace {

constexpr int kCountBlock = 256;
constexpr int kFinalBlock = 1024;

__global__ __launch_bounds__(256) void or_words_kernel(const u64* __restrict__ words,
                                                       std::size_t count,
                                                       unsigned long long* __restrict__ out) {
  u64 acc = 0;
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    acc |= words[i];
  }
  if (acc) atomicOr(out, static_cast<unsigned long long>(acc));
}

// kW > 0: exactly kW words per mask, kept in registers; kW == 0: any W, read
// from memory per row.
template <int kW>
__global__ __launch_bounds__(kCountBlock) void zcount_partial_kernel(
    const u64* __restrict__ z, std::size_t T, std::size_t W, const u64* __restrict__ rows,
    const double* __restrict__ counts, std::size_t R, std::size_t rows_per_chunk,
    double* __restrict__ partials) {
  const std::size_t t = static_cast<std::size_t>(blockIdx.x) * kCountBlock + threadIdx.x;
  const std::size_t r0 = static_cast<std::size_t>(blockIdx.y) * rows_per_chunk;
  const std::size_t r1 = std::min(R, r0 + rows_per_chunk);
  const std::size_t tt = t < T ? t : 0;
  double acc = 0.0;
  if (kW > 0) {
    u64 zt[kW > 0 ? kW : 1];
#pragma unroll
    for (int w = 0; w < kW; ++w) zt[w] = z[tt * kW + w];
    for (std::size_t r = r0; r < r1; ++r) {
      u32 par = 0;
#pragma unroll
      for (int w = 0; w < kW; ++w) par += static_cast<u32>(__popcll(zt[w] & rows[r * kW + w]));
      const double n = counts[r];
      acc += (par & 1U) ? -n : n;
    }
  } else {
    for (std::size_t r = r0; r < r1; ++r) {
      u32 par = 0;
      for (std::size_t w = 0; w < W; ++w) par += static_cast<u32>(__popcll(z[tt * W + w] & rows[r * W + w]));
      const double n = counts[r];
      acc += (par & 1U) ? -n : n;
    }
  }
  if (t < T) partials[static_cast<std::size_t>(blockIdx.y) * T + t] = acc;
}

__global__ __launch_bounds__(kFinalBlock) void zcount_finalize_kernel(
    const double* __restrict__ partials, std::size_t chunks, std::size_t T,
    const double2* __restrict__ coeffs, double total, double2* __restrict__ out) {
  __shared__ double2 red[kFinalBlock];
  double re = 0.0, im = 0.0;
  for (std::size_t t = threadIdx.x; t < T; t += kFinalBlock) {
    double ws = 0.0;
    for (std::size_t c = 0; c < chunks; ++c) ws += partials[c * T + t];
    const double v = ws / total;
    const double2 ct = coeffs[t];
    re += ct.x * v;
    im += ct.y * v;
  }
  red[threadIdx.x] = make_double2(re, im);
  __syncthreads();
  for (int s = kFinalBlock / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) {
      red[threadIdx.x].x += red[threadIdx.x + s].x;
      red[threadIdx.x].y += red[threadIdx.x + s].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

}  // namespace
}  // namespace wg

using namespace wg;

extern "C" {

int wg_expectation_z_counts(const wg_pauli_sum* op, const char* chars, const size_t* offsets,
                            const double* counts, size_t num_rows, void* stream_ptr,
                            double* out2) {
  return guard([&] {
    require_op(op);
    if (num_rows && (chars == nullptr || offsets == nullptr || counts == nullptr)) {
      fail_invalid("Z-count bitstring and count arrays must be non-null");
    }
    out2[0] = 0.0;
    out2[1] = 0.0;
    DeviceScope scope(op->device);
    hipStream_t stream = as_stream(stream_ptr);
    const std::size_t n = op->num_qubits, W = op->words, T = op->num_terms;
    Scratch scratch(op->device);
    // Diagonal terms only (reference src/expectation.cpp:157-164).
    if (T && W) {
      auto* d_or = scratch.take<unsigned long long>(1);
      check_hip(hipMemsetAsync(d_or, 0, sizeof(unsigned long long), stream), "memset");
      or_words_kernel<<<std::min<unsigned>(grid_for(T * W, 256), 1024), 256, 0, stream>>>(
          op->x, T * W, d_or);
      check_launch("or_words");
      unsigned long long any_x = 0;
      check_hip(hipMemcpyAsync(&any_x, d_or, sizeof(any_x), hipMemcpyDeviceToHost, stream),
                "copy device buffer to host");
      sync(stream);
      if (any_x) fail_invalid("expectation_z_counts requires diagonal Pauli terms");
    }
    // Counts (reference src/expectation.cpp:190-206).
    double total = 0.0;
    for (std::size_t r = 0; r < num_rows; ++r) {
      const double c = counts[r];
      if (c < 0.0) fail_invalid("Z-count values must be non-negative");
      if (!std::isfinite(c)) fail_invalid("Z-count values must be finite");
      total += c;
    }
    if (total <= 0.0) fail_invalid("Z-count total count must be positive");
    // Bitstrings (reference src/expectation.cpp:166-188): the right-most
    // character is qubit 0.
    const std::size_t RW = std::max<std::size_t>(W, 1);
    std::vector<u64> packed(checked_mul(num_rows, RW, "counts"), 0);
    for (std::size_t r = 0; r < num_rows; ++r) {
      const std::size_t len = offsets[r + 1] - offsets[r];
      if (offsets[r + 1] < offsets[r] || len != n) {
        fail_invalid("Z-count bitstring length must equal num_qubits");
      }
      const char* s = chars + offsets[r];
      for (std::size_t i = 0; i < len; ++i) {
        if (s[i] != '0' && s[i] != '1') fail_invalid("Z-count bitstrings may contain only 0 or 1");
        if (s[i] == '1') {
          const std::size_t q = n - 1 - i;
          packed[r * RW + q / 64] |= u64{1} << (q % 64);
        }
      }
    }
    if (T == 0 || W == 0) {
      // No terms, or zero qubits (every term is the identity): E = sum c_t.
      if (T) {
        std::vector<double2> hc(T);
        check_hip(hipMemcpyAsync(hc.data(), op->coeffs, T * sizeof(double2), hipMemcpyDeviceToHost,
                                 stream),
                  "copy device buffer to host");
        sync(stream);
        for (const double2& c : hc) {
          out2[0] += c.x * (total / total);
          out2[1] += c.y * (total / total);
        }
      }
      return;
    }
    u64* d_rows = scratch.take<u64>(packed.size());
    double* d_counts = scratch.take<double>(num_rows);
    check_hip(hipMemcpyAsync(d_rows, packed.data(), packed.size() * sizeof(u64),
                             hipMemcpyHostToDevice, stream),
              "copy host buffer to device");
    check_hip(hipMemcpyAsync(d_counts, counts, num_rows * sizeof(double), hipMemcpyHostToDevice,
                             stream),
              "copy host buffer to device");
    // Grid: term blocks x row chunks, about 2,048 blocks when the rows allow.
    const std::size_t tblocks = (T + kCountBlock - 1) / kCountBlock;
    std::size_t chunks = std::max<std::size_t>(1, 2048 / tblocks);
    chunks = std::min(chunks, std::max<std::size_t>(1, num_rows / 64));
    chunks = std::min<std::size_t>(chunks, 65535);
    const std::size_t per_chunk = (num_rows + chunks - 1) / chunks;
    chunks = (num_rows + per_chunk - 1) / per_chunk;
    if (tblocks > 0x7fffffffULL) fail_invalid("request exceeds HIP launch grid limits");
    double* partials = scratch.take<double>(checked_mul(chunks, T, "z-count partial sums"));
    double2* d_out = scratch.take<double2>(1);
    const dim3 grid(static_cast<unsigned>(tblocks), static_cast<unsigned>(chunks));
    {
      ProfileScope prof("zcount_partial", stream);
      switch (W) {
        case 1:
          zcount_partial_kernel<1><<<grid, kCountBlock, 0, stream>>>(op->z, T, W, d_rows, d_counts,
                                                                     num_rows, per_chunk, partials);
          break;
        case 2:
          zcount_partial_kernel<2><<<grid, kCountBlock, 0, stream>>>(op->z, T, W, d_rows, d_counts,
                                                                     num_rows, per_chunk, partials);
          break;
        case 3:
          zcount_partial_kernel<3><<<grid, kCountBlock, 0, stream>>>(op->z, T, W, d_rows, d_counts,
                                                                     num_rows, per_chunk, partials);
          break;
        case 4:
          zcount_partial_kernel<4><<<grid, kCountBlock, 0, stream>>>(op->z, T, W, d_rows, d_counts,
                                                                     num_rows, per_chunk, partials);
          break;
        default:
          zcount_partial_kernel<0><<<grid, kCountBlock, 0, stream>>>(op->z, T, W, d_rows, d_counts,
                                                                     num_rows, per_chunk, partials);
          break;
      }
      check_launch("zcount_partial");
    }
    zcount_finalize_kernel<<<1, kFinalBlock, 0, stream>>>(partials, chunks, T, op->coeffs, total,
                                                          d_out);
    check_launch("zcount_finalize");
    double2 h;
    check_hip(hipMemcpyAsync(&h, d_out, sizeof(double2), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
    out2[0] = h.x;
    out2[1] = h.y;
  });
}

}  // extern "C"

// Term-pair multiply: lhs @ rhs with rhs acting first.
//
// Product (i, j) is written at p = i*R + j (lhs-outer, rhs-inner order;
// reference docs/architecture/semantic_contracts.md:157-180):
//   out_x = lx ^ rx,  out_z = lz ^ rz                           (per word)
//   k     = |out_x & out_z| - |lx & lz| - |rx & rz| + 2 |lx & rz|  (mod 4)
//   coeff = i^k * (c_l * c_r)
// (reference src/multiply.cpp:267-402). The complex product is (ac - bd,
// ad + bc) with each product rounded separately: this file is compiled with
// -ffp-contract=off so no FMA is formed, which reproduces the x86-64 reference
// bits. i^k is an exact swap/negate (reference src/detail/phase.hpp:24-36).
//
// Kernel shape: a block takes one lhs term (its masks, y-parity and coefficient
// are wave-uniform scalars) and 1024 consecutive rhs terms, 4 per thread; all
// reads and writes are contiguous across lanes.

#include <hip/hip_runtime.h>

#include <string>

#include "wg_internal.hpp"
#include "wg_product_term.hpp"
#include "wg_simplify.hpp"

namespace wg {
namespace {

constexpr int kBlock = 256;
constexpr int kPerThread = 4;
constexpr std::size_t kChunk = static_cast<std::size_t>(kBlock) * kPerThread;

template <int WC>
__global__ __launch_bounds__(kBlock) void product_kernel(
    const u64* __restrict__ lx, const u64* __restrict__ lz, const double2* __restrict__ lc,
    const u64* __restrict__ rx, const u64* __restrict__ rz, const double2* __restrict__ rc,
    std::size_t R, std::size_t Wrt, std::size_t chunks, u64* __restrict__ ox,
    u64* __restrict__ oz, double2* __restrict__ oc) {
  const std::size_t W = WC > 0 ? static_cast<std::size_t>(WC) : Wrt;
  const std::size_t i = blockIdx.x / chunks;
  const std::size_t j0 = (blockIdx.x % chunks) * kChunk;
  int ly = 0;
  for (std::size_t w = 0; w < W; ++w) ly += __popcll(lx[i * W + w] & lz[i * W + w]);
  const double2 a = lc[i];
#pragma unroll
  for (int q = 0; q < kPerThread; ++q) {
    const std::size_t j = j0 + static_cast<std::size_t>(q) * kBlock + threadIdx.x;
    if (j >= R) break;
    const std::size_t p = i * R + j;
    int k = -ly;
    for (std::size_t w = 0; w < W; ++w) {
      const u64 lxw = lx[i * W + w], lzw = lz[i * W + w];
      const u64 rxw = rx[j * W + w], rzw = rz[j * W + w];
      const u64 oxw = lxw ^ rxw, ozw = lzw ^ rzw;
      ox[p * W + w] = oxw;
      oz[p * W + w] = ozw;
      k += __popcll(oxw & ozw) - __popcll(rxw & rzw) + 2 * __popcll(lxw & rzw);
    }
    const double2 b = rc[j];
    oc[p] = apply_phase(complex_product(a, b), k);
  }
}

void launch_products(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, u64* ox, u64* oz,
                     double2* oc, hipStream_t stream) {
  const std::size_t L = lhs->num_terms, R = rhs->num_terms, W = lhs->words;
  const std::size_t chunks = (R + kChunk - 1) / kChunk;
  const unsigned grid = grid_for(checked_mul(L, chunks, "matmul grid"), 1);
  ProfileScope prof("product", stream);
  switch (W) {
    case 1:
      product_kernel<1><<<grid, kBlock, 0, stream>>>(lhs->x, lhs->z, lhs->coeffs, rhs->x, rhs->z,
                                                     rhs->coeffs, R, 1, chunks, ox, oz, oc);
      break;
    case 2:
      product_kernel<2><<<grid, kBlock, 0, stream>>>(lhs->x, lhs->z, lhs->coeffs, rhs->x, rhs->z,
                                                     rhs->coeffs, R, 2, chunks, ox, oz, oc);
      break;
    default:
      product_kernel<0><<<grid, kBlock, 0, stream>>>(lhs->x, lhs->z, lhs->coeffs, rhs->x, rhs->z,
                                                     rhs->coeffs, R, W, chunks, ox, oz, oc);
      break;
  }
  check_launch("product");
}

}  // namespace

int matmul_seed_mode(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, hipStream_t stream) {
  // Reference src/multiply.cpp:107-261: with W in {1, 2} and a
  // duplicate-heavy operand the reference accumulates in a hash map seeded by
  // the first product, with a fixed 1e-12 drop; otherwise it calls simplify()
  // with its defaults on the materialized products.
  const std::size_t W = lhs->words;
  const bool hash = (W == 1 || W == 2) && lhs->num_terms != 0 && rhs->num_terms != 0 &&
                    (operand_duplicate_heavy(lhs, stream) || operand_duplicate_heavy(rhs, stream));
  return hash ? 1 : -1;
}

wg_pauli_sum* matmul_device(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, int seed_mode,
                            Scratch& scratch, hipStream_t stream) {
  const std::size_t W = lhs->words;
  const std::size_t P = checked_mul(lhs->num_terms, rhs->num_terms, "matmul products");
  if (P == 0) return new_pauli_sum(lhs->num_qubits, 0, lhs->device);
  if (W >= 1 && W <= 4 && P < 0xFFFFFFFFULL) {
    // Up to 256 qubits the products are never materialized: the simplify
    // reads them through the operands.
    return simplify_products(lhs, rhs, 1.0e-12, 0.0, seed_mode, scratch, stream);
  }
  u64* px = scratch.take<u64>(checked_mul(P, W, "matmul products"));
  u64* pz = scratch.take<u64>(P * W);
  double2* pc = scratch.take<double2>(P);
  launch_products(lhs, rhs, px, pz, pc, stream);
  return simplify_device(px, pz, pc, lhs->num_qubits, P, lhs->device, 1.0e-12, 0.0, seed_mode,
                         scratch, stream);
}

}  // namespace wg

using namespace wg;

extern "C" {

int wg_matmul(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, int simplify_output,
              size_t max_intermediate_terms, void* stream_ptr, wg_pauli_sum** out) {
  return guard([&] {
    *out = nullptr;
    require_op(lhs);
    require_op(rhs);
    require_same_device("HIP matmul", lhs, rhs);
    if (lhs->num_qubits != rhs->num_qubits) {
      fail_invalid("PauliSum matmul requires the same num_qubits");
    }
    // Reference src/detail/checked_arithmetic.hpp:34-47, before any allocation.
    const std::size_t L = lhs->num_terms, R = rhs->num_terms;
    if (L != 0 && R > SIZE_MAX / L) {
      fail_invalid("matmul intermediate term count overflows size_t");
    }
    const std::size_t P = L * R;
    if (P > max_intermediate_terms) {
      fail_invalid("matmul intermediate term count exceeds max_intermediate_terms");
    }
    DeviceScope scope(lhs->device);
    hipStream_t stream = as_stream(stream_ptr);
    if (P == 0) {
      *out = new_pauli_sum(lhs->num_qubits, 0, lhs->device);
      return;
    }
    const std::size_t W = lhs->words;
    if (!simplify_output) {
      std::unique_ptr<wg_pauli_sum> prod(new_pauli_sum(lhs->num_qubits, P, lhs->device));
      if (W == 0) {
        // Zero-qubit operators carry only coefficients.
        launch_products(lhs, rhs, nullptr, nullptr, prod->coeffs, stream);
      } else {
        launch_products(lhs, rhs, prod->x, prod->z, prod->coeffs, stream);
      }
      sync(stream);
      *out = prod.release();
      return;
    }
    Scratch scratch(lhs->device);
    *out = matmul_device(lhs, rhs, matmul_seed_mode(lhs, rhs, stream), scratch, stream);
  });
}

}  // extern "C"

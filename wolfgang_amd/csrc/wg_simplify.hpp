// Device simplify pipeline shared by simplify and matmul.
#pragma once

#include <vector>

#include "wg_internal.hpp"

namespace wg {

// Sort-reduce simplify of device buffers (x, z: T*W words; c: T coefficients)
// into a new operator on `device`. seed_mode: -1 picks the reference's
// simplify() accumulation seed from the data, 1 seeds with the first
// coefficient (hash accumulation), 0 with 0.0 (sort path).
wg_pauli_sum* simplify_device(const u64* x, const u64* z, const double2* c, std::size_t n,
                              std::size_t T, int device, double atol, double rtol, int seed_mode,
                              Scratch& scratch, hipStream_t stream);

// Simplify of lhs @ rhs (1 to 4 mask words) without materializing the
// products: only sort records are written, masks and coefficients are read
// through the operands. Same result bits as simplify_device on the
// materialized products with the same arguments.
wg_pauli_sum* simplify_products(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, double atol,
                                double rtol, int seed_mode, Scratch& scratch, hipStream_t stream);

// matmul's duplicate-heavy operand test (reference src/multiply.cpp:29-69).
bool operand_duplicate_heavy(const wg_pauli_sum* op, hipStream_t stream);

}  // namespace wg

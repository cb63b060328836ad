// Device simplify pipeline shared by simplify and matmul.
#pragma once

#include <vector>

#include "wg_internal.hpp"
#include "wg_product_term.hpp"

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

// Operand records of the unmaterialized products (product_record_words(W)
// words per term of `op`).
void build_operand_records(const wg_pauli_sum* op, u64* rec, hipStream_t stream);

// The part of simplify_products after the key pass: keys / idx hold the T
// products' word-0 keys and indices (reordered in place by the sort; the
// sorted copy may end up in scratch), block_stats the key pass's per-unit
// statistics (units units, product_keys_unit layout).
wg_pauli_sum* simplify_product_keys(const ProductTerms& P, std::size_t n, std::size_t T,
                                    int device, u64* keys, u32* idx, const u64* block_stats,
                                    std::size_t units, double atol, double rtol, int seed_mode,
                                    Scratch& scratch, hipStream_t stream);

// matmul with simplify (wg_multiply.hip): lhs @ rhs simplified with matmul's
// rule; seed_mode from matmul_seed_mode (the reference's hash-path choice).
int matmul_seed_mode(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, hipStream_t stream);
wg_pauli_sum* matmul_device(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, int seed_mode,
                            Scratch& scratch, hipStream_t stream);

// matmul's duplicate-heavy operand test (reference src/multiply.cpp:29-69).
bool operand_duplicate_heavy(const wg_pauli_sum* op, hipStream_t stream);

}  // namespace wg

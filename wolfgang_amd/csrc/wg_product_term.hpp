// Term-pair product arithmetic shared by the product kernels (wg_multiply.hip,
// wg_products.hip) and the simplify of products that are never materialized
// (wg_simplify.hip).
//
// Product of lhs term i and rhs term j (rhs acting first; reference
// src/multiply.cpp:267-402):
//   out_x = lx ^ rx,  out_z = lz ^ rz                              (per word)
//   k     = |out_x & out_z| - |lx & lz| - |rx & rz| + 2 |lx & rz|   (mod 4)
//   coeff = i^k * (c_l * c_r)
// The complex product is (ac - bd, ad + bc) with every product rounded on its
// own: the library is compiled with -ffp-contract=off, so no FMA is formed and
// the bits are those of the x86-64 reference. i^k is an exact swap/negate
// (reference src/detail/phase.hpp:24-36).
#pragma once

#include <hip/hip_runtime.h>

#include "wg_internal.hpp"

namespace wg {

__device__ inline double2 apply_phase(double2 v, int k) {
  switch (k & 3) {
    case 0: return v;
    case 1: return make_double2(-v.y, v.x);
    case 2: return make_double2(-v.x, -v.y);
    default: return make_double2(v.y, -v.x);
  }
}

__device__ inline double2 complex_product(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Operand term records of the unmaterialized products: the term's x words,
// z words and coefficient bits, padded to a power-of-two size, so that one
// product's operands are two cache lines (the separate mask and coefficient
// arrays would be six reads per product).
__host__ __device__ constexpr u32 product_record_words(u32 W) {
  return W <= 1 ? 4 : W <= 3 ? 8 : 16;
}

// The terms of lhs @ rhs in lhs-outer, rhs-inner order, read through the
// operands' records: term t is lhs term t / R times rhs term t % R. The
// operands of a product are small next to the L*R products (8,192 terms of 2
// words are 512 KB of records), so these reads hit L2 where a materialized
// product would be a random read from HBM.
struct ProductTerms {
  const u64* lrec;
  const u64* rrec;
  u32 R;
  u32 W;

  __device__ void split(u32 t, u32& i, u32& j) const {
    i = t / R;
    j = t - i * R;
  }
  // Key word w of term t in the canonical order [x0, z0, x1, z1, ...].
  __device__ u64 word(u32 t, int w) const {
    u32 i, j;
    split(t, i, j);
    const std::size_t RW = product_record_words(W);
    const std::size_t at = (w & 1) ? W + static_cast<std::size_t>(w >> 1) : static_cast<std::size_t>(w >> 1);
    return lrec[i * RW + at] ^ rrec[j * RW + at];
  }
  // Loads record `r` (RW words) as 16-byte vectors.
  template <int W_>
  __device__ static void load(const u64* __restrict__ r, u64* out) {
    constexpr int RW = static_cast<int>(product_record_words(W_));
#pragma unroll
    for (int k = 0; k < 2 * W_ + 2; k += 2) {
      const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(r + k);
      out[k] = v.x;
      out[k + 1] = v.y;
    }
    static_assert(2 * W_ + 2 <= RW, "record holds masks and coefficient");
  }
  // Masks (ox[W], oz[W]) and coefficient of the product of lhs term i and
  // rhs term j.
  template <int W_>
  __device__ double2 term_ij(u32 i, u32 j, u64* ox, u64* oz) const {
    constexpr u32 RW = product_record_words(W_);
    u64 a[2 * W_ + 2], b[2 * W_ + 2];
    load<W_>(lrec + static_cast<std::size_t>(i) * RW, a);
    load<W_>(rrec + static_cast<std::size_t>(j) * RW, b);
    int k = 0;
#pragma unroll
    for (int w = 0; w < W_; ++w) {
      ox[w] = a[w] ^ b[w];
      oz[w] = a[W_ + w] ^ b[W_ + w];
      k += __popcll(ox[w] & oz[w]) - __popcll(a[w] & a[W_ + w]) - __popcll(b[w] & b[W_ + w]) +
           2 * __popcll(a[w] & b[W_ + w]);
    }
    const double2 ca = make_double2(__longlong_as_double(static_cast<long long>(a[2 * W_])),
                                    __longlong_as_double(static_cast<long long>(a[2 * W_ + 1])));
    const double2 cb = make_double2(__longlong_as_double(static_cast<long long>(b[2 * W_])),
                                    __longlong_as_double(static_cast<long long>(b[2 * W_ + 1])));
    return apply_phase(complex_product(ca, cb), k);
  }
  template <int W_>
  __device__ double2 term(u32 t, u64* ox, u64* oz) const {
    u32 i, j;
    split(t, i, j);
    return term_ij<W_>(i, j, ox, oz);
  }
  template <int W_>
  __device__ double2 coeff(u32 t) const {
    u64 ox[W_], oz[W_];
    return term<W_>(t, ox, oz);
  }
};

}  // namespace wg

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
    return combine<W_>(lrec + static_cast<std::size_t>(i) * RW,
                       rrec + static_cast<std::size_t>(j) * RW, ox, oz);
  }
  // The same from the two operand records wherever they are (global memory or
  // a copy in LDS).
  template <int W_>
  __device__ static double2 combine(const u64* __restrict__ ra, const u64* __restrict__ rb,
                                    u64* ox, u64* oz) {
    u64 a[2 * W_ + 2], b[2 * W_ + 2];
    load<W_>(ra, a);
    load<W_>(rb, b);
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

// Key pass of the unmaterialized products, one unit = lhs term i times rhs
// terms [j0, j0 + kProductChunk) on a kProductKeyBlock-thread block: writes
// each product's word-0 sort key (x << 32 | z when kPacked) and its index
// t = i * R + j, and the unit's statistics — OR of each of the 2W key words
// [x0, z0, x1, z1, ...], their AND, and max |c| as bits (|c| >= 0, so the
// bits order like the values; only with kMax) — at block_stats[k * stride +
// unit] for k < 4W + 1.
constexpr u32 kProductKeyBlock = 256;
constexpr u32 kProductPerThread = 8;
constexpr u32 kProductChunk = kProductKeyBlock * kProductPerThread;
constexpr int kProductMaxStats = 4 * 4 + 1;

template <int W, bool kPacked, bool kMax>
__device__ inline void product_keys_unit(const ProductTerms& P, u32 i, u32 j0,
                                         u64* __restrict__ keys, u32* __restrict__ idx,
                                         u64* __restrict__ block_stats, std::size_t stride,
                                         std::size_t unit,
                                         u64 (*lds)[kProductMaxStats]) {
  static_assert(W >= 1 && W <= 4 && (!kPacked || W == 1), "product key shapes");
  constexpr int NS = 4 * W + 1;
  u64 o[2 * W], a[2 * W];
#pragma unroll
  for (int k = 0; k < 2 * W; ++k) {
    o[k] = 0;
    a[k] = ~0ULL;
  }
  double m = 0.0;
#pragma unroll
  for (u32 q = 0; q < kProductPerThread; ++q) {
    const u32 j = j0 + q * kProductKeyBlock + threadIdx.x;
    if (j >= P.R) break;
    u64 ox[W], oz[W];
    const double2 c = P.term_ij<W>(i, j, ox, oz);
#pragma unroll
    for (int w = 0; w < W; ++w) {
      o[2 * w] |= ox[w];
      a[2 * w] &= ox[w];
      o[2 * w + 1] |= oz[w];
      a[2 * w + 1] &= oz[w];
    }
    if constexpr (kMax) m = fmax(m, hypot(c.x, c.y));
    const std::size_t t = static_cast<std::size_t>(i) * P.R + j;
    keys[t] = kPacked ? ((ox[0] << 32) | oz[0]) : ox[0];
    idx[t] = static_cast<u32>(t);
  }
  // Block partials: waves by shuffles, then the waves through LDS.
  const u32 lane = threadIdx.x & 63U, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 2 * W; ++k) {
    u64 ok = o[k], ak = a[k];
    for (int d = 32; d > 0; d >>= 1) {
      ok |= __shfl_xor(ok, d, 64);
      ak &= __shfl_xor(ak, d, 64);
    }
    if (lane == 0) {
      lds[wave][k] = ok;
      lds[wave][2 * W + k] = ak;
    }
  }
  if constexpr (kMax) {
    for (int d = 32; d > 0; d >>= 1) m = fmax(m, __shfl_xor(m, d, 64));
  }
  if (lane == 0) lds[wave][4 * W] = static_cast<u64>(__double_as_longlong(m));
  __syncthreads();
  if (threadIdx.x < static_cast<u32>(NS)) {
    const u32 k = threadIdx.x;
    u64 v = lds[0][k];
    for (u32 w = 1; w < kProductKeyBlock / 64; ++w) {
      const u64 u = lds[w][k];
      v = k < 2 * W ? (v | u) : k < 4 * W ? (v & u) : (u > v ? u : v);
    }
    block_stats[static_cast<std::size_t>(k) * stride + unit] = v;
  }
  // The LDS partials are read before the block's next unit writes them.
  __syncthreads();
}

}  // namespace wg

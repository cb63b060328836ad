// Expectation plan shared by the plain, adjoint and batched expectation paths.
#pragma once

#include <memory>

#include "wg_internal.hpp"

namespace wg {

// Keys per work item: one x-group's amplitude products feed up to kZ z-masks.
constexpr int kZ = 8;

struct ExpItem {
  u64 x;
  int nz;
  int pad;
  u64 z[kZ];
  u32 key[kZ];
};

struct ExpectationPlan {
  std::size_t n = 0;
  std::size_t T = 0;
  int device = 0;
  std::size_t num_keys = 0;
  std::size_t num_items = 0;
  std::size_t num_groups = 0;
  u64 half_range = 0;
  u32 blocks_per_item = 1;
  ExpItem* d_items = nullptr;     // [num_items]
  u32* d_term_key = nullptr;      // [T] term -> distinct key
  u64* d_group_x = nullptr;       // [num_groups]
  u32* d_group_begin = nullptr;   // [num_groups + 1] into the key arrays below
  u64* d_key_z = nullptr;         // keys in group order
  int* d_key_y = nullptr;         // |x & z| mod 4 per key
  double2* d_key_c = nullptr;     // coefficient sum per key (terms in order)
  ~ExpectationPlan();
};

std::shared_ptr<ExpectationPlan> get_plan(wg_pauli_sum* op, hipStream_t stream);

}  // namespace wg

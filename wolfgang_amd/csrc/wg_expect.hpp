// Expectation plan shared by the plain, adjoint and batched expectation paths.
#pragma once

#include <memory>
#include <vector>

#include "wg_internal.hpp"

namespace wg {

// Keys per work item: one x-group's amplitude products feed up to kZ z-masks.
constexpr int kZ = 8;
// Largest step-class dimension of the tiled off-diagonal sweep.
constexpr int kMaxClassLog = 3;
constexpr unsigned char kNoClasses = 0xFF;

// Largest tile of the tiled pass: 2^kTileMaxLog amplitudes (64 KiB of complex128
// in LDS) and the contiguous low-bit run inside it (2^kTileLowLog amplitudes =
// 128 bytes, one full cache line, per segment). The other kTileMaxLog -
// kTileLowLog local bits (the segment coordinates) map to any directions of the
// pass's subspace, including the thread-lane bits 3..5: every memory
// instruction of a wave still moves whole 128-byte lines.
constexpr int kTileMaxLog = 12;
constexpr int kTileLowLog = 3;
constexpr int kSegBits = kTileMaxLog - kTileLowLog;
constexpr int kMaxSegments = 1 << kSegBits;

// One x-group (or a slice of up to kZ of its keys).
struct ExpItem {
  u64 x;        // global x mask (0 = diagonal group)
  u64 xl;       // x in tile-local coordinates (tiled path)
  u64 pbit;     // local pivot bit of xl (tiled path)
  int nz;
  // Tiled path, off-diagonal items: the step signs of key j are
  // (-1)^{lam[j] . class(s)} with class(s) in {0,1}^cls_log; steps are
  // enumerated as e -> s = XOR of ebasis[l] over the set bits l of e, so that
  // class(s) = e >> (5 - cls_log). cls_log = kNoClasses: not decomposed.
  unsigned char cls_log;
  unsigned char ebasis[5];
  unsigned char lam[kZ];
  unsigned char variant;  // tiled path: the sweep kernel instance (set per pass)
  unsigned char pad;
  u64 z[kZ];    // global z masks
  u64 zl[kZ];   // z in tile-local coordinates: parity(z & b) = parity(z & rep) ^ parity(zl & u)
  u32 key[kZ];  // distinct-key ids
  int y[kZ];    // |x & z| mod 4
};

// A tiled pass: tiles are the cosets rep ^ U of the subspace U spanned by the
// low kTileLowLog bits, `a` further coordinate bits and `m` generator masks.
// Local coordinate u: bits [0, l) -> low index bits, [l, l+a) -> the coordinate
// positions, [l+a, k) -> generators; the global index is
// (rep ^ seg_off[u >> l]) | (u & (2^l - 1)).
struct PassDesc {
  int n;
  int l;
  int k;
  int nfree;
  u64 free_mask;  // index bits not spanned by the tile (coset representative bits)
  u64 seg_off[kMaxSegments];
  unsigned char free_pos[64];
};

// Owner sweep of a tiled pass's off-diagonal items: thread tid of a
// 512-thread block owns the tile points u = tid + 512 i (i < 8, local bits
// 9..11 = i) in registers and evaluates every pair {u, u ^ x} of an item once,
// from the point whose pivot bit is clear, reading only the partner from LDS.
// Each bundle may change the basis of the 9 segment coordinates (local bits
// 3..11) so that every item of a register-pivot bundle has local bit 11 set;
// items inside the low 3 bits use a lane pivot.
constexpr int kOwnerMaxSlots = 11;
constexpr int kOwnerMaxKeys = 4;

struct OwnerKey {
  u32 w;     // z compressed to the tile-index bits (bits 0..29: parity(z & rep(t)) =
             // parity(w & t)) | (z local bit 9) << 30 | (z local bit 10) << 31
  u32 zlt;   // z in bundle-local coordinates, bits 0..8 (the thread part of the sign)
  u32 info;  // bit 2: z local bit 11; bits 4..7: partial column (key index in the
             // item); bit 15: slot key in use
  u32 pad;
};

struct OwnerSlot {
  u32 item;   // item index relative to the pass's first item
  u32 hbit;   // lane pivot: the pivot bit of x (< 3)
  u32 addr0;  // (x local bits 0..8) << 4 | (x local bits 9..11) << 13: the partner of
              // representative j of thread tid is at byte (tid << 4) ^ addr0 ^ (j << 13)
              // (register pivot) or ((tid << 4) ^ addr0 | h << 15) + (j << 13) (lane pivot)
  u32 pad;
  OwnerKey key[kOwnerMaxKeys];
};

struct OwnerBundle {
  u64 seg_basis[kSegBits];  // global offsets of the 9 segment coordinates in this layout
  u32 kind;          // bit 0 odd |x & z|, bit 1 lane pivot, bits 2.. shape
  u32 pad;
  OwnerSlot slot[kOwnerMaxSlots];
};

// One key of the tiled sigma sweep, in the pass's item order: the keys of an
// item are consecutive and the last one carries kSigmaLast. The plan holds
// z, zl, xl and key; every call writes c and cd from its coefficients.
constexpr u32 kSigmaLast = 1U << 31;
constexpr u32 kSigmaPadKey = 0xFFFFFFFFU;  // key of a padding record (zero coefficients)
constexpr u32 kSigmaChunk = 4;             // records per scalar-load chunk; every pass's
                                           // list is padded to a multiple of it

struct SigmaKey {
  u64 z;       // global z mask (tile part of the sign: parity(z & rep))
  u32 zl;      // z in tile-local coordinates
  u32 xl;      // the item's x in tile-local coordinates | kSigmaLast on its last key
  u32 sh[3];   // high words of (-1)^{parity(zl & (I << 10))} as a double, I = 1, 2, 3
  u32 key;     // distinct-key id
  double2 c;   // (-1)^{parity(zl & xl)} c_key i^{y_key} (sigma), per call
  double2 cd;  // the same with conj(c_key) (sigma^dagger), per call
};

// The owner sweeps of all passes run in one launch: each pass with bundles
// takes a contiguous range of blocks (a multiple of 8, so the XCD of a block
// is the same for the pass-relative block index).
struct OwnerLaunch {
  u32 pass;          // pass index (descriptor)
  u32 owner_first;   // first bundle of the pass (plan-wide index)
  u32 groups;        // bundles of the pass
  u32 tile_blocks;
  u32 partial_base;
  u32 block_begin;   // first block of the pass in the launch
};

// Items of one sweep variant inside a pass: one launch.
struct SubLaunch {
  int variant;
  u32 first;   // relative to the pass's first item
  u32 count;
  u32 groups;  // blocks of kTileWaves items
};

struct PassLaunch {
  PassDesc desc;
  u32 first_item;    // items of this pass are [first_item, first_item + num_items)
  u32 num_items;
  u32 item_groups;   // blocks of kTileWaves items, over all sub-launches
  u32 tile_blocks;   // blocks per item group (multiple of 8)
  u32 partial_base;  // first partial row of this pass
  std::vector<SubLaunch> subs;
  u32 owner_first = 0;  // owner bundles of the pass (plan-wide indices), one launch
  u32 owner_count = 0;
  u32 sigma_begin = 0;  // the pass's sigma keys [sigma_begin, sigma_end); pass 0 also
  u32 sigma_end = 0;    // holds the diagonal items the forward sweeps separately
};

struct ExpectationPlan {
  std::size_t n = 0;
  std::size_t T = 0;
  int device = 0;
  std::size_t num_keys = 0;
  std::size_t num_items = 0;
  std::size_t num_groups = 0;
  std::size_t distinct_x = 0;        // G_x: distinct nonzero x + 1 if diagonal terms exist
  bool tiled = false;
  u32 blocks_per_item = 1;           // direct path
  std::vector<PassLaunch> passes;    // tiled path
  std::vector<OwnerBundle> bundles;  // owner sweep bundles of all passes
  OwnerBundle* d_bundles = nullptr;
  std::vector<OwnerLaunch> owner_table;  // passes with bundles, in launch order
  OwnerLaunch* d_owner_table = nullptr;
  u32* d_owner_entry = nullptr;          // [owner_blocks / 8] table entry of each group of 8 blocks
  u64 owner_blocks = 0;                  // blocks of the single owner launch
  u32 sigma_log_blocks = 0;          // tiled sigma grid: 2^sigma_log_blocks blocks (<= CUs)
  std::size_t sigma_keys = 0;        // tiled sigma key records, all passes
  SigmaKey* d_sigma_keys = nullptr;
  u32 diag_items = 0;                // tiled path: leading diagonal items swept by the
                                     // Walsh-Hadamard kernel (0: they ride in pass 0)
  u32 diag_keys = 0;
  u32 diag_blocks = 0;               // its grid; item i's partial rows are i * diag_blocks + c
  std::vector<u32> item_partial_offset;  // host copy; device copy below
  std::size_t partial_slots = 0;     // total partial entries (x kZ)
  ExpItem* d_items = nullptr;        // [num_items]
  u32* d_partial_offset = nullptr;   // [num_items] first partial slot of each item
  u32* d_partial_count = nullptr;    // [num_items] partial slots per item
  u32* d_term_key = nullptr;         // [T] term -> distinct key
  u64* d_group_x = nullptr;          // [num_groups]
  u32* d_group_begin = nullptr;      // [num_groups + 1] into the key arrays below
  u64* d_key_z = nullptr;            // keys in group order
  int* d_key_y = nullptr;            // |x & z| mod 4 per key
  double2* d_key_c = nullptr;        // coefficient sum per key (terms in order)
  u32* d_key_term_begin = nullptr;   // [num_keys + 1] CSR over keys in group order
  u32* d_key_terms = nullptr;        // [T] terms of each key, in term order
  PassDesc* d_passes = nullptr;      // [passes] descriptors read by the tiled kernel
  u32* d_flat_key = nullptr;         // [num_keys] group-order position -> distinct-key id
  // Host copies used to assemble batched launches.
  std::vector<ExpItem> direct_items;  // items before tiling (direct-sweep form)
  std::vector<u64> h_group_x;
  std::vector<u32> h_group_begin;
  std::vector<u64> h_key_z;
  std::vector<int> h_key_y;
  std::vector<double2> h_key_c;
  std::vector<u32> h_key_term_begin;  // [num_keys + 1] CSR over keys in group order
  std::vector<u32> h_key_terms;       // [T] terms of each key, in term order
  std::vector<u32> h_term_key;
  ~ExpectationPlan();
};

std::shared_ptr<ExpectationPlan> get_plan(wg_pauli_sum* op, hipStream_t stream);

// Device bytes the forward sequence of `plan` needs beside its inputs and
// outputs: per-key values (256-byte aligned) followed by the partial sums.
std::size_t forward_scratch_bytes(const ExpectationPlan& plan);

// Enqueues the forward sequence (sweeps, ordered reduction, finalize) of a
// plan on `stream`: E into the device location out_dev, per-term values when
// term_values is non-null; key_vals and partials are the caller's buffers laid
// out as forward_scratch_bytes describes. No allocation, no synchronization.
void forward_into(const double2* coeffs, const ExpectationPlan& plan, const double2* psi,
                  double2* term_values, double2* out_dev, double* key_vals, double2* partials,
                  hipStream_t stream);

// Forward + adjoint of one operator with at least one term (the work of
// wg_expectation_adjoint), enqueued on `stream`: E goes to the device location
// `dout`, per-term values and sigma / sigma^dagger where their pointers are
// non-null. Temporary buffers come from `scratch`; the caller synchronizes the
// stream before `scratch` is destroyed.
void adjoint_enqueue(wg_pauli_sum* op, const double* psi, std::size_t len, const double* coeffs,
                     double* term_values, double* sigma, double* sigma_dagger, hipStream_t stream,
                     Scratch& scratch, double2* dout);

// Device bytes tiled_adjoint_enqueue needs for `plan` (one buffer).
std::size_t tiled_adjoint_bytes(const ExpectationPlan& plan);

// Forward + adjoint of an operator with a tiled plan, all outputs optional:
// E into the device location dout, per-term values, sigma = H psi and
// H^dagger psi (coeffs, if non-null, replaces op_coeffs). Temporary buffers
// come from `buffer` (tiled_adjoint_bytes bytes). No allocation, no memset
// node, no synchronization: the sequence can be captured into a graph.
void tiled_adjoint_enqueue(const ExpectationPlan& plan, const double2* op_coeffs, const double2* psi,
                           const double2* coeffs, double2* term_values, double2* sigma,
                           double2* sigma_dagger, double2* dout, void* buffer, hipStream_t stream);

// Builds the tiled passes for the given items (host only; exposed for tests of
// the planner through wg_expectation_plan_info). With diag_separate the
// diagonal items are moved to the front of `items` (their count in
// *num_diag_front) and left out of the passes.
// Stops early (leaving `items` untouched and max_passes + 1 empty entries in
// `passes`) once the plan needs more than max_passes passes.
void plan_tiled_passes(std::size_t n, std::vector<ExpItem>& items, std::vector<PassLaunch>& passes,
                       bool diag_separate, u32* num_diag_front, std::size_t max_passes);

// Owner-sweep bundles for the off-diagonal items of every pass (appended to
// `bundles`; each pass records its range).
void plan_owner_bundles(const std::vector<ExpItem>& items, std::vector<PassLaunch>& passes,
                        std::vector<OwnerBundle>& bundles);

}  // namespace wg

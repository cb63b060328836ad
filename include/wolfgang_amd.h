/*
 * wolfgang_amd C ABI: the drop-in boundary for the packed symplectic Pauli hot
 * path on MI355X (gfx950).
 *
 * Every entry point takes plain pointers and sizes, returns an int status
 * (WG_OK or one of the WG_ERR_* categories) and leaves a message for
 * wg_last_error() on failure (thread-local). The categories map to the
 * reference's exception types: WG_ERR_INVALID <-> std::invalid_argument
 * (Python ValueError), WG_ERR_RUNTIME <-> std::runtime_error (RuntimeError),
 * WG_ERR_TYPE <-> TypeError, WG_ERR_OVERFLOW <-> std::overflow_error.
 *
 * Operators are device-resident handles mirroring the reference's
 * DevicePauliSum::Impl (reference src/hip/device_pauli_sum.hip.hpp:19-29):
 * SoA packed words x[t*W + w], z[t*W + w] (uint64; qubit q is bit q%64 of word
 * q/64, W = ceil(n/64)) and coefficients as interleaved complex128 (re, im).
 *
 * Unless an entry point says otherwise it runs on the handle's device, on the
 * given stream (NULL = the device's null stream), and synchronizes that stream
 * before returning, as the reference's HIP backend does
 * (reference docs/architecture/rocm_backend.md:441-448).
 */
#ifndef WOLFGANG_AMD_H
#define WOLFGANG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wg_pauli_sum wg_pauli_sum;
typedef struct wg_comm_matrix wg_comm_matrix;

enum {
  WG_OK = 0,
  WG_ERR_INVALID = 1,
  WG_ERR_RUNTIME = 2,
  WG_ERR_TYPE = 3,
  WG_ERR_OVERFLOW = 4
};

enum { WG_DTYPE_COMPLEX64 = 0, WG_DTYPE_COMPLEX128 = 1 };

/* Message of the last failing call on this thread ("" if none). */
const char* wg_last_error(void);

/* ABI revision of this header. */
int wg_abi_version(void);

/* Device discovery; replaces DevicePauliSum::hip_status / hip_available
 * (reference include/wolfgang/device_pauli_sum.hpp:111-112,
 * src/hip/device_pauli_sum.hip.cpp:171-221). A missing device is not an error:
 * count is 0 and wg_last_error() holds the skip reason. */
int wg_device_count(int* count);
int wg_device_info(int ordinal, char* name, size_t name_len, char* gfx_target, size_t gfx_len,
                   size_t* total_memory_bytes, int* compute_units);
int wg_runtime_versions(int* runtime_version, int* driver_version);

/* Device memory cache (no reference counterpart; the reference allocates and
 * frees per call). Device buffers of operators, matrices, plans, batches and
 * scratch are kept after they are freed, up to a byte limit per device
 * (default: a quarter of the device's memory), and reused for requests of the
 * same size class; a block re-enters the cache only after the device has
 * finished its queued work. wg_memory_cache_limit sets the limit for every
 * device (SIZE_MAX: back to the default) and trims the caches to it;
 * wg_memory_cache_release returns every cached block of a device to the
 * runtime; wg_memory_cache_info reports cached bytes and blocks and the bytes
 * handed out and not yet freed. */
int wg_memory_cache_limit(size_t bytes, size_t* previous);
int wg_memory_cache_release(int device);
int wg_memory_cache_info(int device, size_t* cached_bytes, size_t* cached_blocks,
                         size_t* live_bytes);

/* DevicePauliSum::from_host (reference include/wolfgang/device_pauli_sum.hpp:104-108;
 * src/hip/device_pauli_sum.hip.cpp:116-157). coeffs holds 2*num_terms doubles. */
int wg_pauli_sum_create(size_t num_qubits, size_t num_terms, const uint64_t* x, const uint64_t* z,
                        const double* coeffs, int device, wg_pauli_sum** out);

/* DevicePauliSum::to_host (reference src/hip/device_pauli_sum.hip.cpp:239-266).
 * x and z hold num_terms*words words, coeffs 2*num_terms doubles. */
int wg_pauli_sum_to_host(const wg_pauli_sum* op, uint64_t* x, uint64_t* z, double* coeffs);

/* num_qubits/num_terms/words/device accessors
 * (reference include/wolfgang/device_pauli_sum.hpp:152-155). */
int wg_pauli_sum_info(const wg_pauli_sum* op, size_t* num_qubits, size_t* num_terms,
                      size_t* words, int* device);

/* Device addresses of the packed buffers (read-only views, owned by op). */
int wg_pauli_sum_device_buffers(const wg_pauli_sum* op, const uint64_t** x, const uint64_t** z,
                                const double** coeffs);

/* Frees the device buffers on the owning device (reference
 * src/hip/device_pauli_sum.hip.cpp:86-99). NULL is accepted. */
void wg_pauli_sum_destroy(wg_pauli_sum* op);

/* DevicePauliSum::simplify (reference include/wolfgang/device_pauli_sum.hpp:117;
 * semantics of src/simplify.cpp:147-430): duplicate (x, z) keys summed in
 * construction order, |sum| <= atol + rtol*max|c| dropped, canonical
 * lexicographic (x0, z0, x1, z1, ...) output order. */
int wg_simplify(const wg_pauli_sum* op, double atol, double rtol, void* stream,
                wg_pauli_sum** out);

/* DevicePauliSum::matmul (reference include/wolfgang/device_pauli_sum.hpp:148-151;
 * semantics of src/multiply.cpp:73-402): lhs @ rhs with rhs acting first,
 * products in lhs-outer / rhs-inner order, then optionally simplified with
 * atol = 1e-12. max_intermediate_terms is checked before any allocation. */
int wg_matmul(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, int simplify_output,
              size_t max_intermediate_terms, void* stream, wg_pauli_sum** out);

/* DevicePauliSum::commutes_with_into (reference
 * include/wolfgang/device_pauli_sum.hpp:137-140; src/hip/commutation_hip.hip.cpp:72-135):
 * row-major lhs x rhs flags, 1 = commute, into a caller-owned HOST buffer of
 * exactly out_len = L*R bytes. */
int wg_commutes_with_host(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, uint8_t* out,
                          size_t out_len, size_t max_entries, void* stream);

/* DeviceCommutationMatrix::empty (reference
 * include/wolfgang/device_commutation_matrix.hpp:27-35;
 * src/hip/device_commutation_matrix.hip.cpp:177-212). The matrix is stored
 * as packed bits: row i occupies row_words = ceil(cols/64) uint64 words, bit j%64
 * of word j/64 is entry (i, j); bits past cols are zero. */
int wg_comm_matrix_create(size_t rows, size_t cols, int device, wg_comm_matrix** out);
void wg_comm_matrix_destroy(wg_comm_matrix* m);
int wg_comm_matrix_info(const wg_comm_matrix* m, size_t* rows, size_t* cols, size_t* row_words,
                        int* device);
int wg_comm_matrix_device_bits(const wg_comm_matrix* m, const uint64_t** bits);

/* DevicePauliSum::commutes_with_device_into (reference
 * include/wolfgang/device_pauli_sum.hpp:144-147; src/hip/commutation_hip.hip.cpp:137-209). */
int wg_commutes_with_device(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, wg_comm_matrix* out,
                            size_t max_entries, void* stream);

/* DeviceCommutationMatrix::to_host as dense uint8 (reference
 * src/hip/device_commutation_matrix.hip.cpp:225-262): out_len = rows*cols. */
int wg_comm_matrix_to_host_u8(const wg_comm_matrix* m, uint8_t* out, size_t out_len, void* stream);

/* Packed words to host: out_words = rows*row_words. */
int wg_comm_matrix_to_host_packed(const wg_comm_matrix* m, uint64_t* out, size_t out_words,
                                  void* stream);

/* copy_device_commutation_matrix_from_host_for_testing (reference
 * include/wolfgang/device_commutation_matrix.hpp:64-66,
 * src/hip/device_commutation_matrix.hip.cpp:372-391): overwrite the matrix
 * with len = rows*cols host bytes, row-major; a nonzero byte sets its bit.
 * Synchronizes `stream` before returning. */
int wg_comm_matrix_from_host_u8(wg_comm_matrix* m, const uint8_t* values, size_t len,
                                void* stream);

/* count_commuting / count_commuting_rows / count_commuting_cols and
 * conflict_degrees (reference include/wolfgang/device_commutation_matrix.hpp:37-39;
 * src/hip/device_commutation_matrix.hip.cpp:225-321; bindings/python/pauli_sum_py.cpp:598-661). axis -1: total into out[0];
 * axis 1: per-row counts (out_len = rows); axis 0: per-column counts
 * (out_len = cols). conflicts != 0 counts anticommuting entries instead. */
int wg_comm_matrix_count(const wg_comm_matrix* m, int axis, int conflicts, uint64_t* out,
                         size_t out_len, void* stream);

/* expectation_statevector_{complex128,complex64,device_pointer}
 * (reference include/wolfgang/device_pauli_sum.hpp:118-125;
 * src/expectation.cpp:37-155 semantics): E = sum_k c_k <psi|P_k|psi>.
 * psi_on_device != 0: psi is a device address on op's device (no copy).
 * out2 receives (Re E, Im E). */
int wg_expectation(const wg_pauli_sum* op, const void* psi, int dtype, size_t length,
                   int psi_on_device, void* stream, double* out2);

/* Allocation-free variant of wg_expectation. No reference counterpart: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). psi is a
 * device complex128 statevector on op's device (length 2^n); out2_device is a
 * 16-byte aligned DEVICE location receiving (Re E, Im E). The operator's
 * expectation plan must already be built (wg_expectation_plan_info or any
 * expectation call builds it; WG_ERR_INVALID otherwise). Partial sums use a
 * per-device static scratch of 256 MiB (held until the process ends); a plan
 * needing more is rejected with WG_ERR_INVALID. It calls no hipMalloc/hipFree,
 * enqueues on `stream` and returns without synchronizing. Calls that overlap
 * on one device share that device's scratch and race. Results are
 * bit-identical to wg_expectation's. */
int wg_expectation_noalloc(const wg_pauli_sum* op, const double* psi, size_t length,
                           double* out2_device, void* stream);

/* expectation_z_counts (reference include/wolfgang/pauli_sum.hpp:98,
 * src/expectation.cpp:157-247; binding bindings/python/pauli_sum_py.cpp:1614-1631):
 * E = sum_t c_t * sum_r counts[r] (-1)^{|z_t & b_r|} / sum_r counts[r] over
 * computational-basis count rows, for diagonal terms only. Row r's bitstring
 * is chars[offsets[r] .. offsets[r+1]) ('0'/'1', right-most character =
 * qubit 0; offsets has num_rows + 1 entries), counts[r] its count; all three
 * arrays are host memory. Fails with WG_ERR_INVALID and the reference's
 * messages, in its order: a term with an X or Y factor, a negative or
 * non-finite count, a non-positive total, a bitstring of the wrong length or
 * with another character. out2 receives (Re E, Im E). */
int wg_expectation_z_counts(const wg_pauli_sum* op, const char* chars, const size_t* offsets,
                            const double* counts, size_t num_rows, void* stream, double* out2);

/* Forward plus adjoint in one call, all pointers device-resident complex128 on
 * op's device (no reference counterpart: the new adjoint). coeffs (num_terms
 * complex128, may be NULL) replaces the operator's own coefficients for this
 * call, so a caller training the coefficients reuses one cached plan:
 *   out2            <- E = sum_k c_k <P_k>
 *   term_values     <- <psi|P_k|psi> per term (num_terms complex128; real for
 *                      Hermitian P_k), i.e. dE/dc_k          (may be NULL)
 *   sigma           <- H psi (length complex128), i.e. dE/dpsi*  (may be NULL)
 *   sigma_dagger    <- H^dagger psi (coefficients conjugated)    (may be NULL)
 * out2 may be NULL when E is not wanted. For states swept in tiles (18 qubits
 * and up) E and the per-term values come from the forward sweeps and sigma /
 * sigma^dagger from one sigma sweep over the same passes; with out2 and
 * term_values both NULL only the sigma sweep runs. */
int wg_expectation_adjoint(const wg_pauli_sum* op, const double* psi, size_t length,
                           const double* coeffs, double* term_values, double* sigma,
                           double* sigma_dagger, void* stream, double* out2);

/* Batched forward + adjoint over `count` (psi, H) instances on one device (no
 * reference counterpart). Per instance: psi[i] (2^n_i complex128, device),
 * optional coeffs[i] (num_terms complex128, device; NULL = the operator's own),
 * optional outputs term_values[i], sigma[i] = H psi, sigma_dagger[i] =
 * H^dagger psi; energies holds 2*count doubles (device) for E_i. All pointers
 * are fixed at creation; the coefficient buffers are read on every run (a
 * caller may update them between runs). Instances whose plan is direct (small
 * states or few x-groups) are cut into units served by one persistent kernel
 * from a device work queue; instances with a tiled plan run the tiled sweeps of
 * wg_expectation_adjoint with buffers allocated at creation, spread over
 * worker streams forked from and joined back to the caller's stream. Both parts
 * are captured into one HIP graph (or the tiled part is enqueued per run when
 * that timed faster at creation). tunables = {log2 forward chunk, sigma
 * outputs per unit, persistent grid, tiled worker streams (1..4), tiled part
 * inside the graph (0/1)} or NULL to pick them by timing the batch. */
typedef struct wg_batch wg_batch;
int wg_batch_create(size_t count, const wg_pauli_sum* const* ops, const double* const* psi,
                    const double* const* coeffs, double* const* term_values, double* const* sigma,
                    double* const* sigma_dagger, double* energies, const int* tunables,
                    void* stream, wg_batch** out);
/* Enqueues one run (one graph launch) on `stream` and returns without waiting. */
int wg_batch_run(wg_batch* batch, void* stream);
int wg_batch_sync(wg_batch* batch, void* stream);
/* tunables receives the 5 values above. */
int wg_batch_info(const wg_batch* batch, int* tunables, size_t* units, size_t* items);
/* Number of instances that run the tiled sweeps. */
int wg_batch_tiled_count(const wg_batch* batch, size_t* tiled);
void wg_batch_destroy(wg_batch* batch);

/* Batched Hamiltonian products (no reference counterpart; each pair follows
 * DevicePauliSum::matmul, src/hip/matmul_hip.hip.cpp:91-147, and its result is
 * bit-identical to wg_matmul on that pair): lhs[i] @ rhs[i] for `count` pairs
 * of uneven term and qubit counts on one device. Creation validates every pair
 * with matmul's messages and limits (max_intermediate_terms per pair), packs
 * the pairs into chunks that fit budget_bytes (0 = 16 GiB) and captures, per
 * chunk, one persistent kernel fed by a device work queue into a HIP graph.
 * With simplify (pairs of 1 to 256 qubits) the kernel writes each product's
 * sort key and index and the products are never materialized; the chunks are
 * sized for those records. Runs without simplify materialize the products in
 * chunks sized for them (graphs captured on the first such run). The operands
 * must outlive the batch and are read at every run. wg_product_batch_run
 * replays the graphs and writes `count` new operators to outs (simplified with
 * matmul's rule when simplify_output != 0); it synchronizes `stream` before
 * returning. wg_product_batch_info reports the chunks of the simplify path. */
typedef struct wg_product_batch wg_product_batch;
int wg_product_batch_create(size_t count, const wg_pauli_sum* const* lhs,
                            const wg_pauli_sum* const* rhs, size_t max_intermediate_terms,
                            size_t budget_bytes, void* stream, wg_product_batch** out);
int wg_product_batch_run(wg_product_batch* batch, int simplify_output, void* stream,
                         wg_pauli_sum** outs);
int wg_product_batch_info(const wg_product_batch* batch, size_t* pairs, size_t* chunks,
                          size_t* total_products);
void wg_product_batch_destroy(wg_product_batch* batch);

/* PauliSum::group_commuting(mode, "largest_first", max_terms_for_graph)
 * (reference include/wolfgang/pauli_sum.hpp:84-87, src/grouping.cpp:197-290):
 * greedy largest-first commuting groups of op's terms, full commutation when
 * full != 0, else qubit-wise. Writes the group count, the group sizes (host,
 * up to num_terms entries) and the member term indices (host, num_terms
 * entries; groups concatenated in creation order, members in placement
 * order). Identical to the reference's graph and streaming paths. Full mode
 * supports up to 262,144 terms (it builds the packed commutation matrix). */
int wg_group_commuting(const wg_pauli_sum* op, int full, void* stream, size_t* num_groups,
                       uint32_t* sizes, uint32_t* members);

/* Expectation plan summary for an operator (built and cached on first use;
 * no reference counterpart): distinct (x, z) keys, distinct x-masks counting
 * the diagonal group (G_x), sweeps over psi per evaluation (tiled passes, or
 * work items for the direct sweep) and whether the tiled sweep is used. */
int wg_expectation_plan_info(const wg_pauli_sum* op, size_t* distinct_keys, size_t* x_groups,
                             size_t* sweeps, int* tiled, void* stream);

/* Host-only verification of the tiled expectation plan (test hook; no
 * reference counterpart, no GPU needed): replays the tile decomposition of a
 * W = 1 operator on the CPU over psi (length 2^n, complex128) and writes <P>
 * per distinct (x, z) key in first-appearance order. */
int wg_debug_tiled_plan(size_t num_qubits, size_t num_terms, const uint64_t* x, const uint64_t* z,
                        const double* psi, double* key_vals, size_t* num_keys, size_t* num_passes);

/* Host-only: per-pass summary of the tiled plan of a W = 1 operator (test and
 * tuning hook): rows of {k, items, item groups, tile blocks, diagonal items,
 * off-diagonal items, owner bundles, owner slots}, up to max_passes rows;
 * *num_passes gets the count. */
int wg_debug_plan_passes(size_t num_qubits, size_t num_terms, const uint64_t* x, const uint64_t* z,
                         int* rows, size_t max_passes, size_t* num_passes);

/* Host-only: where the tiled plan of a W = 1 operator evaluates each distinct
 * key (first-appearance order; test hook, no reference counterpart):
 * key_pass[k] = its pass (-1: the diagonal Walsh-Hadamard sweep), key_kind[k]
 * = the kind word of the first owner bundle holding it (bit 0 odd |x & z|,
 * bit 1 lane pivot, bits 2.. shape; -1: none). Both arrays hold num_terms
 * entries; *num_keys gets the key count. */
int wg_debug_key_sweeps(size_t num_qubits, size_t num_terms, const uint64_t* x, const uint64_t* z,
                        int* key_pass, int* key_kind, size_t* num_keys);

/* Host-only replay of the tiled sigma sweep of a W = 1 operator with
 * 18 <= n <= 24 (test hook; no reference counterpart): builds the plan's items,
 * passes and sigma key records and evaluates them on the CPU over psi
 * (complex128, 2^n), writing sigma = H psi and H^dagger psi (2^n complex128
 * each). coeffs holds 2*num_terms doubles. *num_records gets the number of key
 * records (padding included). */
int wg_debug_sigma_plan(size_t num_qubits, size_t num_terms, const uint64_t* x, const uint64_t* z,
                        const double* coeffs, const double* psi, double* sigma, double* sigma_dag,
                        size_t* num_records);

/* Kernel timing with HIP events (no reference counterpart; measurement only).
 * While enabled, the library brackets each launch of its hot kernels
 * ("expect_direct", "commute_packed", "product", "radix_scatter", ...) with an
 * event pair on the launch stream; wg_profile_get resolves the pending pairs and
 * returns the accumulated milliseconds and launch count for one kernel name. */
int wg_profile_enable(int enabled);
int wg_profile_reset(void);
int wg_profile_get(const char* kernel, double* total_ms, uint64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* WOLFGANG_AMD_H */

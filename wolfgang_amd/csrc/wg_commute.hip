// Commutation matrix as a packed bit matrix.
//
// Entry (i, j) = 1 iff the parity of popcount((lx_i & rz_j) ^ (lz_i & rx_j))
// over all W words is even (reference src/detail/commutation.hpp:10-46,
// src/simd/commute_kernels_scalar.cpp:29-90). Row i of the output holds
// row_words = ceil(R/64) uint64 words; bit j%64 of word j/64 is entry (i, j).
//
// Kernel shape (gfx950, wave64): a block owns 64 lhs rows (one per lane) and
// kWordsPerBlock output words (64 rhs terms each). Every rhs term a wave visits
// is wave-uniform, so its masks are fetched through scalar loads and enter the
// VALU as SGPR operands; the per-pair cost is a handful of 32-bit AND/XOR ops,
// one popcount and one bit insert. Results are staged through LDS so each
// row's run of output words is written with contiguous stores.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "wg_internal.hpp"

namespace wg {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kWordsPerWave = 4;
constexpr int kWordsPerBlock = kWaves * kWordsPerWave;

// (a & b) ^ c in one VALU op: v_bitop3_b32 with the truth table of
// (S0 & S1) ^ S2 (S0 = 0xF0, S1 = 0xCC, S2 = 0xAA).
__device__ inline u32 and_xor(u32 a, u32 b, u32 c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0x6A); }

// XOR of every 32-bit half of (lx & rz) ^ (lz & rx) over the words: its
// popcount parity is the parity of the whole conflict popcount.
template <int WC>
__device__ inline u32 conflict_fold(const u64* __restrict__ rx, const u64* __restrict__ rz,
                                    std::size_t col, std::size_t W, const u64* lxr,
                                    const u64* lzr) {
  if constexpr (WC > 0) {
    u32 acc = 0;
#pragma unroll
    for (int w = 0; w < WC; ++w) {
      const u64 rzw = rz[col * WC + w], rxw = rx[col * WC + w];
      acc = and_xor(static_cast<u32>(lxr[w]), static_cast<u32>(rzw), acc);
      acc = and_xor(static_cast<u32>(lxr[w] >> 32), static_cast<u32>(rzw >> 32), acc);
      acc = and_xor(static_cast<u32>(lzr[w]), static_cast<u32>(rxw), acc);
      acc = and_xor(static_cast<u32>(lzr[w] >> 32), static_cast<u32>(rxw >> 32), acc);
    }
    return acc;
  } else {
    u64 acc = 0;
    for (std::size_t w = 0; w < W; ++w) {
      acc ^= (lxr[w] & rz[col * W + w]) ^ (lzr[w] & rx[col * W + w]);
    }
    return static_cast<u32>(acc) ^ static_cast<u32>(acc >> 32);
  }
}

// WC = compile-time word count (1, 2, 4) or 0 for a runtime count. For WC == 0
// the lane's row words are re-read from global memory (cached) per pair.
template <int WC>
__global__ __launch_bounds__(kBlock) void commute_packed_kernel(
    const u64* __restrict__ lx, const u64* __restrict__ lz, const u64* __restrict__ rx,
    const u64* __restrict__ rz, std::size_t L, std::size_t R, std::size_t W,
    std::size_t row_words, std::size_t word_blocks, u64* __restrict__ out) {
  __shared__ u64 stage[64][kWordsPerBlock + 1];
  const unsigned lane = threadIdx.x & 63;
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const std::size_t wb = blockIdx.x % word_blocks;
  const std::size_t rb = blockIdx.x / word_blocks;
  const std::size_t row = rb * 64 + lane;
  const std::size_t safe_row = row < L ? row : (L - 1);

  constexpr int kRegWords = WC > 0 ? WC : 1;
  u64 lxr[kRegWords], lzr[kRegWords];
  if constexpr (WC > 0) {
#pragma unroll
    for (int w = 0; w < WC; ++w) {
      lxr[w] = lx[safe_row * WC + w];
      lzr[w] = lz[safe_row * WC + w];
    }
  }

  for (int k = 0; k < kWordsPerWave; ++k) {
    const std::size_t word = wb * kWordsPerBlock + wave * kWordsPerWave + k;
    u64 bits = 0;
    if (word < row_words) {
      const std::size_t c0 = word * 64;
      const std::size_t ncols = (R - c0) < 64 ? (R - c0) : 64;
      if constexpr (WC > 0) {
        if (ncols == 64) {
          // Odd-parity (anticommuting) bits first, inverted once per word.
          u32 lo = 0, hi = 0;
#pragma unroll
          for (int b = 0; b < 32; ++b) {
            const u32 f = conflict_fold<WC>(rx, rz, c0 + b, W, lxr, lzr);
            lo |= (__popc(f) & 1U) << b;
          }
#pragma unroll
          for (int b = 0; b < 32; ++b) {
            const u32 f = conflict_fold<WC>(rx, rz, c0 + 32 + b, W, lxr, lzr);
            hi |= (__popc(f) & 1U) << b;
          }
          bits = ~((static_cast<u64>(hi) << 32) | lo);
        } else {
          for (std::size_t b = 0; b < ncols; ++b) {
            const u32 f = conflict_fold<WC>(rx, rz, c0 + b, W, lxr, lzr);
            bits |= static_cast<u64>((__popc(f) & 1U) ^ 1U) << b;
          }
        }
      } else {
        for (std::size_t b = 0; b < ncols; ++b) {
          u64 acc = 0;
          for (std::size_t w = 0; w < W; ++w) {
            acc ^= (lx[safe_row * W + w] & rz[(c0 + b) * W + w]) ^
                   (lz[safe_row * W + w] & rx[(c0 + b) * W + w]);
          }
          const u32 f = static_cast<u32>(acc) ^ static_cast<u32>(acc >> 32);
          bits |= static_cast<u64>((__popc(f) & 1U) ^ 1U) << b;
        }
      }
    }
    stage[lane][wave * kWordsPerWave + k] = bits;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 64 * kWordsPerBlock; i += kBlock) {
    const int r = i / kWordsPerBlock;
    const int w = i % kWordsPerBlock;
    const std::size_t orow = rb * 64 + r;
    const std::size_t oword = wb * kWordsPerBlock + w;
    if (orow < L && oword < row_words) out[orow * row_words + oword] = stage[r][w];
  }
}

// Dense uint8 view: out[i*cols + j] = bit (i, j). A block covers kUnpackBytes
// consecutive columns of one row (one 64-bit division per block, none per
// byte); each thread expands 16 bits of one packed word into 16 bytes.
constexpr unsigned kUnpackBytes = 256 * 16;
__device__ inline u32 spread4(u32 v) { return ((v & 0xFU) * 0x00204081U) & 0x01010101U; }

__global__ __launch_bounds__(256) void unpack_u8_kernel(const u64* __restrict__ bits,
                                                        std::size_t cols, std::size_t row_words,
                                                        std::size_t chunks_per_row,
                                                        std::uint8_t* __restrict__ out) {
  const std::size_t r = blockIdx.x / chunks_per_row;
  const std::size_t c0 =
      (blockIdx.x - r * chunks_per_row) * kUnpackBytes + static_cast<std::size_t>(threadIdx.x) * 16;
  if (c0 >= cols) return;
  // c0 is a multiple of 16, so its 16 columns share one packed word.
  const u32 v = static_cast<u32>(bits[r * row_words + c0 / 64] >> (c0 % 64)) & 0xFFFFU;
  std::uint8_t* dst = out + r * cols + c0;
  if (c0 + 16 <= cols && (reinterpret_cast<std::uintptr_t>(dst) & 15U) == 0) {
    *reinterpret_cast<uint4*>(dst) =
        make_uint4(spread4(v), spread4(v >> 4), spread4(v >> 8), spread4(v >> 12));
  } else {
    const std::size_t n = cols - c0 < 16 ? cols - c0 : 16;
    for (std::size_t j = 0; j < n; ++j) dst[j] = static_cast<std::uint8_t>((v >> j) & 1U);
  }
}

// Row counts: one wave per row (grid-stride over rows), popcount of the row's
// words; optional per-row output. The total is reduced per block and added
// with one integer atomic per block (exact in any order).
constexpr unsigned kCountBlocks = 2048;
__global__ __launch_bounds__(256) void count_rows_kernel(const u64* __restrict__ bits,
                                                         std::size_t rows, std::size_t row_words,
                                                         unsigned long long* row_counts,
                                                         unsigned long long* total) {
  __shared__ unsigned long long wave_sum[4];
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long mine = 0;
  for (std::size_t row = static_cast<std::size_t>(blockIdx.x) * 4 + wave; row < rows;
       row += static_cast<std::size_t>(gridDim.x) * 4) {
    unsigned long long c = 0;
    for (std::size_t w = lane; w < row_words; w += 64) c += __popcll(bits[row * row_words + w]);
    for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane == 0 && row_counts) row_counts[row] = c;
    mine += c;
  }
  if (!total) return;
  if (lane == 0) wave_sum[wave] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long s = wave_sum[0] + wave_sum[1] + wave_sum[2] + wave_sum[3];
    if (s) atomicAdd(total, s);
  }
}

// Column counts: thread per packed word column over a chunk of rows, with the
// 64 per-bit counters of its word in registers (one coalesced 8-byte load per
// row and thread); one atomic per nonzero counter at the end.
constexpr std::size_t kColRowChunk = 1024;
__global__ __launch_bounds__(64) void count_cols_kernel(const u64* __restrict__ bits,
                                                        std::size_t rows, std::size_t cols,
                                                        std::size_t row_words,
                                                        std::size_t word_groups,
                                                        unsigned long long* col_counts) {
  const std::size_t wg = blockIdx.x % word_groups;
  const std::size_t chunk = blockIdx.x / word_groups;
  const std::size_t word = wg * 64 + threadIdx.x;
  if (word >= row_words) return;
  const std::size_t r0 = chunk * kColRowChunk;
  const std::size_t r1 = (r0 + kColRowChunk) < rows ? (r0 + kColRowChunk) : rows;
  u32 cnt[64];
#pragma unroll
  for (int b = 0; b < 64; ++b) cnt[b] = 0;
  for (std::size_t r = r0; r < r1; ++r) {
    const u64 v = bits[r * row_words + word];
    const u32 lo = static_cast<u32>(v), hi = static_cast<u32>(v >> 32);
#pragma unroll
    for (int b = 0; b < 32; ++b) {
      cnt[b] += (lo >> b) & 1U;
      cnt[32 + b] += (hi >> b) & 1U;
    }
  }
#pragma unroll
  for (int b = 0; b < 64; ++b) {
    const std::size_t col = word * 64 + b;
    if (cnt[b] && col < cols) atomicAdd(&col_counts[col], static_cast<unsigned long long>(cnt[b]));
  }
}

void launch_unpack(const u64* bits, std::size_t rows, std::size_t cols, std::size_t row_words,
                   std::uint8_t* dense, hipStream_t stream) {
  const std::size_t chunks = (cols + kUnpackBytes - 1) / kUnpackBytes;
  ProfileScope prof("unpack_u8", stream);
  unpack_u8_kernel<<<grid_for(checked_mul(rows, chunks, "unpack grid"), 1), 256, 0, stream>>>(
      bits, cols, row_words, chunks, dense);
  check_launch("unpack_u8");
}

void check_entries(std::size_t L, std::size_t R, std::size_t max_entries) {
  // Reference src/detail/checked_arithmetic.hpp:49-63.
  if (L != 0 && R > SIZE_MAX / L) {
    fail_invalid("commutation matrix entry count overflows size_t");
  }
  if (L * R > max_entries) {
    fail_invalid("commutation matrix entry count exceeds max_commutation_matrix_entries");
  }
}

}  // namespace

// Packed commutation bits of (lx, lz) [L x W] against (rx, rz) [R x W] into
// out [L x row_words]: bit c of row r is 1 when the terms commute.
void launch_commute_raw(const u64* lx, const u64* lz, std::size_t L, const u64* rx, const u64* rz,
                        std::size_t R, std::size_t W, u64* out, std::size_t row_words,
                        hipStream_t stream) {
  if (L == 0 || R == 0) return;
  const std::size_t word_blocks = (row_words + kWordsPerBlock - 1) / kWordsPerBlock;
  const std::size_t row_blocks = (L + 63) / 64;
  const unsigned grid = grid_for(checked_mul(word_blocks, row_blocks, "commutation grid"), 1);
  ProfileScope prof("commute_packed", stream);
  switch (W) {
    case 0:
      // Zero-qubit operators: every pair commutes.
      commute_packed_kernel<0><<<grid, kBlock, 0, stream>>>(lx, lz, rx, rz, L, R, 0, row_words,
                                                            word_blocks, out);
      break;
    case 1:
      commute_packed_kernel<1><<<grid, kBlock, 0, stream>>>(lx, lz, rx, rz, L, R, 1, row_words,
                                                            word_blocks, out);
      break;
    case 2:
      commute_packed_kernel<2><<<grid, kBlock, 0, stream>>>(lx, lz, rx, rz, L, R, 2, row_words,
                                                            word_blocks, out);
      break;
    case 4:
      commute_packed_kernel<4><<<grid, kBlock, 0, stream>>>(lx, lz, rx, rz, L, R, 4, row_words,
                                                            word_blocks, out);
      break;
    default:
      commute_packed_kernel<0><<<grid, kBlock, 0, stream>>>(lx, lz, rx, rz, L, R, W, row_words,
                                                            word_blocks, out);
      break;
  }
  check_launch("commute_packed");
}

namespace {

void launch_commute(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, u64* out,
                    std::size_t row_words, hipStream_t stream) {
  launch_commute_raw(lhs->x, lhs->z, lhs->num_terms, rhs->x, rhs->z, rhs->num_terms, lhs->words,
                     out, row_words, stream);
}

void validate_operands(const char* operation, const wg_pauli_sum* lhs, const wg_pauli_sum* rhs) {
  require_op(lhs);
  require_op(rhs);
  require_same_device(operation, lhs, rhs);
  if (lhs->num_qubits != rhs->num_qubits) {
    fail_invalid("PauliSum commutes_with requires the same num_qubits");
  }
}

}  // namespace
}  // namespace wg

using namespace wg;

extern "C" {

int wg_commutes_with_host(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, uint8_t* out,
                          size_t out_len, size_t max_entries, void* stream_ptr) {
  return guard([&] {
    validate_operands("HIP commutes_with", lhs, rhs);
    const std::size_t L = lhs->num_terms, R = rhs->num_terms;
    check_entries(L, R, max_entries);
    if (out_len != L * R) {
      fail_invalid("HIP commutes_with output buffer size does not match entry count");
    }
    if (L * R == 0) return;
    if (out == nullptr) fail_invalid("commutes_with output pointer must be non-null");
    DeviceScope scope(lhs->device);
    hipStream_t stream = as_stream(stream_ptr);
    Scratch scratch(lhs->device);
    const std::size_t row_words = (R + 63) / 64;
    u64* bits = scratch.take<u64>(checked_mul(L, row_words, "commutation matrix"));
    std::uint8_t* dense = scratch.take<std::uint8_t>(L * R);
    launch_commute(lhs, rhs, bits, row_words, stream);
    launch_unpack(bits, L, R, row_words, dense, stream);
    check_hip(hipMemcpyAsync(out, dense, L * R, hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
  });
}

int wg_comm_matrix_create(size_t rows, size_t cols, int device, wg_comm_matrix** out) {
  return guard([&] {
    *out = nullptr;
    validate_device(device);
    (void)checked_mul(rows, cols, "HIP device commutation matrix entries");
    std::unique_ptr<wg_comm_matrix> m(new wg_comm_matrix());
    m->rows = rows;
    m->cols = cols;
    m->row_words = (cols + 63) / 64;
    m->device = device;
    DeviceScope scope(device);
    const std::size_t words = checked_mul(rows, m->row_words, "HIP device commutation matrix");
    m->bits = device_alloc<u64>(words);
    if (words) check_hip(hipMemset(m->bits, 0, words * sizeof(u64)), "clear device buffer");
    *out = m.release();
  });
}

void wg_comm_matrix_destroy(wg_comm_matrix* m) { delete m; }

int wg_comm_matrix_info(const wg_comm_matrix* m, size_t* rows, size_t* cols, size_t* row_words,
                        int* device) {
  return guard([&] {
    if (m == nullptr) fail_runtime("DeviceCommutationMatrix is empty or moved-from");
    if (rows) *rows = m->rows;
    if (cols) *cols = m->cols;
    if (row_words) *row_words = m->row_words;
    if (device) *device = m->device;
  });
}

int wg_comm_matrix_device_bits(const wg_comm_matrix* m, const uint64_t** bits) {
  return guard([&] {
    if (m == nullptr) fail_runtime("DeviceCommutationMatrix is empty or moved-from");
    *bits = m->bits;
  });
}

int wg_commutes_with_device(const wg_pauli_sum* lhs, const wg_pauli_sum* rhs, wg_comm_matrix* out,
                            size_t max_entries, void* stream_ptr) {
  return guard([&] {
    require_op(lhs);
    require_op(rhs);
    if (out == nullptr) fail_runtime("DeviceCommutationMatrix is empty or moved-from");
    require_same_device("HIP commutes_with_device", lhs, rhs);
    if (out->device != lhs->device) {
      fail_invalid(std::string("HIP commutes_with_device output requires operands on the same "
                               "accelerator backend and same device (operation=HIP "
                               "commutes_with_device output, left_backend=hip, left_device=") +
                   std::to_string(lhs->device) + ", right_backend=hip, right_device=" +
                   std::to_string(out->device) + ", failure_stage=pre_allocation)");
    }
    if (lhs->num_qubits != rhs->num_qubits) {
      fail_invalid("PauliSum commutes_with requires the same num_qubits");
    }
    const std::size_t L = lhs->num_terms, R = rhs->num_terms;
    check_entries(L, R, max_entries);
    if (out->rows != L || out->cols != R) {
      fail_invalid("HIP commutes_with_device output shape does not match operand term counts");
    }
    if (L * R == 0) return;
    DeviceScope scope(lhs->device);
    hipStream_t stream = as_stream(stream_ptr);
    launch_commute(lhs, rhs, out->bits, out->row_words, stream);
    sync(stream);
  });
}

int wg_comm_matrix_to_host_u8(const wg_comm_matrix* m, uint8_t* out, size_t out_len,
                              void* stream_ptr) {
  return guard([&] {
    if (m == nullptr) fail_runtime("DeviceCommutationMatrix is empty or moved-from");
    const std::size_t entries = m->rows * m->cols;
    if (out_len != entries) fail_invalid("output buffer size does not match matrix entry count");
    if (entries == 0) return;
    DeviceScope scope(m->device);
    hipStream_t stream = as_stream(stream_ptr);
    Scratch scratch(m->device);
    std::uint8_t* dense = scratch.take<std::uint8_t>(entries);
    launch_unpack(m->bits, m->rows, m->cols, m->row_words, dense, stream);
    check_hip(hipMemcpyAsync(out, dense, entries, hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
  });
}

int wg_comm_matrix_to_host_packed(const wg_comm_matrix* m, uint64_t* out, size_t out_words,
                                  void* stream_ptr) {
  return guard([&] {
    if (m == nullptr) fail_runtime("DeviceCommutationMatrix is empty or moved-from");
    const std::size_t words = m->rows * m->row_words;
    if (out_words != words) fail_invalid("output buffer size does not match packed word count");
    if (words == 0) return;
    DeviceScope scope(m->device);
    hipStream_t stream = as_stream(stream_ptr);
    check_hip(hipMemcpyAsync(out, m->bits, words * sizeof(u64), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
  });
}

int wg_comm_matrix_from_host_u8(wg_comm_matrix* m, const uint8_t* values, size_t len,
                                void* stream_ptr) {
  return guard([&] {
    if (m == nullptr) fail_runtime("DeviceCommutationMatrix is empty or moved-from");
    if (len != m->rows * m->cols) {
      fail_invalid("DeviceCommutationMatrix testing copy size does not match matrix entries");
    }
    if (len == 0) return;
    if (values == nullptr) fail_invalid("testing copy source must be non-null");
    // Pack on the host: entry (i, j) nonzero -> bit j % 64 of word j / 64 of row i.
    std::vector<u64> packed(m->rows * m->row_words, 0);
    for (std::size_t i = 0; i < m->rows; ++i) {
      const uint8_t* row = values + i * m->cols;
      u64* dst = packed.data() + i * m->row_words;
      for (std::size_t j = 0; j < m->cols; ++j) {
        if (row[j] != 0) dst[j >> 6] |= u64{1} << (j & 63);
      }
    }
    DeviceScope scope(m->device);
    hipStream_t stream = as_stream(stream_ptr);
    check_hip(hipMemcpyAsync(m->bits, packed.data(), packed.size() * sizeof(u64),
                             hipMemcpyHostToDevice, stream),
              "DeviceCommutationMatrix testing copy");
    sync(stream);
  });
}

int wg_comm_matrix_count(const wg_comm_matrix* m, int axis, int conflicts, uint64_t* out,
                         size_t out_len, void* stream_ptr) {
  return guard([&] {
    if (m == nullptr) fail_runtime("DeviceCommutationMatrix is empty or moved-from");
    if (axis != -1 && axis != 0 && axis != 1) fail_invalid("axis must be None, 0, or 1");
    const std::size_t want = axis == -1 ? 1 : (axis == 1 ? m->rows : m->cols);
    if (out_len != want) fail_invalid("count output length does not match the requested axis");
    for (std::size_t i = 0; i < want; ++i) out[i] = 0;
    if (m->rows == 0 || m->cols == 0) return;
    DeviceScope scope(m->device);
    hipStream_t stream = as_stream(stream_ptr);
    Scratch scratch(m->device);
    unsigned long long* dev = scratch.take<unsigned long long>(want);
    check_hip(hipMemsetAsync(dev, 0, want * sizeof(unsigned long long), stream), "clear counts");
    if (axis == 0) {
      const std::size_t groups = (m->row_words + 63) / 64;
      const std::size_t chunks = (m->rows + kColRowChunk - 1) / kColRowChunk;
      ProfileScope prof("count_cols", stream);
      count_cols_kernel<<<grid_for(groups * chunks, 1), 64, 0, stream>>>(
          m->bits, m->rows, m->cols, m->row_words, groups, dev);
      check_launch("count_cols");
    } else {
      const unsigned grid =
          static_cast<unsigned>(std::min<std::size_t>((m->rows + 3) / 4, kCountBlocks));
      ProfileScope prof("count_rows", stream);
      count_rows_kernel<<<grid, 256, 0, stream>>>(m->bits, m->rows, m->row_words,
                                                  axis == 1 ? dev : nullptr,
                                                  axis == -1 ? dev : nullptr);
      check_launch("count_rows");
    }
    check_hip(hipMemcpyAsync(out, dev, want * sizeof(u64), hipMemcpyDeviceToHost, stream),
              "copy device buffer to host");
    sync(stream);
    if (conflicts) {
      // Anticommuting entries = line length - commuting entries.
      const std::size_t line = axis == -1 ? m->rows * m->cols : (axis == 1 ? m->cols : m->rows);
      for (std::size_t i = 0; i < want; ++i) out[i] = line - out[i];
    }
  });
}

}  // extern "C"

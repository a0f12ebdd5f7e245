// leo_kernels.hpp — device kernels of the Leopard engines and their launchers.
//
// The work array is a block of rows in HBM with a fixed row pitch; each call
// processes `cols` bytes of every row (a column chunk of the shards).  A
// GF(2^16) element is split over a 64-byte chunk (32 low bytes, then 32 high
// bytes, leopard.go:907-925); a GF(2^8) element is one byte.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace rsx {

struct LeoDev {
    int bits = 0;                     // 16 or 8
    unsigned modulus = 0;
    const uint32_t *tables = nullptr; // (modulus+1) multiply records
    const uint16_t *skew = nullptr;   // fftSkew, `modulus` entries
    const uint16_t *walsh = nullptr;  // logWalsh, modulus + 1 entries
};

struct LeoWork {
    uint8_t *base = nullptr;  // row 0
    size_t pitch = 0;         // bytes between rows
    size_t cols = 0;          // bytes processed per row (multiple of 64)
};

// work[t] = ptrs[t] ? (logs ? ptrs[t]*exp(logs[t]) : ptrs[t]) : 0 for t < nrows,
// reading ptrs[t] + col_off.  ptrs and logs are device arrays.
hipError_t leo_load_rows(const LeoDev &f, const uint8_t *const *ptrs, const uint16_t *logs,
                         size_t col_off, int nrows, const LeoWork &w, hipStream_t s);

// One radix-4 step (two butterfly layers) at distance `dist` over groups
// r < mtrunc (step 4*dist).  `skew` points at the skew used for iend = r+dist:
// encoder IFFT passes fftSkew + m-1 + group*m with bias 0 (skewLUT[iend]),
// decoder IFFT and FFT pass fftSkew with bias -1 (skewLUT[iend-1]).
hipError_t leo_step4(const LeoDev &f, bool inverse, const LeoWork &w, int dist, int mtrunc,
                     const uint16_t *skew, int bias, hipStream_t s);

// Final IFFT radix-2 layer: ifftDIT2(work[i], work[i+dist], log_m) for i < dist.
hipError_t leo_ifft2_last(const LeoDev &f, const LeoWork &w, int dist, uint16_t log_m,
                          hipStream_t s);

// Final FFT radix-2 layer: fftDIT2(work[r], work[r+1], skew[r]) for even r < mtrunc.
hipError_t leo_fft2_last(const LeoDev &f, const LeoWork &w, int mtrunc, hipStream_t s);

// dst[t] ^= src[t] for t < nrows (work arrays with the same pitch/cols).
hipError_t leo_xor_rows(const LeoWork &dst, const LeoWork &src, int nrows, hipStream_t s);

// Formal derivative of leopard.go:619-623, out of place:
// out[t] = in[t] ^ XOR over bits b with (t & 2^b) == 0 of in[t | 2^b].
hipError_t leo_derivative(const LeoWork &in, const LeoWork &out, int n, hipStream_t s);

// dst[q] + col_off = logs ? work[rows[q]]*exp(logs[q]) : work[rows[q]] for q < count.
hipError_t leo_store_rows(const LeoDev &f, const LeoWork &w, const int *rows, const uint16_t *logs,
                          uint8_t *const *dst, size_t col_off, int count, hipStream_t s);

// out[j] ^= exp(logs[j]) * in over the GF(2^16) split layout (LowLevel
// GF16MulSliceXor[8]); len a multiple of 64, pointers 16-byte aligned.
struct Gf16MulXor {
    const uint8_t *in;
    uint8_t *out[8];
    uint16_t logs[8];
    int nout;
    size_t len;
};
hipError_t leo_gf16_mul_xor(const LeoDev &f, const Gf16MulXor &a, hipStream_t s);

// Error locators of one erasure pattern on the device (leopard.go:496-510,
// GF(2^16)): ind holds the mtrunc = m + k 0/1 erasure indicators in work-row
// order (leo_erasure_indicator); the kernel runs the truncated Walsh-Hadamard
// transform, the product with logWalsh and the full transform in LDS (one
// workgroup, 128 KiB), then writes errlocs[0, n) and, for the nrows output
// rows, the reveal logs logs[q] = modulus - errlocs[rows[q]] (leopard.go:640-646).
hipError_t leo_error_locators_device(const LeoDev &f, const uint8_t *ind, int mtrunc, int n, const int *rows,
                                     int nrows, uint16_t *errlocs, uint16_t *logs, hipStream_t s);

// ---- LDS-resident transforms (leo_fused.hip) ----
// Encode keeps one group's m rows of a column slice in LDS and, with several
// groups, their XOR accumulator (m * slice bytes) in registers; reconstruct
// keeps all n rows in LDS and n * slice bytes in registers for the formal
// derivative.  R rows take slice * (R + 1) bytes of LDS.
constexpr size_t kLeoFusedEncodeLds = 160u << 10;
constexpr size_t kLeoFusedDecodeLds = 64u << 10;
// Reconstructs with n * slice within this many bytes run a kernel variant
// that keeps half the derivative registers (more workgroups per CU).
constexpr size_t kLeoDerivSmall = 32u << 10;
// Threads per workgroup of the LDS-resident kernels (512 where one radix-4
// step over the tile has that many items and LDS leaves room for one
// workgroup per CU), and the uint4 slots per thread holding the derivative.
constexpr int kLeoTileThreads = 256;
constexpr int kLeoTileThreadsMax = 512;
constexpr int kLeoDerivMax = static_cast<int>(kLeoFusedDecodeLds / 16 / kLeoTileThreads);
unsigned leo_tile_threads(int bits, int rows, int slice, size_t lds);

struct LeoEncodeArgs {
    const uint8_t *const *data;  // groups*m row pointers, group-major, null past row k
    uint8_t *const *parity;      // p row pointers
    size_t size;                 // bytes per shard (multiple of 64)
    int k, p, m, groups;         // m = ceilPow2(p)
    int slice;                   // bytes per row per workgroup (power of two >= 64; 32 = one
                                 // GF(2^16) unit of a chunk)
    uint32_t *flag;              // Verify: when set, compare the parity rows with `parity`
    uint32_t flag_value;         // instead of writing them; a mismatch stores flag_value
};

struct LeoDecodeArgs {
    const uint8_t *const *rows;  // n row pointers: parity, zeros, data, zeros (null = zero/erased)
    const uint16_t *errlocs;     // n logs multiplied into the rows as they load
    const int *out_rows;         // nout work rows written back ...
    const uint16_t *out_logs;    // ... times exp(out_logs[q]) ...
    uint8_t *const *out;         // ... to out[q]
    size_t size;
    int n, nout, ifft_trunc, fft_trunc;
    int slice;
    const uint32_t *need;  // FFT groups that hold an output row (leo_fft_need), or null: all
};

// FFT pruning (leopard.go:1306-1500 errorBitfield): for every FFT level of an
// n-row transform, one bit per butterfly group, set when a row of the group
// is an output row; groups without one are skipped (their rows are never
// written out).  Levels in FFT order: radix-4 steps at dist = n/4, n/16, ...
// (groups of 4*dist rows), then, for odd log2 n, the radix-2 layer (pairs).
// Each level starts on a 32-bit word.
std::vector<uint32_t> leo_fft_need(int n, const std::vector<int> &out_rows);

// One Leopard item of a batched launch: op 0 = encode (enc), 1 = reconstruct
// (dec); its column slices are global slices [slice0, slice0 + slices).
struct LeoBatchTask {
    LeoEncodeArgs enc;
    LeoDecodeArgs dec;
    uint32_t slice0;
    uint32_t op;
};

// Runs every task's slices (tasks sorted by slice0, all of one field) with
// `grid` persistent blocks of 256 threads and `lds` bytes (the largest task).
// deriv_bytes: the largest n * slice of the reconstruct tasks (0 if none).
// quarter: GF(2^16) tasks whose slices are 16-byte quarter chunks (every
// task of the launch).
hipError_t leo_batch_launch(const LeoDev &f, const LeoBatchTask *tasks, int ntasks, uint32_t nslices, size_t lds,
                            unsigned grid, size_t deriv_bytes, bool quarter, hipStream_t s);

// Slice width for `regions` LDS regions of `rows` rows, radix-4 steps over
// `step_rows` rows and shards of `size` bytes, or 0 when even the narrowest
// slice (64 bytes; 32 for GF(2^16)) exceeds `lds_budget`.
int leo_fused_slice(int bits, int rows, int regions, int step_rows, size_t size, size_t lds_budget);
// 16 when a GF(2^16) reconstruct of n work rows runs as one LDS tile of
// quarter-chunk slices (8 elements of every row; n = 4096), else 0.
int leo_quarter_slice(int n);
// 16 when a GF(2^16) encode with `groups` groups of m rows runs as one LDS
// tile of quarter-chunk slices (m up to 8192; with several groups their XOR
// is held in registers, 16 slots per thread), else 0.
int leo_quarter_encode_slice(int m, int groups);
hipError_t leo_encode_fused(const LeoDev &f, const LeoEncodeArgs &a, hipStream_t s);
hipError_t leo_decode_fused(const LeoDev &f, const LeoDecodeArgs &a, hipStream_t s);

// ---- GF(2^16) reconstruct in three LDS-tiled passes (leo_multipass.hip) ----
// For n >= 8192 rows (one tile holds 4096): work arrays hold `slices` 64-byte column slices of all
// n rows, slice-major.
struct LeoMultipassArgs {
    const uint8_t *const *rows;  // n row pointers (null = zero / erased)
    const uint16_t *errlocs;     // n logs multiplied into the rows as they load
    const int *out_index;        // n: output number of each work row, or -1
    const uint16_t *out_logs;    // nout logs of the reveal multiply
    uint8_t *const *out;         // nout destinations
    uint8_t *work, *work2;       // n * 64 bytes per slice each
    size_t col0;                 // first shard byte of this column chunk
    int slices;                  // 64-byte slices in this chunk
    int n, ifft_trunc, fft_trunc;
    int C, P, A;                 // tile shapes from leo_multipass_shape
    const uint32_t *need;        // leo_fft_need bits, or null
    // Bit b: rows [b*P, (b+1)*P) hold an output row (leo_pass3_blocks).  Pass 3
    // runs only those row blocks and pass 2 stores only their rows.
    uint64_t pass3_blocks;
};

// Bit b set when one of `out_rows` lies in rows [b*P, (b+1)*P) (n / P <= 64).
uint64_t leo_pass3_blocks(int P, const std::vector<int> &out_rows);
// Row blocks of pass 3 to launch: one past the highest set bit.
int leo_pass3_grid(uint64_t blocks);

// Bytes per row per workgroup of the three-pass kernels: one 64-byte chunk.
constexpr int kLeoPassSlice = 64;
// Tile shapes for n rows; false when n is out of the multipass range.
bool leo_multipass_shape(int n, int *C, int *P, int *A);
// Column chunk (bytes) whose two work arrays stay within 512 MiB each.
size_t leo_multipass_chunk(int n, size_t size);
hipError_t leo_decode_multipass(const LeoDev &f, const LeoMultipassArgs &a, hipStream_t s);

// ---- GF(2^16) encode in three LDS-tiled passes (m >= 2048 rows without a tile) ----
struct LeoEncPassArgs {
    const uint8_t *const *rows;  // pass 1: the m row pointers of group g (null past k)
    uint8_t *const *parity;      // pass 3: the p parity shards
    uint8_t *work, *work2;       // m * 64 bytes per slice each: group IFFT / accumulator
    size_t col0;                 // first shard byte of this column chunk
    int slices, m, p;
    int mtrunc;                  // rows of group g (IFFT truncation)
    int skew_off;                // group g's IFFT skew: fftSkew + m - 1 + g * m
    int first, last;             // group g is the first / last one
    int C, P, A;                 // leo_multipass_shape(m)
    uint32_t *flag;              // pass 3: compare with the parity shards instead of storing;
    uint32_t flag_value;         // a mismatch stores flag_value to *flag
};
// pass 1 / 2 once per group in order, then pass 3.
hipError_t leo_encode_multipass(const LeoDev &f, const LeoEncPassArgs &a, int pass, hipStream_t s);
// Batched: one launch per pass and group index over the tasks of that launch
// (sorted by block0; leo_enc_pass_blocks workgroups each); max_p: the largest
// P of the pass-3 tasks.
struct LeoEncPassTask {
    LeoEncPassArgs a;
    uint32_t block0;
};
uint32_t leo_enc_pass_blocks(const LeoEncPassArgs &a, int pass);
hipError_t leo_encode_multipass_batch(const LeoDev &f, const LeoEncPassTask *tasks, int ntasks, uint32_t blocks,
                                      int pass, int max_p, hipStream_t s);

// Batched three-pass reconstructs: every task's whole column (col0 = 0) in
// one launch per pass; block0[p] is the task's first workgroup in pass p
// (tasks sorted by it), its workgroup count leo_pass_blocks.
struct LeoPassTask {
    LeoMultipassArgs a;
    uint32_t block0[3];
};
void leo_pass_blocks(const LeoMultipassArgs &a, uint32_t blocks[3]);
// blocks: total workgroups per pass; max_p: the largest P of the tasks.
hipError_t leo_decode_multipass_batch(const LeoDev &f, const LeoPassTask *tasks, int ntasks, const uint32_t blocks[3],
                                      int max_p, hipStream_t s);

}  // namespace rsx

// leo_fused.hip — Leopard encode / reconstruct with every work row of one
// column slice resident in LDS (gfx950: up to 160 KiB per workgroup).
//
// One workgroup owns a slice of `slice` bytes of every shard.  It reads the
// slice's input rows from HBM once, runs all butterfly layers of the
// transform in LDS (radix-4 steps, one barrier each), and writes the output
// rows once, so HBM sees exactly the algorithmic bytes.
#include <algorithm>
#include <type_traits>

#include "leo_tile.hpp"

namespace rsx {

namespace {

using namespace leo;
using namespace leo_tile;

// Tile of the slice that starts `sc` bytes into every shard: a 32-byte
// GF(2^16) slice is one unit of the 64-byte chunk at sc & ~63, a 16-byte one
// (F16Q) one quarter of it.
template <class F>
__device__ __forceinline__ Tile slice_tile(uint8_t *lds, int rows, int slice, size_t size, size_t sc,
                                           size_t &col) {
    if (std::is_same<F, F16Q>::value) {
        col = sc & ~size_t{63};
        return Tile{lds, rows + 1, 0, 1, static_cast<int>((sc >> 4) & 3), 1};
    }
    const int cols = static_cast<int>(size - sc < static_cast<size_t>(slice) ? size - sc : slice);
    const bool unit = slice < 64;
    col = unit ? (sc & ~size_t{63}) : sc;
    return Tile{lds, rows + 1, log2i(slice / F::kUnitBytes), cols / F::kUnitBytes, unit ? static_cast<int>((sc >> 5) & 1) : 0,
                unit ? 1 : 2};
}

// The encode's end: FFT of the accumulated IFFTs, then the parity rows
// stored (or compared, for Verify).
template <class F>
__device__ __forceinline__ void encode_finish(const LeoDev &f, const LeoEncodeArgs &a, const Tile &w, size_t col) {
    fft<F, true>(w, a.m, a.p, f);
    if (a.flag) compare_rows<F>(w, a.parity, a.p, col, a.flag, a.flag_value);
    else store_rows<F>(w, a.parity, nullptr, nullptr, a.p, col, f);
    __syncthreads();  // the next slice reuses the rows
}

// One column slice (bytes [sc, sc + slice) of every shard) of an encode.
// With several m-row groups, the XOR of their IFFTs accumulates in
// registers (MAXD slots per thread, row-fastest like the derivative) while
// the one LDS tile takes the next group; the last group's IFFT is folded
// into the tile in place before the FFT.
template <class F, int MAXD>
__device__ __forceinline__ void encode_slice(const LeoDev &f, const LeoEncodeArgs &a, uint8_t *lds, size_t sc) {
    size_t col;
    const Tile w = slice_tile<F>(lds, a.m, a.slice, a.size, sc, col);
    const int ms = log2i(a.m);
    const int items = a.m * (w.ucount * F::kUnitBytes / 16);
    uint4 acc[MAXD];
    for (int g = 0; g < a.groups; ++g) {
        const int mtrunc = g == 0 ? (a.k < a.m ? a.k : a.m) : (a.k - g * a.m < a.m ? a.k - g * a.m : a.m);
        load_rows<F>(w, a.data + static_cast<size_t>(g) * a.m, nullptr, a.m, col, f);
        ifft<F, true>(w, a.m, mtrunc, f.skew + a.m - 1 + g * a.m, 0, f);
        if (a.groups == 1) break;
        const bool first = g == 0, last = g == a.groups - 1;
#pragma unroll
        for (int j = 0; j < MAXD; ++j) {
            const int it = threadIdx.x + j * blockDim.x;
            if (it < items) {
                uint4 *p = slot(w, it & (a.m - 1), it >> ms);
                if (first) acc[j] = *p;
                else if (last) *p = xor4(acc[j], *p);
                else acc[j] = xor4(acc[j], *p);
            }
        }
        __syncthreads();  // before the next group's rows (or the FFT) reuse the tile
    }
    encode_finish<F>(f, a, w, col);
}

// The batched GF(2^8) encode: the groups' XOR accumulates in a second LDS
// region of the tile's shape (the batch kernel's LDS is sized for a
// reconstruct tile anyway), which frees the registers for group g + 1's rows
// to load while group g's IFFT runs: no load phase of its own per group.
// The planner sizes LDS for both regions (leo_batch_encode_lds).
template <class F>
__device__ __forceinline__ void encode_slice_pipelined(const LeoDev &f, const LeoEncodeArgs &a, uint8_t *lds,
                                                       size_t sc) {
    constexpr int kPre = 4;  // prefetched slots per thread: m * slice / 16 <= 4 * 256 in every batched plan
    size_t col;
    const Tile w = slice_tile<F>(lds, a.m, a.slice, a.size, sc, col);
    if (a.groups == 1 || !RowPrefetch<F, kPre>::fits(w, a.m)) {
        encode_slice<F, 8>(f, a, lds, sc);
        return;
    }
    Tile acc = w;
    acc.base = lds + static_cast<size_t>(a.slice) * (a.m + 1);
    const int ms = log2i(a.m);
    const int items = a.m * (w.ucount * F::kUnitBytes / 16);
    RowPrefetch<F, kPre> next;
    next.issue(w, a.data, a.m, col);
    for (int g = 0; g < a.groups; ++g) {
        const int mtrunc = g == 0 ? (a.k < a.m ? a.k : a.m) : (a.k - g * a.m < a.m ? a.k - g * a.m : a.m);
        next.put(w, a.m);
        __syncthreads();
        if (g + 1 < a.groups) next.issue(w, a.data + static_cast<size_t>(g + 1) * a.m, a.m, col);
        ifft<F, true>(w, a.m, mtrunc, f.skew + a.m - 1 + g * a.m, 0, f);
        const bool first = g == 0, last = g == a.groups - 1;
        for (int it = threadIdx.x; it < items; it += blockDim.x) {
            uint4 *p = slot(w, it & (a.m - 1), it >> ms), *q = slot(acc, it & (a.m - 1), it >> ms);
            if (first) *q = *p;
            else if (last) *p = xor4(*q, *p);
            else *q = xor4(*q, *p);
        }
        __syncthreads();  // before the next group's rows (or the FFT) reuse the tile
    }
    encode_finish<F>(f, a, w, col);
}

// One column slice of a reconstruct; MAXD bounds the slots per thread the
// formal derivative holds in registers (n * slice / 16 / threads).
template <class F, int MAXD>
__device__ __forceinline__ void decode_slice(const LeoDev &f, const LeoDecodeArgs &a, uint8_t *lds, size_t sc) {
    size_t col;
    const Tile w = slice_tile<F>(lds, a.n, a.slice, a.size, sc, col);
    const int cols = w.ucount * F::kUnitBytes;
    load_rows<F>(w, a.rows, a.errlocs, a.n, col, f);
    ifft<F>(w, a.n, a.ifft_trunc, f.skew, -1, f);
    derivative<MAXD>(w, a.n, cols / 16);
    fft<F>(w, a.n, a.fft_trunc, f, a.need);
    store_rows<F>(w, a.out, a.out_rows, a.out_logs, a.nout, col, f);
    __syncthreads();
}

// GF(2^8): the workgroup's copy of the 256 multiply records and the 255
// fftSkew entries in LDS, read by the butterflies instead of global memory;
// GF(2^16) keeps its 5 MB of records and its skews in global memory.
template <class F>
__device__ __forceinline__ LeoDev lds_tables(const LeoDev &f) {
    if constexpr (std::is_same<F, F8>::value) {
        constexpr int kSkew = 255, kSkewPad = 288;  // the pad reads as zero
        constexpr int kRecords = 256;
        __shared__ uint32_t t8[kRecords * F8::kTW];
        __shared__ uint16_t s8[kSkewPad];
        for (int i = threadIdx.x; i < kRecords * F8::kTW; i += blockDim.x) t8[i] = f.tables[i];
        for (int i = threadIdx.x; i < kSkewPad; i += blockDim.x) s8[i] = i < kSkew ? f.skew[i] : 0;
        __syncthreads();
        LeoDev l = f;
        l.tables = t8;
        l.skew = s8;
        return l;
    } else {
        return f;
    }
}

// Slice index of workgroup-loop index g: the two unit slices of a chunk go
// to one XCD.
__device__ __forceinline__ size_t slice_start(uint32_t g, uint32_t nslices, int slice) {
    const uint32_t s = slice == 16 ? xcd_quarter_slice(g, nslices) : slice < 64 ? xcd_unit_slice(g, nslices) : g;
    return static_cast<size_t>(s) * slice;
}

template <class F, int MAXD>
__global__ __launch_bounds__(kFBMax) void k_encode_fused(const LeoDev f0, const LeoEncodeArgs a) {
    extern __shared__ uint4 lds_raw[];
    const LeoDev f = lds_tables<F>(f0);
    const uint32_t nslices = static_cast<uint32_t>((a.size + a.slice - 1) / a.slice);
    for (uint32_t g = blockIdx.x; g < nslices; g += gridDim.x)
        encode_slice<F, MAXD>(f, a, reinterpret_cast<uint8_t *>(lds_raw), slice_start(g, nslices, a.slice));
}

template <class F, int MAXD>
__global__ __launch_bounds__(kFBMax) void k_decode_fused(const LeoDev f0, const LeoDecodeArgs a) {
    extern __shared__ uint4 lds_raw[];
    const LeoDev f = lds_tables<F>(f0);
    const uint32_t nslices = static_cast<uint32_t>((a.size + a.slice - 1) / a.slice);
    for (uint32_t g = blockIdx.x; g < nslices; g += gridDim.x)
        decode_slice<F, MAXD>(f, a, reinterpret_cast<uint8_t *>(lds_raw), slice_start(g, nslices, a.slice));
}

// Batched: blocks stride over the global slice index of all tasks, find
// their task by binary search over the first-slice offsets, and run that
// task's slice.  The task is uniform per block.
template <class F, int MAXD, bool PIPE = false>
__device__ __forceinline__ void leo_batch_slices(const LeoDev &f0, const LeoBatchTask *tasks, int ntasks,
                                                 uint32_t nslices) {
    extern __shared__ uint4 lds_raw[];
    const LeoDev f = lds_tables<F>(f0);
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    int lo = 0;  // task of the previous slice: g only grows, so tasks only move forward
    for (uint32_t g = blockIdx.x; g < nslices; g += gridDim.x) {
        if (lo + 1 < ntasks && tasks[lo + 1].slice0 <= g) {
            int hi = ntasks - 1;
            ++lo;  // tasks[lo].slice0 <= g: search [lo, hi] for the last such task
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (tasks[mid].slice0 <= g) lo = mid;
                else hi = mid - 1;
            }
        }
        const uint32_t end = lo + 1 < ntasks ? tasks[lo + 1].slice0 : nslices;
        const uint32_t l = g - tasks[lo].slice0, count = end - tasks[lo].slice0;
        if (tasks[lo].op == 0) {
            const LeoEncodeArgs a = tasks[lo].enc;
            if constexpr (PIPE) encode_slice_pipelined<F>(f, a, lds, slice_start(l, count, a.slice));
            else encode_slice<F, MAXD>(f, a, lds, slice_start(l, count, a.slice));
        } else {
            const LeoDecodeArgs a = tasks[lo].dec;
            decode_slice<F, MAXD>(f, a, lds, slice_start(l, count, a.slice));
        }
    }
}

template <class F, int MAXD>
__global__ __launch_bounds__(kFB) void k_leo_batch(const LeoDev f0, const LeoBatchTask *tasks, int ntasks,
                                                   uint32_t nslices) {
    leo_batch_slices<F, MAXD>(f0, tasks, ntasks, nslices);
}

// GF(2^8) with the smaller register budget: four 256-thread workgroups per
// CU are what LDS allows for a reconstruct tile, and four waves per SIMD
// keep the kernel within 128 VGPRs (121, no spills; unbounded the
// compiler takes 130, three waves per SIMD).
template <class F, int MAXD>
__global__ __launch_bounds__(kFB) __attribute__((amdgpu_waves_per_eu(4))) void k_leo_batch_w4(
    const LeoDev f0, const LeoBatchTask *tasks, int ntasks, uint32_t nslices) {
    leo_batch_slices<F, MAXD, true>(f0, tasks, ntasks, nslices);
}

unsigned grid_for(size_t size, int slice) {
    const size_t slices = (size + slice - 1) / slice;
    return static_cast<unsigned>(std::min<size_t>(slices, 65535));
}

}  // namespace

template <class K>
static hipError_t allow_lds(K kernel, size_t lds) {
    if (lds <= (64u << 10)) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(lds));
}

hipError_t leo_encode_fused(const LeoDev &f, const LeoEncodeArgs &a, hipStream_t s) {
    // Several groups accumulate m * slice bytes in registers (kDerivMax
    // 16-byte slots per thread).
    const size_t acc = a.groups > 1 ? static_cast<size_t>(a.m) * a.slice : 0;
    const size_t lds = static_cast<size_t>(a.slice) * (a.m + 1);
    const unsigned threads = leo_tile_threads(f.bits, a.m, a.slice, lds);
    if (acc > static_cast<size_t>(kDerivMax) * 16 * threads) return hipErrorInvalidValue;
    if (a.slice == 16 && (f.bits != 16 || lds > kLeoFusedEncodeLds)) return hipErrorInvalidValue;
    const bool small = acc <= kLeoDerivSmall;
    hipError_t e;
#define RSX_LAUNCH(FT, MD)                                                                              \
    do {                                                                                                 \
        if ((e = allow_lds(k_encode_fused<FT, MD>, lds)) != hipSuccess) return e;                        \
        hipLaunchKernelGGL((k_encode_fused<FT, MD>), dim3(grid_for(a.size, a.slice)), dim3(threads), lds, s, f, a); \
    } while (0)
    if (a.slice == 16) {
        RSX_LAUNCH(F16Q, kDerivMax);
    } else if (f.bits == 16) {
        if (small) RSX_LAUNCH(F16, kDerivMax / 2);
        else RSX_LAUNCH(F16, kDerivMax);
    } else {
        if (small) RSX_LAUNCH(F8, kDerivMax / 2);
        else RSX_LAUNCH(F8, kDerivMax);
    }
#undef RSX_LAUNCH
    return hipGetLastError();
}

hipError_t leo_batch_launch(const LeoDev &f, const LeoBatchTask *tasks, int ntasks, uint32_t nslices, size_t lds,
                            unsigned grid, size_t deriv_bytes, bool quarter, hipStream_t s) {
    if (ntasks <= 0 || nslices == 0) return hipSuccess;
    // The smaller register budget when every task's derivative (reconstruct)
    // or group accumulator (encode) fits it.
    const bool small = deriv_bytes <= kLeoDerivSmall;
    hipError_t e;
#define RSX_LAUNCH(FT, MD)                                                                          \
    do {                                                                                             \
        if ((e = allow_lds(k_leo_batch<FT, MD>, lds)) != hipSuccess) return e;                       \
        hipLaunchKernelGGL((k_leo_batch<FT, MD>), dim3(grid), dim3(kFB), lds, s, f, tasks, ntasks, nslices); \
    } while (0)
    if (quarter) {
        if (f.bits != 16) return hipErrorInvalidValue;
        RSX_LAUNCH(F16Q, kDerivMax);
    } else if (f.bits == 16) {
        if (small) RSX_LAUNCH(F16, kDerivMax / 2);
        else RSX_LAUNCH(F16, kDerivMax);
    } else {
        if (small) {
            if ((e = allow_lds(k_leo_batch_w4<F8, kDerivMax / 2>, lds)) != hipSuccess) return e;
            hipLaunchKernelGGL((k_leo_batch_w4<F8, kDerivMax / 2>), dim3(grid), dim3(kFB), lds, s, f, tasks, ntasks,
                               nslices);
        } else {
            RSX_LAUNCH(F8, kDerivMax);
        }
    }
#undef RSX_LAUNCH
    return hipGetLastError();
}

hipError_t leo_decode_fused(const LeoDev &f, const LeoDecodeArgs &a, hipStream_t s) {
    // The derivative holds n * slice / 16 slots in registers, kDerivMax per thread.
    const size_t deriv = static_cast<size_t>(a.n) * a.slice;
    const size_t lds = static_cast<size_t>(a.slice) * (a.n + 1);
    const unsigned threads = leo_tile_threads(f.bits, a.n, a.slice, lds);
    if (deriv > static_cast<size_t>(kDerivMax) * 16 * threads) return hipErrorInvalidValue;
    const bool small = deriv <= kLeoDerivSmall;
    hipError_t e;
#define RSX_LAUNCH(FT, MD)                                                                              \
    do {                                                                                                 \
        if ((e = allow_lds(k_decode_fused<FT, MD>, lds)) != hipSuccess) return e;                        \
        hipLaunchKernelGGL((k_decode_fused<FT, MD>), dim3(grid_for(a.size, a.slice)), dim3(threads), lds, s, f, a); \
    } while (0)
    if (f.bits == 16 && a.slice == 16) {
        RSX_LAUNCH(F16Q, kDerivMax);
    } else if (f.bits == 16) {
        if (small) RSX_LAUNCH(F16, kDerivMax / 2);
        else RSX_LAUNCH(F16, kDerivMax);
    } else {
        if (small) RSX_LAUNCH(F8, kDerivMax / 2);
        else RSX_LAUNCH(F8, kDerivMax);
    }
#undef RSX_LAUNCH
    return hipGetLastError();
}

}  // namespace rsx

// batch_kernels.hip — one persistent launch over a device work list of
// heterogeneous GF(2^8) matrix passes (stripes of any (k, m), shard size and
// erasure pattern), the batched form of the per-stripe codeSomeShards calls
// (reedsolomon.go:882).
//
// Work list: tasks of at most M outputs; each task is cut into column tiles
// of `tile` bytes, numbered by a prefix sum.  Blocks stride over the global
// tile index, find their task (one load from a tile -> task table, or a
// binary search over the prefix sums for very large plans), and stage the
// task's coefficient tables into LDS only when the task changes.  Every tile is a
// complete sub-problem (the product is column-separable), so tiles need no
// ordering.
#include "batch.hpp"
#include "gf8_device.hpp"

#include <algorithm>

namespace rsx {

namespace {

constexpr int kBlock = 256;

using namespace gf8dev;

// Upper bound search: last task whose first tile is <= t.
__device__ __forceinline__ int find_task(const BatchTask *tasks, int ntasks, uint32_t t) {
    int lo = 0, hi = ntasks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tasks[mid].tile0 <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Two workgroups per CU at least: caps the 7- and 8-output classes at 256
// VGPRs (280 unbounded, one wave per SIMD); the few spills cost less than
// the lost occupancy (mixed plans 3.61 -> 4.10 TB/s).
template <int M>
__global__ __launch_bounds__(kBlock, 2) void batch_pass(const BatchLaunch L) {
    constexpr int R = Shape<M>::R;
    constexpr uint64_t per_chunk = static_cast<uint64_t>(R) * 64;  // vectors per wave-chunk
    extern __shared__ uint4 lds[];
    uint4 *lo = lds;
    uint32_t *hi = reinterpret_cast<uint32_t *>(lds + L.maxk * M);
    const Tables tab{lo, hi};
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // The staged tables depend only on the task's table array and first
    // output (which fix k, mtot and nout), so tasks that share them (every
    // encode of one codec) stage once per block.
    const uint32_t *cached_tables = nullptr;
    int cached_o0 = -1;
    for (uint32_t t = blockIdx.x; t < L.ntiles; t += gridDim.x) {
        const int ti = L.tile_task ? static_cast<int>(L.tile_task[t]) : find_task(L.tasks, L.ntasks, t);
        const BatchTask task = L.tasks[ti];
        const int k = task.k;
        if (task.tables != cached_tables || task.o0 != cached_o0) {
            __syncthreads();
            stage_tables<M>(lo, hi, task.tables, k, task.mtot, task.o0, task.nout);
            __syncthreads();
            cached_tables = task.tables;
            cached_o0 = task.o0;
        }
        const uint8_t *const *in = L.ptrs + task.in_off;
        uint8_t *const *out = (uint8_t *const *)(L.ptrs + task.out_off);
        const int mode = task.mode;
        uint32_t *flag = mode == kGf8Compare ? L.flags + task.item : nullptr;
        const uint64_t b0 = static_cast<uint64_t>(t - task.tile0) * L.tile;
        const uint64_t b1 = b0 + L.tile < task.len ? b0 + L.tile : task.len;
        uint64_t vend = b0 >> 4;  // first vector not done by the vector paths
        if (task.aligned) {
            const uint64_t v0 = b0 >> 4, v1 = b1 >> 4;
            const uint64_t nchunk = (v1 - v0) / per_chunk;
            for (uint64_t c = wave; c < nchunk; c += kBlock / 64)
                chunk<M, R, Shape<M>::G>(in, out, k, task.nout, mode, tab, v0 + c * per_chunk, lane, flag, 1u);
            for (uint64_t v = v0 + nchunk * per_chunk + threadIdx.x; v < v1; v += kBlock)
                vector1<M>(in, out, k, task.nout, mode, tab, v, flag, 1u);
            vend = v1;
        }
        // Bytes past the last whole vector (every byte of an unaligned task).
        for (uint64_t b = (vend << 4) + threadIdx.x; b < b1; b += kBlock) {
            for (int i = 0; i < task.nout; ++i) {
                uint32_t a = mode == kGf8Accumulate ? out[i][b] : 0u;
                for (int j = 0; j < k; ++j) a ^= mul_byte(in[j][b], lo[j * M + i], hi[j * M + i]);
                if (mode != kGf8Compare) out[i][b] = static_cast<uint8_t>(a);
                else if (out[i][b] != static_cast<uint8_t>(a)) *flag = 1u;
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void batch_compare(const CompareEntry *entries, int n,
                                                        uint32_t *flags) {
    // blockIdx.y strides over entries, blockIdx.x over 64 KiB pieces of one.
    constexpr uint64_t kPiece = 64u << 10;
    for (int e = blockIdx.y; e < n; e += gridDim.y) {
        const CompareEntry c = entries[e];
        const bool aligned = ((reinterpret_cast<uintptr_t>(c.a) | reinterpret_cast<uintptr_t>(c.b)) & 15) == 0;
        uint32_t diff = 0;
        for (uint64_t p0 = blockIdx.x * kPiece; p0 < c.len; p0 += gridDim.x * kPiece) {
            const uint64_t p1 = p0 + kPiece < c.len ? p0 + kPiece : c.len;
            uint64_t b = p0;
            if (aligned) {
                const uint64_t v0 = p0 >> 4, v1 = p1 >> 4;
                for (uint64_t v = v0 + threadIdx.x; v < v1; v += kBlock) {
                    const uint4 x = reinterpret_cast<const uint4 *>(c.a)[v];
                    const uint4 y = reinterpret_cast<const uint4 *>(c.b)[v];
                    diff |= (x.x ^ y.x) | (x.y ^ y.y) | (x.z ^ y.z) | (x.w ^ y.w);
                }
                b = v1 << 4;
            }
            for (uint64_t i = b + threadIdx.x; i < p1; i += kBlock) diff |= c.a[i] ^ c.b[i];
        }
        if (diff) flags[c.item] = 1;
    }
}

}  // namespace

hipError_t launch_batch_compare(const CompareEntry *entries, int n, uint32_t *flags, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_compare, dim3(16, static_cast<unsigned>(std::min(n, 65535))), dim3(kBlock), 0, s,
                       entries, n, flags);
    return hipGetLastError();
}

hipError_t launch_batch_pass(const BatchLaunch &L, int m_class, unsigned grid, hipStream_t s) {
    if (L.ntiles == 0) return hipSuccess;
    const size_t lds = static_cast<size_t>(L.maxk) * m_class * 20;
    switch (m_class) {
#define RSX_CASE(M) \
    case M: hipLaunchKernelGGL(batch_pass<M>, dim3(grid), dim3(kBlock), lds, s, L); break;
        RSX_CASE(1) RSX_CASE(2) RSX_CASE(3) RSX_CASE(4)
        RSX_CASE(5) RSX_CASE(6) RSX_CASE(7) RSX_CASE(8)
#undef RSX_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace rsx

// batch_kernels.hip — one persistent launch over a device work list of
// heterogeneous GF(2^8) matrix passes (stripes of any (k, m), shard size and
// erasure pattern), the batched form of the per-stripe codeSomeShards calls
// (reedsolomon.go:882).
//
// Work list: tasks of at most M outputs; each task is cut into column tiles
// of `tile` bytes, numbered by a prefix sum.  Blocks stride over the global
// tile index, find their task by binary search, and stage the task's
// coefficient tables into LDS only when the task changes.  Every tile is a
// complete sub-problem (the product is column-separable), so tiles need no
// ordering.
#include "batch.hpp"

namespace rsx {

namespace {

constexpr int kBlock = 256;

struct Split3 {
    uint32_t g0, g1, g2;
};
__device__ __forceinline__ Split3 split(uint32_t x) {
    return {x & 0x07070707u, (x >> 3) & 0x07070707u, (x >> 6) & 0x03030303u};
}
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
__device__ __forceinline__ uint32_t sel(uint32_t hi, uint32_t lo, uint32_t idx) {
    return __builtin_amdgcn_perm(hi, lo, idx);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int M>
__device__ __forceinline__ void mac_pair(uint32_t (&acc)[M][4], const uint4 &xa, const uint4 &xb,
                                         const uint4 *lo, const uint32_t *hi, int ja, int jb) {
    const Split3 a[4] = {split(xa.x), split(xa.y), split(xa.z), split(xa.w)};
    const Split3 b[4] = {split(xb.x), split(xb.y), split(xb.z), split(xb.w)};
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const uint4 ta = lo[ja * M + i], tb = lo[jb * M + i];
        const uint32_t ha = hi[ja * M + i], hb = hi[jb * M + i];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            uint32_t r = acc[i][d];
            r = xor3(r, sel(ta.y, ta.x, a[d].g0), sel(ta.w, ta.z, a[d].g1));
            r = xor3(r, sel(ha, ha, a[d].g2), sel(tb.y, tb.x, b[d].g0));
            acc[i][d] = xor3(r, sel(tb.w, tb.z, b[d].g1), sel(hb, hb, b[d].g2));
        }
    }
}

template <int M>
__device__ __forceinline__ void mac_one(uint32_t (&acc)[M][4], const uint4 &xa, const uint4 *lo,
                                        const uint32_t *hi, int ja) {
    const Split3 a[4] = {split(xa.x), split(xa.y), split(xa.z), split(xa.w)};
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const uint4 ta = lo[ja * M + i];
        const uint32_t ha = hi[ja * M + i];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const uint32_t r = xor3(acc[i][d], sel(ta.y, ta.x, a[d].g0), sel(ta.w, ta.z, a[d].g1));
            acc[i][d] = r ^ sel(ha, ha, a[d].g2);
        }
    }
}

__device__ __forceinline__ uint32_t mul_byte(uint32_t x, const uint4 &t, uint32_t h) {
    const Split3 s = split(x);
    return sel(t.y, t.x, s.g0) ^ sel(t.w, t.z, s.g1) ^ sel(h, h, s.g2);
}

constexpr int kGroup = 8;

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

template <int M>
__global__ __launch_bounds__(kBlock) void batch_pass(const BatchLaunch L) {
    extern __shared__ uint4 lds[];
    uint4 *lo = lds;
    uint32_t *hi = reinterpret_cast<uint32_t *>(lds + L.maxk * M);
    int cached = -1;
    for (uint32_t t = blockIdx.x; t < L.ntiles; t += gridDim.x) {
        const int ti = find_task(L.tasks, L.ntasks, t);
        const BatchTask task = L.tasks[ti];
        const int k = task.k;
        if (ti != cached) {
            __syncthreads();
            for (int q = threadIdx.x; q < k * M; q += kBlock) {
                const int j = q / M, i = q % M;
                const uint32_t *src = task.tables + (static_cast<size_t>(j) * task.mtot + task.o0 + i) * 5;
                lo[q] = i < task.nout ? make_uint4(src[0], src[1], src[2], src[3]) : make_uint4(0, 0, 0, 0);
                hi[q] = i < task.nout ? src[4] : 0u;
            }
            __syncthreads();
            cached = ti;
        }
        const uint8_t *const *in = L.ptrs + task.in_off;
        uint8_t *const *out = (uint8_t *const *)(L.ptrs + task.out_off);
        const uint64_t b0 = static_cast<uint64_t>(t - task.tile0) * L.tile;
        const uint64_t b1 = b0 + L.tile < task.len ? b0 + L.tile : task.len;
        const uint64_t vend = task.aligned ? (b1 >> 4) : (b0 >> 4);
        for (uint64_t v = (b0 >> 4) + threadIdx.x; v < vend; v += kBlock) {
            uint32_t acc[M][4];
#pragma unroll
            for (int i = 0; i < M; ++i) {
                if (task.accumulate && i < task.nout) {
                    const uint4 o = reinterpret_cast<const uint4 *>(out[i])[v];
                    acc[i][0] = o.x;
                    acc[i][1] = o.y;
                    acc[i][2] = o.z;
                    acc[i][3] = o.w;
                } else {
                    acc[i][0] = acc[i][1] = acc[i][2] = acc[i][3] = 0;
                }
            }
            uint4 cur[kGroup], nxt[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; ++u)
                if (u < k) cur[u] = reinterpret_cast<const uint4 *>(in[u])[v];
            for (int j0 = 0; j0 < k; j0 += kGroup) {
                const int jn = j0 + kGroup;
#pragma unroll
                for (int u = 0; u < kGroup; ++u)
                    if (jn + u < k) nxt[u] = reinterpret_cast<const uint4 *>(in[jn + u])[v];
#pragma unroll
                for (int u = 0; u < kGroup; u += 2) {
                    if (j0 + u + 1 < k) mac_pair<M>(acc, cur[u], cur[u + 1], lo, hi, j0 + u, j0 + u + 1);
                    else if (j0 + u < k) mac_one<M>(acc, cur[u], lo, hi, j0 + u);
                }
#pragma unroll
                for (int u = 0; u < kGroup; ++u) cur[u] = nxt[u];
            }
#pragma unroll
            for (int i = 0; i < M; ++i)
                if (i < task.nout)
                    __builtin_nontemporal_store(u32x4{acc[i][0], acc[i][1], acc[i][2], acc[i][3]},
                                                reinterpret_cast<u32x4 *>(out[i]) + v);
        }
        // Bytes past the last full vector (or every byte of an unaligned task).
        for (uint64_t b = (vend << 4) + threadIdx.x; b < b1; b += kBlock) {
            for (int i = 0; i < task.nout; ++i) {
                uint32_t acc = task.accumulate ? out[i][b] : 0u;
                for (int j = 0; j < k; ++j) acc ^= mul_byte(in[j][b], lo[j * M + i], hi[j * M + i]);
                out[i][b] = static_cast<uint8_t>(acc);
            }
        }
    }
}

}  // namespace

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

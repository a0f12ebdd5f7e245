// gf8_kernels.hip — GF(2^8) matrix x shards on CDNA4 (gfx950).
//
// out[i][b] (^)= XOR_j c[i][j] * in[j][b] for every byte column b: the device
// form of the reference's codeSomeShards (reedsolomon.go:882) and of the
// generated mulAvxTwo_KxM / mulGFNI_KxM kernels it dispatches to.
//
// Layout: each lane owns 16 consecutive byte columns (one 128-bit load per
// input), keeps the pass's M outputs in 4*M VGPRs and streams all k inputs
// through them, so every input byte is read from HBM once per pass and every
// output byte is written once.  A product c*x is three v_perm_b32 byte
// selects into 8-entry tables of c*(x & 7), c*(x & 0x38), c*(x & 0xC0)
// (GF(2)-linearity), XORed into the accumulator.
#include "kernels.hpp"

namespace rsx {

namespace {

constexpr int kBlock = 256;

struct Split3 {
    uint32_t g0, g1, g2;
};

__device__ __forceinline__ Split3 split(uint32_t x) {
    return {x & 0x07070707u, (x >> 3) & 0x07070707u, (x >> 6) & 0x03030303u};
}

// Product of 4 packed bytes by one coefficient, given its 5-dword table.
__device__ __forceinline__ uint32_t mul4(const Split3 &s, uint32_t t0, uint32_t t1, uint32_t t2,
                                         uint32_t t3, uint32_t t4) {
    return __builtin_amdgcn_perm(t1, t0, s.g0) ^ __builtin_amdgcn_perm(t3, t2, s.g1) ^
           __builtin_amdgcn_perm(t4, t4, s.g2);
}

// XOR of three words in one VALU op (gfx950 v_bitop3_b32, truth table 0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint4 load16(const uint8_t *base, size_t v) {
    return reinterpret_cast<const uint4 *>(base)[v];
}

// Coefficient tables staged in LDS for the whole block: per (input j,
// output i) one 16-byte record {c*(0..3), c*(4..7), c*(0..3)<<3, c*(4..7)<<3}
// read with one ds_read_b128 straight into VGPRs (v_perm_b32 can read only
// one SGPR, so SGPR-resident tables would cost two v_mov per coefficient),
// plus the 4-entry table of c*(x & 0xC0) in a second array.
struct LdsTables {
    const uint4 *lo;
    const uint32_t *hi;
};

// acc[i] ^= c[i][a]*xa ^ c[i][b]*xb for 4 dwords: 6 byte selects folded into
// the accumulator with 3 three-input XORs.
template <int M>
__device__ __forceinline__ void mac_pair(uint32_t (&acc)[M][4], const uint4 &xa, const uint4 &xb,
                                         const LdsTables &t, int ja, int jb) {
    const Split3 a[4] = {split(xa.x), split(xa.y), split(xa.z), split(xa.w)};
    const Split3 b[4] = {split(xb.x), split(xb.y), split(xb.z), split(xb.w)};
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const uint4 ta = t.lo[ja * M + i], tb = t.lo[jb * M + i];
        const uint32_t ha = t.hi[ja * M + i], hb = t.hi[jb * M + i];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            uint32_t r = acc[i][d];
            r = xor3(r, __builtin_amdgcn_perm(ta.y, ta.x, a[d].g0), __builtin_amdgcn_perm(ta.w, ta.z, a[d].g1));
            r = xor3(r, __builtin_amdgcn_perm(ha, ha, a[d].g2), __builtin_amdgcn_perm(tb.y, tb.x, b[d].g0));
            acc[i][d] = xor3(r, __builtin_amdgcn_perm(tb.w, tb.z, b[d].g1), __builtin_amdgcn_perm(hb, hb, b[d].g2));
        }
    }
}

template <int M>
__device__ __forceinline__ void mac_one(uint32_t (&acc)[M][4], const uint4 &xa, const LdsTables &t, int ja) {
    const Split3 a[4] = {split(xa.x), split(xa.y), split(xa.z), split(xa.w)};
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const uint4 ta = t.lo[ja * M + i];
        const uint32_t ha = t.hi[ja * M + i];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const uint32_t r = xor3(acc[i][d], __builtin_amdgcn_perm(ta.y, ta.x, a[d].g0),
                                    __builtin_amdgcn_perm(ta.w, ta.z, a[d].g1));
            acc[i][d] = r ^ __builtin_amdgcn_perm(ha, ha, a[d].g2);
        }
    }
}

// Inputs are fetched in groups of kGroup 16-byte loads per lane, and the next
// group is issued before the current one is consumed, so each wave keeps up
// to 2*kGroup HBM requests in flight instead of one per input.
constexpr int kGroup = 8;

template <int M>
__global__ __launch_bounds__(kBlock) void gf8_pass_vec16(const Gf8Pass p) {
    extern __shared__ uint4 lds[];
    const int k = p.k;
    uint4 *lo = lds;
    uint32_t *hi = reinterpret_cast<uint32_t *>(lds + k * M);
    for (int t = threadIdx.x; t < k * M; t += kBlock) {
        const uint32_t *src = p.tables + (static_cast<size_t>(t / M) * p.mtot + p.o0 + t % M) * 5;
        lo[t] = make_uint4(src[0], src[1], src[2], src[3]);
        hi[t] = src[4];
    }
    __syncthreads();
    const LdsTables tab{lo, hi};

    const size_t nvec = p.len >> 4;
    const size_t stride = static_cast<size_t>(gridDim.x) * kBlock;
    for (size_t v = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; v < nvec; v += stride) {
        uint32_t acc[M][4];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            if (p.accumulate) {
                const uint4 o = load16(p.out[i], v);
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
            if (u < k) cur[u] = load16(p.in[u], v);
        for (int j0 = 0; j0 < k; j0 += kGroup) {
            const int jn = j0 + kGroup;
#pragma unroll
            for (int u = 0; u < kGroup; ++u)
                if (jn + u < k) nxt[u] = load16(p.in[jn + u], v);
#pragma unroll
            for (int u = 0; u < kGroup; u += 2) {
                if (j0 + u + 1 < k) mac_pair<M>(acc, cur[u], cur[u + 1], tab, j0 + u, j0 + u + 1);
                else if (j0 + u < k) mac_one<M>(acc, cur[u], tab, j0 + u);
            }
#pragma unroll
            for (int u = 0; u < kGroup; ++u) cur[u] = nxt[u];
        }
#pragma unroll
        for (int i = 0; i < M; ++i)
            __builtin_nontemporal_store(u32x4{acc[i][0], acc[i][1], acc[i][2], acc[i][3]},
                                        reinterpret_cast<u32x4 *>(p.out[i]) + v);
    }
}

// Byte-granular form for the tail of a shard (len % 16) and for shards whose
// addresses are not 16-byte aligned.
__global__ __launch_bounds__(kBlock) void gf8_pass_bytes(const Gf8Pass p, size_t start) {
    const size_t stride = static_cast<size_t>(gridDim.x) * kBlock;
    for (size_t b = start + static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; b < p.len;
         b += stride) {
        for (int i = 0; i < p.nout; ++i) {
            uint32_t acc = p.accumulate ? p.out[i][b] : 0u;
            const uint32_t *tab = p.tables + static_cast<size_t>(p.o0 + i) * 5;
            for (int j = 0; j < p.k; ++j, tab += static_cast<size_t>(p.mtot) * 5) {
                const Split3 s = split(p.in[j][b]);
                acc ^= mul4(s, tab[0], tab[1], tab[2], tab[3], tab[4]);
            }
            p.out[i][b] = static_cast<uint8_t>(acc);
        }
    }
}

__global__ __launch_bounds__(kBlock) void compare_bytes(const uint8_t *a, const uint8_t *b,
                                                        size_t n, int *flag) {
    const size_t stride = static_cast<size_t>(gridDim.x) * kBlock;
    int diff = 0;
    for (size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
        diff |= a[i] != b[i];
    if (diff) *flag = 1;
}

__global__ __launch_bounds__(kBlock) void fill_u32(uint32_t *p, size_t n, uint32_t v) {
    const size_t stride = static_cast<size_t>(gridDim.x) * kBlock;
    for (size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) p[i] = v;
}

unsigned grid_for(size_t work_items) {
    size_t g = (work_items + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > 65535u * 16) g = 65535u * 16;
    return static_cast<unsigned>(g);
}

bool aligned16(const Gf8Pass &p) {
    uintptr_t bits = 0;
    for (int j = 0; j < p.k; ++j) bits |= reinterpret_cast<uintptr_t>(p.in[j]);
    for (int i = 0; i < p.nout; ++i) bits |= reinterpret_cast<uintptr_t>(p.out[i]);
    return (bits & 15) == 0;
}

}  // namespace

hipError_t launch_gf8_pass(const Gf8Pass &p, hipStream_t stream) {
    if (p.len == 0 || p.nout == 0) return hipSuccess;
    size_t tail_start = 0;
    if (aligned16(p)) {
        const size_t nvec = p.len >> 4;
        tail_start = nvec << 4;
        if (nvec) {
            const dim3 grid(grid_for(nvec)), block(kBlock);
            const size_t lds_bytes = static_cast<size_t>(p.k) * p.nout * 20;
            switch (p.nout) {
#define RSX_CASE(M) \
    case M: hipLaunchKernelGGL(gf8_pass_vec16<M>, grid, block, lds_bytes, stream, p); break;
                RSX_CASE(1) RSX_CASE(2) RSX_CASE(3) RSX_CASE(4)
                RSX_CASE(5) RSX_CASE(6) RSX_CASE(7) RSX_CASE(8)
#undef RSX_CASE
            default:
                return hipErrorInvalidValue;
            }
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    if (tail_start < p.len) {
        hipLaunchKernelGGL(gf8_pass_bytes, dim3(grid_for(p.len - tail_start)), dim3(kBlock), 0,
                           stream, p, tail_start);
        return hipGetLastError();
    }
    return hipSuccess;
}

hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    size_t g = (n + kBlock - 1) / kBlock;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(fill_u32, dim3(static_cast<unsigned>(g)), dim3(kBlock), 0, stream, p, n, v);
    return hipGetLastError();
}

hipError_t launch_compare(const uint8_t *a, const uint8_t *b, size_t n, int *flag,
                          hipStream_t stream) {
    if (n == 0) return hipSuccess;
    size_t g = (n + kBlock * 16 - 1) / (kBlock * 16);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(compare_bytes, dim3(static_cast<unsigned>(g)), dim3(kBlock), 0, stream, a, b,
                       n, flag);
    return hipGetLastError();
}

}  // namespace rsx

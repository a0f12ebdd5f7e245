// gf8_device.hpp — device building blocks of the GF(2^8) matrix passes,
// shared by the per-stripe kernels (gf8_kernels.hip) and the batched
// persistent kernel (batch_kernels.hip).
//
// Product: a byte x times coefficient c is split into bit groups [0:3),
// [3:6), [6:8); by GF(2)-linearity c*x = T0[x&7] ^ T1[(x>>3)&7] ^ T2[x>>6],
// three v_perm_b32 byte selects on 4 packed bytes, folded into the
// accumulator with v_bitop3_b32 (three-input XOR).
//
// Streaming: a wave owns R consecutive 1 KiB column chunks of a stripe (lane
// l reads the 16 bytes at offset 16*l of each chunk) and walks the k inputs
// in groups of G, so each input is read as R KiB contiguous with
// non-temporal loads (no reuse), and the M outputs stay in VGPRs until one
// non-temporal store per output.  In compare mode (Verify) the outputs are
// read instead and checked against the accumulators; a lane that finds a
// difference stores a caller-chosen value to a flag word.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace rsx {
namespace gf8dev {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

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
__device__ __forceinline__ uint4 load_nt(const uint8_t *base, uint64_t v) {
    const u32x4 t = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(base) + v);
    return make_uint4(t.x, t.y, t.z, t.w);
}
__device__ __forceinline__ void store_nt(uint8_t *base, uint64_t v, const uint32_t (&a)[4]) {
    __builtin_nontemporal_store(u32x4{a[0], a[1], a[2], a[3]}, reinterpret_cast<u32x4 *>(base) + v);
}

// Coefficient tables in LDS: lo[j*M+i] = {c*(0..3), c*(4..7), c*(0..3)<<3,
// c*(4..7)<<3} as one 16-byte record, hi[j*M+i] = c*((0..3)<<6).
struct Tables {
    const uint4 *lo;
    const uint32_t *hi;
};

// acc ^= c[ja]*xa ^ c[jb]*xb for M outputs on 4 dwords.
template <int M>
__device__ __forceinline__ void mac_pair(uint32_t (&acc)[M][4], const uint4 &xa, const uint4 &xb,
                                         const Tables &t, int ja, int jb) {
    const Split3 a[4] = {split(xa.x), split(xa.y), split(xa.z), split(xa.w)};
    const Split3 b[4] = {split(xb.x), split(xb.y), split(xb.z), split(xb.w)};
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const uint4 ta = t.lo[ja * M + i], tb = t.lo[jb * M + i];
        const uint32_t ha = t.hi[ja * M + i], hb = t.hi[jb * M + i];
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
__device__ __forceinline__ void mac_one(uint32_t (&acc)[M][4], const uint4 &xa, const Tables &t, int ja) {
    const Split3 a[4] = {split(xa.x), split(xa.y), split(xa.z), split(xa.w)};
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const uint4 ta = t.lo[ja * M + i];
        const uint32_t ha = t.hi[ja * M + i];
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

// Stages the tables of outputs [o0, o0+nout) of a [k][mtot][5] table array
// into LDS (zero for rows >= nout), for the whole block.
template <int M>
__device__ __forceinline__ void stage_tables(uint4 *lo, uint32_t *hi, const uint32_t *tables, int k,
                                             int mtot, int o0, int nout) {
    for (int q = threadIdx.x; q < k * M; q += blockDim.x) {
        const int j = q / M, i = q % M;
        if (i < nout) {
            const uint32_t *src = tables + (static_cast<size_t>(j) * mtot + o0 + i) * 5;
            lo[q] = make_uint4(src[0], src[1], src[2], src[3]);
            hi[q] = src[4];
        } else {
            lo[q] = make_uint4(0, 0, 0, 0);
            hi[q] = 0;
        }
    }
}

// One wave-chunk: R KiB of every shard starting at 16-byte vector `base`
// (lane-relative vector base + r*64 + lane).  Outputs i < nout are written
// (compared in kGf8Compare mode).
template <int M, int R, int G>
__device__ __forceinline__ void chunk(const uint8_t *const *in, uint8_t *const *out, int k, int nout, int mode,
                                      const Tables &t, uint64_t base, int lane, uint32_t *flag,
                                      uint32_t flag_value) {
    const bool accumulate = mode == kGf8Accumulate;
    uint32_t acc[R][M][4];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int i = 0; i < M; ++i) {
            if (accumulate && i < nout) {
                const uint4 o = reinterpret_cast<const uint4 *>(out[i])[base + r * 64 + lane];
                acc[r][i][0] = o.x;
                acc[r][i][1] = o.y;
                acc[r][i][2] = o.z;
                acc[r][i][3] = o.w;
            } else {
                acc[r][i][0] = acc[r][i][1] = acc[r][i][2] = acc[r][i][3] = 0;
            }
        }
    int j = 0;
    for (; j + G <= k; j += G) {  // whole groups: G*R loads in flight, no guards
        uint4 x[G][R];
#pragma unroll
        for (int u = 0; u < G; ++u)
#pragma unroll
            for (int r = 0; r < R; ++r) x[u][r] = load_nt(in[j + u], base + r * 64 + lane);
#pragma unroll
        for (int u = 0; u < G; u += 2)
#pragma unroll
            for (int r = 0; r < R; ++r) mac_pair<M>(acc[r], x[u][r], x[u + 1][r], t, j + u, j + u + 1);
    }
    for (; j + 2 <= k; j += 2) {
        uint4 x[2][R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            x[0][r] = load_nt(in[j], base + r * 64 + lane);
            x[1][r] = load_nt(in[j + 1], base + r * 64 + lane);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) mac_pair<M>(acc[r], x[0][r], x[1][r], t, j, j + 1);
    }
    if (j < k) {
        uint4 x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = load_nt(in[j], base + r * 64 + lane);
#pragma unroll
        for (int r = 0; r < R; ++r) mac_one<M>(acc[r], x[r], t, j);
    }
    if (mode == kGf8Compare) {
        uint32_t diff = 0;
#pragma unroll
        for (int i = 0; i < M; ++i)
            if (i < nout)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint4 o = load_nt(out[i], base + r * 64 + lane);
                    diff |= (o.x ^ acc[r][i][0]) | (o.y ^ acc[r][i][1]) | (o.z ^ acc[r][i][2]) | (o.w ^ acc[r][i][3]);
                }
        if (diff) *flag = flag_value;  // every writer stores the same value
        return;
    }
#pragma unroll
    for (int i = 0; i < M; ++i)
        if (i < nout)
#pragma unroll
            for (int r = 0; r < R; ++r) store_nt(out[i], base + r * 64 + lane, acc[r][i]);
}

// One 16-byte vector per lane: the R = 1 chunk with lane offset folded into
// v, inputs fetched eight at a time.
template <int M>
__device__ __forceinline__ void vector1(const uint8_t *const *in, uint8_t *const *out, int k, int nout, int mode,
                                        const Tables &t, uint64_t v, uint32_t *flag, uint32_t flag_value) {
    chunk<M, 1, 8>(in, out, k, nout, mode, t, v, 0, flag, flag_value);
}

// Streaming shape per output count: more outputs -> fewer chunks per wave,
// so accumulators + one input group stay within the VGPR budget.
template <int M>
struct Shape {
    static constexpr int R = M <= 4 ? 4 : M <= 8 ? 2 : 1;  // R = 1 beat 2 for no M <= 8 (mixed plans 3.0 vs 3.6 TB/s)
    static constexpr int G = 4;
};

}  // namespace gf8dev
}  // namespace rsx

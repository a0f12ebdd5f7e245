// butterfly_rates.hip — instruction-rate probes for the GF(2^16) butterfly
// of the Leopard transforms (bench.py --only butterfly): how many
// element-butterflies per second each arithmetic formulation can issue when
// nothing else limits it (operands in registers, no memory traffic).  Not
// part of the erasure-coding library; built into libmembench.so.
//
//   kind 0  VALU, the library's formulation: F16::mul_add (v_perm_b32 byte
//           selects over 3/3/2-bit groups, v_bitop3 XOR3) plus the XOR of
//           fft2, on 16 elements per lane (the Leopard kernels' own code).
//   kind 1  MFMA over bit-sliced data, int8: v_mfma_i32_32x32x32_i8.  One
//           butterfly on 32 element pairs is [x';y'] = [[I, M_c]; [I, I+M_c]]
//           [x;y] over GF(2) (32 x 32 bits, K = 32), each bit a 0/1 byte, the
//           sum's low bit the result: 32 element-butterflies per instruction.
//   kind 2  MFMA over bit-sliced data, FP4 (e2m1 0/1 values, exact sums):
//           v_mfma_scale_f32_32x32x64_f8f6f4.  K = 64 holds a radix-4 group
//           (4 elements), two instructions give its 64 output bits: 64
//           element-butterflies per instruction.
//   kind 3  kind 1 as a chain of layers: each instruction's 16 int32 sums per
//           lane turned back into the 0/1 bytes of the next layer's operand
//           (2 v_perm_b32 + 1 v_bitop3 per 4 sums, the least VALU that takes),
//           with A's rows ordered so the sums land in the lanes that need
//           them.  Still optimistic: no A matrix per skew, no data exchange
//           between butterfly distances, no bit packing for LDS.
//
// Every kind keeps 4 independent chains per wave and folds its registers
// into `out` at the end, so nothing is dead code.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../csrc/leo_device.hpp"

namespace {

constexpr int kBlock = 256;

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

using rsx::leo::F16;

__global__ __launch_bounds__(kBlock) void k_valu(uint32_t *out, unsigned iters, uint32_t seed) {
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    uint32_t t[F16::kTW];
#pragma unroll
    for (int i = 0; i < F16::kTW; ++i) t[i] = (seed * (i + 1)) ^ (id * 0x9E3779B9u);
    F16::E x[4], y[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        x[c] = {make_uint4(id, id + c, seed, c), make_uint4(seed ^ c, id * 3, 7, c)};
        y[c] = {make_uint4(id ^ 5, c, seed + c, id), make_uint4(c, seed, id * 5, 9)};
    }
    for (unsigned it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            F16::mul_add(x[c], y[c], t);  // fftDIT2: x ^= c*y, y ^= x
            F16::xor_into(y[c], x[c]);
        }
    }
    uint32_t r = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c)
        r ^= x[c].lo.x ^ x[c].lo.y ^ x[c].lo.z ^ x[c].lo.w ^ x[c].hi.x ^ x[c].hi.y ^ x[c].hi.z ^ x[c].hi.w ^
             y[c].lo.x ^ y[c].hi.w;
    out[id] = r;
}

__global__ __launch_bounds__(kBlock) void k_mfma_i8(uint32_t *out, unsigned iters, uint32_t seed) {
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    const int s = static_cast<int>(seed ^ id);
    const v4i a = {s & 0x01010101, (s >> 1) & 0x01010101, (s >> 2) & 0x01010101, (s >> 3) & 0x01010101};
    v4i b[4];
    v16i c[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        b[q] = v4i{(s >> q) & 0x01010101, (s >> (q + 4)) & 0x01010101, 0x01000100, 0x00010001};
        c[q] = v16i{};
    }
    for (unsigned it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 4; ++q) c[q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b[q], c[q], 0, 0, 0);
    }
    int r = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int i = 0; i < 16; ++i) r ^= c[q][i];
    out[id] = static_cast<uint32_t>(r);
}

__global__ __launch_bounds__(kBlock) void k_mfma_fp4(uint32_t *out, unsigned iters, uint32_t seed) {
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    const int s = static_cast<int>(seed ^ id);
    // FP4 e2m1: 0x2 is 1.0, 0x0 is 0; eight 4-bit values per dword.
    const int one = 0x22222222;
    const v8i a = {s & one, (s >> 1) & one, (s >> 2) & one, (s >> 3) & one, one, 0, s & one, one};
    v8i b[4];
    v16f c[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        b[q] = v8i{(s >> q) & one, one, 0, (s >> (q + 1)) & one, one, (s >> 5) & one, 0, one};
        c[q] = v16f{};
    }
    for (unsigned it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            // cbsz / blgp 4: both operands FP4; scales 127 = 2^0 (E8M0).
            c[q] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b[q], c[q], 4, 4, 0, 0x7f7f7f7f, 0, 0x7f7f7f7f);
    }
    float r = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int i = 0; i < 16; ++i) r += c[q][i];
    out[id] = __float_as_uint(r);
}

// 16 int32 sums -> 4 dwords of 0/1 bytes (the low bit of each sum).
__device__ __forceinline__ void repack(const uint32_t (&c)[16], uint32_t (&b)[4]) {
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const uint32_t lo = __builtin_amdgcn_perm(c[4 * d + 1], c[4 * d], 0x0c0c0400u);
        const uint32_t hi = __builtin_amdgcn_perm(c[4 * d + 3], c[4 * d + 2], 0x04000c0cu);
        b[d] = __builtin_amdgcn_bitop3_b32(lo, hi, 0x01010101u, 0xA8);  // (lo | hi) & mask
    }
}

__global__ __launch_bounds__(kBlock) void k_mfma_i8_chain(uint32_t *out, unsigned iters, uint32_t seed) {
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    const int s = static_cast<int>(seed ^ id);
    const v4i a = {s & 0x01010101, (s >> 1) & 0x01010101, (s >> 2) & 0x01010101, (s >> 3) & 0x01010101};
    uint32_t b[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i) b[q][i] = static_cast<uint32_t>(s >> (q + i)) & 0x01010101u;
    for (unsigned it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const v4i bv = {static_cast<int>(b[q][0]), static_cast<int>(b[q][1]), static_cast<int>(b[q][2]),
                            static_cast<int>(b[q][3])};
            const v16i c = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, bv, v16i{}, 0, 0, 0);
            uint32_t cv[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) cv[i] = static_cast<uint32_t>(c[i]);
            repack(cv, b[q]);
        }
    }
    uint32_t r = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) r ^= b[q][0] ^ b[q][1] ^ b[q][2] ^ b[q][3];
    out[id] = r;
}

}  // namespace

// Runs kind 0-3 (above) on `grid` workgroups of 256 threads, `iters` loop
// iterations each; out holds grid * 256 dwords.
extern "C" int butterfly_rates(int kind, unsigned grid, unsigned iters, void *out, void *stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t *o = static_cast<uint32_t *>(out);
    switch (kind) {
    case 0: hipLaunchKernelGGL(k_valu, dim3(grid), dim3(kBlock), 0, s, o, iters, 0x2545F491u); break;
    case 1: hipLaunchKernelGGL(k_mfma_i8, dim3(grid), dim3(kBlock), 0, s, o, iters, 0x2545F491u); break;
    case 2: hipLaunchKernelGGL(k_mfma_fp4, dim3(grid), dim3(kBlock), 0, s, o, iters, 0x2545F491u); break;
    case 3: hipLaunchKernelGGL(k_mfma_i8_chain, dim3(grid), dim3(kBlock), 0, s, o, iters, 0x2545F491u); break;
    default: return static_cast<int>(hipErrorInvalidValue);
    }
    return static_cast<int>(hipGetLastError());
}

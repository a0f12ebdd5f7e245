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

#include "gf8_device.hpp"

#include <algorithm>

namespace rsx {

namespace {

constexpr int kBlock = 256;

// Streaming pass: each wave owns R consecutive KiB-wide column chunks and
// walks all k inputs (gf8_device.hpp); vectors past the last whole chunk (all
// of them when R == 0) are done one per lane.
template <int M, int R>
__global__ __launch_bounds__(kBlock) void gf8_pass_stream(const Gf8Pass p) {
    extern __shared__ uint4 lds[];
    uint4 *lo = lds;
    uint32_t *hi = reinterpret_cast<uint32_t *>(lds + p.k * M);
    gf8dev::stage_tables<M>(lo, hi, p.tables, p.k, p.mtot, p.o0, p.nout);
    __syncthreads();
    const gf8dev::Tables t{lo, hi};
    const uint64_t nvec = p.len >> 4;
    const uint64_t per_chunk = static_cast<uint64_t>(R > 0 ? R : 1) * 64;
    const uint64_t nchunk = R > 0 ? nvec / per_chunk : 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint64_t c = static_cast<uint64_t>(blockIdx.x) * (kBlock / 64) + wave; c < nchunk;
         c += static_cast<uint64_t>(gridDim.x) * (kBlock / 64))
        gf8dev::chunk<M, (R > 0 ? R : 1), gf8dev::Shape<M>::G>(p.in, p.out, p.k, p.nout, p.mode, t, c * per_chunk,
                                                               lane, p.flag, p.flag_value);
    for (uint64_t v = nchunk * per_chunk + static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; v < nvec;
         v += static_cast<uint64_t>(gridDim.x) * kBlock)
        gf8dev::vector1<M>(p.in, p.out, p.k, p.nout, p.mode, t, v, p.flag, p.flag_value);
}

// Byte-granular form for the tail of a shard (len % 16) and for shards whose
// addresses are not 16-byte aligned.
__global__ __launch_bounds__(kBlock) void gf8_pass_bytes(const Gf8Pass p, size_t start) {
    const size_t stride = static_cast<size_t>(gridDim.x) * kBlock;
    for (size_t b = start + static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; b < p.len;
         b += stride) {
        for (int i = 0; i < p.nout; ++i) {
            uint32_t acc = p.mode == kGf8Accumulate ? p.out[i][b] : 0u;
            const uint32_t *tab = p.tables + static_cast<size_t>(p.o0 + i) * 5;
            for (int j = 0; j < p.k; ++j, tab += static_cast<size_t>(p.mtot) * 5)
                acc ^= gf8dev::mul_byte(p.in[j][b], make_uint4(tab[0], tab[1], tab[2], tab[3]), tab[4]);
            if (p.mode != kGf8Compare) p.out[i][b] = static_cast<uint8_t>(acc);
            else if (p.out[i][b] != static_cast<uint8_t>(acc)) *p.flag = p.flag_value;
        }
    }
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
            // Chunks per wave by shard length: one 16-byte vector per lane
            // below 4 MiB (latency-bound: as many waves as possible), two
            // 1 KiB chunks per wave up to 16 MiB, four beyond (with at most
            // four outputs, so the accumulators fit).  9-16 outputs run the
            // 16-output kernel, one vector per lane (64 accumulator VGPRs).
            const int M = p.nout <= 8 ? p.nout : 16;
            const int R = M == 16 || p.len < (4u << 20) ? 1 : (p.nout <= 4 && p.len >= (16u << 20)) ? 4 : 2;
            const size_t per_block = static_cast<size_t>(R) * 64 * (kBlock / 64);
            const size_t nchunk_blocks = R == 1 ? 0 : nvec / per_block;
            const size_t rest = nvec - nchunk_blocks * per_block;
            size_t g = std::max<size_t>(nchunk_blocks, (rest + kBlock - 1) / kBlock);
            g = std::max<size_t>(1, std::min<size_t>(g, 65535u * 16));
            const dim3 grid(static_cast<unsigned>(g)), block(kBlock);
            const size_t lds_bytes = static_cast<size_t>(p.k) * M * 20;
#define RSX_CASE(M)                                                                                  \
    case M:                                                                                          \
        if (R == 4) hipLaunchKernelGGL((gf8_pass_stream<M, 4>), grid, block, lds_bytes, stream, p);  \
        else if (R == 2) hipLaunchKernelGGL((gf8_pass_stream<M, 2>), grid, block, lds_bytes, stream, p); \
        else hipLaunchKernelGGL((gf8_pass_stream<M, 0>), grid, block, lds_bytes, stream, p);         \
        break;
            switch (M) {
                RSX_CASE(1) RSX_CASE(2) RSX_CASE(3) RSX_CASE(4)
                RSX_CASE(5) RSX_CASE(6) RSX_CASE(7) RSX_CASE(8)
            case 16:
                hipLaunchKernelGGL((gf8_pass_stream<16, 0>), grid, block, lds_bytes, stream, p);
                break;
            default:
                return hipErrorInvalidValue;
            }
#undef RSX_CASE
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

}  // namespace rsx

// membench.hip — HBM streaming microbenchmark (bench.py --only ceiling):
// the bandwidth the device sustains for a given mix of streamed reads and
// writes, with the access pattern of the GF(2^8) pass kernels and none of
// their arithmetic.  Not part of the erasure-coding library (librsx.so);
// built beside it into libmembench.so.
//
// hbm_stream<NIN, NOUT>: NIN input regions and NOUT output regions of `bytes`
// each.  Each wave owns chunks of R = 4 KiB (lane l moves the 16 bytes at
// offset 16*l of every KiB), reads the chunk of every input with
// non-temporal loads, XORs them, and writes the result (XOR the output's
// index) to every output with non-temporal stores: the loads, stores and
// chunk walk of the pass kernels, with an XOR in place of the products.  With
// no output, a lane whose XOR equals an improbable value stores a flag, so
// the reads cannot be dropped.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {

constexpr int kBlock = 256;
constexpr int kR = 4;  // KiB chunks per wave

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct StreamArgs {
    const uint8_t *in[16];
    uint8_t *out[8];
    size_t bytes;     // per region, a multiple of 4 KiB
    uint32_t *flag;   // read-only runs
};

template <int NIN, int NOUT>
__global__ __launch_bounds__(kBlock) void hbm_stream(const StreamArgs a) {
    // Dynamic LDS (unused beyond this word) only limits workgroups per CU.
    extern __shared__ uint32_t lds_reserve[];
    if (threadIdx.x == 0) lds_reserve[0] = 0;
    const int lane = threadIdx.x & 63;
    const uint64_t nchunk = a.bytes / (kR * 1024);
    const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kBlock / 64);
    for (uint64_t c = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); c < nchunk; c += waves) {
        const uint64_t v0 = c * kR * 64 + lane;  // 16-byte vector index
        u32x4 acc[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) acc[r] = u32x4{0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < NIN; ++j)
#pragma unroll
            for (int r = 0; r < kR; ++r)
                acc[r] ^= __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(a.in[j]) + v0 + r * 64);
        if constexpr (NOUT == 0) {
            uint32_t hit = 0;
#pragma unroll
            for (int r = 0; r < kR; ++r) hit |= (acc[r].x == 0x9E3779B9u) & (acc[r].y == 0x7F4A7C15u);
            if (hit) *a.flag = 1;
        } else {
#pragma unroll
            for (int i = 0; i < NOUT; ++i)
#pragma unroll
                for (int r = 0; r < kR; ++r)
                    __builtin_nontemporal_store(acc[r] ^ static_cast<uint32_t>(i),
                                                reinterpret_cast<u32x4 *>(a.out[i]) + v0 + r * 64);
        }
    }
}

}  // namespace

extern "C" {

// Launches one pass of the NIN-in / NOUT-out stream on `stream` with `grid`
// workgroups of 256 threads, each reserving `lds` bytes of LDS (0 or up to
// 64 KiB: fewer workgroups per CU).  Supported mixes: (1,1) copy, (4,0)
// read, (0,1) write, (4,1), (10,4), (16,4).  Returns a hipError_t.
int membench_stream(int nin, int nout, const void *const *in, void *const *out, size_t bytes, uint32_t *flag,
                    unsigned grid, unsigned lds, void *stream) {
    StreamArgs a{};
    if (nin < 0 || nin > 16 || nout < 0 || nout > 8 || bytes % (kR * 1024) || lds > (64u << 10))
        return hipErrorInvalidValue;
    lds = lds < 4 ? 4 : lds;
    for (int j = 0; j < nin; ++j) a.in[j] = static_cast<const uint8_t *>(in[j]);
    for (int i = 0; i < nout; ++i) a.out[i] = static_cast<uint8_t *>(out[i]);
    a.bytes = bytes;
    a.flag = flag;
    hipStream_t s = static_cast<hipStream_t>(stream);
#define RSX_MIX(I, O) \
    if (nin == I && nout == O) { hipLaunchKernelGGL((hbm_stream<I, O>), dim3(grid), dim3(kBlock), lds, s, a); return hipGetLastError(); }
    RSX_MIX(1, 1) RSX_MIX(4, 0) RSX_MIX(0, 1) RSX_MIX(4, 1) RSX_MIX(10, 4) RSX_MIX(16, 4)
#undef RSX_MIX
    return hipErrorInvalidValue;
}

}  // extern "C"

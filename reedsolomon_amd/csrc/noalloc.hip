// noalloc.hip — rsx_reconstruct_noalloc: the classic engine's Reconstruct /
// ReconstructData (reedsolomon.go:1670-1781) on device shards with nothing
// allocated during the call, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid it).
//
// rsx_reconstruct keeps each erasure pattern's decode tables in device
// memory it allocates on first use.  Here the decode rows are inverted on
// the host in a workspace the codec sized at New, handed to the device as
// kernel arguments, expanded into coefficient tables by a small kernel into
// `scratch` (a file-scope device array, one per GPU, kept until the process
// ends), and the GF(2^8) pass reads them from there.  Nothing is copied
// between host and device and no call waits for the device.
//
// Every call on one GPU shares that GPU's scratch array: calls whose work
// overlaps on one GPU (different codecs on different streams) race on it.
// Calls on one codec are ordered by the codec's lock and stream.  Each call
// writes every table entry its pass reads, so nothing needs clearing.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "codec.hpp"
#include "gf256.hpp"
#include "kernels.hpp"

using namespace rsx;

namespace {

constexpr int kBlock = 256;

// Coefficient tables for up to 3276 coefficients (k x erased shards), e.g.
// RS(32,8) with 8 erasures uses 256, (200,16) with 16 uses 3200.
constexpr size_t kScratchBytes = 64u << 10;
constexpr int kScratchCoefs = static_cast<int>(kScratchBytes / (kTableDwords * 4));
__device__ __align__(256) uint32_t scratch[kScratchBytes / 4];

// Device address of `scratch` per GPU, looked up once per device with that
// device current (each GPU holds its own copy of the array).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch[kMaxDevices] = {};

int scratch_for(int device, uint32_t **out) {
    if (device < 0 || device >= kMaxDevices) return fail(RSX_ERR_DEVICE, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                    "hipGetSymbolAddress(scratch)");
        g_scratch[device] = static_cast<uint32_t *>(p);
    }
    *out = g_scratch[device];
    return RSX_OK;
}

// Decode coefficients handed over as kernel arguments, 2048 per launch.
constexpr int kChunkCoefs = 2048;
struct ExpandChunk {
    uint32_t first;  // index of coef[0] in the [k][nout] coefficient array
    uint32_t count;
    uint8_t coef[kChunkCoefs];
};

// Carry-less product reduced by x^8 + x^4 + x^3 + x^2 + 1 (galois.go:26).
__device__ __forceinline__ uint32_t gf_mul(uint32_t a, uint32_t b) {
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if ((b >> i) & 1u) r ^= a << i;
#pragma unroll
    for (int i = 14; i >= 8; --i)
        if ((r >> i) & 1u) r ^= 0x11Du << (i - 8);
    return r;
}

// The 5-dword record of each coefficient c (byte t of the record is c*t for
// t < 8, c*((t-8) << 3) for t < 16, c*((t-16) << 6) after), the layout of
// coefficient_table in gf256.cpp, at tables + 5 * (first + q).
__global__ __launch_bounds__(kBlock) void expand_tables(const ExpandChunk a, uint32_t *tables) {
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= a.count) return;
    const uint32_t c = a.coef[q];
    uint32_t *dst = tables + static_cast<size_t>(a.first + q) * kTableDwords;
#pragma unroll
    for (int d = 0; d < kTableDwords; ++d) {
        uint32_t w = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int t = 4 * d + b;
            const uint32_t x = t < 8 ? t : t < 16 ? (t - 8) << 3 : (t - 16) << 6;
            w |= gf_mul(c, x) << (8 * b);
        }
        dst[d] = w;
    }
}

bool on_device(const void *p) {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
           attr.type == hipMemoryTypeUnified;
}

}  // namespace

extern "C" int rsx_reconstruct_noalloc(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
                                       int data_only, uint8_t *filled) {
    if (nshards > 0) std::memset(filled, 0, static_cast<size_t>(nshards));
    if (c->engine) return fail(RSX_ERR_NOT_SUPPORTED, "rsx_reconstruct_noalloc serves the classic engine only");
    const int k = c->k, total = c->total;
    // reedsolomon.go:1670-1709: argument checks and the early returns.
    if (nshards != total) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    int err = check_shards(lens, nshards, true);
    if (err) return err;
    const size_t len = shard_size(lens, nshards);
    int present = 0, data_present = 0;
    for (int i = 0; i < total; ++i)
        if (lens[i]) {
            ++present;
            if (i < k) ++data_present;
        }
    if (present == total || (data_only && data_present == k)) return RSX_OK;
    if (present < k) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");

    // The first k present shards are the inputs; the erased data shards, and
    // unless data_only the erased parity shards, the outputs (1714-1774).
    int inputs[kMaxIn], outputs[kMaxIn];
    int nin = 0, nout = 0;
    for (int r = 0; r < total && nin < k; ++r)
        if (lens[r]) inputs[nin++] = r;
    const int end = data_only ? k : total;
    for (int i = 0; i < end; ++i)
        if (!lens[i]) outputs[nout++] = i;
    if (k * nout > kScratchCoefs)
        return fail(RSX_ERR_NOT_SUPPORTED, "decode tables of " + std::to_string(k) + " x " + std::to_string(nout) +
                                               " coefficients exceed the 64 KiB scratch (3276 coefficients)");
    for (int s = 0; s < nin; ++s)
        if (!on_device(shards[inputs[s]]))
            return fail(RSX_ERR_MIXED_MEMORY, "rsx_reconstruct_noalloc needs shards in device memory");
    for (int s = 0; s < nout; ++s)
        if (!shards[outputs[s]] || !on_device(shards[outputs[s]]))
            return fail(RSX_ERR_INVALID_INPUT, "no device buffer for an absent shard");

    RSX_HIP(hipSetDevice(c->device));
    uint32_t *tables = nullptr;
    if ((err = scratch_for(c->device, &tables))) return err;

    std::lock_guard<std::mutex> lock(c->mu);
    // Decode rows (reedsolomon.go:620-673) in the codec's workspace:
    // [sub | I] inverted in place, then one k-byte row per output.
    uint8_t *aug = c->noalloc_work.data();
    uint8_t *rows = aug + static_cast<size_t>(2 * k) * k;
    const int w = 2 * k;
    for (int r = 0; r < k; ++r) {
        uint8_t *a = aug + static_cast<size_t>(r) * w;
        std::memcpy(a, c->matrix.row(inputs[r]), k);
        std::memset(a + k, 0, k);
        a[k + r] = 1;
    }
    if (!invert_in_place(aug, k)) return fail(RSX_ERR_SINGULAR, "matrix is singular");
    const Gf256 &f = gf();
    for (int o = 0; o < nout; ++o) {
        const int i = outputs[o];
        uint8_t *dst = rows + static_cast<size_t>(o) * k;
        if (i < k) {
            std::memcpy(dst, aug + static_cast<size_t>(i) * w + k, k);
        } else {
            // multiplyRowWithMatrix(parity row, inverse), reedsolomon.go:620-631.
            const uint8_t *prow = c->matrix.row(i);
            for (int j = 0; j < k; ++j) {
                uint8_t acc = 0;
                for (int t = 0; t < k; ++t) acc ^= f.mul[prow[t]][aug[static_cast<size_t>(t) * w + k + j]];
                dst[j] = acc;
            }
        }
    }
    // Tables [k][nout][5]: coefficient q = j * nout + o is rows[o][j].
    const int ncoef = k * nout;
    ExpandChunk ch;
    for (int first = 0; first < ncoef; first += kChunkCoefs) {
        ch.first = static_cast<uint32_t>(first);
        ch.count = static_cast<uint32_t>(std::min(kChunkCoefs, ncoef - first));
        for (uint32_t q = 0; q < ch.count; ++q) {
            const int g = first + static_cast<int>(q);
            ch.coef[q] = rows[static_cast<size_t>(g % nout) * k + g / nout];
        }
        hipLaunchKernelGGL(expand_tables, dim3((ch.count + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, ch,
                           tables);
        RSX_HIP(hipGetLastError());
    }
    // The product, as run_matmul issues it for rsx_reconstruct.
    Gf8Pass p{};
    for (int j = 0; j < k; ++j) p.in[j] = shards[inputs[j]];
    for (int o0 = 0; o0 < nout; o0 += p.nout) {
        p.nout = pass_outputs(nout - o0, k);
        for (int i = 0; i < p.nout; ++i) p.out[i] = shards[outputs[o0 + i]];
        p.tables = tables;
        p.len = len;
        p.k = k;
        p.mtot = nout;
        p.o0 = o0;
        p.mode = kGf8Clear;
        p.flag = nullptr;
        p.flag_value = 0;
        RSX_HIP(launch_gf8_pass(p, c->stream));
    }
    for (int o = 0; o < nout; ++o) filled[outputs[o]] = 1;
    return RSX_OK;
}

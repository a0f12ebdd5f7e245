// kernels_host.cpp — host versions of the device launchers for the sanitizer
// build.  The GF(2^8) passes (per-call and batched) and the Verify compare
// run their device semantics on the host, byte by byte, reading the plan's
// task lists, tile table and pointer arrays exactly as the kernels do, so
// ASan checks every range the planners hand to a launch and the results can
// be checked.  The Leopard error locators are computed with the host field's
// transforms; the other Leopard launchers only check their arguments: the
// host build covers the Leopard planners, not the transforms.
#include <hip/hip_runtime.h>

#include <cstring>

#include "../batch.hpp"
#include "../kernels.hpp"
#include "../leo_kernels.hpp"
#include "../leopard_field.hpp"

#include <vector>

namespace rsx {

namespace {

// c * x from a coefficient's 20-byte record: bytes [0, 8) = c * v,
// [8, 16) = c * (v << 3), [16, 20) = c * (v << 6).
inline uint8_t mul(const uint32_t *rec, uint8_t x) {
    const uint8_t *t = reinterpret_cast<const uint8_t *>(rec);
    return static_cast<uint8_t>(t[x & 7] ^ t[8 + ((x >> 3) & 7)] ^ t[16 + (x >> 6)]);
}

// Output i of a pass over bytes [b0, b1).
void pass_range(const uint8_t *const *in, uint8_t *const *out, const uint32_t *tables, int k, int mtot, int o0,
                int nout, int mode, uint32_t *flag, uint32_t flag_value, uint64_t b0, uint64_t b1) {
    for (int i = 0; i < nout; ++i) {
        for (uint64_t b = b0; b < b1; ++b) {
            uint8_t a = mode == kGf8Accumulate ? out[i][b] : 0;
            for (int j = 0; j < k; ++j) a ^= mul(tables + (static_cast<size_t>(j) * mtot + o0 + i) * 5, in[j][b]);
            if (mode != kGf8Compare) out[i][b] = a;
            else if (out[i][b] != a) *flag = flag_value;
        }
    }
}

}  // namespace

hipError_t launch_gf8_pass(const Gf8Pass &p, hipStream_t) {
    if (p.k <= 0 || p.k > kMaxIn || p.nout <= 0 || p.nout > kMaxOut || p.o0 + p.nout > p.mtot)
        return hipErrorInvalidValue;
    pass_range(p.in, p.out, p.tables, p.k, p.mtot, p.o0, p.nout, p.mode, p.flag, p.flag_value, 0, p.len);
    return hipSuccess;
}

hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t) {
    for (size_t i = 0; i < n; ++i) p[i] = v;
    return hipSuccess;
}

hipError_t launch_batch_compare(const CompareEntry *entries, int n, uint32_t *flags, hipStream_t) {
    for (int e = 0; e < n; ++e)
        if (std::memcmp(entries[e].a, entries[e].b, entries[e].len) != 0) flags[entries[e].item] = 1;
    return hipSuccess;
}

hipError_t launch_batch_pass(const BatchLaunch &L, int m_class, unsigned grid, hipStream_t) {
    if (m_class < 1 || m_class > 8 || grid == 0) return hipErrorInvalidValue;
    // Tiles in the order a grid-strided persistent launch takes them.
    for (unsigned blk = 0; blk < grid; ++blk) {
        for (uint32_t t = blk; t < L.ntiles; t += grid) {
            int ti;
            if (L.tile_task) {
                ti = static_cast<int>(L.tile_task[t]);
            } else {
                int lo = 0, hi = L.ntasks - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (L.tasks[mid].tile0 <= t) lo = mid;
                    else hi = mid - 1;
                }
                ti = lo;
            }
            if (ti < 0 || ti >= L.ntasks) return hipErrorInvalidValue;
            const BatchTask &task = L.tasks[ti];
            if (task.nout > m_class || task.k > L.maxk || t < task.tile0) return hipErrorInvalidValue;
            const uint64_t b0 = static_cast<uint64_t>(t - task.tile0) * L.tile;
            if (b0 >= task.len) return hipErrorInvalidValue;  // a tile past its task
            const uint64_t b1 = b0 + L.tile < task.len ? b0 + L.tile : task.len;
            const uint8_t *const *in = L.ptrs + task.in_off;
            uint8_t *const *out = const_cast<uint8_t *const *>(L.ptrs + task.out_off);
            uint32_t *flag = task.mode == kGf8Compare ? L.flags + task.item : nullptr;
            pass_range(in, out, task.tables, task.k, task.mtot, task.o0, task.nout, task.mode, flag, 1u, b0, b1);
        }
    }
    return hipSuccess;
}

// ---- Leopard: argument checks only ----
hipError_t leo_load_rows(const LeoDev &, const uint8_t *const *ptrs, const uint16_t *, size_t, int nrows,
                         const LeoWork &, hipStream_t) {
    return ptrs && nrows >= 0 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t leo_step4(const LeoDev &, bool, const LeoWork &, int dist, int, const uint16_t *skew, int, hipStream_t) {
    return skew && dist > 0 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t leo_ifft2_last(const LeoDev &, const LeoWork &, int dist, uint16_t, hipStream_t) {
    return dist > 0 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t leo_fft2_last(const LeoDev &, const LeoWork &, int, hipStream_t) { return hipSuccess; }
hipError_t leo_xor_rows(const LeoWork &, const LeoWork &, int nrows, hipStream_t) {
    return nrows >= 0 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t leo_derivative(const LeoWork &, const LeoWork &, int n, hipStream_t) {
    return n > 0 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t leo_store_rows(const LeoDev &, const LeoWork &, const int *rows, const uint16_t *, uint8_t *const *dst,
                          size_t, int count, hipStream_t) {
    return (rows && dst) || count == 0 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t leo_error_locators_device(const LeoDev &f, const uint8_t *ind, int mtrunc, int n, const int *rows,
                                     int nrows, uint16_t *errlocs, uint16_t *logs, hipStream_t) {
    // The host field's transforms on the same input (the device kernel's
    // arithmetic is LeoField::fwht's).
    if (f.bits != 16 || !f.walsh || mtrunc <= 0 || mtrunc > 65536 || n <= 0 || n > 65536 || nrows < 0)
        return hipErrorInvalidValue;
    const LeoField &lf = leo_field(16);
    std::vector<uint16_t> e(65536, 0);
    for (int i = 0; i < mtrunc; ++i) e[i] = ind[i];
    lf.fwht(e.data(), mtrunc);
    for (int i = 0; i < 65536; ++i) {
        uint32_t x = static_cast<uint32_t>(e[i]) * f.walsh[i];
        x = (x & 0xFFFFu) + (x >> 16);
        x = (x & 0xFFFFu) + (x >> 16);
        if (x >= 0xFFFFu) x -= 0xFFFFu;
        e[i] = static_cast<uint16_t>(x);
    }
    lf.fwht(e.data(), 65536);
    for (int i = 0; i < n; ++i) errlocs[i] = e[i];
    for (int q = 0; q < nrows; ++q) logs[q] = static_cast<uint16_t>(0xFFFFu - e[rows[q]]);
    return hipSuccess;
}

hipError_t leo_gf16_mul_xor(const LeoDev &, const Gf16MulXor &a, hipStream_t) {
    return a.in && a.nout >= 0 && a.nout <= 8 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t leo_batch_launch(const LeoDev &, const LeoBatchTask *tasks, int ntasks, uint32_t, size_t, unsigned grid,
                            size_t, bool, hipStream_t) {
    return (tasks || ntasks == 0) && grid > 0 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t leo_encode_fused(const LeoDev &, const LeoEncodeArgs &a, hipStream_t) {
    return a.data && a.parity && a.slice > 0 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t leo_decode_fused(const LeoDev &, const LeoDecodeArgs &a, hipStream_t) {
    return a.rows && a.errlocs && a.n > 0 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t leo_decode_multipass(const LeoDev &f, const LeoMultipassArgs &a, hipStream_t) {
    return f.bits == 16 && a.slices > 0 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t leo_encode_multipass(const LeoDev &f, const LeoEncPassArgs &, int pass, hipStream_t) {
    return f.bits == 16 && pass >= 1 && pass <= 3 ? hipSuccess : hipErrorInvalidValue;
}
hipError_t leo_encode_multipass_batch(const LeoDev &f, const LeoEncPassTask *tasks, int ntasks, uint32_t blocks,
                                      int pass, int, hipStream_t) {
    return f.bits == 16 && tasks && ntasks > 0 && blocks > 0 && pass >= 1 && pass <= 3 ? hipSuccess
                                                                                        : hipErrorInvalidValue;
}
hipError_t leo_decode_multipass_batch(const LeoDev &f, const LeoPassTask *tasks, int ntasks, const uint32_t blocks[3],
                                      int, hipStream_t) {
    return f.bits == 16 && tasks && ntasks > 0 && blocks ? hipSuccess : hipErrorInvalidValue;
}

}  // namespace rsx

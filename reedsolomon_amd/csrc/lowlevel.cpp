// lowlevel.cpp — the LowLevel helpers of include/rsx.h (mulslice.go, gf16.go):
// GF(2^8) slice multiply(-XOR) on the GF(2^8) pass kernel, GF(2^16)
// slice multiply-XOR on the Leopard field kernel, and the scalar helpers.
//
// Calls are synchronous, like the Go functions.  Device pointers run in
// place (16-byte aligned ones for GF(2^16)); host memory and unaligned
// GF(2^16) buffers are staged through an HBM buffer of the device context.
#include <algorithm>
#include <map>
#include <memory>

#include "codec.hpp"
#include "kernels.hpp"
#include "leo_kernels.hpp"
#include "leopard.hpp"
#include "leopard_field.hpp"

using namespace rsx;

namespace {

// Per-device state: a stream, the 256 GF(2^8) coefficient tables, staging.
struct LowLevelDevice {
    std::mutex mu;
    int device = 0;
    hipStream_t stream = nullptr;
    DeviceBuffer gf8_tables;  // [256][5]: the table of coefficient c at c*5
    DeviceBuffer staging;
};

std::mutex g_mu;
std::map<int, std::unique_ptr<LowLevelDevice>> g_devices;

int get_device(int device, LowLevelDevice **out) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return fail(RSX_ERR_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= ndev) return fail(RSX_ERR_DEVICE, "invalid device ordinal");
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_devices.find(device);
    if (it == g_devices.end()) {
        RSX_HIP(hipSetDevice(device));
        auto d = std::make_unique<LowLevelDevice>();
        d->device = device;
        RSX_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
        std::vector<uint32_t> host(256 * kTableDwords);
        for (int c = 0; c < 256; ++c) coefficient_table(static_cast<uint8_t>(c), &host[c * kTableDwords]);
        int err = d->gf8_tables.reserve(host.size() * 4);
        if (err) return err;
        RSX_HIP(hipMemcpyAsync(d->gf8_tables.ptr, host.data(), host.size() * 4, hipMemcpyHostToDevice, d->stream));
        RSX_HIP(hipStreamSynchronize(d->stream));
        it = g_devices.emplace(device, std::move(d)).first;
    }
    *out = it->second.get();
    return RSX_OK;
}

}  // namespace

extern "C" {

uint8_t rsx_inv(uint8_t e) { return gf().inv(e); }

int rsx_leopard_tables(int bits, uint16_t *exp, uint16_t *log, uint16_t *skew, uint16_t *walsh) {
    if (bits != 16 && bits != 8) return fail(RSX_ERR_INVALID_INPUT, "bits must be 16 or 8");
    const LeoField &f = leo_field(bits);
    if (exp) std::copy(f.exp.begin(), f.exp.end(), exp);
    if (log) std::copy(f.log.begin(), f.log.end(), log);
    if (skew) std::copy(f.skew.begin(), f.skew.end(), skew);
    if (walsh) std::copy(f.walsh.begin(), f.walsh.end(), walsh);
    return RSX_OK;
}

uint16_t rsx_gf16_mul(uint16_t a, uint16_t b) {
    // gf16.go:6-16: exp(addMod(log a, log b)) with the sum reduced once more.
    if (a == 0 || b == 0) return 0;
    const LeoField &f = leo_field(16);
    unsigned s = f.add_mod(f.log[a], f.log[b]);
    if (s >= f.modulus) s -= f.modulus;
    return f.exp[s];
}

// mulslice.go:29-39 GalMulSlice / GalMulSliceXor.
int rsx_gal_mul_slice(int device, uint8_t c, const uint8_t *in, uint8_t *out, size_t len, int do_xor) {
    if (len == 0) return RSX_OK;
    if (!in || !out) return fail(RSX_ERR_INVALID_INPUT, "nil slice");
    LowLevelDevice *d = nullptr;
    int err = get_device(device, &d);
    if (err) return err;
    std::lock_guard<std::mutex> lock(d->mu);
    (void)hipSetDevice(device);
    MemKind kind;
    if ((err = classify_all({in, out}, kind))) return err;
    const uint8_t *din = in;
    uint8_t *dout = out;
    const size_t pitch = (len + 255) / 256 * 256;
    if (kind == MemKind::Host) {
        if ((err = d->staging.reserve(2 * pitch))) return err;
        uint8_t *base = d->staging.as<uint8_t>();
        RSX_HIP(hipMemcpyAsync(base, in, len, hipMemcpyHostToDevice, d->stream));
        if (do_xor) RSX_HIP(hipMemcpyAsync(base + pitch, out, len, hipMemcpyHostToDevice, d->stream));
        din = base;
        dout = base + pitch;
    }
    Gf8Pass p{};
    p.in[0] = din;
    p.out[0] = dout;
    p.tables = d->gf8_tables.as<uint32_t>() + static_cast<size_t>(c) * kTableDwords;
    p.len = len;
    p.k = 1;
    p.mtot = 1;
    p.o0 = 0;
    p.nout = 1;
    p.mode = do_xor ? kGf8Accumulate : kGf8Clear;
    RSX_HIP(launch_gf8_pass(p, d->stream));
    if (kind == MemKind::Host) RSX_HIP(hipMemcpyAsync(out, dout, len, hipMemcpyDeviceToHost, d->stream));
    RSX_HIP(hipStreamSynchronize(d->stream));
    return RSX_OK;
}

// mulslice.go:45-95 GF16MulSliceXor8 / GF16MulSliceXor.
int rsx_gf16_mul_slice_xor(int device, const uint16_t *scalars, int nout, const uint8_t *in, uint8_t *const *outs,
                           const size_t *out_lens, size_t len) {
    if (nout < 0 || nout > 8) return fail(RSX_ERR_INVALID_INPUT, "at most 8 outputs");
    if (len == 0) return RSX_OK;
    // The reference panics on these; here they are errors.
    if (len % 64) return fail(RSX_ERR_INVALID_SHARD_SIZE, "length must be a multiple of 64");
    if (!in) return fail(RSX_ERR_INVALID_INPUT, "nil slice");
    for (int j = 0; j < nout; ++j)
        if (out_lens[j] != len || !outs[j]) return fail(RSX_ERR_INVALID_SHARD_SIZE, "output length differs");
    LowLevelDevice *d = nullptr;
    int err = get_device(device, &d);
    if (err) return err;
    const LeoDev *f = nullptr;
    std::lock_guard<std::mutex> lock(d->mu);
    (void)hipSetDevice(device);
    if ((err = leo_device_tables(device, 16, d->stream, &f))) return err;
    const LeoField &lf = leo_field(16);
    Gf16MulXor a{};
    std::vector<int> which;  // outputs with a non-zero scalar
    for (int j = 0; j < nout; ++j)
        if (scalars[j]) which.push_back(j);
    if (which.empty()) return RSX_OK;
    std::vector<const void *> ptrs{in};
    uintptr_t bits = reinterpret_cast<uintptr_t>(in);
    for (int j : which) {
        ptrs.push_back(outs[j]);
        bits |= reinterpret_cast<uintptr_t>(outs[j]);
    }
    MemKind kind;
    if ((err = classify_all(ptrs, kind))) return err;
    const bool staged = kind == MemKind::Host || (bits & 15);
    const hipMemcpyKind to = kind == MemKind::Host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    const hipMemcpyKind from = kind == MemKind::Host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    const size_t nw = which.size();
    if (staged) {
        if ((err = d->staging.reserve(len * (nw + 1)))) return err;
        uint8_t *base = d->staging.as<uint8_t>();
        RSX_HIP(hipMemcpyAsync(base, in, len, to, d->stream));
        a.in = base;
        for (size_t q = 0; q < nw; ++q) {
            RSX_HIP(hipMemcpyAsync(base + len * (q + 1), outs[which[q]], len, to, d->stream));
            a.out[q] = base + len * (q + 1);
        }
    } else {
        a.in = in;
        for (size_t q = 0; q < nw; ++q) a.out[q] = outs[which[q]];
    }
    for (size_t q = 0; q < nw; ++q) a.logs[q] = lf.log[scalars[which[q]]];
    a.nout = static_cast<int>(nw);
    a.len = len;
    RSX_HIP(leo_gf16_mul_xor(*f, a, d->stream));
    if (staged)
        for (size_t q = 0; q < nw; ++q)
            RSX_HIP(hipMemcpyAsync(outs[which[q]], a.out[q], len, from, d->stream));
    RSX_HIP(hipStreamSynchronize(d->stream));
    return RSX_OK;
}

}  // extern "C"

// leo_kernels.hip — Leopard GF(2^16) / GF(2^8) butterflies on CDNA4 (gfx950).
//
// Thread layout: blockIdx.y picks the butterfly unit (so its skews, and the
// "skew == modulus => xor only" branch, are uniform across the block) and the
// lanes of the block spread over the column chunk: one lane owns one
// 16-element column unit of every row it touches.  GF(2^16) products use the
// split lo/hi byte layout directly: 12 v_perm_b32 byte selects per 4 elements
// (3 bit groups x 2 source bytes x 2 product bytes), folded with 3-input XORs.
#include "leo_kernels.hpp"

#include <algorithm>

namespace rsx {

namespace {

constexpr int kBlock = 256;

using ConstU32 = const __attribute__((address_space(4))) uint32_t;

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
__device__ __forceinline__ uint32_t sel(uint32_t hi, uint32_t lo, uint32_t idx) {
    return __builtin_amdgcn_perm(hi, lo, idx);
}
__device__ __forceinline__ uint4 xor4(const uint4 &a, const uint4 &b) {
    return make_uint4(a.x ^ b.x, a.y ^ b.y, a.z ^ b.z, a.w ^ b.w);
}
__device__ __forceinline__ ConstU32 *as_const(const uint32_t *p) {
    return (ConstU32 *)(reinterpret_cast<uintptr_t>(p));
}

// ---- GF(2^16): 16 elements per unit, lo bytes and hi bytes in two uint4 ----
struct F16 {
    struct E {
        uint4 lo, hi;
    };
    static constexpr int kUnitBytes = 32;
    static constexpr int kTW = 20;
    static constexpr unsigned kMod = 65535;

    __device__ static size_t offset(int u) { return static_cast<size_t>(u >> 1) * 64 + (u & 1) * 16; }
    __device__ static E load(const uint8_t *row, int u) {
        const uint8_t *c = row + offset(u);
        return {*reinterpret_cast<const uint4 *>(c), *reinterpret_cast<const uint4 *>(c + 32)};
    }
    __device__ static void store(uint8_t *row, int u, const E &e) {
        uint8_t *c = row + offset(u);
        *reinterpret_cast<uint4 *>(c) = e.lo;
        *reinterpret_cast<uint4 *>(c + 32) = e.hi;
    }
    __device__ static E zero() { return {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)}; }
    __device__ static void xor_into(E &x, const E &y) {
        x.lo = xor4(x.lo, y.lo);
        x.hi = xor4(x.hi, y.hi);
    }
    // (lo, hi) of one dword pair times the table's element, XORed into (xl, xh).
    __device__ static void mac1(uint32_t &xl, uint32_t &xh, uint32_t L, uint32_t H, ConstU32 *t) {
        const uint32_t a0 = L & 0x07070707u, a1 = (L >> 3) & 0x07070707u, a2 = (L >> 6) & 0x03030303u;
        const uint32_t b0 = H & 0x07070707u, b1 = (H >> 3) & 0x07070707u, b2 = (H >> 6) & 0x03030303u;
        uint32_t r = xor3(xl, sel(t[1], t[0], a0), sel(t[5], t[4], a1));
        r = xor3(r, sel(t[8], t[8], a2), sel(t[11], t[10], b0));
        xl = xor3(r, sel(t[15], t[14], b1), sel(t[18], t[18], b2));
        r = xor3(xh, sel(t[3], t[2], a0), sel(t[7], t[6], a1));
        r = xor3(r, sel(t[9], t[9], a2), sel(t[13], t[12], b0));
        xh = xor3(r, sel(t[17], t[16], b1), sel(t[19], t[19], b2));
    }
    __device__ static void mul_add(E &x, const E &y, ConstU32 *t) {
        mac1(x.lo.x, x.hi.x, y.lo.x, y.hi.x, t);
        mac1(x.lo.y, x.hi.y, y.lo.y, y.hi.y, t);
        mac1(x.lo.z, x.hi.z, y.lo.z, y.hi.z, t);
        mac1(x.lo.w, x.hi.w, y.lo.w, y.hi.w, t);
    }
};

// ---- GF(2^8): 16 one-byte elements per unit ----
struct F8 {
    typedef uint4 E;
    static constexpr int kUnitBytes = 16;
    static constexpr int kTW = 5;
    static constexpr unsigned kMod = 255;

    __device__ static E load(const uint8_t *row, int u) {
        return reinterpret_cast<const uint4 *>(row)[u];
    }
    __device__ static void store(uint8_t *row, int u, const E &e) {
        reinterpret_cast<uint4 *>(row)[u] = e;
    }
    __device__ static E zero() { return make_uint4(0, 0, 0, 0); }
    __device__ static void xor_into(E &x, const E &y) { x = xor4(x, y); }
    __device__ static uint32_t mac1(uint32_t x, uint32_t v, ConstU32 *t) {
        const uint32_t g0 = v & 0x07070707u, g1 = (v >> 3) & 0x07070707u, g2 = (v >> 6) & 0x03030303u;
        return xor3(xor3(x, sel(t[1], t[0], g0), sel(t[3], t[2], g1)), sel(t[4], t[4], g2), 0u);
    }
    __device__ static void mul_add(E &x, const E &y, ConstU32 *t) {
        x.x = mac1(x.x, y.x, t);
        x.y = mac1(x.y, y.y, t);
        x.z = mac1(x.z, y.z, t);
        x.w = mac1(x.w, y.w, t);
    }
};

template <class F>
__device__ __forceinline__ ConstU32 *table(const LeoDev &f, unsigned log_m) {
    return as_const(f.tables) + static_cast<size_t>(log_m) * F::kTW;
}

// galois_noasm.go:94-113: fftDIT2 = mulAdd then xor, ifftDIT2 = xor then
// mulAdd; a skew equal to the modulus means "xor only" (leopard.go:741-760).
template <class F>
__device__ __forceinline__ void fft2(typename F::E &x, typename F::E &y, unsigned lm, const LeoDev &f) {
    if (lm != F::kMod) F::mul_add(x, y, table<F>(f, lm));
    F::xor_into(y, x);
}
template <class F>
__device__ __forceinline__ void ifft2(typename F::E &x, typename F::E &y, unsigned lm, const LeoDev &f) {
    F::xor_into(y, x);
    if (lm != F::kMod) F::mul_add(x, y, table<F>(f, lm));
}

template <class F, bool INV>
__global__ __launch_bounds__(kBlock) void k_step4(const LeoDev f, const LeoWork w, int dist,
                                                  int nunits, const uint16_t *skew, int bias) {
    const int u = blockIdx.x * kBlock + threadIdx.x;
    if (u >= static_cast<int>(w.cols / F::kUnitBytes)) return;
    for (int q = blockIdx.y; q < nunits; q += gridDim.y) {
    const int r = (q / dist) * 4 * dist;
    const int i = r + q % dist;
    const int iend = r + dist;
    const unsigned lm01 = skew[iend + bias], lm02 = skew[iend + dist + bias],
                   lm23 = skew[iend + 2 * dist + bias];
    uint8_t *r0 = w.base + static_cast<size_t>(i) * w.pitch;
    const size_t d = static_cast<size_t>(dist) * w.pitch;
    typename F::E x0 = F::load(r0, u), x1 = F::load(r0 + d, u), x2 = F::load(r0 + 2 * d, u),
                  x3 = F::load(r0 + 3 * d, u);
    if (INV) {  // leopard.go:875-897 ifftDIT4Ref
        ifft2<F>(x0, x1, lm01, f);
        ifft2<F>(x2, x3, lm23, f);
        ifft2<F>(x0, x2, lm02, f);
        ifft2<F>(x1, x3, lm02, f);
    } else {    // leopard.go:739-761 fftDIT4Ref
        fft2<F>(x0, x2, lm02, f);
        fft2<F>(x1, x3, lm02, f);
        fft2<F>(x0, x1, lm01, f);
        fft2<F>(x2, x3, lm23, f);
    }
    F::store(r0, u, x0);
    F::store(r0 + d, u, x1);
    F::store(r0 + 2 * d, u, x2);
    F::store(r0 + 3 * d, u, x3);
    }
}

template <class F>
__global__ __launch_bounds__(kBlock) void k_ifft2_last(const LeoDev f, const LeoWork w, int dist,
                                                       unsigned lm) {
    const int u = blockIdx.x * kBlock + threadIdx.x;
    if (u >= static_cast<int>(w.cols / F::kUnitBytes)) return;
    for (int i = blockIdx.y; i < dist; i += gridDim.y) {
        uint8_t *rx = w.base + static_cast<size_t>(i) * w.pitch;
        uint8_t *ry = rx + static_cast<size_t>(dist) * w.pitch;
        typename F::E x = F::load(rx, u), y = F::load(ry, u);
        ifft2<F>(x, y, lm, f);
        F::store(rx, u, x);
        F::store(ry, u, y);
    }
}

template <class F>
__global__ __launch_bounds__(kBlock) void k_fft2_last(const LeoDev f, const LeoWork w, int pairs) {
    const int u = blockIdx.x * kBlock + threadIdx.x;
    if (u >= static_cast<int>(w.cols / F::kUnitBytes)) return;
    for (int q = blockIdx.y; q < pairs; q += gridDim.y) {
        const int r = 2 * q;
        uint8_t *rx = w.base + static_cast<size_t>(r) * w.pitch;
        uint8_t *ry = rx + w.pitch;
        typename F::E x = F::load(rx, u), y = F::load(ry, u);
        fft2<F>(x, y, f.skew[r], f);
        F::store(rx, u, x);
        F::store(ry, u, y);
    }
}

template <class F>
__global__ __launch_bounds__(kBlock) void k_load_rows(const LeoDev f, const uint8_t *const *ptrs,
                                                      const uint16_t *logs, size_t col_off,
                                                      const LeoWork w, int nrows) {
    const int u = blockIdx.x * kBlock + threadIdx.x;
    if (u >= static_cast<int>(w.cols / F::kUnitBytes)) return;
    for (int t = blockIdx.y; t < nrows; t += gridDim.y) {
        const uint8_t *src = ptrs[t];
        typename F::E v = F::zero();
        if (src) {
            const typename F::E x = F::load(src + col_off, u);
            if (logs) F::mul_add(v, x, table<F>(f, logs[t]));
            else v = x;
        }
        F::store(w.base + static_cast<size_t>(t) * w.pitch, u, v);
    }
}

template <class F>
__global__ __launch_bounds__(kBlock) void k_store_rows(const LeoDev f, const LeoWork w,
                                                       const int *rows, const uint16_t *logs,
                                                       uint8_t *const *dst, size_t col_off, int count) {
    const int u = blockIdx.x * kBlock + threadIdx.x;
    if (u >= static_cast<int>(w.cols / F::kUnitBytes)) return;
    for (int q = blockIdx.y; q < count; q += gridDim.y) {
        const typename F::E x = F::load(w.base + static_cast<size_t>(rows[q]) * w.pitch, u);
        typename F::E v = F::zero();
        if (logs) F::mul_add(v, x, table<F>(f, logs[q]));
        else v = x;
        F::store(dst[q] + col_off, u, v);
    }
}

__global__ __launch_bounds__(kBlock) void k_xor_rows(const LeoWork dst, const LeoWork src, int nrows) {
    const int u = blockIdx.x * kBlock + threadIdx.x;
    if (u >= static_cast<int>(dst.cols / 16)) return;
    for (int t = blockIdx.y; t < nrows; t += gridDim.y) {
        uint4 *d = reinterpret_cast<uint4 *>(dst.base + static_cast<size_t>(t) * dst.pitch) + u;
        const uint4 *s = reinterpret_cast<const uint4 *>(src.base + static_cast<size_t>(t) * src.pitch) + u;
        *d = xor4(*d, *s);
    }
}

__global__ __launch_bounds__(kBlock) void k_derivative(const LeoWork in, const LeoWork out, int n) {
    const int u = blockIdx.x * kBlock + threadIdx.x;
    if (u >= static_cast<int>(in.cols / 16)) return;
    for (int t = blockIdx.y; t < n; t += gridDim.y) {
        uint4 acc = reinterpret_cast<const uint4 *>(in.base + static_cast<size_t>(t) * in.pitch)[u];
        for (int b = 1; b < n; b <<= 1)
            if (!(t & b))
                acc = xor4(acc, reinterpret_cast<const uint4 *>(in.base + static_cast<size_t>(t | b) * in.pitch)[u]);
        reinterpret_cast<uint4 *>(out.base + static_cast<size_t>(t) * out.pitch)[u] = acc;
    }
}

// Lanes over column units in x; rows (or butterfly units) in y, grid-strided
// past 32768 so every launch stays within the y-dimension limit.
dim3 grid2(size_t units, int rows) {
    return dim3(static_cast<unsigned>((units + kBlock - 1) / kBlock),
                static_cast<unsigned>(std::min(rows, 32768)));
}

}  // namespace

hipError_t leo_load_rows(const LeoDev &f, const uint8_t *const *ptrs, const uint16_t *logs,
                         size_t col_off, int nrows, const LeoWork &w, hipStream_t s) {
    if (nrows <= 0) return hipSuccess;
    if (f.bits == 16)
        hipLaunchKernelGGL(k_load_rows<F16>, grid2(w.cols / F16::kUnitBytes, nrows), dim3(kBlock), 0, s,
                           f, ptrs, logs, col_off, w, nrows);
    else
        hipLaunchKernelGGL(k_load_rows<F8>, grid2(w.cols / F8::kUnitBytes, nrows), dim3(kBlock), 0, s,
                           f, ptrs, logs, col_off, w, nrows);
    return hipGetLastError();
}

hipError_t leo_step4(const LeoDev &f, bool inverse, const LeoWork &w, int dist, int mtrunc,
                     const uint16_t *skew, int bias, hipStream_t s) {
    const int groups = (mtrunc + 4 * dist - 1) / (4 * dist);
    const int units = groups * dist;
    if (units <= 0) return hipSuccess;
    if (f.bits == 16) {
        const dim3 g = grid2(w.cols / F16::kUnitBytes, units);
        if (inverse) hipLaunchKernelGGL((k_step4<F16, true>), g, dim3(kBlock), 0, s, f, w, dist, units, skew, bias);
        else hipLaunchKernelGGL((k_step4<F16, false>), g, dim3(kBlock), 0, s, f, w, dist, units, skew, bias);
    } else {
        const dim3 g = grid2(w.cols / F8::kUnitBytes, units);
        if (inverse) hipLaunchKernelGGL((k_step4<F8, true>), g, dim3(kBlock), 0, s, f, w, dist, units, skew, bias);
        else hipLaunchKernelGGL((k_step4<F8, false>), g, dim3(kBlock), 0, s, f, w, dist, units, skew, bias);
    }
    return hipGetLastError();
}

hipError_t leo_ifft2_last(const LeoDev &f, const LeoWork &w, int dist, uint16_t log_m,
                          hipStream_t s) {
    if (f.bits == 16)
        hipLaunchKernelGGL(k_ifft2_last<F16>, grid2(w.cols / F16::kUnitBytes, dist), dim3(kBlock), 0, s,
                           f, w, dist, static_cast<unsigned>(log_m));
    else
        hipLaunchKernelGGL(k_ifft2_last<F8>, grid2(w.cols / F8::kUnitBytes, dist), dim3(kBlock), 0, s,
                           f, w, dist, static_cast<unsigned>(log_m));
    return hipGetLastError();
}

hipError_t leo_fft2_last(const LeoDev &f, const LeoWork &w, int mtrunc, hipStream_t s) {
    const int pairs = (mtrunc + 1) / 2;
    if (pairs <= 0) return hipSuccess;
    if (f.bits == 16)
        hipLaunchKernelGGL(k_fft2_last<F16>, grid2(w.cols / F16::kUnitBytes, pairs), dim3(kBlock), 0, s, f, w, pairs);
    else
        hipLaunchKernelGGL(k_fft2_last<F8>, grid2(w.cols / F8::kUnitBytes, pairs), dim3(kBlock), 0, s, f, w, pairs);
    return hipGetLastError();
}

hipError_t leo_xor_rows(const LeoWork &dst, const LeoWork &src, int nrows, hipStream_t s) {
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_xor_rows, grid2(dst.cols / 16, nrows), dim3(kBlock), 0, s, dst, src, nrows);
    return hipGetLastError();
}

hipError_t leo_derivative(const LeoWork &in, const LeoWork &out, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_derivative, grid2(in.cols / 16, n), dim3(kBlock), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t leo_store_rows(const LeoDev &f, const LeoWork &w, const int *rows, const uint16_t *logs,
                          uint8_t *const *dst, size_t col_off, int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (f.bits == 16)
        hipLaunchKernelGGL(k_store_rows<F16>, grid2(w.cols / F16::kUnitBytes, count), dim3(kBlock), 0, s,
                           f, w, rows, logs, dst, col_off, count);
    else
        hipLaunchKernelGGL(k_store_rows<F8>, grid2(w.cols / F8::kUnitBytes, count), dim3(kBlock), 0, s,
                           f, w, rows, logs, dst, col_off, count);
    return hipGetLastError();
}

}  // namespace rsx

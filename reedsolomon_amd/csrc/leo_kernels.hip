// leo_kernels.hip — Leopard GF(2^16) / GF(2^8) butterflies on CDNA4 (gfx950).
//
// Thread layout: blockIdx.y picks the butterfly unit (so its skews, and the
// "skew == modulus => xor only" branch, are uniform across the block) and the
// lanes of the block spread over the column chunk: one lane owns one
// 16-element column unit of every row it touches.
#include "leo_kernels.hpp"
#include "leo_device.hpp"

#include <algorithm>

namespace rsx {

namespace {

constexpr int kBlock = 256;

using namespace leo;

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

// One 16-element unit of `in` times each scalar, XORed into each output.
__global__ __launch_bounds__(kBlock) void k_gf16_mul_xor(const LeoDev f, const Gf16MulXor a) {
    const size_t units = a.len / F16::kUnitBytes;
    for (size_t u = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; u < units;
         u += static_cast<size_t>(gridDim.x) * kBlock) {
        const F16::E y = F16::load(a.in, static_cast<int>(u));
        for (int j = 0; j < a.nout; ++j) {
            F16::E x = F16::load(a.out[j], static_cast<int>(u));
            F16::mul_add(x, y, table<F16>(f, a.logs[j]));
            F16::store(a.out[j], static_cast<int>(u), x);
        }
    }
}

// Lanes over column units in x; rows (or butterfly units) in y, grid-strided
// past 32768 so every launch stays within the y-dimension limit.
dim3 grid2(size_t units, int rows) {
    return dim3(static_cast<unsigned>((units + kBlock - 1) / kBlock),
                static_cast<unsigned>(std::min(rows, 32768)));
}

}  // namespace

namespace {

constexpr int kErrThreads = 1024;
constexpr unsigned kOrder16 = 65536;

// In-place Walsh-Hadamard transform mod 65535 over e[0, 65536), radix-4 steps
// (dist 1, 4, 16, ...), only blocks that start below mtrunc: the arithmetic of
// LeoField::fwht (32-bit sums folded to 16 bits), so the results are equal.
__device__ void fwht_lds(uint16_t *e, int mtrunc) {
    constexpr uint32_t mask = 0xFFFFu;
    for (int dist = 1; dist * 4 <= static_cast<int>(kOrder16); dist *= 4) {
        const int dist4 = dist * 4;
        const int groups = (mtrunc + dist4 - 1) / dist4 * dist;
        for (int g = threadIdx.x; g < groups; g += kErrThreads) {
            const int r = g / dist * dist4 + g % dist;
            const uint32_t t0 = e[r], t1 = e[r + dist], t2 = e[r + 2 * dist], t3 = e[r + 3 * dist];
            uint32_t s01 = t0 + t1, d01 = t0 - t1, s23 = t2 + t3, d23 = t2 - t3;
            s01 = (s01 + (s01 >> 16)) & mask;
            d01 = (d01 + (d01 >> 16)) & mask;
            s23 = (s23 + (s23 >> 16)) & mask;
            d23 = (d23 + (d23 >> 16)) & mask;
            const uint32_t a = s01 + s23, c = s01 - s23, b = d01 + d23, d = d01 - d23;
            e[r] = static_cast<uint16_t>((a + (a >> 16)) & mask);
            e[r + 2 * dist] = static_cast<uint16_t>((c + (c >> 16)) & mask);
            e[r + dist] = static_cast<uint16_t>((b + (b >> 16)) & mask);
            e[r + 3 * dist] = static_cast<uint16_t>((d + (d >> 16)) & mask);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kErrThreads) void k_error_locators(const uint16_t *walsh, const uint8_t *ind,
                                                                int mtrunc, int n, const int *rows, int nrows,
                                                                uint16_t *errlocs, uint16_t *logs) {
    extern __shared__ uint16_t e[];  // kOrder16 entries
    for (int i = threadIdx.x; i < static_cast<int>(kOrder16); i += kErrThreads) e[i] = i < mtrunc ? ind[i] : 0;
    __syncthreads();
    fwht_lds(e, mtrunc);
    // (e * logWalsh) mod 65535 by folding: 2^16 is 1 modulo 65535.
    for (int i = threadIdx.x; i < static_cast<int>(kOrder16); i += kErrThreads) {
        uint32_t x = static_cast<uint32_t>(e[i]) * walsh[i];
        x = (x & 0xFFFFu) + (x >> 16);
        x = (x & 0xFFFFu) + (x >> 16);
        if (x >= 0xFFFFu) x -= 0xFFFFu;
        e[i] = static_cast<uint16_t>(x);
    }
    __syncthreads();
    fwht_lds(e, static_cast<int>(kOrder16));
    for (int i = threadIdx.x; i < n; i += kErrThreads) errlocs[i] = e[i];
    for (int q = threadIdx.x; q < nrows; q += kErrThreads) logs[q] = static_cast<uint16_t>(0xFFFFu - e[rows[q]]);
}

}  // namespace

hipError_t leo_error_locators_device(const LeoDev &f, const uint8_t *ind, int mtrunc, int n, const int *rows,
                                     int nrows, uint16_t *errlocs, uint16_t *logs, hipStream_t s) {
    if (f.bits != 16 || !f.walsh || mtrunc <= 0 || mtrunc > static_cast<int>(kOrder16) || n <= 0 ||
        n > static_cast<int>(kOrder16) || nrows < 0)
        return hipErrorInvalidValue;
    const size_t lds = kOrder16 * sizeof(uint16_t);
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_error_locators),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_error_locators, dim3(1), dim3(kErrThreads), lds, s, f.walsh, ind, mtrunc, n, rows, nrows,
                       errlocs, logs);
    return hipGetLastError();
}

hipError_t leo_gf16_mul_xor(const LeoDev &f, const Gf16MulXor &a, hipStream_t s) {
    const size_t units = a.len / F16::kUnitBytes;
    if (units == 0 || a.nout <= 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((units + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_gf16_mul_xor, dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s, f, a);
    return hipGetLastError();
}

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

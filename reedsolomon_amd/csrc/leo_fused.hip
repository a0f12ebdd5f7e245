// leo_fused.hip — Leopard encode / reconstruct with every work row of one
// column slice resident in LDS (gfx950: up to 160 KiB per workgroup).
//
// One workgroup owns a slice of `slice` bytes of every shard.  It reads the
// slice's input rows from HBM once, runs all butterfly layers of the
// transform in LDS (radix-4 steps, one barrier each), and writes the output
// rows once, so HBM sees exactly the algorithmic bytes.  The butterfly
// schedule, skews and truncations are those of the layer-per-launch kernels
// (leopard.go:696-897 fftDIT / ifftDIT, leopard8.go likewise); only where the
// work rows live differs.
//
// Lanes cover (butterfly unit, column unit) pairs: consecutive lanes take
// consecutive 16-byte column units of the same rows, so LDS accesses of a
// wave are contiguous.  Skews and multiply records are per lane.
#include <algorithm>

#include "leo_device.hpp"
#include "leo_kernels.hpp"

namespace rsx {

namespace {

using namespace leo;

constexpr int kFB = 256;        // threads per workgroup
constexpr int kDerivMax = kLeoFusedDecodeLds / 16 / kFB;  // uint4 per thread in the derivative

// Work rows of one slice in LDS: row r at base + r * pitch, a row laid out
// exactly as the slice of the shard it came from.
struct Rows {
    uint8_t *base;
    int pitch;   // bytes per row (the slice width)
    int ushift;  // log2(column units per row)
    int ucount;  // column units of this slice (the last slice may be narrower)
};

__device__ __forceinline__ int log2i(int x) {
    int s = 0;
    while ((1 << s) < x) ++s;
    return s;
}

__device__ __forceinline__ uint8_t *row(const Rows &w, int r) { return w.base + r * w.pitch; }

// One radix-4 step (two layers) at distance 2^dshift over butterfly groups
// r < mtrunc, as in k_step4.
template <class F, bool INV>
__device__ __forceinline__ void step4(const Rows &w, int dshift, int mtrunc, const uint16_t *skew, int bias,
                                      const LeoDev &f) {
    const int dist = 1 << dshift;
    const int groups = (mtrunc + 4 * dist - 1) >> (dshift + 2);
    const int items = (groups << dshift) << w.ushift;
    const int umask = (1 << w.ushift) - 1;
    for (int it = threadIdx.x; it < items; it += kFB) {
        const int u = it & umask;
        if (u >= w.ucount) continue;
        const int q = it >> w.ushift;
        const int r = (q >> dshift) << (dshift + 2);
        const int i = r + (q & (dist - 1));
        const int iend = r + dist;
        const unsigned lm01 = skew[iend + bias], lm02 = skew[iend + dist + bias],
                       lm23 = skew[iend + 2 * dist + bias];
        uint8_t *r0 = row(w, i);
        const int d = dist * w.pitch;
        typename F::E x0 = F::load(r0, u), x1 = F::load(r0 + d, u), x2 = F::load(r0 + 2 * d, u),
                      x3 = F::load(r0 + 3 * d, u);
        if (INV) {
            ifft2<F>(x0, x1, lm01, f);
            ifft2<F>(x2, x3, lm23, f);
            ifft2<F>(x0, x2, lm02, f);
            ifft2<F>(x1, x3, lm02, f);
        } else {
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
    __syncthreads();
}

// IFFT over `n` rows truncated to `mtrunc`; `skew` + `bias` as in leo_step4,
// the final radix-2 layer (odd log2 n) uses skew[dist + bias].
template <class F>
__device__ __forceinline__ void ifft(const Rows &w, int n, int mtrunc, const uint16_t *skew, int bias,
                                     const LeoDev &f) {
    int dshift = 0;
    while ((4 << dshift) <= n) {
        step4<F, true>(w, dshift, mtrunc, skew, bias, f);
        dshift += 2;
    }
    const int dist = 1 << dshift;
    if (dist < n) {
        const unsigned lm = skew[dist + bias];
        const int items = dist << w.ushift;
        const int umask = (1 << w.ushift) - 1;
        for (int it = threadIdx.x; it < items; it += kFB) {
            const int u = it & umask;
            if (u >= w.ucount) continue;
            const int i = it >> w.ushift;
            uint8_t *rx = row(w, i), *ry = row(w, i + dist);
            typename F::E x = F::load(rx, u), y = F::load(ry, u);
            ifft2<F>(x, y, lm, f);
            F::store(rx, u, x);
            F::store(ry, u, y);
        }
        __syncthreads();
    }
}

// FFT over `n` rows truncated to `mtrunc` (skew = fftSkew, bias -1).
template <class F>
__device__ __forceinline__ void fft(const Rows &w, int n, int mtrunc, const LeoDev &f) {
    int dist4 = n, dist = n >> 2;
    while (dist != 0) {
        step4<F, false>(w, log2i(dist), mtrunc, f.skew, -1, f);
        dist4 = dist;
        dist >>= 2;
    }
    if (dist4 == 2) {
        const int pairs = (mtrunc + 1) / 2;
        const int items = pairs << w.ushift;
        const int umask = (1 << w.ushift) - 1;
        for (int it = threadIdx.x; it < items; it += kFB) {
            const int u = it & umask;
            if (u >= w.ucount) continue;
            const int r = 2 * (it >> w.ushift);
            uint8_t *rx = row(w, r), *ry = row(w, r + 1);
            typename F::E x = F::load(rx, u), y = F::load(ry, u);
            fft2<F>(x, y, f.skew[r], f);
            F::store(rx, u, x);
            F::store(ry, u, y);
        }
        __syncthreads();
    }
}

// rows [0, nrows) of w <- src[t] (+col) times exp(logs[t]) when logs, zero
// where src[t] is null.
template <class F>
__device__ __forceinline__ void load_rows(const Rows &w, const uint8_t *const *src, const uint16_t *logs,
                                          int nrows, size_t col, const LeoDev &f) {
    const int items = nrows << w.ushift;
    const int umask = (1 << w.ushift) - 1;
    for (int it = threadIdx.x; it < items; it += kFB) {
        const int u = it & umask;
        if (u >= w.ucount) continue;
        const int t = it >> w.ushift;
        const uint8_t *s = src[t];
        typename F::E v = F::zero();
        if (s) {
            const typename F::E x = F::gload(s + col, u);
            if (logs) F::mul_add(v, x, table<F>(f, logs[t]));
            else v = x;
        }
        F::store(row(w, t), u, v);
    }
    __syncthreads();
}

// dst[q] (+col) <- row rows[q] (or row q) times exp(logs[q]) when logs.
template <class F>
__device__ __forceinline__ void store_rows(const Rows &w, uint8_t *const *dst, const int *rows,
                                           const uint16_t *logs, int count, size_t col, const LeoDev &f) {
    const int items = count << w.ushift;
    const int umask = (1 << w.ushift) - 1;
    for (int it = threadIdx.x; it < items; it += kFB) {
        const int u = it & umask;
        if (u >= w.ucount) continue;
        const int q = it >> w.ushift;
        const typename F::E x = F::load(row(w, rows ? rows[q] : q), u);
        typename F::E v = F::zero();
        if (logs) F::mul_add(v, x, table<F>(f, logs[q]));
        else v = x;
        F::gstore(dst[q] + col, u, v);
    }
}

// acc ^= tmp over nrows rows of `chunks` 16-byte chunks (2^cshift per row).
__device__ __forceinline__ void xor_rows(const Rows &acc, const Rows &tmp, int nrows, int chunks, int cshift) {
    const int items = nrows << cshift;
    const int cmask = (1 << cshift) - 1;
    for (int it = threadIdx.x; it < items; it += kFB) {
        const int c = it & cmask;
        if (c >= chunks) continue;
        const int t = it >> cshift;
        uint4 *a = reinterpret_cast<uint4 *>(acc.base + t * acc.pitch) + c;
        const uint4 *b = reinterpret_cast<const uint4 *>(tmp.base + t * tmp.pitch) + c;
        *a = xor4(*a, *b);
    }
    __syncthreads();
}

// Formal derivative in place (leopard.go:619-623): out[t] = in[t] ^ XOR over
// zero bits b of t (t | 2^b < n) of in[t | 2^b].  Every thread computes its
// chunks into registers, then all store after a barrier.
__device__ __forceinline__ void derivative(const Rows &w, int n, int chunks, int cshift) {
    const int items = n << cshift;
    const int cmask = (1 << cshift) - 1;
    uint4 res[kDerivMax];
#pragma unroll
    for (int j = 0; j < kDerivMax; ++j) {
        const int it = threadIdx.x + j * kFB;
        if (it < items && (it & cmask) < chunks) {
            const int t = it >> cshift, c = it & cmask;
            uint4 acc = reinterpret_cast<const uint4 *>(w.base + t * w.pitch)[c];
            for (int b = 1; b < n; b <<= 1)
                if (!(t & b)) acc = xor4(acc, reinterpret_cast<const uint4 *>(w.base + (t | b) * w.pitch)[c]);
            res[j] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kDerivMax; ++j) {
        const int it = threadIdx.x + j * kFB;
        if (it < items && (it & cmask) < chunks)
            reinterpret_cast<uint4 *>(w.base + (it >> cshift) * w.pitch)[it & cmask] = res[j];
    }
    __syncthreads();
}

template <class F>
__global__ __launch_bounds__(kFB) void k_encode_fused(const LeoDev f, const LeoEncodeArgs a) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    const int ushift_full = log2i(a.slice / F::kUnitBytes);
    const int cshift = log2i(a.slice / 16);
    for (size_t col = static_cast<size_t>(blockIdx.x) * a.slice; col < a.size;
         col += static_cast<size_t>(gridDim.x) * a.slice) {
        const int cols = static_cast<int>(a.size - col < static_cast<size_t>(a.slice) ? a.size - col : a.slice);
        const Rows acc{lds, a.slice, ushift_full, cols / F::kUnitBytes};
        const Rows tmp{lds + a.m * a.slice, a.slice, ushift_full, cols / F::kUnitBytes};
        for (int g = 0; g < a.groups; ++g) {
            const Rows &w = g == 0 ? acc : tmp;
            const int mtrunc = g == 0 ? (a.k < a.m ? a.k : a.m) : (a.k - g * a.m < a.m ? a.k - g * a.m : a.m);
            load_rows<F>(w, a.data + static_cast<size_t>(g) * a.m, nullptr, a.m, col, f);
            ifft<F>(w, a.m, mtrunc, f.skew + a.m - 1 + g * a.m, 0, f);
            if (g > 0) xor_rows(acc, tmp, a.m, cols / 16, cshift);
        }
        fft<F>(acc, a.m, a.p, f);
        store_rows<F>(acc, a.parity, nullptr, nullptr, a.p, col, f);
        __syncthreads();  // the next slice reuses the rows
    }
}

template <class F>
__global__ __launch_bounds__(kFB) void k_decode_fused(const LeoDev f, const LeoDecodeArgs a) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    const int ushift_full = log2i(a.slice / F::kUnitBytes);
    const int cshift = log2i(a.slice / 16);
    for (size_t col = static_cast<size_t>(blockIdx.x) * a.slice; col < a.size;
         col += static_cast<size_t>(gridDim.x) * a.slice) {
        const int cols = static_cast<int>(a.size - col < static_cast<size_t>(a.slice) ? a.size - col : a.slice);
        const Rows w{lds, a.slice, ushift_full, cols / F::kUnitBytes};
        load_rows<F>(w, a.rows, a.errlocs, a.n, col, f);
        ifft<F>(w, a.n, a.ifft_trunc, f.skew, -1, f);
        derivative(w, a.n, cols / 16, cshift);
        fft<F>(w, a.n, a.fft_trunc, f);
        store_rows<F>(w, a.out, a.out_rows, a.out_logs, a.nout, col, f);
        __syncthreads();
    }
}

unsigned grid_for(size_t size, int slice) {
    const size_t slices = (size + slice - 1) / slice;
    return static_cast<unsigned>(std::min<size_t>(slices, 65535));
}

}  // namespace

int leo_fused_slice(int bits, int rows, int step_rows, size_t lds_budget) {
    const int unit = bits == 16 ? 32 : 16;
    // Enough column units that one radix-4 step gives every lane an item,
    // at least 128 bytes per row read (64 where nothing wider fits) ...
    int slice = 128;
    while ((step_rows / 4) * (slice / unit) < kFB && slice < 4096) slice *= 2;
    // ... within the LDS budget.
    while (slice > 64 && static_cast<size_t>(rows) * slice > lds_budget) slice /= 2;
    return static_cast<size_t>(rows) * slice <= lds_budget ? slice : 0;
}

template <class K>
static hipError_t allow_lds(K kernel, size_t lds) {
    if (lds <= (64u << 10)) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(lds));
}

hipError_t leo_encode_fused(const LeoDev &f, const LeoEncodeArgs &a, hipStream_t s) {
    const int rows = a.groups > 1 ? 2 * a.m : a.m;
    const size_t lds = static_cast<size_t>(rows) * a.slice;
    hipError_t e;
    if (f.bits == 16) {
        if ((e = allow_lds(k_encode_fused<F16>, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_encode_fused<F16>, dim3(grid_for(a.size, a.slice)), dim3(kFB), lds, s, f, a);
    } else {
        if ((e = allow_lds(k_encode_fused<F8>, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_encode_fused<F8>, dim3(grid_for(a.size, a.slice)), dim3(kFB), lds, s, f, a);
    }
    return hipGetLastError();
}

hipError_t leo_decode_fused(const LeoDev &f, const LeoDecodeArgs &a, hipStream_t s) {
    const size_t lds = static_cast<size_t>(a.n) * a.slice;
    if (lds > kLeoFusedDecodeLds) return hipErrorInvalidValue;  // derivative register budget
    if (f.bits == 16)
        hipLaunchKernelGGL(k_decode_fused<F16>, dim3(grid_for(a.size, a.slice)), dim3(kFB), lds, s, f, a);
    else
        hipLaunchKernelGGL(k_decode_fused<F8>, dim3(grid_for(a.size, a.slice)), dim3(kFB), lds, s, f, a);
    return hipGetLastError();
}

}  // namespace rsx

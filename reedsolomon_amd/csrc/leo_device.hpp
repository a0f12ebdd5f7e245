// leo_device.hpp — device-side field arithmetic of the Leopard engines,
// shared by the layer-per-launch kernels and the LDS-resident kernels.
//
// GF(2^16) products use the split lo/hi byte layout directly: 12
// v_perm_b32 byte selects per 4 elements (3 bit groups x 2 source bytes x 2
// product bytes), folded with 3-input XORs.  GF(2^8) products take 3 selects
// per 4 elements.  A multiply record holds the selects' 8-entry byte tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>

#include "leo_kernels.hpp"

namespace rsx {
namespace leo {

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
typedef uint32_t gu32x4 __attribute__((ext_vector_type(4)));
using GlobalV4 = __attribute__((address_space(1))) gu32x4;
// Shard rows are global memory: typed so the compiler emits global (not
// flat) accesses.
__device__ __forceinline__ uint4 gload4(const uint8_t *p) {
    const gu32x4 v = *(const GlobalV4 *)(reinterpret_cast<uintptr_t>(p));
    return make_uint4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void gstore4(uint8_t *p, const uint4 &v) {
    *(GlobalV4 *)(reinterpret_cast<uintptr_t>(p)) = gu32x4{v.x, v.y, v.z, v.w};
}
typedef uint32_t gu32x2 __attribute__((ext_vector_type(2)));
using GlobalV2 = __attribute__((address_space(1))) gu32x2;
__device__ __forceinline__ uint2 gload2(const uint8_t *p) {
    const gu32x2 v = *(const GlobalV2 *)(reinterpret_cast<uintptr_t>(p));
    return make_uint2(v[0], v[1]);
}
__device__ __forceinline__ void gstore2(uint8_t *p, const uint2 &v) {
    *(GlobalV2 *)(reinterpret_cast<uintptr_t>(p)) = gu32x2{v.x, v.y};
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
    __device__ static E gload(const uint8_t *row, int u) {
        const uint8_t *c = row + offset(u);
        return {gload4(c), gload4(c + 32)};
    }
    __device__ static void gstore(uint8_t *row, int u, const E &e) {
        uint8_t *c = row + offset(u);
        gstore4(c, e.lo);
        gstore4(c + 32, e.hi);
    }
    __device__ static E zero() { return {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)}; }
    __device__ static void xor_into(E &x, const E &y) {
        x.lo = xor4(x.lo, y.lo);
        x.hi = xor4(x.hi, y.hi);
    }
    __device__ static uint32_t differ(const E &x, const E &y) {
        const uint4 a = xor4(x.lo, y.lo), b = xor4(x.hi, y.hi);
        return a.x | a.y | a.z | a.w | b.x | b.y | b.z | b.w;
    }
    // (lo, hi) of one dword pair times the table's element, XORed into (xl, xh).
    // T: the multiply record in memory (ConstU32 *) or in registers.
    template <class T>
    __device__ static void mac1(uint32_t &xl, uint32_t &xh, uint32_t L, uint32_t H, const T &t) {
        const uint32_t a0 = L & 0x07070707u, a1 = (L >> 3) & 0x07070707u, a2 = (L >> 6) & 0x03030303u;
        const uint32_t b0 = H & 0x07070707u, b1 = (H >> 3) & 0x07070707u, b2 = (H >> 6) & 0x03030303u;
        uint32_t r = xor3(xl, sel(t[1], t[0], a0), sel(t[5], t[4], a1));
        r = xor3(r, sel(t[8], t[8], a2), sel(t[11], t[10], b0));
        xl = xor3(r, sel(t[15], t[14], b1), sel(t[18], t[18], b2));
        r = xor3(xh, sel(t[3], t[2], a0), sel(t[7], t[6], a1));
        r = xor3(r, sel(t[9], t[9], a2), sel(t[13], t[12], b0));
        xh = xor3(r, sel(t[17], t[16], b1), sel(t[19], t[19], b2));
    }
    template <class T>
    __device__ static void mul_add(E &x, const E &y, const T &t) {
        mac1(x.lo.x, x.hi.x, y.lo.x, y.hi.x, t);
        mac1(x.lo.y, x.hi.y, y.lo.y, y.hi.y, t);
        mac1(x.lo.z, x.hi.z, y.lo.z, y.hi.z, t);
        mac1(x.lo.w, x.hi.w, y.lo.w, y.hi.w, t);
    }
};

// ---- GF(2^16), quarter chunks: 8 elements per unit, the 8 low bytes at
// offset 8q of a 64-byte chunk and their 8 high bytes 32 further ----
struct F16Q {
    struct E {
        uint2 lo, hi;
    };
    static constexpr int kUnitBytes = 16;  // row bytes one unit covers
    static constexpr int kTW = F16::kTW;
    static constexpr unsigned kMod = F16::kMod;

    __device__ static size_t offset(int u) { return static_cast<size_t>(u >> 2) * 64 + (u & 3) * 8; }
    __device__ static E gload(const uint8_t *row, int u) {
        const uint8_t *c = row + offset(u);
        return {gload2(c), gload2(c + 32)};
    }
    __device__ static void gstore(uint8_t *row, int u, const E &e) {
        uint8_t *c = row + offset(u);
        gstore2(c, e.lo);
        gstore2(c + 32, e.hi);
    }
    __device__ static E zero() { return {make_uint2(0, 0), make_uint2(0, 0)}; }
    __device__ static void xor_into(E &x, const E &y) {
        x.lo = make_uint2(x.lo.x ^ y.lo.x, x.lo.y ^ y.lo.y);
        x.hi = make_uint2(x.hi.x ^ y.hi.x, x.hi.y ^ y.hi.y);
    }
    __device__ static uint32_t differ(const E &x, const E &y) {
        return (x.lo.x ^ y.lo.x) | (x.lo.y ^ y.lo.y) | (x.hi.x ^ y.hi.x) | (x.hi.y ^ y.hi.y);
    }
    template <class T>
    __device__ static void mul_add(E &x, const E &y, const T &t) {
        F16::mac1(x.lo.x, x.hi.x, y.lo.x, y.hi.x, t);
        F16::mac1(x.lo.y, x.hi.y, y.lo.y, y.hi.y, t);
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
    __device__ static E gload(const uint8_t *row, int u) { return gload4(row + 16 * u); }
    __device__ static void gstore(uint8_t *row, int u, const E &e) { gstore4(row + 16 * u, e); }
    __device__ static E zero() { return make_uint4(0, 0, 0, 0); }
    __device__ static void xor_into(E &x, const E &y) { x = xor4(x, y); }
    __device__ static uint32_t differ(const E &x, const E &y) {
        const uint4 a = xor4(x, y);
        return a.x | a.y | a.z | a.w;
    }
    template <class T>
    __device__ static uint32_t mac1(uint32_t x, uint32_t v, const T &t) {
        const uint32_t g0 = v & 0x07070707u, g1 = (v >> 3) & 0x07070707u, g2 = (v >> 6) & 0x03030303u;
        return xor3(xor3(x, sel(t[1], t[0], g0), sel(t[3], t[2], g1)), sel(t[4], t[4], g2), 0u);
    }
    template <class T>
    __device__ static void mul_add(E &x, const E &y, const T &t) {
        x.x = mac1(x.x, y.x, t);
        x.y = mac1(x.y, y.y, t);
        x.z = mac1(x.z, y.z, t);
        x.w = mac1(x.w, y.w, t);
    }
};

// GF(2^16) records are read from global memory (the constant address space:
// wave-uniform reads become scalar loads).  GF(2^8)'s 256 records (5 KiB) are
// staged in LDS by the LDS-resident kernels (lds_tables), so their pointer is
// kept generic.
// Logs are below 2^16, so the record offset is a (full-rate) 24-bit multiply.
template <class F>
__device__ __forceinline__ auto table(const LeoDev &f, unsigned log_m) {
    const unsigned off = __umul24(log_m, static_cast<unsigned>(F::kTW));
    if constexpr (std::is_same<F, F8>::value) return f.tables + off;
    else return as_const(f.tables) + off;
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
// The same with the multiply record of lm already in registers.
template <class F>
__device__ __forceinline__ void fft2r(typename F::E &x, typename F::E &y, unsigned lm, const uint32_t *t) {
    if (lm != F::kMod) F::mul_add(x, y, t);
    F::xor_into(y, x);
}
template <class F>
__device__ __forceinline__ void ifft2r(typename F::E &x, typename F::E &y, unsigned lm, const uint32_t *t) {
    F::xor_into(y, x);
    if (lm != F::kMod) F::mul_add(x, y, t);
}

}  // namespace leo
}  // namespace rsx

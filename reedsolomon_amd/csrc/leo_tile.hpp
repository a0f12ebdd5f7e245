// leo_tile.hpp — device building blocks of the LDS-resident Leopard
// kernels: the LDS tile layout, lane-to-work mappings, radix-4 / radix-2
// butterfly steps over a tile, row loads and stores, and the formal
// derivative.  The butterfly schedule, skews and truncations are those of the
// layer-per-launch kernels (leopard.go:696-897 fftDIT / ifftDIT,
// leopard8.go likewise); only where the work rows live differs.
#pragma once

#include <type_traits>

#include "leo_device.hpp"
#include "leo_kernels.hpp"

namespace rsx {
namespace leo_tile {

using namespace leo;

constexpr int kFB = kLeoTileThreads;      // threads per workgroup (512 where a step has that many items)
constexpr int kFBMax = kLeoTileThreadsMax;
constexpr int kDerivMax = kLeoDerivMax;    // uint4 per thread in the derivative

// Work rows of one slice in LDS, slot-major: 16-byte slot s of row r sits at
// base + (s * r1 + pos(r)) * 16 with r1 = rows + 1.  A GF(2^8) column unit
// is one slot.  A GF(2^16) unit is a lo slot and a hi slot `hio` further:
// slices of whole 64-byte chunks keep the chunk layout (hio = 2); a 32-byte
// slice holds one unit of a chunk (unit `ubase` of it, hio = 1), which halves
// the LDS of a tile.
//
// GF(2^16) butterfly lanes take consecutive rows of one slot; pos() permutes
// the low four row bits by bits 4 and 5 so that the 16 rows a quarter-wave
// touches in any radix-4 or radix-2 step (consecutive rows, rows 4q+j, rows
// 16a+b+4j, rows 2q) fall in 16 distinct bank groups.  Copy lanes, and the
// GF(2^8) butterfly lanes (kUnitsFastest), take consecutive slots of one
// row; the +1 row of padding keeps those distinct.
struct Tile {
    uint8_t *base;
    int r1;         // rows + 1
    int ushift;     // log2(column units in a full slice)
    int ucount;     // column units of this slice (the last slice may be narrower)
    int ubase = 0;  // unit of the shard chunk that local unit 0 is
    int hio = 2;    // GF(2^16): slots from a unit's lo slot to its hi slot
};

__device__ __forceinline__ int log2i(int x) {
    int s = 0;
    while ((1 << s) < x) ++s;
    return s;
}

// (b4 * 5) ^ (b5 * 10) == ((r >> 4) & 3) * 5: the two terms share no bits.
// pos is linear over XOR (pos(r ^ d) == pos(r) ^ pos(d)), so rows r and
// r + d with no common bits, as a butterfly's partners and the derivative's
// r | b are, sit at pos(r) ^ pos(d), one XOR from the first row's position.
__device__ __forceinline__ int pos(int r) { return r ^ (((r >> 4) & 3) * 5); }

// Workgroup index g -> unit slice index when a 64-byte GF(2^16) chunk is
// split into two 32-byte unit slices (slice 2c + u is unit u of chunk c).
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2, so g and g + 8 share an XCD: they take the two units of one chunk, and
// the chunk's cache lines are fetched from HBM once.  The tail past the last
// multiple of 16 keeps g itself.
__device__ __forceinline__ uint32_t xcd_unit_slice(uint32_t g, uint32_t count) {
    if (g >= (count & ~15u)) return g;
    return (((g >> 4) << 4) | ((g & 7) << 1)) | ((g >> 3) & 1);
}

// The same for four 16-byte quarter slices per chunk: g, g + 8, g + 16 and
// g + 24 take the quarters of one chunk.
__device__ __forceinline__ uint32_t xcd_quarter_slice(uint32_t g, uint32_t count) {
    if (g >= (count & ~31u)) return g;
    return (((g >> 5) << 5) | ((g & 7) << 2)) | ((g >> 3) & 3);
}

// Byte offset of row r within a slot column.
__device__ __forceinline__ int pos_off(int r) { return pos(r) << 4; }

// Slot column s (its row-position-0 slot); slot counts and row counts stay
// far below 2^24, so 24-bit multiplies (full rate) suffice for the offsets.
__device__ __forceinline__ uint8_t *column(const Tile &t, int s) { return t.base + (__mul24(s, t.r1) << 4); }
__device__ __forceinline__ uint4 *cell(uint8_t *col, int off) { return reinterpret_cast<uint4 *>(col + off); }

__device__ __forceinline__ uint4 *slot(const Tile &t, int r, int s) { return cell(column(t, s), pos_off(r)); }

// Element access by unit column (col) and row byte offset (pos_off of the
// row), so a butterfly's rows derive their offsets from the first one's.
template <class F>
struct Lds;
template <>
struct Lds<F8> {
    __device__ static uint8_t *col(const Tile &t, int u) { return column(t, u); }
    __device__ static uint4 at(const Tile &, uint8_t *c, int off) { return *cell(c, off); }
    __device__ static void put(const Tile &, uint8_t *c, int off, const uint4 &e) { *cell(c, off) = e; }
    __device__ static uint4 load(const Tile &t, int r, int u) { return at(t, col(t, u), pos_off(r)); }
    __device__ static void store(const Tile &t, int r, int u, const uint4 &e) { put(t, col(t, u), pos_off(r), e); }
};
template <>
struct Lds<F16> {
    __device__ static int lo(const Tile &t, int u) { return t.hio == 2 ? (((u >> 1) << 2) | (u & 1)) : (u << 1); }
    __device__ static uint8_t *col(const Tile &t, int u) { return column(t, lo(t, u)); }
    __device__ static int hoff(const Tile &t) { return (t.hio * t.r1) << 4; }  // lo slot -> hi slot
    __device__ static F16::E at(const Tile &t, uint8_t *c, int off) {
        return {*cell(c, off), *cell(c + hoff(t), off)};
    }
    __device__ static void put(const Tile &t, uint8_t *c, int off, const F16::E &e) {
        *cell(c, off) = e.lo;
        *cell(c + hoff(t), off) = e.hi;
    }
    __device__ static F16::E load(const Tile &t, int r, int u) { return at(t, col(t, u), pos_off(r)); }
    __device__ static void store(const Tile &t, int r, int u, const F16::E &e) { put(t, col(t, u), pos_off(r), e); }
};

// A quarter-chunk unit is one slot: its 8 low bytes, then its 8 high bytes.
template <>
struct Lds<F16Q> {
    __device__ static uint8_t *col(const Tile &t, int u) { return column(t, u); }
    __device__ static F16Q::E at(const Tile &, uint8_t *c, int off) {
        const uint4 v = *cell(c, off);
        return {make_uint2(v.x, v.y), make_uint2(v.z, v.w)};
    }
    __device__ static void put(const Tile &, uint8_t *c, int off, const F16Q::E &e) {
        *cell(c, off) = make_uint4(e.lo.x, e.lo.y, e.hi.x, e.hi.y);
    }
    __device__ static F16Q::E load(const Tile &t, int r, int u) { return at(t, col(t, u), pos_off(r)); }
    __device__ static void store(const Tile &t, int r, int u, const F16Q::E &e) { put(t, col(t, u), pos_off(r), e); }
};

// fn(q, u) for q < nq, u < ucount; q varies fastest across lanes.
template <class Fn>
__device__ __forceinline__ void for_rows(const Tile &t, int nq, Fn fn) {
    const int qs = log2i(nq);
    const int items = (1 << qs) << t.ushift;
    const int qmask = (1 << qs) - 1;
    for (int it = threadIdx.x; it < items; it += blockDim.x) {
        const int q = it & qmask, u = it >> qs;
        if (q < nq && u < t.ucount) fn(q, u);
    }
}

// fn(r, u) for r < nrows, u < ucount; u varies fastest (coalesced HBM rows).
template <class Fn>
__device__ __forceinline__ void for_units(const Tile &t, int nrows, Fn fn) {
    const int items = nrows << t.ushift;
    const int umask = (1 << t.ushift) - 1;
    for (int it = threadIdx.x; it < items; it += blockDim.x) {
        const int u = it & umask, r = it >> t.ushift;
        if (u < t.ucount) fn(r, u);
    }
}

// Lane order of the butterfly steps.  GF(2^8) tiles run the column units of
// one butterfly item fastest: the lanes of an item share its skews and
// multiply records (LDS broadcasts instead of a record per lane), and a
// wave's 64 >> ushift items are consecutive, so the skews are wave-uniform
// from distance 64 >> ushift.  Data reads stay conflict-free: consecutive
// slot columns are r1 = rows + 1 (odd) cells apart.  GF(2^16) tiles run
// rows fastest (the row swizzle keeps a quarter-wave's rows in distinct bank
// groups), which matters there because a unit is two slots.
template <class F>
constexpr bool kUnitsFastest = std::is_same<F, F8>::value;

template <class F, class Fn>
__device__ __forceinline__ void for_items(const Tile &t, int nq, Fn fn) {
    if constexpr (kUnitsFastest<F>) for_units(t, nq, fn);
    else for_rows(t, nq, fn);
}

// p[i] for a wave-uniform i, by a scalar load of the aligned dword holding
// it (the skew tables are whole allocations, so every entry is 2-byte
// aligned inside a 4-byte aligned buffer).
__device__ __forceinline__ unsigned uniform_u16(const uint16_t *p, int i) {
    i = __builtin_amdgcn_readfirstlane(i);
    const uintptr_t a = reinterpret_cast<uintptr_t>(p + i);
    const uint32_t d = *as_const(reinterpret_cast<const uint32_t *>(a & ~uintptr_t{3}));
    return (d >> ((a & 2) << 3)) & 0xFFFFu;
}

// A wave-uniform skew: GF(2^16) by a scalar load from global memory, GF(2^8)
// from the workgroup's LDS copy (lds_tables).
template <class F>
__device__ __forceinline__ unsigned uniform_skew(const uint16_t *p, int i) {
    if constexpr (std::is_same<F, F8>::value) return p[__builtin_amdgcn_readfirstlane(i)];
    else return uniform_u16(p, i);
}

// The three multiply records of a wave-uniform butterfly group, read
// together (scalar loads, one wait) before any product uses them.  A skew
// equal to the modulus reads a record that is never used (that product is
// skipped, leopard.go:741-760).
template <class F>
__device__ __forceinline__ void uniform_records(const LeoDev &f, unsigned lm01, unsigned lm02, unsigned lm23,
                                                uint32_t (&rec)[3][F::kTW]) {
    const auto t0 = table<F>(f, lm01), t1 = table<F>(f, lm02), t2 = table<F>(f, lm23);
#pragma unroll
    for (int o = 0; o < F::kTW; ++o) {
        rec[0][o] = t0[o];
        rec[1][o] = t1[o];
        rec[2][o] = t2[o];
    }
}

// Lane `SRC` of each aligned group of G lanes (a DPP quad for G = 4, a DPP
// row for G = 16) broadcast to the group.
template <int G, int SRC>
__device__ __forceinline__ uint32_t group_bcast(uint32_t v) {
    if constexpr (G == 4) return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), SRC * 0x55, 0xF, 0xF, false));
    else return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x150 + SRC, 0xF, 0xF, false));
}

// The three multiply records of a butterfly group whose G lanes are aligned
// and all active: lanes 0-2 of the group load one record each, the group
// shares them by broadcasts (one record load per lane instead of three).
template <class F, int G>
__device__ __forceinline__ void group_records(const LeoDev &f, unsigned lm01, unsigned lm02, unsigned lm23,
                                              uint32_t (&rec)[3][F::kTW]) {
    const int r = threadIdx.x & (G - 1);
    uint32_t v[F::kTW];
#pragma unroll
    for (int o = 0; o < F::kTW; ++o) v[o] = 0;
    if (r < 3) {
        const auto t = table<F>(f, r == 0 ? lm01 : r == 1 ? lm02 : lm23);
#pragma unroll
        for (int o = 0; o < F::kTW; ++o) v[o] = t[o];
    }
#pragma unroll
    for (int o = 0; o < F::kTW; ++o) {
        rec[0][o] = group_bcast<G, 0>(v[o]);
        rec[1][o] = group_bcast<G, 1>(v[o]);
        rec[2][o] = group_bcast<G, 2>(v[o]);
    }
}

// One radix-4 step (two layers) at distance 2^dshift over the butterfly
// groups r < mtrunc of a tile of `rows` consecutive rows whose local row 0 is
// row `base` of the transform (skews and truncation use the global row), as
// in k_step4.
// `need` (one bit per group of 4*dist rows of the whole transform, or null)
// skips the groups no output row depends on.
// UNI: dist >= 64, so the 64 consecutive q of a wave lie in one group and its
// three skews are wave-uniform; taking them from the first lane puts the
// multiply records in scalar registers (scalar loads) instead of 64 per-lane
// vector loads of the same 80 bytes.
// SHARED (G = dist = 4 or 16 lanes per group): records shared by group_records.
template <class F, bool INV, bool UNI, int SHARED = 0>
__device__ __forceinline__ void step4_groups(const Tile &w, int base, int dshift, int groups,
                                             const uint16_t *skew, int bias, const LeoDev &f,
                                             const uint32_t *need) {
    const int dist = 1 << dshift;
    // Rows i + j*dist == i ^ (j*dist): i has zero bits dshift and dshift + 1.
    const int d1 = pos_off(dist), d2 = pos_off(2 * dist), d3 = pos_off(3 * dist);
    for_items<F>(w, groups << dshift, [&](int q, int u) {
        const int r = (q >> dshift) << (dshift + 2);
        if (need) {
            const int g = (base + r) >> (dshift + 2);
            if (!((need[g >> 5] >> (g & 31)) & 1u)) return;
        }
        const int i = r + (q & (dist - 1));
        const int iend = base + r + dist;
        uint8_t *c = Lds<F>::col(w, u);
        const int o0 = pos_off(i), o1 = o0 ^ d1, o2 = o0 ^ d2, o3 = o0 ^ d3;
        typename F::E x0 = Lds<F>::at(w, c, o0), x1 = Lds<F>::at(w, c, o1), x2 = Lds<F>::at(w, c, o2),
                      x3 = Lds<F>::at(w, c, o3);
        if constexpr (UNI) {
            const unsigned lm01 = uniform_skew<F>(skew, iend + bias), lm02 = uniform_skew<F>(skew, iend + dist + bias),
                           lm23 = uniform_skew<F>(skew, iend + 2 * dist + bias);
            uint32_t rec[3][F::kTW];
            uniform_records<F>(f, lm01, lm02, lm23, rec);
            if (INV) {
                ifft2r<F>(x0, x1, lm01, rec[0]);
                ifft2r<F>(x2, x3, lm23, rec[2]);
                ifft2r<F>(x0, x2, lm02, rec[1]);
                ifft2r<F>(x1, x3, lm02, rec[1]);
            } else {
                fft2r<F>(x0, x2, lm02, rec[1]);
                fft2r<F>(x1, x3, lm02, rec[1]);
                fft2r<F>(x0, x1, lm01, rec[0]);
                fft2r<F>(x2, x3, lm23, rec[2]);
            }
        } else if constexpr (SHARED != 0) {
            const unsigned lm01 = skew[iend + bias], lm02 = skew[iend + dist + bias],
                           lm23 = skew[iend + 2 * dist + bias];
            uint32_t rec[3][F::kTW];
            group_records<F, SHARED>(f, lm01, lm02, lm23, rec);
            if (INV) {
                ifft2r<F>(x0, x1, lm01, rec[0]);
                ifft2r<F>(x2, x3, lm23, rec[2]);
                ifft2r<F>(x0, x2, lm02, rec[1]);
                ifft2r<F>(x1, x3, lm02, rec[1]);
            } else {
                fft2r<F>(x0, x2, lm02, rec[1]);
                fft2r<F>(x1, x3, lm02, rec[1]);
                fft2r<F>(x0, x1, lm01, rec[0]);
                fft2r<F>(x2, x3, lm23, rec[2]);
            }
        } else {
            const unsigned lm01 = skew[iend + bias], lm02 = skew[iend + dist + bias],
                           lm23 = skew[iend + 2 * dist + bias];
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
        }
        Lds<F>::put(w, c, o0, x0);
        Lds<F>::put(w, c, o1, x1);
        Lds<F>::put(w, c, o2, x2);
        Lds<F>::put(w, c, o3, x3);
    });
}

// SHR: groups of 4 / 16 lanes share their multiply records (group_records).
// It pays in the encoder (C3 encode 0.38 -> 0.37 ms, (2000,50) 64 KiB 0.31 ->
// 0.26 ms); in the decoders the extra registers cost about as much as it saves.
template <class F, bool INV, bool SHR = false>
__device__ __forceinline__ void step4(const Tile &w, int rows, int base, int dshift, int mtrunc,
                                      const uint16_t *skew, int bias, const LeoDev &f,
                                      const uint32_t *need = nullptr) {
    const int left = mtrunc - base;
    if (left <= 0) return;  // uniform across the block: no barrier is skipped by a subset
    const int groups = min(rows >> (dshift + 2), (left + (4 << dshift) - 1) >> (dshift + 2));
    // A wave spans 64 items (rows fastest) or 64 >> ushift (units fastest).
    const int wave_items_log2 = kUnitsFastest<F> ? max(6 - w.ushift, 0) : 6;
    if (dshift >= wave_items_log2) step4_groups<F, INV, true>(w, base, dshift, groups, skew, bias, f, need);
    // (Units fastest: the lanes of a quad or row may include units past a
    // narrow last slice, which group_records cannot have; their items share
    // records by LDS broadcast instead.)
    else if (SHR && !kUnitsFastest<F> && dshift == 4)
        step4_groups<F, INV, false, 16>(w, base, dshift, groups, skew, bias, f, need);
    else if (SHR && !kUnitsFastest<F> && dshift == 2)
        step4_groups<F, INV, false, 4>(w, base, dshift, groups, skew, bias, f, need);
    else step4_groups<F, INV, false>(w, base, dshift, groups, skew, bias, f, need);
    __syncthreads();
}

// IFFT over `n` rows truncated to `mtrunc`; `skew` + `bias` as in leo_step4,
// the final radix-2 layer (odd log2 n) uses skew[dist + bias].
template <class F, bool SHR = false>
__device__ __forceinline__ void ifft(const Tile &w, int n, int mtrunc, const uint16_t *skew, int bias,
                                     const LeoDev &f) {
    int dshift = 0;
    while ((4 << dshift) <= n) {
        step4<F, true, SHR>(w, n, 0, dshift, mtrunc, skew, bias, f);
        dshift += 2;
    }
    const int dist = 1 << dshift;
    if (dist < n) {
        const unsigned lm = skew[dist + bias];
        for_items<F>(w, dist, [&](int i, int u) {
            typename F::E x = Lds<F>::load(w, i, u), y = Lds<F>::load(w, i + dist, u);
            ifft2<F>(x, y, lm, f);
            Lds<F>::store(w, i, u, x);
            Lds<F>::store(w, i + dist, u, y);
        });
        __syncthreads();
    }
}

// Words of one level of leo_fft_need bits with `groups` groups.
__device__ __forceinline__ int need_words(int groups) { return (groups + 31) >> 5; }

// FFT over `n` rows truncated to `mtrunc` (skew = fftSkew, bias -1); `need`
// (leo_fft_need bits, or null) prunes groups without output rows.
template <class F, bool SHR = false>
__device__ __forceinline__ void fft(const Tile &w, int n, int mtrunc, const LeoDev &f,
                                    const uint32_t *need = nullptr) {
    int dist4 = n, dist = n >> 2;
    while (dist != 0) {
        step4<F, false, SHR>(w, n, 0, log2i(dist), mtrunc, f.skew, -1, f, need);
        if (need) need += need_words(n / (4 * dist));
        dist4 = dist;
        dist >>= 2;
    }
    if (dist4 == 2) {
        for_items<F>(w, (mtrunc + 1) / 2, [&](int q, int u) {
            if (need && !((need[q >> 5] >> (q & 31)) & 1u)) return;
            const int r = 2 * q;
            typename F::E x = Lds<F>::load(w, r, u), y = Lds<F>::load(w, r + 1, u);
            fft2<F>(x, y, f.skew[r], f);
            Lds<F>::store(w, r, u, x);
            Lds<F>::store(w, r + 1, u, y);
        });
        __syncthreads();
    }
}

// rows [0, nrows) of w <- src[t] (+col) times exp(logs[t]) when logs, zero
// where src[t] is null.  Each thread first loads the row pointers (and, for
// a reconstruct, the logs) of its kBatch items, from indices clamped into
// range so those loads carry no branch, then issues the kBatch data loads,
// and only then stores any: a data load never waits for an earlier item's
// data (item j's pointer check waits for pointer j alone).  LOGS is a
// template flag so the log loads are unconditional too.
template <class F, bool LOGS>
__device__ __forceinline__ void load_rows_impl(const Tile &w, const uint8_t *const *src, const uint16_t *logs,
                                               int nrows, size_t col, const LeoDev &f) {
    constexpr int kBatch = 32 / F::kUnitBytes * 4;  // 8 GF(2^8) / 4 GF(2^16) units in flight
    const int items = nrows << w.ushift;
    const int umask = (1 << w.ushift) - 1;
    for (int it0 = threadIdx.x; it0 < items; it0 += kBatch * blockDim.x) {
        const uint8_t *sp[kBatch];
        unsigned lg[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int t = min((it0 + j * static_cast<int>(blockDim.x)) >> w.ushift, nrows - 1);
            sp[j] = src[t];
            if constexpr (LOGS) lg[j] = logs[t];
        }
        typename F::E v[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int it = it0 + j * blockDim.x;
            v[j] = F::zero();
            if (it < items && (it & umask) < w.ucount && sp[j]) v[j] = F::gload(sp[j] + col, (it & umask) + w.ubase);
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int it = it0 + j * blockDim.x;
            if (it < items && (it & umask) < w.ucount) {
                if constexpr (LOGS) {
                    typename F::E y = F::zero();
                    F::mul_add(y, v[j], table<F>(f, lg[j]));
                    v[j] = y;
                }
                Lds<F>::store(w, it >> w.ushift, it & umask, v[j]);
            }
        }
    }
    __syncthreads();
}

template <class F>
__device__ __forceinline__ void load_rows(const Tile &w, const uint8_t *const *src, const uint16_t *logs,
                                          int nrows, size_t col, const LeoDev &f) {
    if (logs) load_rows_impl<F, true>(w, src, logs, nrows, col, f);
    else load_rows_impl<F, false>(w, src, nullptr, nrows, col, f);
}

// A row load split in two halves so that other work overlaps it: issue()
// starts the data loads of the thread's items (at most N per thread: the
// caller checks fits()), put() stores them into the tile once they arrive.
// No logs; rows with a null pointer read as zero.
template <class F, int N>
struct RowPrefetch {
    typename F::E v[N];

    __device__ static bool fits(const Tile &w, int nrows) {
        return (nrows << w.ushift) <= N * static_cast<int>(blockDim.x);
    }
    __device__ void issue(const Tile &w, const uint8_t *const *src, int nrows, size_t col) {
        const int items = nrows << w.ushift;
        const int umask = (1 << w.ushift) - 1;
        const uint8_t *sp[N];
#pragma unroll
        for (int j = 0; j < N; ++j)
            sp[j] = src[min((static_cast<int>(threadIdx.x) + j * static_cast<int>(blockDim.x)) >> w.ushift, nrows - 1)];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const int it = threadIdx.x + j * blockDim.x;
            v[j] = F::zero();
            if (it < items && (it & umask) < w.ucount && sp[j]) v[j] = F::gload(sp[j] + col, (it & umask) + w.ubase);
        }
    }
    __device__ void put(const Tile &w, int nrows) const {
        const int items = nrows << w.ushift;
        const int umask = (1 << w.ushift) - 1;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const int it = threadIdx.x + j * blockDim.x;
            if (it < items && (it & umask) < w.ucount) Lds<F>::store(w, it >> w.ushift, it & umask, v[j]);
        }
    }
};

// dst[q] (+col) <- row rows[q] times exp(logs[q]) (MAPPED: a reconstruct's
// erased rows) or row q (an encode's parity).  As in load_rows, each thread
// loads the metadata of its kBatch items first (indices clamped into range),
// then reads, multiplies and stores, so no item waits for another's loads.
template <class F, bool MAPPED>
__device__ __forceinline__ void store_rows_impl(const Tile &w, uint8_t *const *dst, const int *rows,
                                                const uint16_t *logs, int count, size_t col, const LeoDev &f) {
    constexpr int kBatch = 4;
    const int items = count << w.ushift;
    const int umask = (1 << w.ushift) - 1;
    for (int it0 = threadIdx.x; it0 < items; it0 += kBatch * blockDim.x) {
        uint8_t *dp[kBatch];
        int rw[kBatch];
        unsigned lg[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int q = min((it0 + j * static_cast<int>(blockDim.x)) >> w.ushift, count - 1);
            dp[j] = dst[q];
            if constexpr (MAPPED) {
                rw[j] = rows[q];
                lg[j] = logs[q];
            }
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int it = it0 + j * blockDim.x;
            if (it < items && (it & umask) < w.ucount) {
                const int u = it & umask;
                const typename F::E x = Lds<F>::load(w, MAPPED ? rw[j] : it >> w.ushift, u);
                typename F::E v = x;
                if constexpr (MAPPED) {
                    v = F::zero();
                    F::mul_add(v, x, table<F>(f, lg[j]));
                }
                F::gstore(dp[j] + col, u + w.ubase, v);
            }
        }
    }
}

template <class F>
__device__ __forceinline__ void store_rows(const Tile &w, uint8_t *const *dst, const int *rows,
                                           const uint16_t *logs, int count, size_t col, const LeoDev &f) {
    if (rows) store_rows_impl<F, true>(w, dst, rows, logs, count, col, f);
    else store_rows_impl<F, false>(w, dst, nullptr, nullptr, count, col, f);
}

// Verify: compares rows [0, count) of w with dst[q] (+col); a thread that
// finds a difference stores `value` to *flag (every writer stores the same).
template <class F>
__device__ __forceinline__ void compare_rows(const Tile &w, uint8_t *const *dst, int count, size_t col,
                                             uint32_t *flag, uint32_t value) {
    uint32_t diff = 0;
    for_units(w, count, [&](int q, int u) {
        const typename F::E x = Lds<F>::load(w, q, u);
        const typename F::E y = F::gload(dst[q] + col, u + w.ubase);
        diff |= F::differ(x, y);
    });
    if (diff) *flag = value;
}

// Formal derivative in place (leopard.go:619-623): out[t] = in[t] ^ XOR over
// zero bits b of t (t | 2^b < n) of in[t | 2^b].  Every thread computes its
// slots into registers, then all store after a barrier.  The bit loop is
// outermost and branch-free: for each bit, every item reads row t ^ 2^b
// (in range either way) and keeps it under a mask when bit b of t is zero,
// so one bit's reads of all items issue together instead of one divergent,
// latency-bound read at a time.
template <int MAXD = kDerivMax>
__device__ __forceinline__ void derivative(const Tile &w, int n, int slots) {
    const int ns = log2i(n);
    const int items = n * slots;
    uint4 res[MAXD];
    uint8_t *col[MAXD];  // the item's slot column
    int po[MAXD];        // pos_off of its row
#pragma unroll
    for (int j = 0; j < MAXD; ++j) {
        const int it = threadIdx.x + j * blockDim.x;
        const int ic = it < items ? it : 0;  // spare items read item 0 (in range)
        col[j] = column(w, ic >> ns);
        po[j] = pos_off(ic & (n - 1));
        res[j] = *cell(col[j], po[j]);
    }
    for (int s = 0; s < ns; ++s) {
        const int ob = pos_off(1 << s);
#pragma unroll
        for (int j = 0; j < MAXD; ++j) {
            const int t = (threadIdx.x + j * blockDim.x) & (n - 1);
            const uint32_t keep = ((t >> s) & 1) ? 0u : ~0u;
            const uint4 v = *cell(col[j], po[j] ^ ob);  // row t ^ 2^s of the same slot
            res[j].x ^= v.x & keep;
            res[j].y ^= v.y & keep;
            res[j].z ^= v.z & keep;
            res[j].w ^= v.w & keep;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < MAXD; ++j) {
        const int it = threadIdx.x + j * blockDim.x;
        if (it < items) *slot(w, it & (n - 1), it >> ns) = res[j];
    }
    __syncthreads();
}

}  // namespace leo_tile
}  // namespace rsx

// leo_multipass.hip — Leopard GF(2^16) reconstruct and encode for transforms
// taller than one LDS tile (reconstruct n >= 8192 rows, encode m >= 8192 or
// small shards; the passes hold from 2048 rows).  The encoder's passes are
// described at enc_pass1; the reconstruct's here.
//
// The n-row transform (leopard.go:421-647: scale by error locators, IFFT,
// formal derivative, FFT, reveal) runs as three passes over 64-byte column
// slices, each holding a tile of rows in LDS.  Let P be the largest FFT step
// distance <= 1024 and C = 1024.
//
//   pass 1  tiles of C consecutive rows: load the shards scaled by the error
//           locators, IFFT layers with distance < C.
//   pass 2  tiles of the rows lo + j*P (j < n/P) for a run of lo: the IFFT
//           layers >= C, then the FFT steps with distance >= P, applied to
//           both in and y = in + H(in), where H is the part of the formal
//           derivative over row bits >= P.  Writes u = FFT_B(y), v = FFT_B(in).
//   pass 3  tiles of P consecutive rows: z = u + L(v), where L is the part of
//           the derivative over row bits < P, then the FFT steps with
//           distance < P, and the erased rows are scaled back and stored.
//
// The derivative out[t] = in[t] ^ XOR over zero bits b of t of in[t | b]
// splits as D = I + L + H.  FFT_B (distances >= P) mixes rows that differ in
// bits >= P with skews that depend on those bits only, so it commutes with L:
// FFT_B(D in) = FFT_B(in + H in) + L(FFT_B in) = u + L(v).
//
// Work arrays are slice-major (row r of slice s at (s*n + r)*64) so every
// tile reads and writes contiguous runs.
#include <algorithm>

#include "leo_tile.hpp"

namespace rsx {

namespace {

using namespace leo;
using namespace leo_tile;

constexpr int kMB = 256;     // threads per workgroup of pass 3 (z = u + L(v) in registers)
// Passes 1 and 2 hold 64 KiB tiles (two workgroups per CU): 512 threads
// keep 16 waves per CU busy (5% faster than 256 at n = 8192).
constexpr int kMB1 = 512;    // threads per workgroup of pass 1
constexpr int kMB2 = 512;    // threads per workgroup of pass 2
constexpr int kSlice = kLeoPassSlice;

__device__ __forceinline__ uint8_t *wrow(uint8_t *work, int n, int s, int r) {
    return work + (static_cast<size_t>(s) * n + r) * kSlice;
}

// Copies rows [0, nrows) of a tile from global rows rp(l) (64 bytes each; a
// null row reads as zero), several loads in flight per thread.
template <class F, class RowPtr>
__device__ __forceinline__ void load_tile(const Tile &w, int nrows, RowPtr rp) {
    constexpr int kBatch = 4;
    const int items = nrows << w.ushift;
    const int umask = (1 << w.ushift) - 1;
    for (int it0 = threadIdx.x; it0 < items; it0 += kBatch * blockDim.x) {
        typename F::E v[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int it = it0 + j * blockDim.x;
            if (it < items) {
                const uint8_t *row = rp(it >> w.ushift);
                v[j] = row ? F::gload(row, it & umask) : F::zero();
            }
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int it = it0 + j * blockDim.x;
            if (it < items) Lds<F>::store(w, it >> w.ushift, it & umask, v[j]);
        }
    }
}

// Stores rows [0, nrows) of a tile to global rows rp(l); null rows are skipped.
template <class F, class RowPtr>
__device__ __forceinline__ void store_tile(const Tile &w, int nrows, RowPtr rp) {
    for_units(w, nrows, [&](int l, int u) {
        uint8_t *row = rp(l);
        if (row) F::gstore(row, u, Lds<F>::load(w, l, u));
    });
}

// One radix-4 step over tiles whose local row l = j*A + a is global row
// lo0 + a + j*P: butterflies at global distance `dist` (>= P) connect local
// rows j, j + dj, j + 2dj, j + 3dj with dj = dist / P.  `w2`, when given, gets
// the same butterflies (same skews).  Lanes take a (the lo run) fastest, then
// the unit, then the quad q; UNI: a wave's 64 lanes share one q (A << ushift
// >= 64), so the skews are wave-uniform and the multiply records are scalar
// loads.
template <class F, bool INV, bool UNI>
__device__ __forceinline__ void strided_quads(const Tile &w, const Tile *w2, int ashift, int P, int dist,
                                              int djs, int nq, const uint16_t *sk, const LeoDev &f) {
    const int A = 1 << ashift, dj = 1 << djs;
    const int items = (A << log2i(nq)) << w.ushift;
    // Rows l0 + k*ld == l0 ^ (k*ld): l0 has zero bits at ld and 2*ld.
    const int ld = dj << ashift;
    const int d1 = pos_off(ld), d2 = pos_off(2 * ld), d3 = pos_off(3 * ld);
    const int umask = (1 << w.ushift) - 1;
    // With w2 the two tiles are separate items (the second tile's after the
    // first's): a step over both keeps twice the lanes busy.  items is a
    // multiple of 64 where UNI holds, so a wave stays within one tile.
    for (int it2 = threadIdx.x; it2 < (w2 ? 2 * items : items); it2 += blockDim.x) {
        const int t = it2 >= items;
        const int it = it2 - t * items;
        const int a = it & (A - 1);
        const int u = (it >> ashift) & umask;
        const int q = it >> (ashift + w.ushift);
        if (q >= nq || u >= w.ucount) continue;
        const int rj = (q >> djs) << (djs + 2);
        const int j = rj + (q & (dj - 1));
        const int iend = rj * P + dist;  // global r + dist
        const int l0 = (j << ashift) + a;
        const Tile &x = t ? *w2 : w;
        uint8_t *c = Lds<F>::col(x, u);
        const int o0 = pos_off(l0), o1 = o0 ^ d1, o2 = o0 ^ d2, o3 = o0 ^ d3;
        typename F::E x0 = Lds<F>::at(x, c, o0), x1 = Lds<F>::at(x, c, o1), x2 = Lds<F>::at(x, c, o2),
                      x3 = Lds<F>::at(x, c, o3);
        if constexpr (UNI) {
            const unsigned lm01 = uniform_u16(sk, iend), lm02 = uniform_u16(sk, iend + dist),
                           lm23 = uniform_u16(sk, iend + 2 * dist);
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
        } else {
            const unsigned lm01 = sk[iend], lm02 = sk[iend + dist], lm23 = sk[iend + 2 * dist];
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
        Lds<F>::put(x, c, o0, x0);
        Lds<F>::put(x, c, o1, x1);
        Lds<F>::put(x, c, o2, x2);
        Lds<F>::put(x, c, o3, x3);
    }
}

// sk: skew table with its bias applied (the butterfly group ending at global
// row r + dist reads sk[r + dist]): fftSkew - 1 for the decoder's IFFT and
// every FFT, fftSkew + m - 1 + g*m for encoder group g's IFFT.
template <class F, bool INV>
__device__ __forceinline__ void strided_step4(const Tile &w, const Tile *w2, int ashift, int J, int P, int dist,
                                              int mtrunc, const uint16_t *sk, const LeoDev &f) {
    const int dj = dist / P, djs = log2i(dj);  // dist and P are powers of two
    const int gcount = min(J / (4 * dj), (mtrunc + 4 * dist - 1) / (4 * dist));
    if (gcount <= 0) return;
    const int nq = gcount * dj;  // quads per lo
    if (ashift + w.ushift >= 6) strided_quads<F, INV, true>(w, w2, ashift, P, dist, djs, nq, sk, f);
    else strided_quads<F, INV, false>(w, w2, ashift, P, dist, djs, nq, sk, f);
    __syncthreads();
}

// Pass bodies: workgroup (s, by) = 64-byte column slice s, row block by.
template <class F>
__device__ __forceinline__ void pass1(const LeoDev &f, const LeoMultipassArgs &a, int s, int by) {
    extern __shared__ uint4 lds_raw[];
    const int base = by * a.C;
    // Rows from inputCount on are zero (leopard.go:488-495) and the IFFT
    // layers of this pass stay within the block: a block past it stays zero
    // and pass 2 reads it as zero without loading it.
    if (base >= a.ifft_trunc) return;
    const Tile w{reinterpret_cast<uint8_t *>(lds_raw), a.C + 1, 1, 2};
    load_rows<F>(w, a.rows + base, a.errlocs + base, a.C, a.col0 + static_cast<size_t>(s) * kSlice, f);
    for (int dshift = 0; (4 << dshift) <= a.C; dshift += 2)
        step4<F, true>(w, a.C, base, dshift, a.ifft_trunc, f.skew, -1, f);
    store_tile<F>(w, a.C, [&](int l) { return wrow(a.work, a.n, s, base + l); });
}

template <class F>
__device__ __forceinline__ void pass2(const LeoDev &f, const LeoMultipassArgs &a, int s, int by) {
    extern __shared__ uint4 lds_raw[];
    const int lo0 = by * a.A;
    const int J = a.n / a.P;
    const int ashift = log2i(a.A), jshift = log2i(J);
    const int R = a.A * J;
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    const Tile in{lds, R + 1, 1, 2};
    const Tile y{lds + kSlice * (R + 1), R + 1, 1, 2};
    auto grow = [&](int l) { return lo0 + (l & (a.A - 1)) + (l >> ashift) * a.P; };
    // Pass 1 left the blocks of C rows from ceil(inputCount / C) * C zero
    // without writing them.
    const int zero_from = (a.ifft_trunc + a.C - 1) / a.C * a.C;
    load_tile<F>(in, R, [&](int l) -> const uint8_t * {
        const int r = grow(l);
        return r < zero_from ? wrow(a.work, a.n, s, r) : nullptr;
    });
    __syncthreads();
    // IFFT layers with distance >= C.
    int dist = a.C;
    for (; 4 * dist <= a.n; dist *= 4)
        strided_step4<F, true>(in, nullptr, ashift, J, a.P, dist, a.ifft_trunc, f.skew - 1, f);
    if (dist < a.n) {
        // final radix-2 layer: rows j, j + dj for j < dj, skew[dist - 1]
        const int dj = dist / a.P;
        const unsigned lm = f.skew[dist - 1];
        const int items = (R / 2) << in.ushift;
        for (int it = threadIdx.x; it < items; it += blockDim.x) {
            const int u = it & 1, p = it >> 1;  // p < A * dj
            const int l0 = p, l1 = p + (dj << ashift);
            typename F::E x0 = Lds<F>::load(in, l0, u), x1 = Lds<F>::load(in, l1, u);
            ifft2<F>(x0, x1, lm, f);
            Lds<F>::store(in, l0, u, x0);
            Lds<F>::store(in, l1, u, x1);
        }
        __syncthreads();
    }
    // y = in + H(in): H over the row bits >= P, i.e. all bits of j.
    {
        const int slots = kSlice / 16;
        const int items = R * slots;
        const int rshift = ashift + jshift;
        for (int it = threadIdx.x; it < items; it += blockDim.x) {
            const int l = it & (R - 1), sl = it >> rshift;
            const int j = l >> ashift;
            uint8_t *c = column(in, sl);
            const int o = pos_off(l);
            uint4 acc = *cell(c, o);
            for (int b = 0; (1 << b) < J; ++b) {  // branch-free: row l ^ (2^b << ashift), kept where bit b of j is 0
                const uint32_t keep = ((j >> b) & 1) ? 0u : ~0u;
                const uint4 v = *cell(c, o ^ pos_off((1 << b) << ashift));
                acc.x ^= v.x & keep;
                acc.y ^= v.y & keep;
                acc.z ^= v.z & keep;
                acc.w ^= v.w & keep;
            }
            *cell(column(y, sl), o) = acc;
        }
        __syncthreads();
    }
    // FFT steps with distance >= P on both tiles.
    for (int d = a.n >> 2; d >= a.P; d >>= 2)
        strided_step4<F, false>(in, &y, ashift, J, a.P, d, a.fft_trunc, f.skew - 1, f);
    // Only the row blocks pass 3 runs are read again (row block = l >> ashift).
    auto kept = [&](int l) { return ((a.pass3_blocks >> (l >> ashift)) & 1u) != 0; };
    store_tile<F>(in, R, [&](int l) { return kept(l) ? wrow(a.work2, a.n, s, grow(l)) : nullptr; });  // v
    store_tile<F>(y, R, [&](int l) { return kept(l) ? wrow(a.work, a.n, s, grow(l)) : nullptr; });    // u
}

template <class F, int MAXD>
__device__ __forceinline__ void pass3(const LeoDev &f, const LeoMultipassArgs &a, int s, int by) {
    extern __shared__ uint4 lds_raw[];
    // A row block without an output row writes nothing (uniform per workgroup).
    if (!((a.pass3_blocks >> by) & 1u)) return;
    const int base = by * a.P;
    const Tile w{reinterpret_cast<uint8_t *>(lds_raw), a.P + 1, 1, 2};
    load_tile<F>(w, a.P, [&](int l) -> const uint8_t * { return wrow(a.work2, a.n, s, base + l); });  // v
    __syncthreads();
    // z = u + L(v): L over the row bits < P.  Registers first, then LDS.  All
    // of a thread's u loads issue before any is used, then the bit loop runs
    // outermost and branch-free (row t ^ 2^b is read for every item and kept
    // where bit b of t is zero), as in the tile derivative.
    {
        const int slots = kSlice / 16;
        const int ps = log2i(a.P);
        const int items = a.P * slots;
        uint4 res[MAXD];
        uint8_t *cc[MAXD];
        int po[MAXD];
#pragma unroll
        for (int jj = 0; jj < MAXD; ++jj) {
            const int it = threadIdx.x + jj * blockDim.x;
            const int ic = it < items ? it : 0;  // spare items read item 0 (in range)
            const int t = ic & (a.P - 1), sl = ic >> ps;
            res[jj] = gload4(wrow(a.work, a.n, s, base + t) + sl * 16);
            cc[jj] = column(w, sl);
            po[jj] = pos_off(t);
        }
        for (int b = 0; (1 << b) < a.P; ++b) {
            const int ob = pos_off(1 << b);
#pragma unroll
            for (int jj = 0; jj < MAXD; ++jj) {
                const int t = (threadIdx.x + jj * blockDim.x) & (a.P - 1);
                const uint32_t keep = ((t >> b) & 1) ? 0u : ~0u;
                const uint4 v = *cell(cc[jj], po[jj] ^ ob);  // row t ^ 2^b
                res[jj].x ^= v.x & keep;
                res[jj].y ^= v.y & keep;
                res[jj].z ^= v.z & keep;
                res[jj].w ^= v.w & keep;
            }
        }
        __syncthreads();
#pragma unroll
        for (int jj = 0; jj < MAXD; ++jj) {
            const int it = threadIdx.x + jj * blockDim.x;
            if (it < items) *slot(w, it & (a.P - 1), it >> ps) = res[jj];
        }
        __syncthreads();
    }
    // FFT steps with distance < P, then the final radix-2 layer; `need` levels
    // are walked from the first FFT step so their offsets match.
    int dist4 = a.n, dist = a.n >> 2;
    const uint32_t *need = a.need;
    for (; dist != 0; dist4 = dist, dist >>= 2) {
        if (dist < a.P) step4<F, false>(w, a.P, base, log2i(dist), a.fft_trunc, f.skew, -1, f, need);
        if (need) need += need_words(a.n / (4 * dist));
    }
    if (dist4 == 2 && a.fft_trunc > base) {
        const int pairs = min(a.P / 2, (a.fft_trunc - base + 1) / 2);
        for_rows(w, pairs, [&](int q, int u) {
            const int pg = (base >> 1) + q;
            if (need && !((need[pg >> 5] >> (pg & 31)) & 1u)) return;
            const int r = 2 * q;
            typename F::E x = Lds<F>::load(w, r, u), yv = Lds<F>::load(w, r + 1, u);
            fft2<F>(x, yv, f.skew[base + r], f);
            Lds<F>::store(w, r, u, x);
            Lds<F>::store(w, r + 1, u, yv);
        });
        __syncthreads();
    }
    // Reveal the erased rows of this tile.
    const size_t col = a.col0 + static_cast<size_t>(s) * kSlice;
    for_units(w, a.P, [&](int t, int u) {
        const int q = a.out_index[base + t];
        if (q < 0) return;
        typename F::E v = F::zero();
        F::mul_add(v, Lds<F>::load(w, t, u), table<F>(f, a.out_logs[q]));
        F::gstore(a.out[q] + col, u, v);
    });
}

// ---- Encoder (leopard.go:169-260) for m taller than one LDS tile ----
// parity = FFT_m(XOR over groups g of IFFT_m^g(data rows g*m ...)), the same
// three-way split: pass 1 tiles C consecutive rows of group g (IFFT layers <
// C) into work; pass 2 tiles rows lo + j*P (IFFT layers >= C), XORs into the
// accumulator work2 and, after the last group, runs the FFT steps >= P on
// it; pass 3 tiles P consecutive accumulator rows (FFT steps < P) and stores
// rows < p to the parity shards.
template <class F>
__device__ __forceinline__ void enc_pass1(const LeoDev &f, const LeoEncPassArgs &a, int s, int by) {
    extern __shared__ uint4 lds_raw[];
    const int base = by * a.C;
    const Tile w{reinterpret_cast<uint8_t *>(lds_raw), a.C + 1, 1, 2};
    load_rows<F>(w, a.rows + base, nullptr, a.C, a.col0 + static_cast<size_t>(s) * kSlice, f);
    for (int dshift = 0; (4 << dshift) <= a.C; dshift += 2)
        step4<F, true, true>(w, a.C, base, dshift, a.mtrunc, f.skew + a.skew_off, 0, f);
    store_tile<F>(w, a.C, [&](int l) { return wrow(a.work, a.m, s, base + l); });
}

template <class F>
__device__ __forceinline__ void enc_pass2(const LeoDev &f, const LeoEncPassArgs &a, int s, int by) {
    extern __shared__ uint4 lds_raw[];
    const int lo0 = by * a.A;
    const int J = a.m / a.P;
    const int ashift = log2i(a.A);
    const int R = a.A * J;
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    const Tile in{lds, R + 1, 1, 2};
    const Tile acc{lds + kSlice * (R + 1), R + 1, 1, 2};
    auto grow = [&](int l) { return lo0 + (l & (a.A - 1)) + (l >> ashift) * a.P; };
    load_tile<F>(in, R, [&](int l) { return wrow(a.work, a.m, s, grow(l)); });
    if (!a.first) load_tile<F>(acc, R, [&](int l) { return wrow(a.work2, a.m, s, grow(l)); });
    __syncthreads();
    const uint16_t *isk = f.skew + a.skew_off;
    int dist = a.C;
    for (; 4 * dist <= a.m; dist *= 4) strided_step4<F, true>(in, nullptr, ashift, J, a.P, dist, a.mtrunc, isk, f);
    if (dist < a.m) {
        // final radix-2 layer: rows j, j + dj for j < dj, skew isk[dist]
        const int dj = dist / a.P;
        const unsigned lm = isk[dist];
        const int items = (R / 2) << in.ushift;
        for (int it = threadIdx.x; it < items; it += blockDim.x) {
            const int u = it & 1, p = it >> 1;
            const int l0 = p, l1 = p + (dj << ashift);
            typename F::E x0 = Lds<F>::load(in, l0, u), x1 = Lds<F>::load(in, l1, u);
            ifft2<F>(x0, x1, lm, f);
            Lds<F>::store(in, l0, u, x0);
            Lds<F>::store(in, l1, u, x1);
        }
        __syncthreads();
    }
    if (!a.first) {
        const int rshift = log2i(R);  // R = A * J, a power of two
        for (int it = threadIdx.x; it < R * (kSlice / 16); it += blockDim.x) {
            const int l = it & (R - 1), sl = it >> rshift;
            *slot(in, l, sl) = xor4(*slot(in, l, sl), *slot(acc, l, sl));
        }
        __syncthreads();
    }
    if (a.last)
        for (int d = a.m >> 2; d >= a.P; d >>= 2)
            strided_step4<F, false>(in, nullptr, ashift, J, a.P, d, a.p, f.skew - 1, f);
    store_tile<F>(in, R, [&](int l) { return wrow(a.work2, a.m, s, grow(l)); });
}

template <class F>
__device__ __forceinline__ void enc_pass3(const LeoDev &f, const LeoEncPassArgs &a, int s, int by) {
    extern __shared__ uint4 lds_raw[];
    const int base = by * a.P;
    if (base >= a.p) return;  // no parity row in this tile (uniform per workgroup)
    const Tile w{reinterpret_cast<uint8_t *>(lds_raw), a.P + 1, 1, 2};
    load_tile<F>(w, a.P, [&](int l) { return wrow(a.work2, a.m, s, base + l); });
    __syncthreads();
    int dist4 = a.m, dist = a.m >> 2;
    for (; dist != 0; dist4 = dist, dist >>= 2)
        if (dist < a.P) step4<F, false, true>(w, a.P, base, log2i(dist), a.p, f.skew, -1, f);
    if (dist4 == 2) {
        const int pairs = min(a.P / 2, (a.p - base + 1) / 2);
        for_rows(w, pairs, [&](int q, int u) {
            const int r = 2 * q;
            typename F::E x = Lds<F>::load(w, r, u), yv = Lds<F>::load(w, r + 1, u);
            fft2<F>(x, yv, f.skew[base + r], f);
            Lds<F>::store(w, r, u, x);
            Lds<F>::store(w, r + 1, u, yv);
        });
        __syncthreads();
    }
    const int rows = min(a.P, a.p - base);
    const size_t col = a.col0 + static_cast<size_t>(s) * kSlice;
    if (a.flag) {
        compare_rows<F>(w, a.parity + base, rows, col, a.flag, a.flag_value);  // Verify
        return;
    }
    for_units(w, rows, [&](int t, int u) { F::gstore(a.parity[base + t] + col, u, Lds<F>::load(w, t, u)); });
}

template <class F>
__global__ __launch_bounds__(kMB1) void k_enc_pass1(const LeoDev f, const LeoEncPassArgs a) {
    enc_pass1<F>(f, a, blockIdx.x, blockIdx.y);
}
template <class F>
__global__ __launch_bounds__(kMB2) void k_enc_pass2(const LeoDev f, const LeoEncPassArgs a) {
    enc_pass2<F>(f, a, blockIdx.x, blockIdx.y);
}
template <class F>
__global__ __launch_bounds__(kMB) void k_enc_pass3(const LeoDev f, const LeoEncPassArgs a) {
    enc_pass3<F>(f, a, blockIdx.x, blockIdx.y);
}

template <class F>
__global__ __launch_bounds__(kMB1) void k_dec_pass1(const LeoDev f, const LeoMultipassArgs a) {
    pass1<F>(f, a, blockIdx.x, blockIdx.y);
}
template <class F>
__global__ __launch_bounds__(kMB2) void k_dec_pass2(const LeoDev f, const LeoMultipassArgs a) {
    pass2<F>(f, a, blockIdx.x, blockIdx.y);
}
template <class F, int MAXD>
__global__ __launch_bounds__(kMB) void k_dec_pass3(const LeoDev f, const LeoMultipassArgs a) {
    pass3<F, MAXD>(f, a, blockIdx.x, blockIdx.y);
}

// Batched: one 1-D grid per pass over the workgroups of every task; a block
// finds its task by binary search over the tasks' first blocks of this pass,
// then its (slice, row block) with the slice fastest, as in the 2-D grids.
template <int PASS>
__device__ __forceinline__ int pass_task(const LeoPassTask *tasks, int ntasks, uint32_t g) {
    int lo = 0, hi = ntasks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tasks[mid].block0[PASS] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
template <class F>
__global__ __launch_bounds__(kMB1) void k_dec_pass1_batch(const LeoDev f, const LeoPassTask *tasks, int ntasks) {
    const int t = pass_task<0>(tasks, ntasks, blockIdx.x);
    const LeoMultipassArgs a = tasks[t].a;
    const uint32_t l = blockIdx.x - tasks[t].block0[0];
    pass1<F>(f, a, static_cast<int>(l % a.slices), static_cast<int>(l / a.slices));
}
template <class F>
__global__ __launch_bounds__(kMB2) void k_dec_pass2_batch(const LeoDev f, const LeoPassTask *tasks, int ntasks) {
    const int t = pass_task<1>(tasks, ntasks, blockIdx.x);
    const LeoMultipassArgs a = tasks[t].a;
    const uint32_t l = blockIdx.x - tasks[t].block0[1];
    pass2<F>(f, a, static_cast<int>(l % a.slices), static_cast<int>(l / a.slices));
}
template <class F, int MAXD>
__global__ __launch_bounds__(kMB) void k_dec_pass3_batch(const LeoDev f, const LeoPassTask *tasks, int ntasks) {
    const int t = pass_task<2>(tasks, ntasks, blockIdx.x);
    const LeoMultipassArgs a = tasks[t].a;
    const uint32_t l = blockIdx.x - tasks[t].block0[2];
    pass3<F, MAXD>(f, a, static_cast<int>(l % a.slices), static_cast<int>(l / a.slices));
}

// Batched encoder passes: one 1-D grid over the workgroups of every task
// of the launch (tasks sorted by block0).
__device__ __forceinline__ int enc_task(const LeoEncPassTask *tasks, int ntasks, uint32_t g) {
    int lo = 0, hi = ntasks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tasks[mid].block0 <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
template <class F>
__global__ __launch_bounds__(kMB1) void k_enc_pass1_batch(const LeoDev f, const LeoEncPassTask *tasks, int ntasks) {
    const int t = enc_task(tasks, ntasks, blockIdx.x);
    const LeoEncPassArgs a = tasks[t].a;
    const uint32_t l = blockIdx.x - tasks[t].block0;
    enc_pass1<F>(f, a, static_cast<int>(l % a.slices), static_cast<int>(l / a.slices));
}
template <class F>
__global__ __launch_bounds__(kMB2) void k_enc_pass2_batch(const LeoDev f, const LeoEncPassTask *tasks, int ntasks) {
    const int t = enc_task(tasks, ntasks, blockIdx.x);
    const LeoEncPassArgs a = tasks[t].a;
    const uint32_t l = blockIdx.x - tasks[t].block0;
    enc_pass2<F>(f, a, static_cast<int>(l % a.slices), static_cast<int>(l / a.slices));
}
template <class F>
__global__ __launch_bounds__(kMB) void k_enc_pass3_batch(const LeoDev f, const LeoEncPassTask *tasks, int ntasks) {
    const int t = enc_task(tasks, ntasks, blockIdx.x);
    const LeoEncPassArgs a = tasks[t].a;
    const uint32_t l = blockIdx.x - tasks[t].block0;
    enc_pass3<F>(f, a, static_cast<int>(l % a.slices), static_cast<int>(l / a.slices));
}

}  // namespace

hipError_t leo_encode_multipass_batch(const LeoDev &f, const LeoEncPassTask *tasks, int ntasks, uint32_t blocks,
                                      int pass, int max_p, hipStream_t s) {
    if (f.bits != 16 || ntasks <= 0 || blocks == 0 || pass < 1 || pass > 3) return hipErrorInvalidValue;
    const size_t lds = pass == 1 ? static_cast<size_t>(kSlice) * (1024 + 1)
                       : pass == 2 ? 2 * static_cast<size_t>(kSlice) * (512 + 1)
                                   : static_cast<size_t>(kSlice) * (max_p + 1);
    const void *k = pass == 1 ? reinterpret_cast<const void *>(k_enc_pass1_batch<F16>)
                    : pass == 2 ? reinterpret_cast<const void *>(k_enc_pass2_batch<F16>)
                                : reinterpret_cast<const void *>(k_enc_pass3_batch<F16>);
    if (lds > (64u << 10)) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    if (pass == 1) hipLaunchKernelGGL(k_enc_pass1_batch<F16>, dim3(blocks), dim3(kMB1), lds, s, f, tasks, ntasks);
    else if (pass == 2) hipLaunchKernelGGL(k_enc_pass2_batch<F16>, dim3(blocks), dim3(kMB2), lds, s, f, tasks, ntasks);
    else hipLaunchKernelGGL(k_enc_pass3_batch<F16>, dim3(blocks), dim3(kMB), lds, s, f, tasks, ntasks);
    return hipGetLastError();
}

hipError_t leo_decode_multipass(const LeoDev &f, const LeoMultipassArgs &a, hipStream_t s) {
    if (f.bits != 16 || a.slices <= 0) return hipErrorInvalidValue;
    const size_t lds1 = static_cast<size_t>(kSlice) * (a.C + 1);
    const size_t lds2 = 2 * static_cast<size_t>(kSlice) * (a.A * (a.n / a.P) + 1);
    const size_t lds3 = static_cast<size_t>(kSlice) * (a.P + 1);
    const unsigned slices = static_cast<unsigned>(a.slices);
    hipError_t e;
    auto allow = [](const void *k, size_t lds) {
        return lds <= (64u << 10) ? hipSuccess
                                   : hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                         static_cast<int>(lds));
    };
    if ((e = allow(reinterpret_cast<const void *>(k_dec_pass1<F16>), lds1)) != hipSuccess) return e;
    if ((e = allow(reinterpret_cast<const void *>(k_dec_pass2<F16>), lds2)) != hipSuccess) return e;
    // z = u + L(v) holds P * 4 slots in registers: 8 per thread up to P = 512.
    const bool small = a.P * (kSlice / 16) <= (kDerivMax / 2) * kMB;
    if ((e = allow(small ? reinterpret_cast<const void *>(k_dec_pass3<F16, kDerivMax / 2>)
                         : reinterpret_cast<const void *>(k_dec_pass3<F16, kDerivMax>),
                   lds3)) != hipSuccess)
        return e;
    if (a.pass3_blocks == 0) return hipSuccess;  // no output row
    // Row blocks of pass 1 past inputCount and of pass 3 past the last one
    // with an output row have nothing to do and are not launched.
    const unsigned by1 = static_cast<unsigned>((std::min(a.ifft_trunc, a.n) + a.C - 1) / a.C);
    const unsigned by3 = static_cast<unsigned>(leo_pass3_grid(a.pass3_blocks));
    hipLaunchKernelGGL(k_dec_pass1<F16>, dim3(slices, by1), dim3(kMB1), lds1, s, f, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_dec_pass2<F16>, dim3(slices, a.P / a.A), dim3(kMB2), lds2, s, f, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (small) hipLaunchKernelGGL((k_dec_pass3<F16, kDerivMax / 2>), dim3(slices, by3), dim3(kMB), lds3, s, f, a);
    else hipLaunchKernelGGL((k_dec_pass3<F16, kDerivMax>), dim3(slices, by3), dim3(kMB), lds3, s, f, a);
    return hipGetLastError();
}

hipError_t leo_encode_multipass(const LeoDev &f, const LeoEncPassArgs &a, int pass, hipStream_t s) {
    if (f.bits != 16 || a.slices <= 0 || pass < 1 || pass > 3) return hipErrorInvalidValue;
    const size_t lds1 = static_cast<size_t>(kSlice) * (a.C + 1);
    const size_t lds2 = 2 * static_cast<size_t>(kSlice) * (a.A * (a.m / a.P) + 1);
    const size_t lds3 = static_cast<size_t>(kSlice) * (a.P + 1);
    auto allow = [](const void *k, size_t lds) {
        return lds <= (64u << 10) ? hipSuccess
                                   : hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                         static_cast<int>(lds));
    };
    hipError_t e;
    const unsigned slices = static_cast<unsigned>(a.slices);
    if (pass == 1) {
        if ((e = allow(reinterpret_cast<const void *>(k_enc_pass1<F16>), lds1)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_enc_pass1<F16>, dim3(slices, a.m / a.C), dim3(kMB1), lds1, s, f, a);
    } else if (pass == 2) {
        if ((e = allow(reinterpret_cast<const void *>(k_enc_pass2<F16>), lds2)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_enc_pass2<F16>, dim3(slices, a.P / a.A), dim3(kMB2), lds2, s, f, a);
    } else {
        if ((e = allow(reinterpret_cast<const void *>(k_enc_pass3<F16>), lds3)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_enc_pass3<F16>, dim3(slices, a.m / a.P), dim3(kMB), lds3, s, f, a);
    }
    return hipGetLastError();
}

hipError_t leo_decode_multipass_batch(const LeoDev &f, const LeoPassTask *tasks, int ntasks, const uint32_t blocks[3],
                                      int max_p, hipStream_t s) {
    if (f.bits != 16 || ntasks <= 0) return hipErrorInvalidValue;
    const size_t lds1 = static_cast<size_t>(kSlice) * (1024 + 1);
    const size_t lds2 = 2 * static_cast<size_t>(kSlice) * (512 + 1);
    const size_t lds3 = static_cast<size_t>(kSlice) * (max_p + 1);
    hipError_t e;
    auto allow = [](const void *k, size_t lds) {
        return lds <= (64u << 10) ? hipSuccess
                                   : hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                         static_cast<int>(lds));
    };
    const bool small = max_p * (kSlice / 16) <= (kDerivMax / 2) * kMB;
    if ((e = allow(reinterpret_cast<const void *>(k_dec_pass1_batch<F16>), lds1)) != hipSuccess) return e;
    if ((e = allow(reinterpret_cast<const void *>(k_dec_pass2_batch<F16>), lds2)) != hipSuccess) return e;
    if ((e = allow(small ? reinterpret_cast<const void *>(k_dec_pass3_batch<F16, kDerivMax / 2>)
                         : reinterpret_cast<const void *>(k_dec_pass3_batch<F16, kDerivMax>),
                   lds3)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_dec_pass1_batch<F16>, dim3(blocks[0]), dim3(kMB1), lds1, s, f, tasks, ntasks);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_dec_pass2_batch<F16>, dim3(blocks[1]), dim3(kMB2), lds2, s, f, tasks, ntasks);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (small) hipLaunchKernelGGL((k_dec_pass3_batch<F16, kDerivMax / 2>), dim3(blocks[2]), dim3(kMB), lds3, s, f, tasks, ntasks);
    else hipLaunchKernelGGL((k_dec_pass3_batch<F16, kDerivMax>), dim3(blocks[2]), dim3(kMB), lds3, s, f, tasks, ntasks);
    return hipGetLastError();
}

}  // namespace rsx

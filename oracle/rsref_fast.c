/*
 * rsref_fast.c — the reference's fast CPU path for the classic engine, for
 * timing the CPU baseline.  TEST INFRASTRUCTURE ONLY (bench.py's CPU leg and
 * the tests that pin it to the scalar restatement).
 *
 * Restates the semantics of the generated kernels the reference dispatches
 * to from codeSomeShardsP (reedsolomon.go:979-1068):
 *   - mulAvxTwo_KxM (galois_gen_amd64.s, generated by _gen/gen.go:499-591):
 *     per 32-byte block, each input is loaded once, split into low/high
 *     nibbles, and two PSHUFB lookups per (input, output) pair are XORed into
 *     M register accumulators, with the tables of genCodeGenMatrix
 *     (galois.go:914-946: mulTableLow / mulTableHigh per coefficient);
 *   - mulGFNI_KxM_64 (AVX-512 + GFNI): one VGF2P8AFFINEQB per pair with the
 *     8x8 bit matrices of gf2p811dMulMatrices (galois.go:948-964);
 *   - the byte range split into max(bytes/threads, minSplit) chunks rounded
 *     to 64 bytes, one thread per chunk (reedsolomon.go:1007, 1055-1067),
 *     and inputs/outputs tiled by at most 10 x 10 (codeGenMaxInputs/Outputs).
 * Outputs are identical to rsref_code_some_shards (checked by the tests).
 */
#include <immintrin.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "rsref.h"

#define MAXI 10
#define MAXO 10

static uint8_t lowt[256][16], hight[256][16];
static uint64_t gfni_m[256];
static int fast_ready;

static void fast_init(void) {
    if (fast_ready) return;
    uint8_t mul[65536];
    rsref_gf8_tables(NULL, NULL, NULL, mul, &lowt[0][0], &hight[0][0], gfni_m);
    fast_ready = 1;
}

/* One 10x10 (or smaller) tile over [start, stop), AVX2 PSHUFB form. */
__attribute__((target("avx2"))) static void tile_avx2(const uint8_t *rows, int ldr, int nout, int nin,
                                                      const uint8_t *const *in, uint8_t *const *out,
                                                      size_t start, size_t stop, int xor_out) {
    __m256i tl[MAXO][MAXI], th[MAXO][MAXI];
    for (int i = 0; i < nout; i++)
        for (int j = 0; j < nin; j++) {
            uint8_t c = rows[i * ldr + j];
            tl[i][j] = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i *)lowt[c]));
            th[i][j] = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i *)hight[c]));
        }
    const __m256i mask = _mm256_set1_epi8(0x0f);
    size_t b = start;
    for (; b + 32 <= stop; b += 32) {
        __m256i acc[MAXO];
        for (int i = 0; i < nout; i++)
            acc[i] = xor_out ? _mm256_loadu_si256((const __m256i *)(out[i] + b)) : _mm256_setzero_si256();
        for (int j = 0; j < nin; j++) {
            __m256i x = _mm256_loadu_si256((const __m256i *)(in[j] + b));
            __m256i lo = _mm256_and_si256(x, mask);
            __m256i hi = _mm256_and_si256(_mm256_srli_epi64(x, 4), mask);
            for (int i = 0; i < nout; i++)
                acc[i] = _mm256_xor_si256(acc[i], _mm256_xor_si256(_mm256_shuffle_epi8(tl[i][j], lo),
                                                                   _mm256_shuffle_epi8(th[i][j], hi)));
        }
        for (int i = 0; i < nout; i++) _mm256_storeu_si256((__m256i *)(out[i] + b), acc[i]);
    }
    for (; b < stop; b++)
        for (int i = 0; i < nout; i++) {
            uint8_t a = xor_out ? out[i][b] : 0;
            for (int j = 0; j < nin; j++) {
                uint8_t c = rows[i * ldr + j], x = in[j][b];
                a ^= lowt[c][x & 15] ^ hight[c][x >> 4];
            }
            out[i][b] = a;
        }
}

/* Same tile with AVX-512 + GFNI affine transforms on 64-byte blocks. */
__attribute__((target("avx512f,avx512bw,gfni"))) static void tile_gfni(
    const uint8_t *rows, int ldr, int nout, int nin, const uint8_t *const *in, uint8_t *const *out,
    size_t start, size_t stop, int xor_out) {
    __m512i mt[MAXO][MAXI];
    for (int i = 0; i < nout; i++)
        for (int j = 0; j < nin; j++) mt[i][j] = _mm512_set1_epi64((long long)gfni_m[rows[i * ldr + j]]);
    size_t b = start;
    for (; b + 64 <= stop; b += 64) {
        __m512i acc[MAXO];
        for (int i = 0; i < nout; i++)
            acc[i] = xor_out ? _mm512_loadu_si512((const void *)(out[i] + b)) : _mm512_setzero_si512();
        for (int j = 0; j < nin; j++) {
            __m512i x = _mm512_loadu_si512((const void *)(in[j] + b));
            for (int i = 0; i < nout; i++)
                acc[i] = _mm512_xor_si512(acc[i], _mm512_gf2p8affine_epi64_epi8(x, mt[i][j], 0));
        }
        for (int i = 0; i < nout; i++) _mm512_storeu_si512((void *)(out[i] + b), acc[i]);
    }
    if (b < stop) tile_avx2(rows, ldr, nout, nin, in, out, b, stop, xor_out);
}

int rsref_fast_isa(void) {
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512bw") && __builtin_cpu_supports("gfni")) return 2;
    if (__builtin_cpu_supports("avx2")) return 1;
    return 0;
}

/* codeSomeShardsP over `threads` threads: out = rows * in (clear) for len bytes. */
void rsref_code_some_shards_fast(const uint8_t *rows, int nout, int nin, const uint8_t *const *in,
                                 uint8_t *const *out, size_t len, int threads, int isa) {
    fast_init();
    if (isa < 0) isa = rsref_fast_isa();
    if (threads < 1) threads = 1;
    size_t per = len / (size_t)threads;
    if (per < 4096) per = 4096;
    per = (per + 63) & ~(size_t)63;
    size_t nchunks = (len + per - 1) / per;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (size_t ci = 0; ci < nchunks; ci++) {
        size_t start = ci * per, stop = start + per < len ? start + per : len;
        for (int o0 = 0; o0 < nout; o0 += MAXO) {
            int no = nout - o0 < MAXO ? nout - o0 : MAXO;
            for (int i0 = 0; i0 < nin; i0 += MAXI) {
                int ni = nin - i0 < MAXI ? nin - i0 : MAXI;
                const uint8_t *r = rows + (size_t)o0 * nin + i0;
                if (isa == 2)
                    tile_gfni(r, nin, no, ni, in + i0, out + o0, start, stop, i0 > 0);
                else if (isa == 1)
                    tile_avx2(r, nin, no, ni, in + i0, out + o0, start, stop, i0 > 0);
                else
                    for (size_t b = start; b < stop; b++)
                        for (int i = 0; i < no; i++) {
                            uint8_t a = i0 > 0 ? out[o0 + i][b] : 0;
                            for (int j = 0; j < ni; j++) {
                                uint8_t c = r[i * nin + j], x = in[i0 + j][b];
                                a ^= lowt[c][x & 15] ^ hight[c][x >> 4];
                            }
                            out[o0 + i][b] = a;
                        }
            }
        }
    }
}

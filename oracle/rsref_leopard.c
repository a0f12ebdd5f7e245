/*
 * rsref_leopard.c — CPU restatement of klauspost/reedsolomon's Leopard
 * engines: leopard.go (GF(2^16)) and leopard8.go (GF(2^8)), with the
 * butterfly operation order of galois_noasm.go:80-135.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker, never the product path.
 *
 * Line-faithful on purpose: the skew indexing (encoder skewLUT[iend] on
 * fftSkew[m-1:], decoder/FFT skewLUT[iend-1] on fftSkew), the
 * "log_m == modulus => xor only" branches, and fftDIT2 = mulAdd then xor vs
 * ifftDIT2 = xor then mulAdd are all kept as the reference writes them.
 * The reference walks the shard columns in cache-sized chunks; every step is
 * column-separable, so this restatement walks them in fixed 4 KiB chunks.
 * The reference's errorBitfield pruning of the final FFT only skips groups
 * whose outputs are not revealed, so the full truncated FFT is used here.
 *
 * Pinned (tests/test_leopard_field_pin.py): the field of both engines by
 * brute-force carry-less products over the reference's Cantor basis and
 * polynomial literals; fftSkew by the subspace-polynomial closed form;
 * logWalsh by the Walsh-Hadamard definition; encode parity by Lagrange
 * interpolation on small shapes; reconstruct by encode -> erase ->
 * reconstruct = original (tests/test_oracle_leopard.py).
 */
#include "rsref.h"

#include <stdlib.h>
#include <string.h>

#define CHUNK 4096 /* column chunk of this restatement (multiple of 64) */

static size_t shard_size(const size_t *lens, int n) {
    for (int i = 0; i < n; i++)
        if (lens[i]) return lens[i];
    return 0;
}
/* reedsolomon.go:1404-1418 checkShards */
static int check_shards(const size_t *lens, int n, int nilok) {
    size_t size = shard_size(lens, n);
    if (size == 0) return RSREF_ERR_SHARD_NO_DATA;
    for (int i = 0; i < n; i++)
        if (lens[i] != size && (lens[i] != 0 || !nilok)) return RSREF_ERR_SHARD_SIZE;
    return RSREF_OK;
}
/* leopard.go:1017-1020 ceilPow2 */
static int ceil_pow2(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}
static void xor_into(uint8_t *dst, const uint8_t *src, size_t n) {
    for (size_t i = 0; i < n; i++) dst[i] ^= src[i];
}

/* ======================================================================== */
/* GF(2^16): leopard.go                                                     */
/* ======================================================================== */

typedef uint16_t ffe;
#define BITS16 16
#define ORDER16 65536
#define MOD16 65535
#define POLY16 0x1002D

static ffe *exp16, *log16, *skew16, *walsh16;
static ffe (*lo16)[256], (*hi16)[256]; /* mul16LUTs[log_m].Lo / .Hi */
static int ready16;

/* leopard.go:1001-1006 addMod (partial reduction, may return modulus). */
static ffe add_mod16(ffe a, ffe b) {
    uint64_t sum = (uint64_t)a + b;
    return (ffe)(sum + (sum >> BITS16));
}
/* leopard.go:984-996 mulLog. */
static ffe mul_log16(ffe a, ffe log_b) {
    if (a == 0) return 0;
    return exp16[add_mod16(log16[a], log_b)];
}
/* leopard.go:1025-1058 fwht with fwht2alt (1082-1089). */
static void fwht16(ffe *data, int mtrunc) {
    int dist = 1, dist4 = 4;
    while (dist4 <= ORDER16) {
        for (int r = 0; r < mtrunc; r += dist4) {
            for (int off = r; off < r + dist; off++) {
                uint64_t t0 = data[off], t1 = data[off + dist], t2 = data[off + dist * 2],
                         t3 = data[off + dist * 3];
#define FWHT2(a, b)                                              \
    do {                                                         \
        uint64_t s_ = (a) + (b), d_ = (a) - (b);                 \
        s_ = s_ + (s_ >> BITS16);                                \
        d_ = d_ + (d_ >> BITS16);                                \
        (a) = (ffe)s_;                                           \
        (b) = (ffe)d_;                                           \
    } while (0)
                FWHT2(t0, t1);
                FWHT2(t2, t3);
                FWHT2(t0, t2);
                FWHT2(t1, t3);
#undef FWHT2
                data[off] = (ffe)t0;
                data[off + dist] = (ffe)t1;
                data[off + dist * 2] = (ffe)t2;
                data[off + dist * 3] = (ffe)t3;
            }
        }
        dist = dist4;
        dist4 <<= 2;
    }
}

static void init16(void) {
    if (ready16) return;
    exp16 = calloc(ORDER16, sizeof(ffe));
    log16 = calloc(ORDER16, sizeof(ffe));
    skew16 = calloc(MOD16, sizeof(ffe));
    walsh16 = calloc(ORDER16, sizeof(ffe));
    lo16 = calloc(ORDER16, sizeof *lo16);
    hi16 = calloc(ORDER16, sizeof *hi16);
    /* leopard.go:1104-1146 initLUTs */
    static const ffe cantor[BITS16] = {0x0001, 0xACCA, 0x3C0E, 0x163E, 0xC582, 0xED2E,
                                       0x914C, 0x4012, 0x6C98, 0x10D8, 0x6A72, 0xB900,
                                       0xFDB8, 0xFB34, 0xFF38, 0x991E};
    unsigned state = 1;
    for (unsigned i = 0; i < MOD16; i++) {
        exp16[state] = (ffe)i;
        state <<= 1;
        if (state >= ORDER16) state ^= POLY16;
    }
    exp16[0] = MOD16;
    log16[0] = 0;
    for (int i = 0; i < BITS16; i++) {
        int width = 1 << i;
        for (int j = 0; j < width; j++) log16[j + width] = log16[j] ^ cantor[i];
    }
    for (int i = 0; i < ORDER16; i++) log16[i] = exp16[log16[i]];
    for (int i = 0; i < ORDER16; i++) exp16[log16[i]] = (ffe)i;
    exp16[MOD16] = exp16[0];

    /* leopard.go:1150-1195 initFFTSkew */
    ffe temp[BITS16 - 1];
    for (int i = 1; i < BITS16; i++) temp[i - 1] = (ffe)(1u << i);
    for (int m = 0; m < BITS16 - 1; m++) {
        int step = 1 << (m + 1);
        skew16[(1 << m) - 1] = 0;
        for (int i = m; i < BITS16 - 1; i++) {
            int s = 1 << (i + 1);
            for (int j = (1 << m) - 1; j < s; j += step) skew16[j + s] = skew16[j] ^ temp[i];
        }
        temp[m] = (ffe)(MOD16 - log16[mul_log16(temp[m], log16[temp[m] ^ 1])]);
        for (int i = m + 1; i < BITS16 - 1; i++) {
            ffe sum = add_mod16(log16[temp[i] ^ 1], temp[m]);
            temp[i] = mul_log16(temp[i], sum);
        }
    }
    for (int i = 0; i < MOD16; i++) skew16[i] = log16[skew16[i]];
    for (int i = 0; i < ORDER16; i++) walsh16[i] = log16[i];
    walsh16[0] = 0;
    fwht16(walsh16, ORDER16);

    /* leopard.go:1197-1218 initMul16LUT (Lo/Hi product tables). */
    for (int lm = 0; lm < ORDER16; lm++) {
        ffe tmp[64];
        for (int nib = 0, shift = 0; nib < 4; nib++, shift += 4)
            for (int x = 0; x < 16; x++) tmp[nib * 16 + x] = mul_log16((ffe)(x << shift), (ffe)lm);
        for (int i = 0; i < 256; i++) {
            lo16[lm][i] = tmp[i & 15] ^ tmp[(i >> 4) + 16];
            hi16[lm][i] = tmp[(i & 15) + 32] ^ tmp[(i >> 4) + 48];
        }
    }
    ready16 = 1;
}

void rsref_leo16_tables(uint16_t *exp_lut, uint16_t *log_lut, uint16_t *fft_skew,
                        uint16_t *log_walsh) {
    init16();
    if (exp_lut) memcpy(exp_lut, exp16, ORDER16 * sizeof(ffe));
    if (log_lut) memcpy(log_lut, log16, ORDER16 * sizeof(ffe));
    if (fft_skew) memcpy(fft_skew, skew16, MOD16 * sizeof(ffe));
    if (log_walsh) memcpy(log_walsh, walsh16, ORDER16 * sizeof(ffe));
}
uint16_t rsref_leo16_mul_log(uint16_t a, uint16_t log_b) {
    init16();
    return mul_log16(a, log_b);
}
uint16_t rsref_leo16_add_mod(uint16_t a, uint16_t b) { return add_mod16(a, b); }

/* gf16.go:6-16 GF16Mul. */
uint16_t rsref_gf16_mul(uint16_t a, uint16_t b) {
    init16();
    if (a == 0 || b == 0) return 0;
    ffe l = add_mod16(log16[a], log16[b]);
    if (l >= MOD16) l -= MOD16;
    return exp16[l];
}

/* leopard.go:907-925 refMulAdd: x ^= y * log_m over 64-byte split chunks
 * (bytes [0,32) are element low bytes, [32,64) high bytes). */
static void mul_add16(uint8_t *x, const uint8_t *y, ffe log_m, size_t n) {
    const ffe *lo = lo16[log_m], *hi = hi16[log_m];
    for (size_t off = 0; off + 64 <= n; off += 64)
        for (int i = 0; i < 32; i++) {
            ffe prod = lo[y[off + i]] ^ hi[y[off + 32 + i]];
            x[off + i] ^= (uint8_t)prod;
            x[off + 32 + i] ^= (uint8_t)(prod >> 8);
        }
}
/* leopard.go:966-981 refMul: x = y * log_m. */
static void mul16(uint8_t *x, const uint8_t *y, ffe log_m, size_t n) {
    const ffe *lo = lo16[log_m], *hi = hi16[log_m];
    for (size_t off = 0; off < n; off += 64)
        for (int i = 0; i < 32; i++) {
            ffe prod = lo[y[off + i]] ^ hi[y[off + 32 + i]];
            x[off + i] = (uint8_t)prod;
            x[off + 32 + i] = (uint8_t)(prod >> 8);
        }
}
/* galois_noasm.go:94-99 fftDIT2: mulAdd then xor. */
static void fft_dit2_16(uint8_t *x, uint8_t *y, ffe log_m, size_t n) {
    mul_add16(x, y, log_m, n);
    xor_into(y, x, n);
}
/* galois_noasm.go:108-113 ifftDIT2: xor then mulAdd. */
static void ifft_dit2_16(uint8_t *x, uint8_t *y, ffe log_m, size_t n) {
    xor_into(y, x, n);
    mul_add16(x, y, log_m, n);
}
/* leopard.go:875-897 ifftDIT4Ref. */
static void ifft_dit4_16(uint8_t **w, int dist, ffe lm01, ffe lm23, ffe lm02, size_t n) {
    if (lm01 == MOD16) xor_into(w[dist], w[0], n);
    else ifft_dit2_16(w[0], w[dist], lm01, n);
    if (lm23 == MOD16) xor_into(w[dist * 3], w[dist * 2], n);
    else ifft_dit2_16(w[dist * 2], w[dist * 3], lm23, n);
    if (lm02 == MOD16) {
        xor_into(w[dist * 2], w[0], n);
        xor_into(w[dist * 3], w[dist], n);
    } else {
        ifft_dit2_16(w[0], w[dist * 2], lm02, n);
        ifft_dit2_16(w[dist], w[dist * 3], lm02, n);
    }
}
/* leopard.go:739-761 fftDIT4Ref. */
static void fft_dit4_16(uint8_t **w, int dist, ffe lm01, ffe lm23, ffe lm02, size_t n) {
    if (lm02 == MOD16) {
        xor_into(w[dist * 2], w[0], n);
        xor_into(w[dist * 3], w[dist], n);
    } else {
        fft_dit2_16(w[0], w[dist * 2], lm02, n);
        fft_dit2_16(w[dist], w[dist * 3], lm02, n);
    }
    if (lm01 == MOD16) xor_into(w[dist], w[0], n);
    else fft_dit2_16(w[0], w[dist], lm01, n);
    if (lm23 == MOD16) xor_into(w[dist * 3], w[dist * 2], n);
    else fft_dit2_16(w[dist * 2], w[dist * 3], lm23, n);
}

/* leopard.go:764-873 ifftDITEncoder (copy + clear + in-place layers). */
static void ifft_dit_encoder16(uint8_t *const *data, int mtrunc, uint8_t **work, uint8_t **xor_res,
                               int m, const ffe *skew, size_t n) {
    int dist = 1, dist4 = 4;
    for (int i = 0; i < mtrunc; i++) memcpy(work[i], data[i], n);
    for (int i = mtrunc; i < m; i++) memset(work[i], 0, n);
    while (dist4 <= m) {
        for (int r = 0; r < mtrunc; r += dist4) {
            int iend = r + dist;
            ffe lm01 = skew[iend], lm02 = skew[iend + dist], lm23 = skew[iend + dist * 2];
            for (int i = r; i < iend; i++) ifft_dit4_16(work + i, dist, lm01, lm23, lm02, n);
        }
        dist = dist4;
        dist4 <<= 2;
    }
    if (dist < m) {
        ffe lm = skew[dist];
        if (lm == MOD16)
            for (int i = 0; i < dist; i++) xor_into(work[dist + i], work[i], n);
        else
            for (int i = 0; i < dist; i++) ifft_dit2_16(work[i], work[i + dist], lm, n);
    }
    if (xor_res)
        for (int i = 0; i < m; i++) xor_into(xor_res[i], work[i], n);
}

/* leopard.go:650-693 ifftDITDecoder. */
static void ifft_dit_decoder16(int mtrunc, uint8_t **work, int m, const ffe *skew, size_t n) {
    int dist = 1, dist4 = 4;
    while (dist4 <= m) {
        for (int r = 0; r < mtrunc; r += dist4) {
            int iend = r + dist;
            ffe lm01 = skew[iend - 1], lm02 = skew[iend + dist - 1], lm23 = skew[iend + dist * 2 - 1];
            for (int i = r; i < iend; i++) ifft_dit4_16(work + i, dist, lm01, lm23, lm02, n);
        }
        dist = dist4;
        dist4 <<= 2;
    }
    if (dist < m) {
        ffe lm = skew[dist - 1];
        if (lm == MOD16)
            for (int i = 0; i < dist; i++) xor_into(work[dist + i], work[i], n);
        else
            for (int i = 0; i < dist; i++) ifft_dit2_16(work[i], work[i + dist], lm, n);
    }
}

/* leopard.go:696-735 fftDIT. */
static void fft_dit16(uint8_t **work, int mtrunc, int m, const ffe *skew, size_t n) {
    int dist4 = m, dist = m >> 2;
    while (dist != 0) {
        for (int r = 0; r < mtrunc; r += dist4) {
            int iend = r + dist;
            ffe lm01 = skew[iend - 1], lm02 = skew[iend + dist - 1], lm23 = skew[iend + dist * 2 - 1];
            for (int i = r; i < iend; i++) fft_dit4_16(work + i, dist, lm01, lm23, lm02, n);
        }
        dist4 = dist;
        dist >>= 2;
    }
    if (dist4 == 2)
        for (int r = 0; r < mtrunc; r += 2) {
            ffe lm = skew[r];
            if (lm == MOD16) xor_into(work[r + 1], work[r], n);
            else fft_dit2_16(work[r], work[r + 1], lm, n);
        }
}

static uint8_t **alloc_rows(int rows, size_t n) {
    uint8_t **w = malloc(sizeof(uint8_t *) * rows);
    for (int i = 0; i < rows; i++) w[i] = calloc(n, 1);
    return w;
}
static void free_rows(uint8_t **w, int rows) {
    for (int i = 0; i < rows; i++) free(w[i]);
    free(w);
}

/* newFF16 (leopard.go:58) only bounds k + m by 65536; with k + ceilPow2(m) >
 * 65536 the codeword positions leave the field and the encoder indexes past
 * fftSkew (a panic in the reference).  This restatement reports it instead;
 * within the bound every skew index the encoder reads is < 65535. */
static int leo16_shape_ok(int k, int p) { return (long)k + ceil_pow2(p) <= ORDER16; }

/* leopard.go:139-150 Encode + 169-214 encode + 216-266 encodeChunk. */
int rsref_leo16_encode(int k, int p, uint8_t *const *shards, const size_t *lens, int nshards) {
    init16();
    if (nshards != k + p) return RSREF_ERR_TOO_FEW_SHARDS;
    int err = check_shards(lens, nshards, 0);
    if (err) return err;
    size_t size = shard_size(lens, nshards);
    if (size % 64) return RSREF_ERR_INVALID_SHARD_SIZE;
    if (!leo16_shape_ok(k, p)) return RSREF_ERR_MAX_SHARD_NUM;
    int m = ceil_pow2(p);
    int mtrunc = k < m ? k : m;
    int last_count = k % m;
    uint8_t **work = alloc_rows(2 * m, CHUNK);
    uint8_t **wv = malloc(sizeof(uint8_t *) * 2 * m);
    uint8_t **sh = malloc(sizeof(uint8_t *) * nshards);
    const ffe *skew = skew16 + (m - 1);
    for (size_t off = 0; off < size; off += CHUNK) {
        size_t n = size - off < CHUNK ? size - off : CHUNK;
        for (int i = 0; i < 2 * m; i++) wv[i] = work[i];
        for (int i = 0; i < nshards; i++) sh[i] = shards[i] + off;
        for (int i = 0; i < p; i++) wv[i] = shards[k + i] + off; /* parity aliased into work */
        ifft_dit_encoder16(sh, mtrunc, wv, NULL, m, skew, n);
        if (m < k) {
            uint8_t *const *g = sh;
            const ffe *sk = skew;
            for (int i = m; i + m <= k; i += m) {
                g += m;
                sk += m;
                ifft_dit_encoder16(g, m, wv + m, wv, m, sk, n);
            }
            if (last_count) {
                g += m;
                sk += m;
                ifft_dit_encoder16(g, last_count, wv + m, wv, m, sk, n);
            }
        }
        fft_dit16(wv, p, m, skew16, n);
    }
    free_rows(work, 2 * m);
    free(wv);
    free(sh);
    return RSREF_OK;
}

/* leopard.go:393-419 Verify: re-encode to temporaries and compare. */
int rsref_leo16_verify(int k, int p, uint8_t *const *shards, const size_t *lens, int nshards,
                       int *ok) {
    *ok = 0;
    if (nshards != k + p) return RSREF_ERR_TOO_FEW_SHARDS;
    int err = check_shards(lens, nshards, 0);
    if (err) return err;
    size_t size = lens[0];
    uint8_t **out = malloc(sizeof(uint8_t *) * nshards);
    size_t *ol = malloc(sizeof(size_t) * nshards);
    for (int i = 0; i < nshards; i++) {
        out[i] = i < k ? shards[i] : calloc(size, 1);
        ol[i] = size;
    }
    err = rsref_leo16_encode(k, p, out, ol, nshards);
    if (!err) {
        *ok = 1;
        for (int i = k; i < nshards; i++)
            if (memcmp(out[i], shards[i], size)) *ok = 0;
    }
    for (int i = k; i < nshards; i++) free(out[i]);
    free(out);
    free(ol);
    return err;
}

/* leopard.go:459-510: error locator logs.  erased has k+p entries in shard
 * order (data first).  out receives the full 65536-entry errLocs. */
static int errlocs16(int k, int p, const uint8_t *erased, ffe *errlocs, int *input_count) {
    int m = ceil_pow2(p);
    memset(errlocs, 0, ORDER16 * sizeof(ffe));
    *input_count = p;
    for (int i = 0; i < p; i++)
        if (erased[k + i]) errlocs[i] = 1;
    for (int i = p; i < m; i++) errlocs[i] = 1;
    for (int i = 0; i < k; i++) {
        if (erased[i]) errlocs[i + m] = 1;
        else *input_count = m + i + 1;
    }
    fwht16(errlocs, m + k);
    for (int i = 0; i < ORDER16; i++)
        errlocs[i] = (ffe)(((uint64_t)errlocs[i] * walsh16[i]) % MOD16);
    fwht16(errlocs, ORDER16);
    return RSREF_OK;
}

int rsref_leo16_errlocs(int k, int p, const uint8_t *erased, int recover_all, uint16_t *out) {
    (void)recover_all;
    init16();
    if (!leo16_shape_ok(k, p)) return RSREF_ERR_MAX_SHARD_NUM;
    int ic;
    return errlocs16(k, p, erased, out, &ic);
}

/* leopard.go:385-391 Reconstruct/ReconstructData + 421-580 reconstruct +
 * 582-647 reconstructChunk. */
int rsref_leo16_reconstruct(int k, int p, uint8_t *const *shards, const size_t *lens, int nshards,
                            int recover_all, uint8_t *filled) {
    init16();
    int total = k + p;
    for (int i = 0; i < nshards; i++) filled[i] = 0;
    if (nshards != total) return RSREF_ERR_TOO_FEW_SHARDS;
    int err = check_shards(lens, nshards, 1);
    if (err) return err;
    int present = 0, data_present = 0;
    for (int i = 0; i < total; i++)
        if (lens[i]) {
            present++;
            if (i < k) data_present++;
        }
    if (present == total || (!recover_all && data_present == k)) return RSREF_OK;
    if (present < k) return RSREF_ERR_TOO_FEW_SHARDS;
    size_t size = shard_size(lens, nshards);
    if (size % 64) return RSREF_ERR_INVALID_SHARD_SIZE;
    if (!leo16_shape_ok(k, p)) return RSREF_ERR_MAX_SHARD_NUM;

    int m = ceil_pow2(p), n = ceil_pow2(m + k);
    uint8_t *erased = malloc(total);
    for (int i = 0; i < total; i++) erased[i] = lens[i] == 0;
    ffe *errlocs = malloc(ORDER16 * sizeof(ffe));
    int input_count;
    errlocs16(k, p, erased, errlocs, &input_count);

    uint8_t **work = alloc_rows(n, CHUNK);
    int end = recover_all ? total : k;
    for (int i = 0; i < end; i++)
        if (erased[i]) filled[i] = 1;
    for (size_t off = 0; off < size; off += CHUNK) {
        size_t len = size - off < CHUNK ? size - off : CHUNK;
        for (int i = 0; i < p; i++) {
            if (!erased[k + i]) mul16(work[i], shards[k + i] + off, errlocs[i], len);
            else memset(work[i], 0, len);
        }
        for (int i = p; i < m; i++) memset(work[i], 0, len);
        for (int i = 0; i < k; i++) {
            if (!erased[i]) mul16(work[m + i], shards[i] + off, errlocs[m + i], len);
            else memset(work[m + i], 0, len);
        }
        for (int i = m + k; i < n; i++) memset(work[i], 0, len);
        ifft_dit_decoder16(input_count, work, n, skew16, len);
        /* leopard.go:619-623 formal derivative */
        for (int i = 1; i < n; i++) {
            int width = ((i ^ (i - 1)) + 1) >> 1;
            for (int j = 0; j < width; j++) xor_into(work[i - width + j], work[i + j], len);
        }
        fft_dit16(work, m + k, n, skew16, len);
        for (int i = 0; i < end; i++) {
            if (!erased[i]) continue;
            if (i >= k) mul16(shards[i] + off, work[i - k], (ffe)(MOD16 - errlocs[i - k]), len);
            else mul16(shards[i] + off, work[i + m], (ffe)(MOD16 - errlocs[i + m]), len);
        }
    }
    free_rows(work, n);
    free(erased);
    free(errlocs);
    return RSREF_OK;
}

/* ======================================================================== */
/* GF(2^8): leopard8.go                                                     */
/* ======================================================================== */

typedef uint8_t ffe8;
#define BITS8 8
#define ORDER8 256
#define MOD8 255
#define POLY8 0x11D
#define WORK8 (32 << 10) /* leopard8.go:103 workSize8 */

static ffe8 exp8[ORDER8], log8[ORDER8], skew8[MOD8], walsh8[ORDER8];
static ffe8 mul8t[ORDER8][256]; /* mul8LUTs[log_m].Value */
static int ready8;

/* leopard8.go:931-936 addMod8 */
static ffe8 add_mod8(ffe8 a, ffe8 b) {
    uint64_t sum = (uint64_t)a + b;
    return (ffe8)(sum + (sum >> BITS8));
}
/* leopard8.go:919-929 mulLog8 */
static ffe8 mul_log8(ffe8 a, ffe8 log_b) {
    if (a == 0) return 0;
    return exp8[add_mod8(log8[a], log_b)];
}
/* leopard8.go:954-987 fwht8 with fwht2alt8 (1012-1014). */
static void fwht8(ffe8 *data, int mtrunc) {
    int dist = 1, dist4 = 4;
    while (dist4 <= ORDER8) {
        for (int r = 0; r < mtrunc; r += dist4)
            for (int off = r; off < r + dist; off++) {
                ffe8 t[4] = {data[off], data[off + dist], data[off + dist * 2], data[off + dist * 3]};
#define FWHT2_8(a, b)                                         \
    do {                                                      \
        uint64_t s_ = (uint64_t)(a) + (b), d_ = (uint64_t)(a) - (b); \
        (a) = (ffe8)(s_ + (s_ >> BITS8));                     \
        (b) = (ffe8)(d_ + (d_ >> BITS8));                     \
    } while (0)
                FWHT2_8(t[0], t[1]);
                FWHT2_8(t[2], t[3]);
                FWHT2_8(t[0], t[2]);
                FWHT2_8(t[1], t[3]);
#undef FWHT2_8
                data[off] = t[0];
                data[off + dist] = t[1];
                data[off + dist * 2] = t[2];
                data[off + dist * 3] = t[3];
            }
        dist = dist4;
        dist4 <<= 2;
    }
}

static void init8(void) {
    if (ready8) return;
    /* leopard8.go:1029-1069 initLUTs8 */
    static const ffe8 cantor[BITS8] = {1, 214, 152, 146, 86, 200, 88, 230};
    unsigned state = 1;
    for (unsigned i = 0; i < MOD8; i++) {
        exp8[state] = (ffe8)i;
        state <<= 1;
        if (state >= ORDER8) state ^= POLY8;
    }
    exp8[0] = MOD8;
    log8[0] = 0;
    for (int i = 0; i < BITS8; i++) {
        int width = 1 << i;
        for (int j = 0; j < width; j++) log8[j + width] = log8[j] ^ cantor[i];
    }
    for (int i = 0; i < ORDER8; i++) log8[i] = exp8[log8[i]];
    for (int i = 0; i < ORDER8; i++) exp8[log8[i]] = (ffe8)i;
    exp8[MOD8] = exp8[0];
    /* leopard8.go:1072-1117 initFFTSkew8 */
    ffe8 temp[BITS8 - 1];
    for (int i = 1; i < BITS8; i++) temp[i - 1] = (ffe8)(1u << i);
    memset(skew8, 0, sizeof skew8);
    for (int m = 0; m < BITS8 - 1; m++) {
        int step = 1 << (m + 1);
        skew8[(1 << m) - 1] = 0;
        for (int i = m; i < BITS8 - 1; i++) {
            int s = 1 << (i + 1);
            for (int j = (1 << m) - 1; j < s; j += step) skew8[j + s] = skew8[j] ^ temp[i];
        }
        temp[m] = (ffe8)(MOD8 - log8[mul_log8(temp[m], log8[temp[m] ^ 1])]);
        for (int i = m + 1; i < BITS8 - 1; i++) {
            ffe8 sum = add_mod8(log8[temp[i] ^ 1], temp[m]);
            temp[i] = mul_log8(temp[i], sum);
        }
    }
    for (int i = 0; i < MOD8; i++) skew8[i] = log8[skew8[i]];
    for (int i = 0; i < ORDER8; i++) walsh8[i] = log8[i];
    walsh8[0] = 0;
    fwht8(walsh8, ORDER8);
    /* leopard8.go:1119-1139 initMul8LUT */
    for (int lm = 0; lm < ORDER8; lm++) {
        ffe8 tmp[64];
        for (int nib = 0, shift = 0; nib < 4; nib++, shift += 4)
            for (int x = 0; x < 16; x++) tmp[nib * 16 + x] = mul_log8((ffe8)(x << shift), (ffe8)lm);
        for (int i = 0; i < 256; i++) mul8t[lm][i] = tmp[i & 15] ^ tmp[(i >> 4) + 16];
    }
    ready8 = 1;
}

void rsref_leo8_tables(uint8_t *exp_lut, uint8_t *log_lut, uint8_t *fft_skew, uint8_t *log_walsh) {
    init8();
    if (exp_lut) memcpy(exp_lut, exp8, ORDER8);
    if (log_lut) memcpy(log_lut, log8, ORDER8);
    if (fft_skew) memcpy(fft_skew, skew8, MOD8);
    if (log_walsh) memcpy(log_walsh, walsh8, ORDER8);
}
uint8_t rsref_leo8_mul_log(uint8_t a, uint8_t log_b) {
    init8();
    return mul_log8(a, log_b);
}
uint8_t rsref_leo8_add_mod(uint8_t a, uint8_t b) { return add_mod8(a, b); }

/* leopard8.go:891-905 refMulAdd8 (64-byte blocks). */
static void mul_add8(uint8_t *x, const uint8_t *y, ffe8 lm, size_t n) {
    const ffe8 *t = mul8t[lm];
    for (size_t off = 0; off + 64 <= n; off += 64)
        for (int i = 0; i < 64; i++) x[off + i] ^= t[y[off + i]];
}
/* leopard8.go:907-917 refMul8. */
static void mul8(uint8_t *x, const uint8_t *y, ffe8 lm, size_t n) {
    const ffe8 *t = mul8t[lm];
    for (size_t i = 0; i < n; i++) x[i] = t[y[i]];
}
static void fft_dit2_8(uint8_t *x, uint8_t *y, ffe8 lm, size_t n) {
    mul_add8(x, y, lm, n);
    xor_into(y, x, n);
}
static void ifft_dit2_8(uint8_t *x, uint8_t *y, ffe8 lm, size_t n) {
    xor_into(y, x, n);
    mul_add8(x, y, lm, n);
}
/* leopard8.go:858-880 ifftDIT4Ref8 */
static void ifft_dit4_8(uint8_t **w, int dist, ffe8 lm01, ffe8 lm23, ffe8 lm02, size_t n) {
    if (lm01 == MOD8) xor_into(w[dist], w[0], n);
    else ifft_dit2_8(w[0], w[dist], lm01, n);
    if (lm23 == MOD8) xor_into(w[dist * 3], w[dist * 2], n);
    else ifft_dit2_8(w[dist * 2], w[dist * 3], lm23, n);
    if (lm02 == MOD8) {
        xor_into(w[dist * 2], w[0], n);
        xor_into(w[dist * 3], w[dist], n);
    } else {
        ifft_dit2_8(w[0], w[dist * 2], lm02, n);
        ifft_dit2_8(w[dist], w[dist * 3], lm02, n);
    }
}
/* leopard8.go:726-748 fftDIT4Ref8 */
static void fft_dit4_8(uint8_t **w, int dist, ffe8 lm01, ffe8 lm23, ffe8 lm02, size_t n) {
    if (lm02 == MOD8) {
        xor_into(w[dist * 2], w[0], n);
        xor_into(w[dist * 3], w[dist], n);
    } else {
        fft_dit2_8(w[0], w[dist * 2], lm02, n);
        fft_dit2_8(w[dist], w[dist * 3], lm02, n);
    }
    if (lm01 == MOD8) xor_into(w[dist], w[0], n);
    else fft_dit2_8(w[0], w[dist], lm01, n);
    if (lm23 == MOD8) xor_into(w[dist * 3], w[dist * 2], n);
    else fft_dit2_8(w[dist * 2], w[dist * 3], lm23, n);
}
/* leopard8.go:755-856 ifftDITEncoder8 */
static void ifft_dit_encoder8(uint8_t *const *data, int mtrunc, uint8_t **work, uint8_t **xor_res,
                              int m, const ffe8 *skew, size_t n) {
    int dist = 1, dist4 = 4;
    for (int i = 0; i < mtrunc; i++) memcpy(work[i], data[i], n);
    for (int i = mtrunc; i < m; i++) memset(work[i], 0, n);
    while (dist4 <= m) {
        for (int r = 0; r < mtrunc; r += dist4) {
            int iend = r + dist;
            ffe8 lm01 = skew[iend], lm02 = skew[iend + dist], lm23 = skew[iend + dist * 2];
            for (int i = r; i < iend; i++) ifft_dit4_8(work + i, dist, lm01, lm23, lm02, n);
        }
        dist = dist4;
        dist4 <<= 2;
    }
    if (dist < m) {
        ffe8 lm = skew[dist];
        if (lm == MOD8)
            for (int i = 0; i < dist; i++) xor_into(work[dist + i], work[i], n);
        else
            for (int i = 0; i < dist; i++) ifft_dit2_8(work[i], work[i + dist], lm, n);
    }
    if (xor_res)
        for (int i = 0; i < m; i++) xor_into(xor_res[i], work[i], n);
}
/* leopard8.go:637-680 ifftDITDecoder8 */
static void ifft_dit_decoder8(int mtrunc, uint8_t **work, int m, const ffe8 *skew, size_t n) {
    int dist = 1, dist4 = 4;
    while (dist4 <= m) {
        for (int r = 0; r < mtrunc; r += dist4) {
            int iend = r + dist;
            ffe8 lm01 = skew[iend - 1], lm02 = skew[iend + dist - 1], lm23 = skew[iend + dist * 2 - 1];
            for (int i = r; i < iend; i++) ifft_dit4_8(work + i, dist, lm01, lm23, lm02, n);
        }
        dist = dist4;
        dist4 <<= 2;
    }
    if (dist < m) {
        ffe8 lm = skew[dist - 1];
        if (lm == MOD8)
            for (int i = 0; i < dist; i++) xor_into(work[dist + i], work[i], n);
        else
            for (int i = 0; i < dist; i++) ifft_dit2_8(work[i], work[i + dist], lm, n);
    }
}
/* leopard8.go:683-723 fftDIT8 */
static void fft_dit8(uint8_t **work, int mtrunc, int m, const ffe8 *skew, size_t n) {
    int dist4 = m, dist = m >> 2;
    while (dist != 0) {
        for (int r = 0; r < mtrunc; r += dist4) {
            int iend = r + dist;
            ffe8 lm01 = skew[iend - 1], lm02 = skew[iend + dist - 1], lm23 = skew[iend + dist * 2 - 1];
            for (int i = r; i < iend; i++) fft_dit4_8(work + i, dist, lm01, lm23, lm02, n);
        }
        dist4 = dist;
        dist >>= 2;
    }
    if (dist4 == 2)
        for (int r = 0; r < mtrunc; r += 2) {
            ffe8 lm = skew[r];
            if (lm == MOD8) xor_into(work[r + 1], work[r], n);
            else fft_dit2_8(work[r], work[r + 1], lm, n);
        }
}

/* The reference indexes past fftSkew8 / errLocs (and panics) when
 * k + ceilPow2(m) > 256; this restatement reports it instead. */
static int leo8_shape_ok(int k, int p) { return k + ceil_pow2(p) <= ORDER8; }

/* leopard8.go:120-130 Encode + 139-243 encode. */
int rsref_leo8_encode(int k, int p, uint8_t *const *shards, const size_t *lens, int nshards) {
    init8();
    if (nshards != k + p) return RSREF_ERR_TOO_FEW_SHARDS;
    int err = check_shards(lens, nshards, 0);
    if (err) return err;
    size_t size = shard_size(lens, nshards);
    if (size % 64) return RSREF_ERR_INVALID_SHARD_SIZE;
    if (!leo8_shape_ok(k, p)) return RSREF_ERR_MAX_SHARD_NUM;
    int m = ceil_pow2(p);
    int mtrunc = k < m ? k : m;
    int last_count = k % m;
    uint8_t **work = alloc_rows(2 * m, CHUNK);
    uint8_t **wv = malloc(sizeof(uint8_t *) * 2 * m);
    uint8_t **sh = malloc(sizeof(uint8_t *) * nshards);
    const ffe8 *skew = skew8 + (m - 1);
    for (size_t off = 0; off < size; off += CHUNK) {
        size_t n = size - off < CHUNK ? size - off : CHUNK;
        for (int i = 0; i < 2 * m; i++) wv[i] = work[i];
        for (int i = 0; i < nshards; i++) sh[i] = shards[i] + off;
        for (int i = 0; i < p; i++) wv[i] = shards[k + i] + off;
        ifft_dit_encoder8(sh, mtrunc, wv, NULL, m, skew, n);
        if (m < k) {
            uint8_t *const *g = sh;
            const ffe8 *sk = skew;
            for (int i = m; i + m <= k; i += m) {
                g += m;
                sk += m;
                ifft_dit_encoder8(g, m, wv + m, wv, m, sk, n);
            }
            if (last_count) {
                g += m;
                sk += m;
                ifft_dit_encoder8(g, last_count, wv + m, wv, m, sk, n);
            }
        }
        fft_dit8(wv, p, m, skew8, n);
    }
    free_rows(work, 2 * m);
    free(wv);
    free(sh);
    return RSREF_OK;
}

/* leopard8.go:365-391 Verify */
int rsref_leo8_verify(int k, int p, uint8_t *const *shards, const size_t *lens, int nshards,
                      int *ok) {
    *ok = 0;
    if (nshards != k + p) return RSREF_ERR_TOO_FEW_SHARDS;
    int err = check_shards(lens, nshards, 0);
    if (err) return err;
    size_t size = lens[0];
    uint8_t **out = malloc(sizeof(uint8_t *) * nshards);
    size_t *ol = malloc(sizeof(size_t) * nshards);
    for (int i = 0; i < nshards; i++) {
        out[i] = i < k ? shards[i] : calloc(size, 1);
        ol[i] = size;
    }
    err = rsref_leo8_encode(k, p, out, ol, nshards);
    if (!err) {
        *ok = 1;
        for (int i = k; i < nshards; i++)
            if (memcmp(out[i], shards[i], size)) *ok = 0;
    }
    for (int i = k; i < nshards; i++) free(out[i]);
    free(out);
    free(ol);
    return err;
}

/* leopard8.go:393-634 reconstruct (errLocs computed fresh: the reference's
 * errLocs cache only memoises them). */
int rsref_leo8_reconstruct(int k, int p, uint8_t *const *shards, const size_t *lens, int nshards,
                           int recover_all, uint8_t *filled) {
    init8();
    int total = k + p;
    for (int i = 0; i < nshards; i++) filled[i] = 0;
    if (nshards != total) return RSREF_ERR_TOO_FEW_SHARDS;
    int err = check_shards(lens, nshards, 1);
    if (err) return err;
    int present = 0, data_present = 0;
    for (int i = 0; i < total; i++)
        if (lens[i]) {
            present++;
            if (i < k) data_present++;
        }
    if (present == total || (!recover_all && data_present == k)) return RSREF_OK;
    if (present < k) return RSREF_ERR_TOO_FEW_SHARDS;
    size_t size = shard_size(lens, nshards);
    if (size % 64) return RSREF_ERR_INVALID_SHARD_SIZE;
    if (!leo8_shape_ok(k, p)) return RSREF_ERR_MAX_SHARD_NUM;

    int m = ceil_pow2(p), n = ceil_pow2(m + k);
    ffe8 errlocs[ORDER8];
    memset(errlocs, 0, sizeof errlocs);
    for (int i = 0; i < p; i++)
        if (!lens[k + i]) errlocs[i] = 1;
    for (int i = p; i < m; i++) errlocs[i] = 1;
    for (int i = 0; i < k; i++)
        if (!lens[i]) errlocs[i + m] = 1;
    fwht8(errlocs, m + k);
    for (int i = 0; i < ORDER8; i++) errlocs[i] = (ffe8)(((unsigned)errlocs[i] * walsh8[i]) % MOD8);
    fwht8(errlocs, ORDER8);

    uint8_t **work = alloc_rows(n, CHUNK);
    int end = recover_all ? total : k;
    for (int i = 0; i < end; i++)
        if (!lens[i]) filled[i] = 1;
    for (size_t off = 0; off < size; off += CHUNK) {
        size_t len = size - off < CHUNK ? size - off : CHUNK;
        for (int i = 0; i < p; i++) {
            if (lens[k + i]) mul8(work[i], shards[k + i] + off, errlocs[i], len);
            else memset(work[i], 0, len);
        }
        for (int i = p; i < m; i++) memset(work[i], 0, len);
        for (int i = 0; i < k; i++) {
            if (lens[i]) mul8(work[m + i], shards[i] + off, errlocs[m + i], len);
            else memset(work[m + i], 0, len);
        }
        for (int i = m + k; i < n; i++) memset(work[i], 0, len);
        ifft_dit_decoder8(m + k, work, n, skew8, len);
        for (int i = 1; i < n; i++) {
            int width = ((i ^ (i - 1)) + 1) >> 1;
            for (int j = 0; j < width; j++) xor_into(work[i - width + j], work[i + j], len);
        }
        fft_dit8(work, m + k, n, skew8, len);
        for (int i = 0; i < end; i++) {
            if (lens[i]) continue;
            if (i >= k) mul8(shards[i] + off, work[i - k], (ffe8)(MOD8 - errlocs[i - k]), len);
            else mul8(shards[i] + off, work[i + m], (ffe8)(MOD8 - errlocs[i + m]), len);
        }
    }
    free_rows(work, n);
    return RSREF_OK;
}

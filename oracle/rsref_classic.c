/*
 * rsref_classic.c — CPU restatement of klauspost/reedsolomon's classic
 * GF(2^8) matrix code: galois.go, gentables.go, matrix.go and the
 * Encode / EncodeIdx / Verify / reconstruct paths of reedsolomon.go.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker, never the product path.
 * Scalar loops on purpose: clarity over speed.
 */
#include "rsref.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* GF(2^8): field of 256 elements generated by x^8 + 29 (galois.go:14-27).    */
/* ------------------------------------------------------------------------ */

static uint8_t g_exp[256];       /* galois.go:71 expTable (entry 255 == entry 0) */
static int16_t g_log[256];       /* gentables.go:8 logTable, log(0) = -1 there    */
static uint8_t g_exp2[510];      /* gentables.go:generateExpTable (2*256-2 long)  */
static uint8_t g_mul[256][256];  /* galois.go:80 mulTable                         */
static int g_ready;

static void gf8_init(void) {
    if (g_ready) return;
    /* LFSR: successive powers of the generator 2 modulo x^8+29 (galois.go:19-27). */
    unsigned x = 1;
    for (int i = 0; i < 255; i++) {
        g_exp[i] = (uint8_t)x;
        g_log[x] = (int16_t)i;
        x <<= 1;
        if (x & 0x100) x ^= 0x100 | 29;
    }
    g_exp[255] = g_exp[0];
    g_log[0] = -1;
    /* gentables.go:generateExpTable: result[log] = result[log+255] = i. */
    for (int i = 1; i < 256; i++) {
        g_exp2[g_log[i]] = (uint8_t)i;
        g_exp2[g_log[i] + 255] = (uint8_t)i;
    }
    /* gentables.go:generateMulTableSplit. */
    for (int a = 0; a < 256; a++)
        for (int b = 0; b < 256; b++)
            g_mul[a][b] = (a == 0 || b == 0) ? 0 : g_exp2[g_log[a] + g_log[b]];
    g_ready = 1;
}

/* galois.go:852 galMultiply = mulTable[a][b]. */
uint8_t rsref_gal_mul(uint8_t a, uint8_t b) {
    gf8_init();
    return g_mul[a][b];
}

/* galois.go:870-885 galDivide (b == 0 panics there; returns 0 here). */
uint8_t rsref_gal_divide(uint8_t a, uint8_t b) {
    gf8_init();
    if (a == 0 || b == 0) return 0;
    int l = g_log[a] - g_log[b];
    if (l < 0) l += 255;
    return g_exp[(uint8_t)l];
}

/* galois.go:887-893 galOneOver: expTable[logTable[a] ^ 255]. */
static uint8_t gal_one_over(uint8_t a) {
    gf8_init();
    return g_exp[(uint8_t)(g_log[a] ^ 255)];
}

/* galois.go:898-912 galExp: a**n with galExp(a,0)=1 and galExp(0,n>0)=0. */
uint8_t rsref_gal_exp(uint8_t a, int n) {
    gf8_init();
    if (n == 0) return 1;
    if (a == 0) return 0;
    int l = g_log[a] * n;
    while (l >= 255) l -= 255;
    return g_exp[(uint8_t)l];
}

/*
 * Tables in the form the reference holds as literals (galois.go:29 logTable
 * with log(0) stored as 0, :71 expTable, :78 invTable, :80 mulTable,
 * :337 mulTableLow, :593 mulTableHigh, :948 gf2p811dMulMatrices).
 * The GFNI matrix of c is the 8x8 GF(2) matrix of x -> c*x in the
 * VGF2P8AFFINEQB layout: output bit j reads byte (7-j), input bit i is bit i
 * of that byte.
 */
void rsref_gf8_tables(uint8_t *exp_tab, uint8_t *log_tab, uint8_t *inv_tab, uint8_t *mul_tab,
                      uint8_t *mul_low, uint8_t *mul_high, uint64_t *gfni) {
    gf8_init();
    for (int i = 0; i < 256; i++) {
        if (exp_tab) exp_tab[i] = g_exp[i];
        if (log_tab) log_tab[i] = (uint8_t)(i == 0 ? 0 : g_log[i]);
        if (inv_tab) inv_tab[i] = (uint8_t)(i == 0 ? 0 : gal_one_over((uint8_t)i));
    }
    if (mul_tab) memcpy(mul_tab, g_mul, sizeof g_mul);
    for (int c = 0; c < 256; c++) {
        for (int n = 0; n < 16; n++) {
            if (mul_low) mul_low[c * 16 + n] = g_mul[c][n];
            if (mul_high) mul_high[c * 16 + n] = g_mul[c][n << 4];
        }
        if (gfni) {
            uint64_t mtx = 0;
            for (int in_bit = 0; in_bit < 8; in_bit++) {
                uint8_t r = g_mul[c][1u << in_bit];
                for (int out_bit = 0; out_bit < 8; out_bit++)
                    if ((r >> out_bit) & 1) mtx |= 1ull << ((7 - out_bit) * 8 + in_bit);
            }
            gfni[c] = mtx;
        }
    }
}

/* galois_noasm.go:9-67 galMulSlice / galMulSliceXor: out (^)= c*in. */
void rsref_gal_mul_slice(uint8_t c, const uint8_t *in, uint8_t *out, size_t n, int do_xor) {
    gf8_init();
    for (size_t i = 0; i < n; i++) {
        uint8_t p = g_mul[c][in[i]];
        out[i] = do_xor ? (uint8_t)(out[i] ^ p) : p;
    }
}

/* ------------------------------------------------------------------------ */
/* matrix.go                                                                */
/* ------------------------------------------------------------------------ */

/* matrix.go:103-118 Multiply. */
int rsref_matrix_multiply(const uint8_t *a, int ar, int ac, const uint8_t *b, int br, int bc,
                          uint8_t *out) {
    gf8_init();
    if (ac != br) return RSREF_ERR_MATRIX;
    for (int r = 0; r < ar; r++)
        for (int c = 0; c < bc; c++) {
            uint8_t v = 0;
            for (int i = 0; i < ac; i++) v ^= g_mul[a[r * ac + i]][b[i * bc + c]];
            out[r * bc + c] = v;
        }
    return RSREF_OK;
}

/* matrix.go:209-259 gaussianElimination on a rows x cols work matrix. */
static int gaussian_elimination(uint8_t *m, int rows, int cols) {
    for (int r = 0; r < rows; r++) {
        if (m[r * cols + r] == 0) {
            for (int below = r + 1; below < rows; below++) {
                if (m[below * cols + r] != 0) {
                    for (int c = 0; c < cols; c++) {
                        uint8_t t = m[r * cols + c];
                        m[r * cols + c] = m[below * cols + c];
                        m[below * cols + c] = t;
                    }
                    break;
                }
            }
        }
        if (m[r * cols + r] == 0) return RSREF_ERR_SINGULAR;
        if (m[r * cols + r] != 1) {
            uint8_t scale = gal_one_over(m[r * cols + r]);
            for (int c = 0; c < cols; c++) m[r * cols + c] = g_mul[m[r * cols + c]][scale];
        }
        for (int below = r + 1; below < rows; below++) {
            uint8_t scale = m[below * cols + r];
            if (scale != 0)
                for (int c = 0; c < cols; c++) m[below * cols + c] ^= g_mul[scale][m[r * cols + c]];
        }
    }
    for (int d = 0; d < rows; d++)
        for (int above = 0; above < d; above++) {
            uint8_t scale = m[above * cols + d];
            if (scale != 0)
                for (int c = 0; c < cols; c++) m[above * cols + c] ^= g_mul[scale][m[d * cols + c]];
        }
    return RSREF_OK;
}

/* matrix.go:192-207 Invert = augment with identity, eliminate, take the right half. */
int rsref_matrix_invert(const uint8_t *m, int rows, int cols, uint8_t *out) {
    gf8_init();
    if (rows != cols) return RSREF_ERR_NOT_SQUARE;
    int n = rows, w = 2 * n;
    uint8_t *work = calloc((size_t)n * w, 1);
    for (int r = 0; r < n; r++) {
        memcpy(work + r * w, m + r * n, n);
        work[r * w + n + r] = 1;
    }
    int err = gaussian_elimination(work, n, w);
    if (err == RSREF_OK)
        for (int r = 0; r < n; r++) memcpy(out + r * n, work + r * w + n, n);
    free(work);
    return err;
}

/* matrix.go:270-281 vandermonde: vm[r][c] = galExp(byte(r), c). */
static void vandermonde(int rows, int cols, uint8_t *out) {
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < cols; c++) out[r * cols + c] = rsref_gal_exp((uint8_t)r, c);
}

/* reedsolomon.go:232-256 buildMatrix: vm * inverse(top k x k of vm). */
static int build_default(int k, int total, uint8_t *out) {
    uint8_t *vm = malloc((size_t)total * k), *inv = malloc((size_t)k * k);
    vandermonde(total, k, vm);
    int err = rsref_matrix_invert(vm, k, k, inv);
    if (err == RSREF_OK) err = rsref_matrix_multiply(vm, total, k, inv, k, k, out);
    free(vm);
    free(inv);
    return err;
}

/* reedsolomon.go:263-342 buildMatrixJerasure. */
static int build_jerasure(int k, int total, uint8_t *vm) {
    vandermonde(total, k, vm);
#define VM(r, c) vm[(r) * k + (c)]
    VM(0, 0) = 1;
    for (int i = 1; i < k; i++) VM(0, i) = 0;
    for (int i = 0; i < k - 1; i++) VM(total - 1, i) = 0;
    VM(total - 1, k - 1) = 1;
    uint8_t *tmp = malloc(k);
    for (int i = 0; i < k; i++) {
        int r = i;
        while (r < total && VM(r, i) == 0) r++;
        if (r != i) {
            memcpy(tmp, &VM(r, 0), k);
            memcpy(&VM(r, 0), &VM(i, 0), k);
            memcpy(&VM(i, 0), tmp, k);
        }
        if (VM(i, i) != 1) {
            uint8_t t = gal_one_over(VM(i, i));
            for (int j = 0; j < total; j++) VM(j, i) = g_mul[VM(j, i)][t];
        }
        for (int j = 0; j < k; j++) {
            uint8_t t = VM(i, j);
            if (j != i && t != 0)
                for (int rr = 0; rr < total; rr++) VM(rr, j) ^= g_mul[t][VM(rr, i)];
        }
    }
    for (int j = 0; j < k; j++) {
        uint8_t t = VM(k, j);
        if (t != 1) {
            t = gal_one_over(t);
            for (int i = k; i < total; i++) VM(i, j) = g_mul[VM(i, j)][t];
        }
    }
    for (int i = k + 1; i < total; i++) {
        uint8_t t = VM(i, 0);
        if (t != 1) {
            t = gal_one_over(t);
            for (int j = 0; j < k; j++) VM(i, j) = g_mul[VM(i, j)][t];
        }
    }
#undef VM
    free(tmp);
    return RSREF_OK;
}

/* reedsolomon.go:232-413: the matrix builders. */
int rsref_build_matrix(int kind, int k, int total, uint8_t *out) {
    gf8_init();
    switch (kind) {
    case RSREF_MATRIX_DEFAULT:
        return build_default(k, total, out);
    case RSREF_MATRIX_JERASURE:
        return build_jerasure(k, total, out);
    default:
        break;
    }
    memset(out, 0, (size_t)total * k);
    for (int r = 0; r < total; r++)
        for (int c = 0; c < k; c++) {
            if (r < k) {
                out[r * k + c] = (uint8_t)(r == c);
                continue;
            }
            switch (kind) {
            case RSREF_MATRIX_CAUCHY: /* reedsolomon.go:369-389: invTable[byte(r^c)] */
                out[r * k + c] = (uint8_t)((r ^ c) & 0xff) ? gal_one_over((uint8_t)(r ^ c)) : 0;
                break;
            case RSREF_MATRIX_PAR1: /* reedsolomon.go:348-367: galExp(byte(c+1), r-k) */
                out[r * k + c] = rsref_gal_exp((uint8_t)(c + 1), r - k);
                break;
            case RSREF_MATRIX_XOR: /* reedsolomon.go:391-413 buildXorMatrix */
                if (total != k + 1) return RSREF_ERR_MATRIX;
                out[r * k + c] = 1;
                break;
            default:
                return RSREF_ERR_MATRIX;
            }
        }
    return RSREF_OK;
}

/* reedsolomon.go:461-488: matrix of New(k, m, ...). kind XOR means
 * WithFastOneParityMatrix, which only applies when m == 1. */
int rsref_codec_matrix(int k, int m, int kind, const uint8_t *custom, int custom_rows,
                       int custom_cols, uint8_t *out) {
    gf8_init();
    int total = k + m;
    if (custom) {
        if (custom_rows < m) return RSREF_ERR_MATRIX;
        if (custom_cols < k) return RSREF_ERR_MATRIX;
        memset(out, 0, (size_t)total * k);
        for (int i = 0; i < k; i++) out[i * k + i] = 1;
        for (int r = 0; r < m; r++) memcpy(out + (k + r) * k, custom + r * custom_cols, k);
        return RSREF_OK;
    }
    if (kind == RSREF_MATRIX_XOR && m != 1) kind = RSREF_MATRIX_DEFAULT;
    return rsref_build_matrix(kind, k, total, out);
}

/* ------------------------------------------------------------------------ */
/* reedsolomon.go: coding                                                   */
/* ------------------------------------------------------------------------ */

/* reedsolomon.go:882 codeSomeShards: out[i] (^)= sum_j rows[i][j] * in[j]. */
void rsref_code_some_shards(const uint8_t *rows, int nout, int nin, const uint8_t *const *in,
                            uint8_t *const *out, size_t len, int clear) {
    gf8_init();
    for (int i = 0; i < nout; i++) {
        uint8_t *o = out[i];
        if (clear) memset(o, 0, len);
        for (int j = 0; j < nin; j++) {
            const uint8_t *mt = g_mul[rows[i * nin + j]];
            const uint8_t *x = in[j];
            for (size_t b = 0; b < len; b++) o[b] ^= mt[x[b]];
        }
    }
}

/* reedsolomon.go:1404-1428 checkShards / shardSize. */
static size_t shard_size(const size_t *lens, int n) {
    for (int i = 0; i < n; i++)
        if (lens[i]) return lens[i];
    return 0;
}
static int check_shards(const size_t *lens, int n, int nilok) {
    size_t size = shard_size(lens, n);
    if (size == 0) return RSREF_ERR_SHARD_NO_DATA;
    for (int i = 0; i < n; i++)
        if (lens[i] != size && (lens[i] != 0 || !nilok)) return RSREF_ERR_SHARD_SIZE;
    return RSREF_OK;
}

/* reedsolomon.go:686-702 Encode. */
int rsref_encode(const uint8_t *matrix, int k, int m, uint8_t *const *shards, const size_t *lens,
                 int nshards) {
    if (nshards != k + m) return RSREF_ERR_TOO_FEW_SHARDS;
    int err = check_shards(lens, nshards, 0);
    if (err) return err;
    rsref_code_some_shards(matrix + (size_t)k * k, m, k, (const uint8_t *const *)shards,
                           shards + k, lens[0], 1);
    return RSREF_OK;
}

/* reedsolomon.go:708-757 EncodeIdx: parity ^= column idx of the parity rows * data. */
int rsref_encode_idx(const uint8_t *matrix, int k, int m, const uint8_t *data, size_t len, int idx,
                     uint8_t *const *parity, const size_t *plens, int nparity) {
    if (nparity != m) return RSREF_ERR_TOO_FEW_SHARDS;
    if (nparity == 0) return RSREF_OK;
    if (idx < 0 || idx >= k) return RSREF_ERR_INV_SHARD_NUM;
    int err = check_shards(plens, nparity, 0);
    if (err) return err;
    if (plens[0] != len) return RSREF_ERR_SHARD_SIZE;
    for (int i = 0; i < m; i++)
        rsref_gal_mul_slice(matrix[(size_t)(k + i) * k + idx], data, parity[i], len, 1);
    return RSREF_OK;
}

/* reedsolomon.go:857-871 Verify + 1373-1387 checkSomeShards. */
int rsref_verify(const uint8_t *matrix, int k, int m, uint8_t *const *shards, const size_t *lens,
                 int nshards, int *ok) {
    *ok = 0;
    if (nshards != k + m) return RSREF_ERR_TOO_FEW_SHARDS;
    int err = check_shards(lens, nshards, 0);
    if (err) return err;
    size_t len = lens[0];
    uint8_t **tmp = malloc(sizeof(uint8_t *) * (m ? m : 1));
    for (int i = 0; i < m; i++) tmp[i] = malloc(len);
    rsref_code_some_shards(matrix + (size_t)k * k, m, k, (const uint8_t *const *)shards, tmp, len, 1);
    *ok = 1;
    for (int i = 0; i < m; i++) {
        if (*ok && memcmp(tmp[i], shards[k + i], len) != 0) *ok = 0;
        free(tmp[i]);
    }
    free(tmp);
    return RSREF_OK;
}

/* reedsolomon.go:1670-1781 reconstruct (+ getDecodeMatrix 634-673,
 * multiplyRowWithMatrix 620-631).  The inversion cache only memoises the
 * inverse, so it is not restated. */
int rsref_reconstruct(const uint8_t *matrix, int k, int m, uint8_t *const *shards,
                      const size_t *lens, int nshards, int data_only, const uint8_t *required,
                      int nrequired, uint8_t *filled) {
    int total = k + m;
    for (int i = 0; i < nshards; i++) filled[i] = 0;
    if (nshards != total || (required && nrequired < k)) return RSREF_ERR_TOO_FEW_SHARDS;
    int err = check_shards(lens, nshards, 1);
    if (err) return err;
    size_t size = shard_size(lens, nshards);
    int present = 0, data_present = 0, missing_required = 0;
    for (int i = 0; i < total; i++) {
        if (lens[i]) {
            present++;
            if (i < k) data_present++;
        } else if (required && required[i]) {
            missing_required++;
        }
    }
    if (present == total || (data_only && data_present == k) || (required && missing_required == 0))
        return RSREF_OK;
    if (present < k) return RSREF_ERR_TOO_FEW_SHARDS;

    int *valid = malloc(sizeof(int) * k);
    const uint8_t **sub = malloc(sizeof(uint8_t *) * k);
    int nv = 0;
    for (int r = 0; r < total && nv < k; r++)
        if (lens[r]) {
            sub[nv] = shards[r];
            valid[nv++] = r;
        }
    uint8_t *subm = malloc((size_t)k * k), *inv = malloc((size_t)k * k);
    for (int r = 0; r < k; r++) memcpy(subm + r * k, matrix + (size_t)valid[r] * k, k);
    err = rsref_matrix_invert(subm, k, k, inv);
    if (err) goto out;

    uint8_t *rows = malloc((size_t)total * k);
    uint8_t **outs = malloc(sizeof(uint8_t *) * total);
    int nout = 0;
    for (int i = 0; i < k; i++)
        if (!lens[i] && (!required || required[i])) {
            memcpy(rows + nout * k, inv + i * k, k);
            outs[nout++] = shards[i];
            filled[i] = 1;
        }
    if (!data_only)
        for (int i = k; i < total; i++)
            if (!lens[i] && (!required || required[i])) {
                /* multiplyRowWithMatrix: result[j] = sum_i row[i] * inv[i][j] */
                const uint8_t *prow = matrix + (size_t)i * k;
                for (int j = 0; j < k; j++) {
                    uint8_t v = 0;
                    for (int t = 0; t < k; t++) v ^= g_mul[prow[t]][inv[t * k + j]];
                    rows[nout * k + j] = v;
                }
                outs[nout++] = shards[i];
                filled[i] = 1;
            }
    if (nout) rsref_code_some_shards(rows, nout, k, sub, outs, size, 1);
    free(rows);
    free(outs);
out:
    free(valid);
    free(sub);
    free(subm);
    free(inv);
    return err;
}

/* ------------------------------------------------------------------------ */
/* reedsolomon.go:762-818 Update / updateParityShards.  A NULL pointer is    */
/* Go's nil.  Each changed old data shard becomes old XOR new, as there.    */
/* ------------------------------------------------------------------------ */
int rsref_update(const uint8_t *matrix, int k, int m, uint8_t *const *shards, const size_t *lens,
                 int nshards, uint8_t *const *newd, const size_t *nlens, int nnew) {
    if (nshards != k + m) return RSREF_ERR_TOO_FEW_SHARDS;
    if (nnew != k) return RSREF_ERR_TOO_FEW_SHARDS;
    int err = check_shards(lens, nshards, 1);
    if (err) return err;
    err = check_shards(nlens, nnew, 1);
    if (err) return err;
    for (int i = 0; i < k; i++)
        if (newd[i] && !shards[i]) return RSREF_ERR_INVALID_INPUT;
    for (int i = k; i < k + m; i++)
        if (!shards[i]) return RSREF_ERR_INVALID_INPUT;
    size_t size = shard_size(lens, nshards);
    /* The reference slices past shorter buffers here; the restatement
     * rejects mismatched lengths instead. */
    for (int i = 0; i < k; i++)
        if (newd[i] && (nlens[i] != size || lens[i] != size)) return RSREF_ERR_SHARD_SIZE;
    for (int i = k; i < k + m; i++)
        if (lens[i] != size) return RSREF_ERR_SHARD_SIZE;
    for (int c = 0; c < k; c++) {
        if (!newd[c]) continue;
        for (size_t b = 0; b < size; b++) shards[c][b] ^= newd[c][b];
        for (int r = 0; r < m; r++)
            rsref_gal_mul_slice(matrix[(size_t)(k + r) * k + c], shards[c], shards[k + r], size, 1);
    }
    return RSREF_OK;
}

/* ------------------------------------------------------------------------ */
/* reedsolomon.go:1464-1617 DecodeIdx.  A NULL pointer is Go's nil;          */
/* expect == NULL is the merge mode.                                         */
/* ------------------------------------------------------------------------ */
int rsref_decode_idx(const uint8_t *matrix, int k, int m, uint8_t *const *dst, const size_t *dlens,
                     int ndst, const uint8_t *expect, int nexpect, uint8_t *const *input,
                     const size_t *ilens, int ninput) {
    int total = k + m;
    if (!expect) {
        if (ndst != ninput) return RSREF_ERR_INVALID_INPUT;
        for (int i = 0; i < ndst; i++) {
            if (!input[i]) continue;
            if (!dst[i]) return RSREF_ERR_INVALID_INPUT;
            if (dlens[i] != ilens[i]) return RSREF_ERR_INVALID_SHARD_SIZE;
            for (size_t b = 0; b < ilens[i]; b++) dst[i][b] ^= input[i][b];
        }
        return RSREF_OK;
    }
    if (nexpect != total || ndst != total || ninput != total) return RSREF_ERR_INVALID_INPUT;
    for (int i = 0; i < total; i++)
        if (input[i] && !expect[i]) return RSREF_ERR_INVALID_INPUT;
    for (int i = 0; i < total; i++)
        if (expect[i] && dst[i]) return RSREF_ERR_INVALID_INPUT;
    int *valid = malloc(sizeof(int) * total);
    int nvalid = 0;
    for (int i = 0; i < total; i++)
        if (expect[i]) valid[nvalid++] = i;
    if (nvalid < k) {
        free(valid);
        return RSREF_ERR_TOO_FEW_SHARDS;
    }
    /* getDecodeMatrix: inverse of the rows of the first k valid indices. */
    uint8_t *subm = malloc((size_t)k * k), *inv = malloc((size_t)k * k);
    for (int r = 0; r < k; r++) memcpy(subm + r * k, matrix + (size_t)valid[r] * k, k);
    int err = rsref_matrix_invert(subm, k, k, inv);
    free(subm);
    size_t size = 0;
    int have = 0;
    if (!err) {
        for (int i = 0; i < total && !err; i++) {
            if (!input[i]) continue;
            if (!have) {
                size = ilens[i];
                have = size != 0;
            } else if (ilens[i] != size) {
                err = RSREF_ERR_INVALID_SHARD_SIZE;
            }
        }
        for (int i = 0; i < total && !err; i++)
            if (dst[i] && dlens[i] != size) err = RSREF_ERR_INVALID_SHARD_SIZE;
    }
    if (err) {
        free(valid);
        free(inv);
        return err;
    }
    /* rows: one per non-nil dst, nvalid columns (zero past column k). */
    uint8_t *rows = calloc((size_t)total * nvalid, 1);
    uint8_t **outs = malloc(sizeof(uint8_t *) * total);
    int nout = 0;
    for (int d = 0; d < total; d++) {
        if (!dst[d]) continue;
        uint8_t *row = rows + (size_t)nout * nvalid;
        if (d < k) {
            memcpy(row, inv + (size_t)d * k, k);
        } else {
            const uint8_t *prow = matrix + (size_t)d * k;
            for (int col = 0; col < nvalid && col < k; col++) {
                uint8_t v = 0;
                for (int i = 0; i < k; i++) v ^= g_mul[prow[i]][inv[i * k + col]];
                row[col] = v;
            }
        }
        outs[nout++] = dst[d];
    }
    if (nout) {
        const uint8_t **ins = malloc(sizeof(uint8_t *) * nvalid);
        uint8_t *adj = malloc((size_t)nout * nvalid + 1);
        int nin = 0;
        int *cols = malloc(sizeof(int) * nvalid);
        for (int col = 0; col < nvalid; col++)
            if (input[valid[col]]) {
                ins[nin] = input[valid[col]];
                cols[nin++] = col;
            }
        for (int j = 0; j < nin; j++)  /* an empty input beside sized ones (a panic in the reference) */
            if (ilens[valid[cols[j]]] != size) err = RSREF_ERR_INVALID_SHARD_SIZE;
        if (nin && !err) {
            for (int o = 0; o < nout; o++)
                for (int j = 0; j < nin; j++) adj[o * nin + j] = rows[(size_t)o * nvalid + cols[j]];
            rsref_code_some_shards(adj, nout, nin, ins, outs, size, 0);
        }
        free(ins);
        free(adj);
        free(cols);
    }
    free(rows);
    free(outs);
    free(valid);
    free(inv);
    return err;
}

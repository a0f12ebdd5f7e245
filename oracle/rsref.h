/*
 * rsref — CPU restatement of klauspost/reedsolomon's coding path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * MI355X engine: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product library (reedsolomon_amd/librsx.so)
 * never links or calls it.
 *
 * Every function cites the reference file:line it restates.  All buffers are
 * host memory.  Return codes use the same numbering as include/rsx.h so the
 * tests can compare error outcomes directly.
 */
#ifndef RSREF_H
#define RSREF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSREF_OK 0
#define RSREF_ERR_INV_SHARD_NUM -1
#define RSREF_ERR_MAX_SHARD_NUM -2
#define RSREF_ERR_NOT_SUPPORTED -3
#define RSREF_ERR_TOO_FEW_SHARDS -4
#define RSREF_ERR_INVALID_INPUT -5
#define RSREF_ERR_SHARD_NO_DATA -6
#define RSREF_ERR_SHARD_SIZE -7
#define RSREF_ERR_INVALID_SHARD_SIZE -8
#define RSREF_ERR_SINGULAR -11
#define RSREF_ERR_MATRIX -12
#define RSREF_ERR_NOT_SQUARE -13

/* Matrix kinds (options.go WithCauchyMatrix / WithPAR1Matrix / WithJerasureMatrix /
 * WithFastOneParityMatrix; default = Vandermonde·inv(top)). */
#define RSREF_MATRIX_DEFAULT 0
#define RSREF_MATRIX_CAUCHY 1
#define RSREF_MATRIX_PAR1 2
#define RSREF_MATRIX_JERASURE 3
#define RSREF_MATRIX_XOR 4

/* ---- GF(2^8), classic representation (galois.go) ---- */
void rsref_gf8_tables(uint8_t *exp_tab, uint8_t *log_tab, uint8_t *inv_tab,
                      uint8_t *mul_tab /*256*256*/, uint8_t *mul_low /*256*16*/,
                      uint8_t *mul_high /*256*16*/, uint64_t *gfni /*256*/);
uint8_t rsref_gal_mul(uint8_t a, uint8_t b);
uint8_t rsref_gal_exp(uint8_t a, int n);
uint8_t rsref_gal_divide(uint8_t a, uint8_t b);
void rsref_gal_mul_slice(uint8_t c, const uint8_t *in, uint8_t *out, size_t n, int do_xor);

/* ---- matrix.go ---- */
int rsref_matrix_multiply(const uint8_t *a, int ar, int ac, const uint8_t *b, int br, int bc,
                          uint8_t *out);
int rsref_matrix_invert(const uint8_t *m, int rows, int cols, uint8_t *out);
int rsref_build_matrix(int kind, int data_shards, int total_shards, uint8_t *out);

/* ---- classic codec (reedsolomon.go) ---- */
/* rows: nout x nin coefficient matrix, row-major. */
void rsref_code_some_shards(const uint8_t *rows, int nout, int nin, const uint8_t *const *in,
                            uint8_t *const *out, size_t len, int clear);
/* Encoding matrix (k+m) x k for New(k, m, ...); custom != NULL replaces the
 * parity rows (custom_rows x k). */
int rsref_codec_matrix(int k, int m, int kind, const uint8_t *custom, int custom_rows,
                       int custom_cols, uint8_t *out);
/* lens follow Go's len(): 0 = absent shard.  For reconstruct, every absent
 * shard pointer must reference a buffer of at least the shard size; filled[i]
 * is set to 1 for every shard written. */
int rsref_encode(const uint8_t *matrix, int k, int m, uint8_t *const *shards, const size_t *lens,
                 int nshards);
int rsref_encode_idx(const uint8_t *matrix, int k, int m, const uint8_t *data, size_t len, int idx,
                     uint8_t *const *parity, const size_t *plens, int nparity);
int rsref_verify(const uint8_t *matrix, int k, int m, uint8_t *const *shards, const size_t *lens,
                 int nshards, int *ok);
int rsref_reconstruct(const uint8_t *matrix, int k, int m, uint8_t *const *shards,
                      const size_t *lens, int nshards, int data_only, const uint8_t *required,
                      int nrequired, uint8_t *filled);
/* Update / DecodeIdx: a NULL shard pointer is Go's nil. */
int rsref_update(const uint8_t *matrix, int k, int m, uint8_t *const *shards, const size_t *lens,
                 int nshards, uint8_t *const *newd, const size_t *nlens, int nnew);
int rsref_decode_idx(const uint8_t *matrix, int k, int m, uint8_t *const *dst, const size_t *dlens,
                     int ndst, const uint8_t *expect, int nexpect, uint8_t *const *input,
                     const size_t *ilens, int ninput);

/* Fast path of the classic product for the CPU baseline (rsref_fast.c):
 * isa 0 scalar, 1 AVX2 PSHUFB, 2 AVX-512+GFNI, -1 = best available. */
int rsref_fast_isa(void);
void rsref_code_some_shards_fast(const uint8_t *rows, int nout, int nin, const uint8_t *const *in,
                                 uint8_t *const *out, size_t len, int threads, int isa);

/* ---- Leopard GF(2^16) (leopard.go) ---- */
void rsref_leo16_tables(uint16_t *exp_lut, uint16_t *log_lut, uint16_t *fft_skew /*65535*/,
                        uint16_t *log_walsh);
uint16_t rsref_leo16_mul_log(uint16_t a, uint16_t log_b);
uint16_t rsref_leo16_add_mod(uint16_t a, uint16_t b);
uint16_t rsref_gf16_mul(uint16_t a, uint16_t b);
int rsref_leo16_encode(int k, int m, uint8_t *const *shards, const size_t *lens, int nshards);
int rsref_leo16_verify(int k, int m, uint8_t *const *shards, const size_t *lens, int nshards,
                       int *ok);
int rsref_leo16_reconstruct(int k, int m, uint8_t *const *shards, const size_t *lens,
                            int nshards, int recover_all, uint8_t *filled);
/* error locator logs for a given erasure mask (1 = erased), k+m entries */
int rsref_leo16_errlocs(int k, int m, const uint8_t *erased, int recover_all, uint16_t *out);

/* ---- Leopard GF(2^8) (leopard8.go) ---- */
void rsref_leo8_tables(uint8_t *exp_lut, uint8_t *log_lut, uint8_t *fft_skew /*255*/,
                       uint8_t *log_walsh);
uint8_t rsref_leo8_mul_log(uint8_t a, uint8_t log_b);
uint8_t rsref_leo8_add_mod(uint8_t a, uint8_t b);
int rsref_leo8_encode(int k, int m, uint8_t *const *shards, const size_t *lens, int nshards);
int rsref_leo8_verify(int k, int m, uint8_t *const *shards, const size_t *lens, int nshards,
                      int *ok);
int rsref_leo8_reconstruct(int k, int m, uint8_t *const *shards, const size_t *lens, int nshards,
                           int recover_all, uint8_t *filled);

#ifdef __cplusplus
}
#endif
#endif

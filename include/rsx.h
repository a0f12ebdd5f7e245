/*
 * rsx.h — C ABI of the MI355X erasure-coding engine (librsx.so).
 *
 * The entry points are the drop-in boundary for klauspost/reedsolomon's
 * Encoder / Extensions API (reedsolomon.go:25-165): a thin cgo shim (or any
 * FFI) binds them one to one.  Semantics, argument meaning, validation order
 * and error outcomes follow the Go methods cited on each declaration.
 *
 * Shards are passed Go-style: an array of pointers plus an array of lengths,
 * where length 0 marks an absent shard (a nil or zero-length slice).
 * All shards of one call live either in device memory (HBM; the call is then
 * stream-ordered on the codec's stream and returns without waiting) or in
 * host memory (staged through HBM; the call returns when the results are in
 * host memory).  Mixing the two in one call is rejected.
 *
 * The library never allocates caller memory.  For Reconstruct*, the caller
 * passes, for every absent shard, a buffer of at least the shard size (the Go
 * shim applies the reference's "reuse capacity else allocate" rule before the
 * call); the library writes the shards it reconstructs and reports them in
 * `filled`.
 */
#ifndef RSX_H
#define RSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes: 0 or one negative code per reference sentinel error. */
#define RSX_OK 0
#define RSX_ERR_INV_SHARD_NUM -1        /* ErrInvShardNum       reedsolomon.go:216 */
#define RSX_ERR_MAX_SHARD_NUM -2        /* ErrMaxShardNum       reedsolomon.go:221 */
#define RSX_ERR_NOT_SUPPORTED -3        /* ErrNotSupported      reedsolomon.go:224 */
#define RSX_ERR_TOO_FEW_SHARDS -4       /* ErrTooFewShards      reedsolomon.go:678 */
#define RSX_ERR_INVALID_INPUT -5        /* ErrInvalidInput      reedsolomon.go:760 */
#define RSX_ERR_SHARD_NO_DATA -6        /* ErrShardNoData       reedsolomon.go:1391 */
#define RSX_ERR_SHARD_SIZE -7           /* ErrShardSize         reedsolomon.go:1395 */
#define RSX_ERR_INVALID_SHARD_SIZE -8   /* ErrInvalidShardSize  reedsolomon.go:1399 */
#define RSX_ERR_SHORT_DATA -9           /* ErrShortData         reedsolomon.go:1785 */
#define RSX_ERR_RECONSTRUCT_REQUIRED -10 /* ErrReconstructRequired reedsolomon.go:1862 */
#define RSX_ERR_SINGULAR -11            /* errSingular          matrix.go:184 */
#define RSX_ERR_MATRIX -12              /* custom-matrix shape errors of New, reedsolomon.go:466-474 */
#define RSX_ERR_RECONSTRUCT_MISMATCH -14 /* ErrReconstructMismatch streaming.go:461 */
#define RSX_ERR_STREAM_READ -15         /* StreamReadError  streaming.go:98 (stream index reported) */
#define RSX_ERR_STREAM_WRITE -16        /* StreamWriteError streaming.go:116 (stream index reported) */
#define RSX_ERR_DEVICE -20              /* HIP runtime error (no reference counterpart) */
#define RSX_ERR_NOMEM -21               /* device allocation failed (no reference counterpart) */
#define RSX_ERR_MIXED_MEMORY -22        /* host and device shards in one call (no counterpart) */
#define RSX_ERR_CANCELLED -23           /* request-loop request cancelled (no counterpart) */
#define RSX_ERR_OVER_BUDGET -24         /* request-loop admission rejected (no counterpart) */

/* Engine selection, options.go:283-305 (WithLeopardGF16 / WithLeopardGF). */
#define RSX_LEOPARD_AS_NEEDED 0
#define RSX_LEOPARD_GF16 1
#define RSX_LEOPARD_GF8 2

/* Matrix selection, options.go:225-276. */
#define RSX_MATRIX_DEFAULT 0  /* Vandermonde * inv(top), reedsolomon.go:232 */
#define RSX_MATRIX_CAUCHY 1   /* WithCauchyMatrix */
#define RSX_MATRIX_PAR1 2     /* WithPAR1Matrix */
#define RSX_MATRIX_JERASURE 3 /* WithJerasureMatrix */

typedef struct rsx_codec rsx_codec;

/* The options of options.go that change engine, matrix or output bytes, plus
 * the new device option.  Options that only tune CPU goroutines/caches
 * (WithMaxGoroutines, WithMinSplitSize, WithAutoGoroutines, WithSSSE3, ...)
 * have no meaning here and are accepted by the shim as no-ops. */
typedef struct rsx_options {
    int leopard;               /* RSX_LEOPARD_*                                   */
    int matrix;                /* RSX_MATRIX_*                                    */
    int fast_one_parity;       /* WithFastOneParityMatrix      options.go:260     */
    int inversion_cache;       /* WithInversionCache (default 1) options.go:152;
                                  0: every reconstruct inverts afresh, no decode
                                  plan is kept in HBM (classic engine)           */
    const uint8_t *custom_matrix; /* WithCustomMatrix, row-major custom_rows x custom_cols */
    int custom_rows;
    int custom_cols;
    int device;                /* HIP device ordinal: the new WithMI355X option   */
} rsx_options;

/* Fills *o with the reference defaults (options.go:46-62) on device 0. */
void rsx_options_default(rsx_options *o);

/* New(dataShards, parityShards, opts...) reedsolomon.go:425-601.
 * Engine routing as reedsolomon.go:434-443; Leopard GF(2^8) additionally
 * rejects k + ceilPow2(m) > 256 with RSX_ERR_MAX_SHARD_NUM where the
 * reference indexes past its tables. */
int rsx_new(int data_shards, int parity_shards, const rsx_options *opt, rsx_codec **out);
void rsx_free(rsx_codec *c);

/* Extensions, reedsolomon.go:137-165. */
int rsx_shard_size_multiple(const rsx_codec *c); /* 1 classic, 64 Leopard */
int rsx_data_shards(const rsx_codec *c);
int rsx_parity_shards(const rsx_codec *c);
int rsx_total_shards(const rsx_codec *c);
/* 0 classic, RSX_LEOPARD_GF16 or RSX_LEOPARD_GF8: the engine New chose. */
int rsx_engine(const rsx_codec *c);
/* Copies the (k+m) x k coding matrix (classic engine only). */
int rsx_coding_matrix(const rsx_codec *c, uint8_t *out);

/* Encoder.Encode, reedsolomon.go:686 (classic), leopard.go:139, leopard8.go:120. */
int rsx_encode(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards);

/* Encoder.EncodeIdx, reedsolomon.go:708 (Leopard: ErrNotSupported). */
int rsx_encode_idx(rsx_codec *c, const uint8_t *data, size_t len, int idx, uint8_t *const *parity,
                   const size_t *parity_lens, int nparity);

/* Encoder.Verify, reedsolomon.go:857, leopard.go:393, leopard8.go:365.
 * Always synchronous: *ok is valid on return. */
int rsx_verify(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards, int *ok);

/* Encoder.Reconstruct / ReconstructData / ReconstructSome,
 * reedsolomon.go:1446 / 1634 / 1655, leopard.go:378-391, leopard8.go:350-363.
 * filled[i] (nshards entries) is set to 1 for each shard written. */
int rsx_reconstruct(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
                    uint8_t *filled);
int rsx_reconstruct_data(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
                         uint8_t *filled);
/* required has nrequired entries: k (ReconstructData of the required data
 * shards) or TotalShards (any shard), else ErrInvalidInput.  With k entries a
 * missing parity shard counts as not required (the reference indexes past a
 * k-entry slice there and panics). */
int rsx_reconstruct_some(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
                         const uint8_t *required, int nrequired, uint8_t *filled);

/* No counterpart in the reference.  Reconstruct (data_only = 0) or
 * ReconstructData (data_only = 1) of the classic engine, same results as
 * rsx_reconstruct / rsx_reconstruct_data, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the call allocates nothing and copies nothing between host and
 * device.  The decode tables go to a fixed 64 KiB device array per GPU (held
 * until the process ends), so k x (shards rebuilt) must be <= 3276, else
 * ErrNotSupported.  Shards must be device memory (else ErrMixedMemory) and
 * the call is stream-ordered on the codec's stream, like rsx_reconstruct.
 * Calls on one GPU whose work overlaps (different codecs on different
 * streams) share that array and race: order them.  Leopard codecs:
 * ErrNotSupported. */
int rsx_reconstruct_noalloc(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
                            int data_only, uint8_t *filled);

/* Encoder.Update, reedsolomon.go:762-818 (Leopard: ErrNotSupported,
 * leopard.go:316).  shards: old data (NULL where unchanged) then the parity,
 * updated in place; new_data: k entries, NULL where unchanged.  A NULL
 * pointer is Go's nil.  As in the reference, each changed old data shard is
 * overwritten with old XOR new.  Lengths of the changed shards and of the
 * parity must equal the shard size (the reference reads past shorter ones;
 * here that is ErrShardSize). */
int rsx_update(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
               uint8_t *const *new_data, const size_t *new_lens, int nnew);

/* Extensions.DecodeIdx, reedsolomon.go:1464-1617 (Leopard: ErrNotSupported).
 * Progressive decode: dst[i] ^= (row i of the decode matrix) * input for the
 * inputs given in this call, the decode matrix built from the first k shards
 * marked in expect_input.  expect_input == NULL is the merge mode
 * (dst[i] ^= input[i]).  NULL pointers are nil.  The reference's joined
 * errors map to their sentinel (ErrInvalidInput, ErrInvalidShardSize,
 * ErrTooFewShards). */
int rsx_decode_idx(rsx_codec *c, uint8_t *const *dst, const size_t *dst_lens, int ndst,
                   const uint8_t *expect_input, int nexpect, uint8_t *const *input,
                   const size_t *input_lens, int ninput);

/* Stream control for device-resident calls (no reference counterpart: the Go
 * methods are synchronous; the shim calls rsx_synchronize before returning
 * when it hands over device buffers it must read). */
int rsx_set_stream(rsx_codec *c, void *hip_stream);
void *rsx_get_stream(rsx_codec *c);
int rsx_synchronize(rsx_codec *c);

/* Number of erasure patterns held in the device-resident decode cache
 * (inversion_tree.go analogue). */
int rsx_decode_cache_size(const rsx_codec *c);

/* ---- Batched path -------------------------------------------------------
 * Many Encode / Verify / Reconstruct / ReconstructData calls planned once
 * into a device work list and run as persistent launches captured in a HIP
 * graph that rsx_batch_launch replays.  Classic codecs (any (k, m), shard
 * size and erasure pattern): one launch per output-count class.  Leopard
 * codecs (leopard.go:139 / 393 / 421, leopard8.go likewise): one LDS-resident
 * launch per field, for 16-byte aligned shards whose rows fit one LDS tile
 * (GF(2^16) encodes and Verifies above the tiles and reconstructs with
 * n >= 8192 work rows run three-pass launches over all such items, their
 * work arrays within 4 GiB per plan); other Leopard items get RSX_ERR_NOT_SUPPORTED and
 * take the per-stripe call.  There is no single reference counterpart: each
 * item stands for one call of reedsolomon.go:686 / 857 / 1446 / 1634 and is
 * validated as that call is; status[i] receives its result.  Items of one
 * plan must be independent (no item may read a shard another item writes),
 * shards must be device memory, and the codecs must outlive the plan.
 * Outputs are overwritten, so a plan may be launched repeatedly. */
#define RSX_OP_ENCODE 1
#define RSX_OP_RECONSTRUCT 2
#define RSX_OP_RECONSTRUCT_DATA 3
#define RSX_OP_VERIFY 4           /* parity recomputed in the pass and compared with the item's
                                     parity shards; nothing is written */

typedef struct rsx_batch_item {
    rsx_codec *codec;       /* codec from rsx_new */
    int op;                 /* RSX_OP_* */
    int nshards;            /* entries of shards / lens (the codec's TotalShards) */
    uint8_t *const *shards; /* device memory; absent shards carry an output buffer */
    const size_t *lens;     /* Go-style lengths, 0 = absent */
} rsx_batch_item;

typedef struct rsx_batch rsx_batch;

int rsx_batch_create(int device, rsx_batch **out);
void rsx_batch_free(rsx_batch *b);
/* filled receives, item after item, TotalShards flags per item.  Re-planning
 * waits for launches on rsx_batch_stream(b); launches made with
 * rsx_batch_launch_on must have completed before rsx_batch_plan or
 * rsx_batch_free is called. */
int rsx_batch_plan(rsx_batch *b, const rsx_batch_item *items, int nitems, int *status,
                   uint8_t *filled);
int rsx_batch_launch(rsx_batch *b);       /* stream-ordered on rsx_batch_stream(b) */
/* Replays the plan on another stream (e.g. to order two plans). */
int rsx_batch_launch_on(rsx_batch *b, void *hip_stream);
int rsx_batch_synchronize(rsx_batch *b);
void *rsx_batch_stream(rsx_batch *b);
int rsx_batch_num_tasks(const rsx_batch *b);
/* Result of a verify item, valid once the launch has completed. */
int rsx_batch_verify_result(const rsx_batch *b, int item, int *ok);
/* 1 (default): plans are captured into a HIP graph and launches replay it;
 * 0: launches enqueue the kernels directly (no per-plan instantiate cost). */
int rsx_batch_set_graph(rsx_batch *b, int enabled);
/* Tile bytes per work unit and resident blocks per CU; 0, 0 = autotune at
 * the next plan (the choice is kept for later plans of this batch).  An
 * autotuning rsx_batch_plan times the plan by running it a few times, so
 * during that call it reads the items' inputs and writes their outputs (the
 * bytes a launch writes); set the tuning first where that is not wanted. */
int rsx_batch_set_tuning(rsx_batch *b, int tile_bytes, int blocks_per_cu);
int rsx_batch_get_tuning(const rsx_batch *b, int *tile_bytes, int *blocks_per_cu);

/* ---- Serving ------------------------------------------------------------
 * A request loop with continuous batching: requests (Encode, Verify,
 * Reconstruct, ReconstructData of any codec, shards in device memory) are
 * admitted in arrival order while the bytes in flight stay within an HBM
 * budget; a scheduler thread plans whatever has been admitted into a new
 * batch whenever one of its launch slots (each with its own stream) is free,
 * so batches overlap on the device; Leopard requests the batch kernel does
 * not take run through the per-stripe call on a worker thread.  Latency is measured from arrival to completion.  Requests in
 * flight at the same time must not share shards.  No reference counterpart:
 * each request is the call of reedsolomon.go it names, with that call's
 * validation and result. */
typedef struct rsx_request {
    rsx_codec *codec;
    int op;                 /* RSX_OP_* */
    int nshards;
    uint8_t *const *shards; /* copied at submit; buffers stay the caller's */
    const size_t *lens;
} rsx_request;

typedef struct rsx_server rsx_server;

/* hbm_budget: bytes of shards in flight (0 = 4 GiB; a request larger than
 * the budget is rejected at submit with RSX_ERR_OVER_BUDGET; one batch takes
 * at most budget / 8); max_batch: requests per launch (0 = 4096). */
int rsx_server_create(int device, size_t hbm_budget, int max_batch, rsx_server **out);
void rsx_server_free(rsx_server *s); /* drains first */
int rsx_server_submit(rsx_server *s, const rsx_request *req, uint64_t *ticket);
/* *done = 1 once complete; then *status, *verify_ok and *latency_ns are set
 * and the ticket is released (report each ticket once: poll or wait). */
int rsx_server_poll(rsx_server *s, uint64_t ticket, int *done, int *status, int *verify_ok,
                    uint64_t *latency_ns);
int rsx_server_wait(rsx_server *s, uint64_t ticket, int *status, int *verify_ok,
                    uint64_t *latency_ns);
/* Waits until every submitted request has completed.  On a paused server
 * with requests queued it returns RSX_ERR_INVALID_INPUT instead of waiting. */
int rsx_server_drain(rsx_server *s);
/* Cancels a request that has not been admitted yet: *cancelled = 1 and the
 * ticket completes with status RSX_ERR_CANCELLED (its shards untouched; it
 * is still reported once by poll / wait).  A request already admitted into a
 * batch, or complete, runs to its normal result: *cancelled = 0. */
int rsx_server_cancel(rsx_server *s, uint64_t ticket, int *cancelled);
/* paused != 0: the scheduler admits nothing new (batches already planned
 * finish); 0 resumes.  Lets a caller hold requests queued, e.g. to cancel
 * them.  rsx_server_free resumes before draining. */
int rsx_server_pause(rsx_server *s, int paused);
/* Rejects, with RSX_ERR_OVER_BUDGET at submit, a request that would take the
 * bytes queued and not yet admitted past max_queued_bytes (0 = no limit,
 * the default). */
int rsx_server_set_queue_limit(rsx_server *s, size_t max_queued_bytes);
/* Requests rejected at submit, cancelled, and shard bytes queued now. */
int rsx_server_admission(const rsx_server *s, uint64_t *rejected, uint64_t *cancelled,
                         uint64_t *queued_bytes);
/* Replays a trace: request i arrives arrival_ns[i] after the start (sorted);
 * fills status / verify_ok / latency_ns per request and the makespan.  A
 * request rejected at submit gets that error as its status (latency 0). */
int rsx_server_run_trace(rsx_server *s, const rsx_request *reqs, const uint64_t *arrival_ns, int n,
                         int *status, int *verify_ok, uint64_t *latency_ns, uint64_t *makespan_ns);
/* Counters: batched launches, requests in them, highest shard bytes in
 * flight, host time spent planning and launching batches, and device time of
 * the batches (sum over batches, first kernel start to last kernel end;
 * batches that overlap count their overlap twice).  plan_ns / device_ns may
 * be NULL. */
int rsx_server_stats(const rsx_server *s, uint64_t *batches, uint64_t *requests,
                     uint64_t *max_inflight_bytes, uint64_t *plan_ns, uint64_t *device_ns);

/* ---- Streaming (streaming.go) ---------------------------------------------
 * StreamEncoder of NewStream (streaming.go:147): shards are read and written
 * through callbacks, block by block (block_size, 0 = 4 MiB), with reads of
 * the next block overlapping the GPU work on the current one.
 * A read callback fills buf with up to n bytes of shard `shard` and returns
 * the count: fewer than n only at the end of the stream (io.ReadFull), < 0
 * on error.  A write callback returns the bytes written (< n or < 0 is an
 * error).  On RSX_ERR_STREAM_READ / _WRITE, *err_stream is the stream index
 * (StreamReadError / StreamWriteError.Stream).  Classic engine only. */
typedef long long (*rsx_stream_read_fn)(void *ctx, int shard, uint8_t *buf, size_t n);
typedef long long (*rsx_stream_write_fn)(void *ctx, int shard, const uint8_t *buf, size_t n);
typedef struct rsx_stream rsx_stream;

int rsx_stream_create(int data_shards, int parity_shards, const rsx_options *opt, size_t block_size,
                      rsx_stream **out);
void rsx_stream_free(rsx_stream *s);
rsx_codec *rsx_stream_codec(rsx_stream *s);
/* WithConcurrentStreamReads / WithConcurrentStreamWrites (options.go:122-146,
 * NewStreamC streaming.go:190): the read (write) callbacks of a block run for
 * all shards at once on a pool of threads, so they must be safe to call for
 * different shards concurrently.  Of several failing streams the lowest index
 * is reported.  Off by default, as in the reference. */
int rsx_stream_set_concurrency(rsx_stream *s, int reads, int writes);
/* rsStream.Encode, streaming.go:199: reads shards 0..k-1, writes parity 0..m-1. */
int rsx_stream_encode(rsx_stream *s, rsx_stream_read_fn rd, void *rctx, rsx_stream_write_fn wr, void *wctx,
                      int *err_stream);
/* rsStream.Verify, streaming.go:430: reads all TotalShards streams. */
int rsx_stream_verify(rsx_stream *s, rsx_stream_read_fn rd, void *rctx, int *ok, int *err_stream);
/* rsStream.Reconstruct, streaming.go:477: valid[i] = stream i is read, fill[i] =
 * stream i is written (TotalShards flags each; both set is
 * ErrReconstructMismatch). */
int rsx_stream_reconstruct(rsx_stream *s, const uint8_t *valid, const uint8_t *fill, rsx_stream_read_fn rd,
                           void *rctx, rsx_stream_write_fn wr, void *wctx, int *err_stream);

/* ---- LowLevel (mulslice.go, gf16.go) -------------------------------------
 * Synchronous, like the Go functions.  Buffers may be device or host memory
 * (host memory is staged through HBM). */
/* LowLevel.GalMulSlice / GalMulSliceXor, mulslice.go:29-39:
 * out[i] = c * in[i] (do_xor = 0) or out[i] ^= c * in[i], for i < len. */
int rsx_gal_mul_slice(int device, uint8_t c, const uint8_t *in, uint8_t *out, size_t len, int do_xor);
/* Inv, mulslice.go:42 (Inv(0) = 0). */
uint8_t rsx_inv(uint8_t e);
/* LowLevel.GF16Mul, gf16.go:6 (Leopard GF(2^16), polynomial 0x1002D). */
uint16_t rsx_gf16_mul(uint16_t a, uint16_t b);
/* The Leopard field tables the library computes and uploads: expLUT / logLUT
 * (order entries), fftSkew (order - 1) and logWalsh (order) of
 * leopard.go:1104-1195 (bits = 16, order 65536) or leopard8.go:1029-1117
 * (bits = 8, order 256; values widened to 16 bits).  Any pointer may be
 * NULL.  No reference counterpart (the Go tables are package-private): it
 * lets a caller check the field representation the engine computes in. */
int rsx_leopard_tables(int bits, uint16_t *exp, uint16_t *log, uint16_t *skew, uint16_t *walsh);
/* LowLevel.GF16MulSliceXor (nout = 1) / GF16MulSliceXor8 (nout = 8),
 * mulslice.go:45-95: outs[j] ^= scalars[j] * in over the Leopard split layout;
 * zero scalars are skipped.  len must be a multiple of 64 and every output
 * len bytes long: the reference panics otherwise, here ErrInvalidShardSize. */
int rsx_gf16_mul_slice_xor(int device, const uint16_t *scalars, int nout, const uint8_t *in,
                           uint8_t *const *outs, const size_t *out_lens, size_t len);

/* Message of the last error on the calling thread. */
const char *rsx_last_error(void);
int rsx_device_count(void);

#ifdef __cplusplus
}
#endif
#endif

// stream.cpp — StreamEncoder (streaming.go) on the GPU: shards arrive through
// read callbacks block by block (4 MiB by default) and leave through write
// callbacks.  Two pinned host blocks and two HBM blocks form a pipeline: while
// the GPU encodes / verifies / reconstructs block b (H2D, kernel, D2H on the
// codec's stream), the host reads block b+1 into the other pinned block.
//
// Block semantics follow streaming.go:211-330: every reader is read with
// ReadFull semantics; readers that end early must agree on their length
// (ErrShardSize otherwise); a common length of 0 is the end of the stream;
// read / write callback failures are StreamReadError / StreamWriteError with
// the stream index.  Output of block b is written before an error found
// while reading block b+1 is reported, as the sequential reference does.
#include <algorithm>
#include <cstring>

#include "codec.hpp"
#include "trace.hpp"

using namespace rsx;

struct rsx_stream {
    rsx_codec *c = nullptr;
    size_t bs = 0;
    uint8_t *host[2] = {nullptr, nullptr};  // pinned, total * bs each
    DeviceBuffer dev[2];                    // total * bs each
    hipEvent_t done[2] = {nullptr, nullptr};
};

namespace {

struct Block {
    std::vector<size_t> lens;  // bytes read per shard (0 for nil readers)
    size_t size = 0;           // the block's shard size
    bool eof = false;
};

// streaming.go:211-246 readShards (ReadFull per reader).
int read_block(rsx_stream *s, int slot, const uint8_t *use, int nuse_first, int n, rsx_stream_read_fn rd, void *ctx,
               Block &b, int *err_stream) {
    b.lens.assign(static_cast<size_t>(n), 0);
    long long common = -1;
    for (int i = 0; i < n; ++i) {
        if (use && !use[i]) continue;
        uint8_t *buf = s->host[slot] + static_cast<size_t>(i + nuse_first) * s->bs;
        const long long got = rd(ctx, i, buf, s->bs);
        if (got < 0) {
            *err_stream = i;
            return fail(RSX_ERR_STREAM_READ, "error reading stream");
        }
        b.lens[i] = static_cast<size_t>(got);
        if (static_cast<size_t>(got) == s->bs) continue;
        if (common < 0) common = got;
        else if (got != common) return fail(RSX_ERR_SHARD_SIZE, "shard sizes do not match");
    }
    b.eof = common == 0;
    b.size = 0;
    for (size_t l : b.lens)
        if (l) {
            b.size = l;
            break;
        }
    return RSX_OK;
}

// streaming.go:248-267 writeShards.
int write_shards(rsx_stream *s, int slot, const uint8_t *which, int first, int n, size_t len,
                 rsx_stream_write_fn wr, void *ctx, int *err_stream) {
    for (int i = 0; i < n; ++i) {
        if (which && !which[i]) continue;
        const uint8_t *buf = s->host[slot] + static_cast<size_t>(first + i) * s->bs;
        const long long put = wr(ctx, i, buf, len);
        if (put < 0 || static_cast<size_t>(put) != len) {
            *err_stream = i;
            return fail(RSX_ERR_STREAM_WRITE, put < 0 ? "error writing stream" : "short write");
        }
    }
    return RSX_OK;
}

uint8_t *dev_shard(rsx_stream *s, int slot, int i) { return s->dev[slot].as<uint8_t>() + static_cast<size_t>(i) * s->bs; }
uint8_t *host_shard(rsx_stream *s, int slot, int i) { return s->host[slot] + static_cast<size_t>(i) * s->bs; }

}  // namespace

extern "C" {

int rsx_stream_create(int data_shards, int parity_shards, const rsx_options *opt, size_t block_size,
                      rsx_stream **out) {
    *out = nullptr;
    // streaming.go:147-150: at most 256 shards; classic engine only.
    if (data_shards + parity_shards > 256) return fail(RSX_ERR_MAX_SHARD_NUM, "too many shards");
    rsx_options o;
    if (opt) o = *opt;
    else rsx_options_default(&o);
    if (o.leopard != RSX_LEOPARD_AS_NEEDED) return fail(RSX_ERR_NOT_SUPPORTED, "streams use the classic engine");
    auto s = new rsx_stream;
    s->bs = block_size ? block_size : (size_t{4} << 20);
    int err = rsx_new(data_shards, parity_shards, &o, &s->c);
    if (err) {
        delete s;
        return err;
    }
    const size_t bytes = static_cast<size_t>(s->c->total) * s->bs;
    (void)hipSetDevice(s->c->device);
    for (int b = 0; b < 2; ++b) {
        if (hipHostMalloc(reinterpret_cast<void **>(&s->host[b]), bytes) != hipSuccess ||
            hipEventCreateWithFlags(&s->done[b], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            rsx_stream_free(s);
            return fail(RSX_ERR_NOMEM, "pinned stream buffers");
        }
        if ((err = s->dev[b].reserve(bytes))) {
            rsx_stream_free(s);
            return err;
        }
    }
    *out = s;
    return RSX_OK;
}

void rsx_stream_free(rsx_stream *s) {
    if (!s) return;
    if (s->c) {
        (void)hipSetDevice(s->c->device);
        (void)rsx_synchronize(s->c);
    }
    for (int b = 0; b < 2; ++b) {
        if (s->host[b]) (void)hipHostFree(s->host[b]);
        if (s->done[b]) (void)hipEventDestroy(s->done[b]);
    }
    rsx_codec *c = s->c;
    delete s;  // device blocks are freed before their codec
    rsx_free(c);
}

rsx_codec *rsx_stream_codec(rsx_stream *s) { return s->c; }

// streaming.go:199-210 Encode.
int rsx_stream_encode(rsx_stream *s, rsx_stream_read_fn rd, void *rctx, rsx_stream_write_fn wr, void *wctx,
                      int *err_stream) {
    RSX_TRACE("rsx_stream_encode");
    *err_stream = -1;
    rsx_codec *c = s->c;
    const int k = c->k, m = c->m, total = c->total;
    (void)hipSetDevice(c->device);
    Block cur, next;
    int err = read_block(s, 0, nullptr, 0, k, rd, rctx, cur, err_stream);
    if (err) return err;
    if (cur.eof) return fail(RSX_ERR_SHARD_NO_DATA, "no shard data");
    int slot = 0;
    for (;;) {
        // GPU: block `slot`.
        const size_t len = cur.size;
        std::vector<uint8_t *> ptrs(static_cast<size_t>(total));
        std::vector<size_t> lens(static_cast<size_t>(total), len);
        for (int i = 0; i < k; ++i) lens[i] = cur.lens[i];
        for (int i = 0; i < total; ++i) ptrs[i] = dev_shard(s, slot, i);
        for (int i = 0; i < k; ++i)
            RSX_HIP(hipMemcpyAsync(ptrs[i], host_shard(s, slot, i), cur.lens[i], hipMemcpyHostToDevice, c->stream));
        if ((err = rsx_encode(c, ptrs.data(), lens.data(), total))) return err;
        for (int i = k; i < total; ++i)
            RSX_HIP(hipMemcpyAsync(host_shard(s, slot, i), ptrs[i], len, hipMemcpyDeviceToHost, c->stream));
        RSX_HIP(hipEventRecord(s->done[slot], c->stream));
        // Host: the next block, while the GPU works.
        const int rerr = read_block(s, slot ^ 1, nullptr, 0, k, rd, rctx, next, err_stream);
        const int rstream = *err_stream;
        RSX_HIP(hipEventSynchronize(s->done[slot]));
        *err_stream = -1;
        if ((err = write_shards(s, slot, nullptr, k, m, len, wr, wctx, err_stream))) return err;
        if (rerr) {
            *err_stream = rstream;
            return rerr;
        }
        if (next.eof) return RSX_OK;
        std::swap(cur, next);
        slot ^= 1;
    }
}

// streaming.go:430-455 Verify.
int rsx_stream_verify(rsx_stream *s, rsx_stream_read_fn rd, void *rctx, int *ok, int *err_stream) {
    RSX_TRACE("rsx_stream_verify");
    *ok = 0;
    *err_stream = -1;
    rsx_codec *c = s->c;
    const int total = c->total;
    (void)hipSetDevice(c->device);
    Block cur, next;
    int err = read_block(s, 0, nullptr, 0, total, rd, rctx, cur, err_stream);
    if (err) return err;
    if (cur.eof) return fail(RSX_ERR_SHARD_NO_DATA, "no shard data");
    int slot = 0;
    for (;;) {
        std::vector<uint8_t *> ptrs(static_cast<size_t>(total));
        for (int i = 0; i < total; ++i) {
            ptrs[i] = dev_shard(s, slot, i);
            RSX_HIP(hipMemcpyAsync(ptrs[i], host_shard(s, slot, i), cur.lens[i], hipMemcpyHostToDevice, c->stream));
        }
        // The read of the next block overlaps the copies; Verify itself
        // synchronises (its result is needed before going on).
        const int rerr = read_block(s, slot ^ 1, nullptr, 0, total, rd, rctx, next, err_stream);
        const int rstream = *err_stream;
        int good = 0;
        if ((err = rsx_verify(c, ptrs.data(), cur.lens.data(), total, &good))) return err;
        if (!good) return RSX_OK;
        if (rerr) {
            *err_stream = rstream;
            return rerr;
        }
        if (next.eof) {
            *ok = 1;
            return RSX_OK;
        }
        std::swap(cur, next);
        slot ^= 1;
    }
}

// streaming.go:477-527 Reconstruct.
int rsx_stream_reconstruct(rsx_stream *s, const uint8_t *valid, const uint8_t *fill, rsx_stream_read_fn rd,
                           void *rctx, rsx_stream_write_fn wr, void *wctx, int *err_stream) {
    RSX_TRACE("rsx_stream_reconstruct");
    *err_stream = -1;
    rsx_codec *c = s->c;
    const int k = c->k, total = c->total;
    bool data_only = true;
    for (int i = 0; i < total; ++i) {
        if (valid[i] && fill[i]) return fail(RSX_ERR_RECONSTRUCT_MISMATCH, "valid and fill on one shard");
        if (i >= k && fill[i]) data_only = false;
    }
    (void)hipSetDevice(c->device);
    Block cur, next;
    int err = read_block(s, 0, valid, 0, total, rd, rctx, cur, err_stream);
    if (err) return err;
    if (cur.eof) return fail(RSX_ERR_SHARD_NO_DATA, "no shard data");
    int slot = 0;
    std::vector<uint8_t> filled(static_cast<size_t>(total));
    for (;;) {
        const size_t len = cur.size;
        std::vector<uint8_t *> ptrs(static_cast<size_t>(total));
        for (int i = 0; i < total; ++i) {
            ptrs[i] = dev_shard(s, slot, i);
            if (cur.lens[i])
                RSX_HIP(hipMemcpyAsync(ptrs[i], host_shard(s, slot, i), cur.lens[i], hipMemcpyHostToDevice,
                                       c->stream));
        }
        err = data_only ? rsx_reconstruct_data(c, ptrs.data(), cur.lens.data(), total, filled.data())
                        : rsx_reconstruct(c, ptrs.data(), cur.lens.data(), total, filled.data());
        if (err) return err;
        for (int i = 0; i < total; ++i)
            if (fill[i]) RSX_HIP(hipMemcpyAsync(host_shard(s, slot, i), ptrs[i], len, hipMemcpyDeviceToHost, c->stream));
        RSX_HIP(hipEventRecord(s->done[slot], c->stream));
        const int rerr = read_block(s, slot ^ 1, valid, 0, total, rd, rctx, next, err_stream);
        const int rstream = *err_stream;
        RSX_HIP(hipEventSynchronize(s->done[slot]));
        *err_stream = -1;
        if ((err = write_shards(s, slot, fill, 0, total, len, wr, wctx, err_stream))) return err;
        if (rerr) {
            *err_stream = rstream;
            return rerr;
        }
        if (next.eof) return RSX_OK;
        std::swap(cur, next);
        slot ^= 1;
    }
}

}  // extern "C"

// stream.cpp — StreamEncoder (streaming.go) on the GPU: shards arrive through
// read callbacks block by block (4 MiB by default) and leave through write
// callbacks.  Two pinned host blocks and two HBM blocks form a pipeline: while
// the GPU encodes / verifies / reconstructs block b (H2D, kernel, D2H on the
// codec's stream), the host reads block b+1 into the other pinned block.
//
// With concurrent reads or writes (WithConcurrentStreamReads / Writes,
// options.go:122-146; streaming.go:334-420 cReadShards / cWriteShards) the
// callbacks of one block run on a pool of threads, one shard per task, so a
// caller's callbacks must then be safe to call for different shards at once;
// of several failing streams the lowest index is reported (the reference
// reports whichever goroutine it collects first).
//
// Block semantics follow streaming.go:211-330: every reader is read with
// ReadFull semantics; readers that end early must agree on their length
// (ErrShardSize otherwise); a common length of 0 is the end of the stream;
// read / write callback failures are StreamReadError / StreamWriteError with
// the stream index.  Output of block b is written before an error found
// while reading block b+1 is reported, as the sequential reference does.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>

#include "codec.hpp"
#include "trace.hpp"

using namespace rsx;

namespace {

// Runs fn(0 .. n-1) on a fixed set of threads plus the caller; returns when
// every task has finished.
class ShardPool {
  public:
    explicit ShardPool(int threads) {
        for (int t = 0; t < threads; ++t) th_.emplace_back([this] { loop(); });
    }
    ~ShardPool() {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void run(int n, const std::function<void(int)> &fn) {
        std::unique_lock<std::mutex> lock(mu_);
        fn_ = &fn;
        next_ = 0;
        n_ = n;
        left_ = n;
        ++gen_;
        cv_.notify_all();
        work(lock);
        done_cv_.wait(lock, [&] { return left_ == 0; });
        fn_ = nullptr;
    }

  private:
    void work(std::unique_lock<std::mutex> &lock) {
        while (fn_ && next_ < n_) {
            const int i = next_++;
            const std::function<void(int)> *fn = fn_;
            lock.unlock();
            (*fn)(i);
            lock.lock();
            if (--left_ == 0) done_cv_.notify_all();
        }
    }
    void loop() {
        std::unique_lock<std::mutex> lock(mu_);
        uint64_t seen = 0;
        for (;;) {
            cv_.wait(lock, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            work(lock);
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)> *fn_ = nullptr;
    int next_ = 0, n_ = 0, left_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

}  // namespace

struct rsx_stream {
    rsx_codec *c = nullptr;
    size_t bs = 0;
    uint8_t *host[2] = {nullptr, nullptr};  // pinned, total * bs each
    DeviceBuffer dev[2];                    // total * bs each
    hipEvent_t done[2] = {nullptr, nullptr};
    bool conc_reads = false, conc_writes = false;
    std::unique_ptr<ShardPool> pool;  // while either is set
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
    std::vector<long long> got_all;
    if (s->conc_reads) {
        // streaming.go:334-380 cReadShards: every reader at once.
        got_all.assign(static_cast<size_t>(n), 0);
        s->pool->run(n, [&](int i) {
            if (use && !use[i]) return;
            got_all[i] = rd(ctx, i, s->host[slot] + static_cast<size_t>(i + nuse_first) * s->bs, s->bs);
        });
    }
    long long common = -1;
    for (int i = 0; i < n; ++i) {
        if (use && !use[i]) continue;
        uint8_t *buf = s->host[slot] + static_cast<size_t>(i + nuse_first) * s->bs;
        const long long got = s->conc_reads ? got_all[i] : rd(ctx, i, buf, s->bs);
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
    std::vector<long long> put_all;
    if (s->conc_writes) {
        // streaming.go:382-420 cWriteShards: every writer at once.
        put_all.assign(static_cast<size_t>(n), 0);
        s->pool->run(n, [&](int i) {
            if (which && !which[i]) return;
            put_all[i] = wr(ctx, i, s->host[slot] + static_cast<size_t>(first + i) * s->bs, len);
        });
    }
    for (int i = 0; i < n; ++i) {
        if (which && !which[i]) continue;
        const uint8_t *buf = s->host[slot] + static_cast<size_t>(first + i) * s->bs;
        const long long put = s->conc_writes ? put_all[i] : wr(ctx, i, buf, len);
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

int rsx_stream_set_concurrency(rsx_stream *s, int reads, int writes) {
    s->conc_reads = reads != 0;
    s->conc_writes = writes != 0;
    if ((s->conc_reads || s->conc_writes) && !s->pool)
        s->pool = std::make_unique<ShardPool>(std::min(s->c->total, 16) - 1);
    return RSX_OK;
}

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

// hostloops.hip — host-side helpers for bench.py that keep Python out of
// timed loops (host code only; built into libmembench.so):
//   membench_mem_read / membench_mem_write: stream callbacks
//     (include/rsx.h rsx_stream_read_fn / rsx_stream_write_fn) over host
//     memory, so the stream encoder's pipeline is timed with memcpy-speed
//     sources and sinks instead of interpreter callbacks;
//   membench_call_loop: `calls` back-to-back calls of one C-ABI shard
//     operation (rsx_encode's signature), the per-call cost a C or cgo caller
//     pays.
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace {

struct MemStreams {
    uint8_t *const *buf;  // one host buffer per stream
    const size_t *len;    // bytes readable per stream (read side)
    size_t *pos;          // read or write position per stream
    const size_t *cap;    // bytes writable per stream (write side)
};

}  // namespace

extern "C" long long membench_mem_read(void *ctx, int shard, uint8_t *buf, size_t n) {
    MemStreams *m = static_cast<MemStreams *>(ctx);
    const size_t left = m->len[shard] - m->pos[shard];
    const size_t c = left < n ? left : n;
    std::memcpy(buf, m->buf[shard] + m->pos[shard], c);
    m->pos[shard] += c;
    return static_cast<long long>(c);
}

extern "C" long long membench_mem_write(void *ctx, int shard, const uint8_t *buf, size_t n) {
    MemStreams *m = static_cast<MemStreams *>(ctx);
    if (m->pos[shard] + n > m->cap[shard]) return -1;
    std::memcpy(m->buf[shard] + m->pos[shard], buf, n);
    m->pos[shard] += n;
    return static_cast<long long>(n);
}

typedef int (*shard_op_fn)(void *codec, uint8_t *const *shards, const size_t *lens, int nshards);

// Returns the first nonzero status, or 0.
extern "C" int membench_call_loop(void *fn, void *codec, uint8_t *const *shards, const size_t *lens, int nshards,
                                  int calls) {
    const shard_op_fn f = reinterpret_cast<shard_op_fn>(fn);
    for (int i = 0; i < calls; ++i) {
        const int rc = f(codec, shards, lens, nshards);
        if (rc) return rc;
    }
    return 0;
}

// batch.cpp — the batched path of include/rsx.h: many Encode / Reconstruct
// calls of classic codecs (any (k, m), shard size and erasure pattern) run as
// one persistent launch per output-count class, captured in a HIP graph.
//
// Each item is validated exactly as its per-stripe method validates it
// (reedsolomon.go:686-702 Encode, 1670-1781 reconstruct), so an item's status
// equals the error the plain call would return.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>

#include "batch.hpp"
#include "codec.hpp"
#include "kernels.hpp"
#include "leopard.hpp"
#include "trace.hpp"

namespace rsx {
int check_shards(const size_t *lens, int n, bool nilok);
size_t shard_size(const size_t *lens, int n);
int decode_plan(rsx_codec *c, const size_t *lens, bool data_only, const uint8_t *required,
                std::shared_ptr<const DecodePlan> *plan);
}  // namespace rsx

using namespace rsx;

namespace {

// Page-locked host staging for a plan's uploads: copies from it are
// asynchronous, so planning does not wait for them.  Grows, never shrinks.
struct PinnedBuffer {
    uint8_t *ptr = nullptr;
    size_t bytes = 0;
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer &) = delete;
    PinnedBuffer &operator=(const PinnedBuffer &) = delete;
    ~PinnedBuffer() {
        if (ptr) (void)hipHostFree(ptr);
    }
    int reserve(size_t n) {
        if (n <= bytes) return RSX_OK;
        n = std::max(n, bytes + bytes / 2);  // few regrowths (each frees: a device-wide wait)
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
        if (hipHostMalloc(reinterpret_cast<void **>(&ptr), n) != hipSuccess) {
            (void)hipGetLastError();
            ptr = nullptr;
            return fail(RSX_ERR_NOMEM, "pinned staging of " + std::to_string(n) + " bytes failed");
        }
        bytes = n;
        return RSX_OK;
    }
};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Grows a plan buffer by at least half its size, so plans of varying size
// settle after a few regrowths (each one waits for the whole device).
int reserve_grow(DeviceBuffer &b, size_t n) {
    return b.reserve(n <= b.bytes ? n : std::max(n, b.bytes + b.bytes / 2));
}

}  // namespace

struct rsx_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    // Leopard launches run on `side`, forked from and joined back to the
    // launch stream, so they overlap the classic passes (also inside graphs).
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // Output-count classes too small to keep the device busy for long (fewer
    // tiles than two grids) run on these streams, forked from and joined back
    // to the launch stream, so their tails overlap other classes instead of
    // idling the device; large classes run in order on the launch stream.
    static constexpr int kClassStreams = 3;
    // Work arrays of a plan's three-pass Leopard items (two of n x shard
    // bytes each); items past the cap get ErrNotSupported.
    static constexpr size_t kPassWorkCap = size_t{4} << 30;
    bool spread_classes = true;  // off: every class in order on `stream`
    hipStream_t cls[kClassStreams] = {};
    hipEvent_t cls_join[kClassStreams] = {};
    int cus = 0;
    // tuning: 0 = autotune on first launch
    uint32_t tile = 0;
    int blocks_per_cu = 0;
    bool tuned = false;

    // plan
    struct Task {
        const uint32_t *tables;  // device tables (batch-owned copy)
        uint64_t len;
        uint32_t in_off, out_off;
        uint16_t k, nout, mtot, o0;
        uint8_t aligned;
        uint8_t mode;   // Gf8Mode
        uint32_t item;  // compare mode: the Verify item
    };
    std::vector<Task> tasks[9];          // by output-count class 1..8
    std::vector<const uint8_t *> pool;   // pointer pool
    // Everything a plan uploads goes through two pinned staging buffers and
    // two copies on `stream`: d_blob (coefficient tables, the pointer pool,
    // Leopard aux arrays and task lists) and d_launch (the task list and the
    // tile -> task table of the chosen tile size).
    DeviceBuffer d_blob, d_launch;
    PinnedBuffer h_blob, h_launch;
    const uint8_t *const *dev_ptrs = nullptr;  // the pool in d_blob
    hipEvent_t uploaded = nullptr;             // after the plan's copies on `stream`
    std::vector<BatchLaunch> launches;   // one per non-empty class
    std::vector<int> classes;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool planned = false;
    bool use_graph = true;

    // Verify items set d_flags[item] on a mismatch, copied to the pinned
    // h_flags at the end: the passes recompute the parity in registers (classic)
    // or LDS (Leopard) and compare it with the item's parity shards.
    bool has_verify = false;
    DeviceBuffer d_flags;
    uint32_t *h_flags = nullptr;
    size_t h_flags_cap = 0;
    int nitems = 0;
    std::vector<uint8_t> is_verify;

    // Leopard items, per field (0: GF(2^8), 1: GF(2^16)), run by the
    // LDS-resident batch kernel after the classic passes.
    // Leopard items by launch kind: GF(2^8), GF(2^16), GF(2^16) quarter-chunk tiles.
    std::vector<rsx::LeoItemPlan> leo[3];
    std::vector<rsx::LeoBatchTask> leo_tasks[3];
    const rsx::LeoBatchTask *dev_leo_tasks[3] = {};  // in d_blob
    size_t leo_lds[3] = {0, 0, 0};
    size_t leo_deriv[3] = {0, 0, 0};  // largest register bytes (LeoItemPlan::regs) of the tasks
    uint32_t leo_slices[3] = {0, 0, 0};
    // Workgroups per CU of the GF(2^8) Leopard launch (set per plan).
    int leo8_per_cu = 8;

    // GF(2^16) reconstructs with 8192 or more work rows: three launches (one
    // per pass) over all such items, each with its two work arrays in d_mp_work.
    std::vector<rsx::LeoItemPlan> leo_mp;
    std::vector<rsx::LeoPassTask> mp_tasks;
    const rsx::LeoPassTask *dev_mp_tasks = nullptr;  // in d_blob
    DeviceBuffer d_mp_work;
    size_t mp_work_bytes = 0;
    uint32_t mp_blocks[3] = {0, 0, 0};
    int mp_max_p = 0;

    // GF(2^16) encodes above the tiles: per group index g, a pass-1 and a
    // pass-2 launch over the items with more than g groups, then one pass-3
    // launch over all; their work arrays also in d_mp_work.
    struct EncLaunch {
        size_t off;  // first task in d_emp_tasks
        int ntasks, pass, max_p;
        uint32_t blocks;
    };
    std::vector<rsx::LeoItemPlan> leo_emp;
    std::vector<rsx::LeoEncPassTask> emp_tasks;
    std::vector<EncLaunch> emp_launches;
    const rsx::LeoEncPassTask *dev_emp_tasks = nullptr;  // in d_blob
};

namespace {

void destroy_graph(rsx_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Tiles above which a plan finds tasks by binary search instead of keeping
// a tile -> task table (4 bytes per tile) on the device.  RSX_BATCH_TILE_TABLE_MAX
// (read at each plan) lowers it, so tests can run the search on small plans.
constexpr uint64_t kTileTableMax = uint64_t{1} << 24;

uint64_t tile_table_max() {
    const char *v = std::getenv("RSX_BATCH_TILE_TABLE_MAX");
    if (!v || !*v) return kTileTableMax;
    return std::min<uint64_t>(kTileTableMax, std::strtoull(v, nullptr, 10));
}

// Output-count classes of the batched GF(2^8) pass (one kernel each).  Plans
// keep 8-output tasks: a 16-output class measured slower in mixed plans
// (9-15-output tasks waste accumulators at one wave per SIMD).
constexpr int kClasses[] = {1, 2, 3, 4, 5, 6, 7, 8};

// Lays the tasks out for one tile size and uploads them.
int build_launches(rsx_batch *b, uint32_t tile) {
    std::vector<BatchTask> all;
    std::vector<uint32_t> tile_task;  // per class, concatenated: class-local task index of every tile
    std::vector<size_t> tt_off;
    uint64_t total_tiles = 0;
    for (int mc : kClasses)
        for (const auto &t : b->tasks[mc]) total_tiles += (t.len + tile - 1) / tile;
    const bool use_table = total_tiles <= tile_table_max();
    b->launches.clear();
    b->classes.clear();
    for (int mc : kClasses) {
        if (b->tasks[mc].empty()) continue;
        BatchLaunch L{};
        L.ntasks = static_cast<int>(b->tasks[mc].size());
        L.tile = tile;
        uint64_t t0 = 0;
        int maxk = 1;
        const size_t first = all.size();
        tt_off.push_back(tile_task.size());
        for (const auto &t : b->tasks[mc]) {
            if (use_table)
                tile_task.insert(tile_task.end(), (t.len + tile - 1) / tile, static_cast<uint32_t>(all.size() - first));
            BatchTask d{};
            d.tables = t.tables;
            d.len = t.len;
            d.tile0 = static_cast<uint32_t>(t0);
            d.in_off = t.in_off;
            d.out_off = t.out_off;
            d.k = t.k;
            d.nout = t.nout;
            d.mtot = t.mtot;
            d.o0 = t.o0;
            d.mode = t.mode;
            d.aligned = t.aligned;
            d.item = t.item;
            all.push_back(d);
            t0 += (t.len + tile - 1) / tile;
            maxk = std::max<int>(maxk, t.k);
        }
        if (t0 > 0xFFFFFFFFull) return fail(RSX_ERR_INVALID_INPUT, "batch too large for one plan");
        L.ntiles = static_cast<uint32_t>(t0);
        L.maxk = maxk;
        L.tasks = reinterpret_cast<const BatchTask *>(first);  // offset, patched after upload
        b->launches.push_back(L);
        b->classes.push_back(mc);
    }
    // Task list then tile table, one pinned copy.  An earlier copy from
    // h_launch has completed: plans and autotune trials wait for the stream
    // before they rebuild.
    const size_t task_bytes = align256(std::max<size_t>(all.size(), 1) * sizeof(BatchTask));
    const size_t tt_bytes = use_table ? tile_task.size() * sizeof(uint32_t) : 0;
    int err = reserve_grow(b->d_launch, task_bytes + std::max<size_t>(tt_bytes, 4));
    if (!err) err = b->h_launch.reserve(task_bytes + std::max<size_t>(tt_bytes, 4));
    if (err) return err;
    if (!all.empty()) std::memcpy(b->h_launch.ptr, all.data(), all.size() * sizeof(BatchTask));
    if (tt_bytes) std::memcpy(b->h_launch.ptr + task_bytes, tile_task.data(), tt_bytes);
    RSX_HIP(hipMemcpyAsync(b->d_launch.ptr, b->h_launch.ptr, task_bytes + tt_bytes, hipMemcpyHostToDevice, b->stream));
    const BatchTask *dev_tasks = b->d_launch.as<BatchTask>();
    const uint32_t *dev_tt = reinterpret_cast<const uint32_t *>(b->d_launch.as<uint8_t>() + task_bytes);
    for (size_t i = 0; i < b->launches.size(); ++i) {
        BatchLaunch &L = b->launches[i];
        L.tasks = dev_tasks + reinterpret_cast<size_t>(L.tasks);
        L.ptrs = b->dev_ptrs;
        L.flags = b->d_flags.as<uint32_t>();
        L.tile_task = use_table ? dev_tt + tt_off[i] : nullptr;
    }
    return RSX_OK;
}

int enqueue(rsx_batch *b, hipStream_t s) {
    const unsigned grid = static_cast<unsigned>(std::max(1, b->cus * b->blocks_per_cu));
    const bool verify = b->has_verify;
    if (verify)
        RSX_HIP(launch_fill_u32(b->d_flags.as<uint32_t>(), static_cast<size_t>(b->nitems), 0u, s));
    const bool leo = !b->leo_tasks[0].empty() || !b->leo_tasks[1].empty() || !b->leo_tasks[2].empty() ||
                     !b->mp_tasks.empty() || !b->emp_launches.empty();
    const bool fork = leo && !b->launches.empty();
    const int spread = b->spread_classes
                           ? std::min<int>(static_cast<int>(b->launches.size()) - 1, rsx_batch::kClassStreams)
                           : 0;
    if (fork || spread > 0) RSX_HIP(hipEventRecord(b->fork, s));
    for (int j = 0; j < spread; ++j) RSX_HIP(hipStreamWaitEvent(b->cls[j], b->fork, 0));
    hipStream_t ls = s;
    if (fork) {
        RSX_HIP(hipStreamWaitEvent(b->side, b->fork, 0));
        ls = b->side;
    }
    for (int fi = 0; fi < 3; ++fi) {
        if (b->leo_tasks[fi].empty()) continue;
        const int per_cu = fi == 0 && fork ? b->leo8_per_cu : 8;
        const unsigned g = std::min<unsigned>(static_cast<unsigned>(std::max(1, b->cus * per_cu)), b->leo_slices[fi]);
        RSX_HIP(leo_batch_launch(*b->leo[fi][0].dev, b->dev_leo_tasks[fi],
                                 static_cast<int>(b->leo_tasks[fi].size()), b->leo_slices[fi], b->leo_lds[fi], g,
                                 b->leo_deriv[fi], fi == 2, ls));
    }
    if (!b->mp_tasks.empty())
        RSX_HIP(leo_decode_multipass_batch(*b->leo_mp[0].dev, b->dev_mp_tasks,
                                           static_cast<int>(b->mp_tasks.size()), b->mp_blocks, b->mp_max_p, ls));
    for (const auto &L : b->emp_launches)
        RSX_HIP(leo_encode_multipass_batch(*b->leo_emp[0].dev, b->dev_emp_tasks + L.off, L.ntasks,
                                           L.blocks, L.pass, L.max_p, ls));
    if (fork) RSX_HIP(hipEventRecord(b->join, b->side));
    int next_side = 0;
    for (size_t i = 0; i < b->launches.size(); ++i) {
        const unsigned g = std::min<unsigned>(grid, std::max<uint32_t>(b->launches[i].ntiles, 1));
        const bool small = static_cast<uint64_t>(b->launches[i].ntiles) < 2ull * grid;
        hipStream_t cs = small && spread > 0 ? b->cls[next_side++ % spread] : s;
        RSX_HIP(launch_batch_pass(b->launches[i], b->classes[i], g, cs));
    }
    for (int j = 0; j < spread; ++j) {
        RSX_HIP(hipEventRecord(b->cls_join[j], b->cls[j]));
        RSX_HIP(hipStreamWaitEvent(s, b->cls_join[j], 0));
    }
    if (fork) RSX_HIP(hipStreamWaitEvent(s, b->join, 0));

    if (verify) {
        RSX_HIP(hipMemcpyAsync(b->h_flags, b->d_flags.ptr, sizeof(uint32_t) * b->nitems,
                               hipMemcpyDeviceToHost, s));
    }
    return RSX_OK;
}

int capture(rsx_batch *b) {
    destroy_graph(b);
    RSX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int err = enqueue(b, b->stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (err) {
        if (g) (void)hipGraphDestroy(g);
        return err;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->graph = g;
    RSX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return RSX_OK;
}

// Times the planned work for a few (tile, blocks/CU) choices and keeps the
// fastest: blocks per CU at 16 KiB tiles first, then the tile size at the
// best blocks per CU (six timed runs).  More blocks than fit at once (up to
// 64 per CU, each taking a few tiles) beat persistent grids on long shards:
// RS(16,4) x 4 MiB encode 5.08 TB/s at 64 KiB x 8 per CU, 5.33 at 16 KiB x 32.
// The work is idempotent (outputs are overwritten), so running it repeatedly
// is harmless.
int autotune(rsx_batch *b) {
    static const uint32_t tiles[] = {16u << 10, 64u << 10, 256u << 10};
    static const int bpcs[] = {2, 8, 32, 64};
    hipEvent_t e0, e1;
    RSX_HIP(hipEventCreate(&e0));
    RSX_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    uint32_t best_tile = tiles[0];
    int best_bpc = 8;
    int err = RSX_OK;
    auto trial = [&](uint32_t tile, int bpc) {
        b->blocks_per_cu = bpc;
        if ((err = enqueue(b, b->stream))) return;  // warm
        (void)hipEventRecord(e0, b->stream);
        if ((err = enqueue(b, b->stream))) return;
        (void)hipEventRecord(e1, b->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
            best = ms;
            best_tile = tile;
            best_bpc = bpc;
        }
    };
    if (!(err = build_launches(b, tiles[0])))
        for (int bpc : bpcs) {
            trial(tiles[0], bpc);
            if (err) break;
        }
    const int bpc = best_bpc;
    for (int i = 1; i < 3 && !err; ++i)
        if (!(err = build_launches(b, tiles[i]))) trial(tiles[i], bpc);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err) return err;
    b->tile = best_tile;
    b->blocks_per_cu = best_bpc;
    b->tuned = true;
    return build_launches(b, b->tile);
}

}  // namespace

namespace rsx {

int batch_reserve(rsx_batch *b, size_t bytes, int max_items) {
    (void)hipSetDevice(b->device);
    RSX_HIP(hipStreamSynchronize(b->stream));
    int err = reserve_grow(b->d_blob, bytes);
    if (!err) err = reserve_grow(b->d_launch, bytes);
    if (!err) err = b->h_blob.reserve(bytes);
    if (!err) err = b->h_launch.reserve(bytes);
    if (!err) err = reserve_grow(b->d_flags, sizeof(uint32_t) * std::max(max_items, 1));
    if (!err && b->h_flags_cap < static_cast<size_t>(max_items)) {
        if (b->h_flags) (void)hipHostFree(b->h_flags);
        b->h_flags = nullptr;
        b->h_flags_cap = 0;
        RSX_HIP(hipHostMalloc(reinterpret_cast<void **>(&b->h_flags), sizeof(uint32_t) * std::max(max_items, 1)));
        b->h_flags_cap = static_cast<size_t>(max_items);
    }
    return err;
}

void batch_set_class_streams(rsx_batch *b, bool on) { b->spread_classes = on; }

}  // namespace rsx

extern "C" {

int rsx_batch_create(int device, rsx_batch **out) {
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return fail(RSX_ERR_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= ndev) return fail(RSX_ERR_DEVICE, "invalid device ordinal");
    RSX_HIP(hipSetDevice(device));
    auto *b = new rsx_batch;
    b->device = device;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->uploaded, hipEventDisableTiming) != hipSuccess ||
        hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
        rsx_batch_free(b);
        return fail(RSX_ERR_DEVICE, "stream creation failed");
    }
    for (int j = 0; j < rsx_batch::kClassStreams; ++j)
        if (hipStreamCreateWithFlags(&b->cls[j], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->cls_join[j], hipEventDisableTiming) != hipSuccess) {
            rsx_batch_free(b);
            return fail(RSX_ERR_DEVICE, "stream creation failed");
        }
    *out = b;
    return RSX_OK;
}

void rsx_batch_free(rsx_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->side) (void)hipStreamSynchronize(b->side);
    for (int j = 0; j < rsx_batch::kClassStreams; ++j)
        if (b->cls[j]) (void)hipStreamSynchronize(b->cls[j]);
    destroy_graph(b);
    if (b->h_flags) (void)hipHostFree(b->h_flags);
    if (b->uploaded) (void)hipEventDestroy(b->uploaded);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->side) (void)hipStreamDestroy(b->side);
    for (int j = 0; j < rsx_batch::kClassStreams; ++j) {
        if (b->cls_join[j]) (void)hipEventDestroy(b->cls_join[j]);
        if (b->cls[j]) (void)hipStreamDestroy(b->cls[j]);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int rsx_batch_set_tuning(rsx_batch *b, int tile_bytes, int blocks_per_cu) {
    if (tile_bytes < 0 || blocks_per_cu < 0 || tile_bytes % 16)
        return fail(RSX_ERR_INVALID_INPUT, "tile must be a multiple of 16 bytes");
    b->tile = static_cast<uint32_t>(tile_bytes);
    b->blocks_per_cu = blocks_per_cu;
    b->tuned = tile_bytes > 0 && blocks_per_cu > 0;
    return RSX_OK;
}

int rsx_batch_get_tuning(const rsx_batch *b, int *tile_bytes, int *blocks_per_cu) {
    *tile_bytes = static_cast<int>(b->tile);
    *blocks_per_cu = b->blocks_per_cu;
    return RSX_OK;
}

void *rsx_batch_stream(rsx_batch *b) { return b->stream; }

int rsx_batch_plan(rsx_batch *b, const rsx_batch_item *items, int n, int *status, uint8_t *filled) {
    RSX_TRACE("rsx_batch_plan");
    (void)hipSetDevice(b->device);
    RSX_HIP(hipStreamSynchronize(b->stream));
    destroy_graph(b);
    for (auto &v : b->tasks) v.clear();
    b->pool.clear();
    b->planned = false;
    b->has_verify = false;
    b->nitems = n;
    b->is_verify.assign(static_cast<size_t>(std::max(n, 0)), 0);
    for (int fi = 0; fi < 3; ++fi) {
        b->leo[fi].clear();
        b->leo_tasks[fi].clear();
        b->leo_lds[fi] = 0;
        b->leo_deriv[fi] = 0;
        b->leo_slices[fi] = 0;
    }
    b->leo_mp.clear();
    b->mp_tasks.clear();
    b->leo_emp.clear();
    b->emp_tasks.clear();
    b->emp_launches.clear();
    b->mp_work_bytes = 0;
    b->mp_max_p = 0;
    for (uint32_t &x : b->mp_blocks) x = 0;
    std::vector<uint8_t> laux;  // Leopard error locators / output rows / logs
    // Every table the plan reads is copied into `htab` (host) when first
    // seen and uploaded once; keyed by (host table, codec, decode-cache
    // epoch) so a table freed by a cache flush is never matched again.
    std::map<std::tuple<const void *, const rsx_codec *, uint64_t>, size_t> table_off;
    const std::pair<const std::tuple<const void *, const rsx_codec *, uint64_t>, size_t> *last_table = nullptr;
    std::vector<std::shared_ptr<const DecodePlan>> held_plans;  // decode plans read by this plan
    std::vector<uint32_t> htab;
    size_t fill_off = 0;

    struct Pending {
        int mc;
        size_t idx;
        size_t toff;  // dword offset of the task's tables in htab
    };
    std::vector<Pending> pend;
    DeviceRanges ranges;
    std::vector<const uint8_t *> in;  // one item's inputs / outputs (reused)
    std::vector<uint8_t *> out;

    for (int it = 0; it < n; ++it) {
        const rsx_batch_item &item = items[it];
        rsx_codec *c = item.codec;
        status[it] = RSX_OK;
        uint8_t *fl = filled ? filled + fill_off : nullptr;
        if (c) fill_off += c->total;
        if (!c) {
            status[it] = fail(RSX_ERR_INVALID_INPUT, "batch item without a codec");
            continue;
        }
        if (fl) std::memset(fl, 0, c->total);
        if (c->engine) {
            // Leopard: one task of the LDS-resident batch kernel; Verify is
            // an encode whose parity rows are compared instead of written.
            if (item.op != RSX_OP_ENCODE && item.op != RSX_OP_VERIFY && item.op != RSX_OP_RECONSTRUCT &&
                item.op != RSX_OP_RECONSTRUCT_DATA) {
                status[it] = fail(RSX_ERR_INVALID_INPUT, "unknown batch op");
                continue;
            }
            LeoItemPlan lp;
            const int e = leopard_plan_item(c, item.op, item.shards, item.lens, item.nshards, b->pool, laux, lp, fl,
                                            ranges);
            if (e) {
                if (fl) std::memset(fl, 0, c->total);
                status[it] = e;
                continue;
            }
            if (item.op == RSX_OP_VERIFY) {
                lp.verify_item = it;
                b->is_verify[it] = 1;
            }
            if (lp.empty) continue;
            if (lp.kind == 3 || lp.kind == 4) {
                // Work arrays of the three-pass items, within a cap per plan.
                const size_t wb = (lp.work_bytes + 255) / 256 * 256;
                if (b->mp_work_bytes + wb > rsx_batch::kPassWorkCap) {
                    if (fl) std::memset(fl, 0, c->total);
                    b->is_verify[it] = 0;
                    status[it] = fail(RSX_ERR_NOT_SUPPORTED, "three-pass work arrays over the plan's cap");
                    continue;
                }
                if (lp.kind == 3) lp.mp.work = reinterpret_cast<uint8_t *>(b->mp_work_bytes);  // offset, patched below
                else lp.emp.work = reinterpret_cast<uint8_t *>(b->mp_work_bytes);
                b->mp_work_bytes += wb;
                (lp.kind == 3 ? b->leo_mp : b->leo_emp).push_back(lp);
            } else {
                b->leo[lp.kind].push_back(lp);
            }
            continue;
        }
        const int k = c->k, total = c->total;
        const size_t *lens = item.lens;
        in.clear();
        out.clear();
        const uint32_t *tables = nullptr;  // host copy of the item's tables
        uint64_t epoch = ~uint64_t{0};
        int mtot = 0;
        int err_item = RSX_OK;
        if (item.op == RSX_OP_ENCODE || item.op == RSX_OP_VERIFY) {
            // reedsolomon.go:686-702 (Encode) / 857-871 (Verify)
            int err = RSX_OK;
            if (item.nshards != total) err = fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
            else err = check_shards(lens, total, false);
            if (err) {
                status[it] = err;
                continue;
            }
            if (c->m == 0) {
                if (item.op == RSX_OP_VERIFY) b->is_verify[it] = 2;  // no parity: trivially true
                continue;
            }
            for (int j = 0; j < k; ++j) in.push_back(item.shards[j]);
            // Verify: the pass recomputes the parity in registers and
            // compares it with these shards (compare mode) instead of writing.
            for (int i = k; i < total; ++i) out.push_back(item.shards[i]);
            if (item.op == RSX_OP_VERIFY) b->is_verify[it] = 1;
            tables = c->parity_tables_host.data();
            mtot = c->m;
        } else if (item.op == RSX_OP_RECONSTRUCT || item.op == RSX_OP_RECONSTRUCT_DATA) {
            // reedsolomon.go:1670-1709
            const bool data_only = item.op == RSX_OP_RECONSTRUCT_DATA;
            int err = RSX_OK;
            if (item.nshards != total) err = fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
            else err = check_shards(lens, total, true);
            if (err) {
                status[it] = err;
                continue;
            }
            int present = 0, data_present = 0;
            for (int i = 0; i < total; ++i)
                if (lens[i]) {
                    ++present;
                    if (i < k) ++data_present;
                }
            if (present == total || (data_only && data_present == k)) continue;
            if (present < k) {
                status[it] = fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
                continue;
            }
            std::shared_ptr<const DecodePlan> plan;
            {
                std::lock_guard<std::mutex> lock(c->mu);
                err = decode_plan(c, lens, data_only, nullptr, &plan);
                epoch = c->decode_epoch;
            }
            if (err) {
                status[it] = err;
                continue;
            }
            for (int i : plan->inputs) in.push_back(item.shards[i]);
            for (int i : plan->outputs) {
                out.push_back(item.shards[i]);
                if (fl) fl[i] = 1;
            }
            tables = plan->host_tables.data();
            mtot = static_cast<int>(plan->outputs.size());
            held_plans.push_back(std::move(plan));  // `tables` stays valid and unique while planning
        } else {
            status[it] = fail(RSX_ERR_INVALID_INPUT, "unknown batch op");
            continue;
        }
        {
            // Batched items run on device memory only; a missing or host
            // buffer becomes this item's status, never a kernel access.
            bool missing = false;
            for (const uint8_t *p : in) missing |= p == nullptr;
            for (const uint8_t *p : out) missing |= p == nullptr;
            bool host = false;
            if (!missing) {
                for (const uint8_t *p : in) host |= !ranges.is_device(p);
                for (const uint8_t *p : out) host |= !ranges.is_device(p);
            }
            if (missing) err_item = fail(RSX_ERR_INVALID_INPUT, "no buffer for a shard of a batched item");
            else if (host) err_item = fail(RSX_ERR_MIXED_MEMORY, "batched items need shards in device memory");
            if (err_item) {
                b->is_verify[it] = 0;
                if (fl) std::memset(fl, 0, c->total);
                status[it] = err_item;
                continue;
            }
        }
        const size_t len = shard_size(lens, total);
        uintptr_t bits = 0;
        for (auto *p : in) bits |= reinterpret_cast<uintptr_t>(p);
        for (auto *p : out) bits |= reinterpret_cast<uintptr_t>(p);
        const auto tkey = std::make_tuple(static_cast<const void *>(tables), static_cast<const rsx_codec *>(c), epoch);
        if (!last_table || tkey != last_table->first) {  // consecutive items often share tables
            auto tf = table_off.find(tkey);
            if (tf == table_off.end()) {
                tf = table_off.emplace(tkey, htab.size()).first;
                htab.insert(htab.end(), tables, tables + static_cast<size_t>(k) * mtot * 5);
            }
            last_table = &*tf;
        }
        const size_t toff = last_table->second;
        const uint32_t in_off = static_cast<uint32_t>(b->pool.size());
        b->pool.insert(b->pool.end(), in.begin(), in.end());
        for (int o0 = 0; o0 < mtot; o0 += 8) {
            const int nout = std::min(8, mtot - o0);
            const int mc = nout;
            const uint32_t out_off = static_cast<uint32_t>(b->pool.size());
            for (int i = 0; i < nout; ++i) b->pool.push_back(out[o0 + i]);
            rsx_batch::Task t{};
            t.tables = nullptr;
            t.len = len;
            t.in_off = in_off;
            t.out_off = out_off;
            t.k = static_cast<uint16_t>(k);
            t.nout = static_cast<uint16_t>(nout);
            t.mtot = static_cast<uint16_t>(mtot);
            t.o0 = static_cast<uint16_t>(o0);
            t.aligned = (bits & 15) == 0;
            t.mode = item.op == RSX_OP_VERIFY ? kGf8Compare : kGf8Clear;
            t.item = static_cast<uint32_t>(it);
            b->tasks[mc].push_back(t);
            pend.push_back({mc, b->tasks[mc].size() - 1, toff});
        }
    }
    for (int it = 0; it < n; ++it) b->has_verify |= b->is_verify[it] == 1;
    if (b->has_verify) {
        int e3 = reserve_grow(b->d_flags, sizeof(uint32_t) * std::max(n, 1));
        if (e3) return e3;
        if (b->h_flags_cap < static_cast<size_t>(n)) {
            if (b->h_flags) (void)hipHostFree(b->h_flags);
            b->h_flags = nullptr;
            RSX_HIP(hipHostMalloc(reinterpret_cast<void **>(&b->h_flags), sizeof(uint32_t) * std::max(n, 1)));
            b->h_flags_cap = static_cast<size_t>(n);
        }
    }
    // Layout of d_blob: tables, pointer pool, Leopard aux arrays, Leopard
    // task lists (sizes are known here; device pointers inside the task
    // lists are filled in against the blob's address below).
    size_t emp_count = 0;
    int max_groups = 0, max_p_emp = 0;
    for (const LeoItemPlan &lp : b->leo_emp) {
        emp_count += 2 * static_cast<size_t>(lp.groups) + 1;  // passes 1 and 2 per group, pass 3
        max_groups = std::max(max_groups, lp.groups);
        max_p_emp = std::max(max_p_emp, lp.emp.P);
    }
    size_t off = 0;
    const size_t o_tab = off;
    off += align256(htab.size() * 4);
    const size_t o_ptrs = off;
    off += align256(b->pool.size() * sizeof(void *));
    const size_t o_aux = off;
    off += align256(laux.size());
    size_t o_leo[3];
    for (int fi = 0; fi < 3; ++fi) {
        o_leo[fi] = off;
        off += align256(b->leo[fi].size() * sizeof(LeoBatchTask));
    }
    const size_t o_emp = off;
    off += align256(emp_count * sizeof(LeoEncPassTask));
    const size_t o_mp = off;
    off += align256(b->leo_mp.size() * sizeof(LeoPassTask));
    const size_t blob_bytes = std::max<size_t>(off, 256);
    int err = reserve_grow(b->d_blob, blob_bytes);
    if (!err) err = b->h_blob.reserve(blob_bytes);
    if (err) return err;
    uint8_t *dev = b->d_blob.as<uint8_t>();
    uint8_t *host = b->h_blob.ptr;
    const uint32_t *dev_tables = reinterpret_cast<const uint32_t *>(dev + o_tab);
    uint8_t *const *dev_pool = reinterpret_cast<uint8_t *const *>(dev + o_ptrs);
    uint8_t *dev_aux = dev + o_aux;
    b->dev_ptrs = reinterpret_cast<const uint8_t *const *>(dev_pool);
    if (!htab.empty()) std::memcpy(host + o_tab, htab.data(), htab.size() * 4);
    if (!b->pool.empty()) std::memcpy(host + o_ptrs, b->pool.data(), b->pool.size() * sizeof(void *));
    if (!laux.empty()) std::memcpy(host + o_aux, laux.data(), laux.size());
    for (auto &p : pend) b->tasks[p.mc][p.idx].tables = dev_tables + p.toff;
    // Leopard tasks: resolve pool / aux offsets to device addresses.
    for (int fi = 0; fi < 3; ++fi) {
        uint32_t slice0 = 0;
        for (const LeoItemPlan &lp : b->leo[fi]) {
            LeoBatchTask t = lp.task;
            t.slice0 = slice0;
            uint8_t *const *pp = dev_pool + lp.ptr_off;
            if (t.op == 0) {
                t.enc.data = pp;
                t.enc.parity = dev_pool + lp.parity_pool;
                if (lp.verify_item >= 0) {
                    t.enc.flag = b->d_flags.as<uint32_t>() + lp.verify_item;
                    t.enc.flag_value = 1;
                }
            } else {
                t.dec.rows = pp;
                t.dec.out = pp + t.dec.n;
                t.dec.errlocs = reinterpret_cast<const uint16_t *>(dev_aux + lp.errl_off);
                t.dec.out_rows = reinterpret_cast<const int *>(dev_aux + lp.rows_off);
                t.dec.out_logs = reinterpret_cast<const uint16_t *>(dev_aux + lp.logs_off);
                t.dec.need = reinterpret_cast<const uint32_t *>(dev_aux + lp.need_off);
            }
            if (static_cast<uint64_t>(slice0) + lp.slices > 0xFFFFFFFFull)
                return fail(RSX_ERR_INVALID_INPUT, "batch too large for one plan");
            slice0 += lp.slices;
            b->leo_lds[fi] = std::max(b->leo_lds[fi], lp.lds);
            b->leo_deriv[fi] = std::max(b->leo_deriv[fi], lp.regs);
            b->leo_tasks[fi].push_back(t);
        }
        b->leo_slices[fi] = slice0;
        b->dev_leo_tasks[fi] = reinterpret_cast<const LeoBatchTask *>(dev + o_leo[fi]);
        if (!b->leo_tasks[fi].empty())
            std::memcpy(host + o_leo[fi], b->leo_tasks[fi].data(), b->leo_tasks[fi].size() * sizeof(LeoBatchTask));
    }
    // GF(2^8) Leopard work (VALU-bound) beside classic passes (HBM-bound) that
    // move most of the plan's bytes: its launch is held to two workgroups per CU
    // so the classic workgroups share every CU with it (one plan of a served
    // request mix with 19% Leopard bytes: 4.98 -> 4.65 ms).  Otherwise it takes
    // the full grid it needs alone (a Leopard-only plan: 2.0 ms, 2.94 held).
    {
        uint64_t classic = 0, leo8 = 0;
        for (int mc = 1; mc <= 8; ++mc)
            for (const auto &t : b->tasks[mc]) classic += t.len * (static_cast<uint64_t>(t.k) + t.nout);
        for (const auto &t : b->leo_tasks[0])
            leo8 += t.op == 0 ? t.enc.size * static_cast<uint64_t>(t.enc.k + t.enc.p)
                              : t.dec.size * static_cast<uint64_t>(t.dec.n);
        b->leo8_per_cu = leo8 && classic >= 3 * leo8 ? 2 : 8;
    }
    if (!b->leo_mp.empty() || !b->leo_emp.empty())
        if ((err = b->d_mp_work.reserve(b->mp_work_bytes))) return err;
    if (!b->leo_emp.empty()) {
        // Task arrays in launch order: (pass 1, pass 2) for g = 0, 1, ..., then pass 3.
        auto add_launch = [&](int pass, int g) -> int {
            rsx_batch::EncLaunch L{b->emp_tasks.size(), 0, pass, max_p_emp, 0};
            for (const LeoItemPlan &lp : b->leo_emp) {
                if (pass != 3 && g >= lp.groups) continue;
                LeoEncPassTask t{};
                t.a = lp.emp;
                t.a.rows = dev_pool + lp.ptr_off + static_cast<size_t>(g) * t.a.m;
                t.a.parity = dev_pool + lp.parity_pool;
                t.a.work = b->d_mp_work.as<uint8_t>() + reinterpret_cast<size_t>(lp.emp.work);
                t.a.work2 = t.a.work + static_cast<size_t>(t.a.m) * static_cast<size_t>(t.a.slices) * 64;
                t.a.col0 = 0;
                t.a.mtrunc = std::min(t.a.m, lp.k - g * t.a.m);
                t.a.skew_off = t.a.m - 1 + g * t.a.m;
                t.a.first = g == 0;
                t.a.last = g == lp.groups - 1;
                if (pass == 3 && lp.verify_item >= 0) {  // Verify: pass 3 compares
                    t.a.flag = b->d_flags.as<uint32_t>() + lp.verify_item;
                    t.a.flag_value = 1;
                }
                t.block0 = L.blocks;
                const uint64_t nb = leo_enc_pass_blocks(t.a, pass);
                if (static_cast<uint64_t>(L.blocks) + nb > 0x7FFFFFFFull) return -1;
                L.blocks += static_cast<uint32_t>(nb);
                b->emp_tasks.push_back(t);
                ++L.ntasks;
            }
            if (L.ntasks) b->emp_launches.push_back(L);
            return 0;
        };
        for (int g = 0; g < max_groups; ++g)
            if (add_launch(1, g) || add_launch(2, g)) return fail(RSX_ERR_INVALID_INPUT, "batch too large for one plan");
        if (add_launch(3, 0)) return fail(RSX_ERR_INVALID_INPUT, "batch too large for one plan");
        if (b->emp_tasks.size() != emp_count) return fail(RSX_ERR_INVALID_INPUT, "encoder pass task count");
        std::memcpy(host + o_emp, b->emp_tasks.data(), emp_count * sizeof(LeoEncPassTask));
    }
    b->dev_emp_tasks = reinterpret_cast<const LeoEncPassTask *>(dev + o_emp);
    if (!b->leo_mp.empty()) {
        for (const LeoItemPlan &lp : b->leo_mp) {
            LeoPassTask t{};
            t.a = lp.mp;
            uint8_t *const *pp = dev_pool + lp.ptr_off;
            t.a.rows = pp;
            t.a.out = pp + t.a.n;
            t.a.errlocs = reinterpret_cast<const uint16_t *>(dev_aux + lp.errl_off);
            t.a.out_index = reinterpret_cast<const int *>(dev_aux + lp.outidx_off);
            t.a.out_logs = reinterpret_cast<const uint16_t *>(dev_aux + lp.logs_off);
            t.a.need = reinterpret_cast<const uint32_t *>(dev_aux + lp.need_off);
            t.a.work = b->d_mp_work.as<uint8_t>() + reinterpret_cast<size_t>(lp.mp.work);
            t.a.work2 = t.a.work + static_cast<size_t>(t.a.n) * static_cast<size_t>(t.a.slices) * 64;
            uint32_t nb[3];
            leo_pass_blocks(t.a, nb);
            for (int p = 0; p < 3; ++p) {
                if (static_cast<uint64_t>(b->mp_blocks[p]) + nb[p] > 0x7FFFFFFFull)
                    return fail(RSX_ERR_INVALID_INPUT, "batch too large for one plan");
                t.block0[p] = b->mp_blocks[p];
                b->mp_blocks[p] += nb[p];
            }
            b->mp_max_p = std::max(b->mp_max_p, t.a.P);
            b->mp_tasks.push_back(t);
        }
        std::memcpy(host + o_mp, b->mp_tasks.data(), b->mp_tasks.size() * sizeof(LeoPassTask));
    }
    b->dev_mp_tasks = reinterpret_cast<const LeoPassTask *>(dev + o_mp);
    if (off) RSX_HIP(hipMemcpyAsync(dev, host, off, hipMemcpyHostToDevice, b->stream));
    if (b->tuned) {
        if ((err = build_launches(b, b->tile))) return err;
    } else {
        if ((err = autotune(b))) return err;
    }
    RSX_HIP(hipEventRecord(b->uploaded, b->stream));
    if (b->use_graph) {
        RSX_HIP(hipStreamSynchronize(b->stream));  // capture starts on an idle stream
        if ((err = capture(b))) return err;
    }
    b->planned = true;
    return RSX_OK;
}

int rsx_batch_launch_on(rsx_batch *b, void *hip_stream) {
    RSX_TRACE("rsx_batch_launch");
    if (!b->planned) return fail(RSX_ERR_INVALID_INPUT, "batch has no plan");
    (void)hipSetDevice(b->device);
    hipStream_t s = hip_stream ? static_cast<hipStream_t>(hip_stream) : b->stream;
    // The plan's copies run on b->stream; another stream waits for them.
    if (s != b->stream) RSX_HIP(hipStreamWaitEvent(s, b->uploaded, 0));
    if (b->use_graph) RSX_HIP(hipGraphLaunch(b->exec, s));
    else return enqueue(b, s);
    return RSX_OK;
}

int rsx_batch_launch(rsx_batch *b) { return rsx_batch_launch_on(b, nullptr); }

int rsx_batch_set_graph(rsx_batch *b, int enabled) {
    b->use_graph = enabled != 0;
    b->planned = false;
    return RSX_OK;
}

int rsx_batch_verify_result(const rsx_batch *b, int item, int *ok) {
    *ok = 0;
    if (item < 0 || item >= b->nitems || !b->is_verify[item])
        return fail(RSX_ERR_INVALID_INPUT, "item is not a verify item of this plan");
    *ok = b->is_verify[item] == 2 || b->h_flags[item] == 0;
    return RSX_OK;
}

int rsx_batch_synchronize(rsx_batch *b) {
    (void)hipSetDevice(b->device);
    RSX_HIP(hipStreamSynchronize(b->stream));
    return RSX_OK;
}

int rsx_batch_num_tasks(const rsx_batch *b) {
    size_t n = 0;
    for (auto &v : b->tasks) n += v.size();
    for (auto &v : b->leo_tasks) n += v.size();
    n += b->mp_tasks.size() + b->leo_emp.size();
    return static_cast<int>(n);
}

}  // extern "C"

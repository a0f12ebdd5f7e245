// codec.cpp — the C ABI of include/rsx.h for the classic GF(2^8) engine and
// the dispatch to the Leopard engines (leopard.cpp).
//
// Argument validation and its order follow the reference methods cited at
// each entry point, so a caller sees the same error for the same input.
#include "codec.hpp"

#include <algorithm>
#include <cstring>

#include "batch.hpp"
#include "kernels.hpp"
#include "leopard.hpp"
#include "trace.hpp"

namespace rsx {

namespace {
thread_local std::string g_last_error;

constexpr size_t kPitchAlign = 256;      // staging slot alignment
constexpr size_t kDecodeCacheCap = 1 << 16;

size_t align_up(size_t n, size_t a) { return (n + a - 1) / a * a; }
}  // namespace

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return RSX_ERR_DEVICE;
}

DeviceBuffer::~DeviceBuffer() {
    if (ptr) (void)hipFree(ptr);
}

int DeviceBuffer::reserve(size_t n) {
    if (n <= bytes) return RSX_OK;
    if (ptr) {
        RSX_HIP(hipDeviceSynchronize());
        (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ptr = nullptr;
        return fail(RSX_ERR_NOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed");
    }
    bytes = n;
    return RSX_OK;
}

// reedsolomon.go:1422-1429 shardSize: first nonzero length.
size_t shard_size(const size_t *lens, int n) {
    for (int i = 0; i < n; ++i)
        if (lens[i]) return lens[i];
    return 0;
}

// reedsolomon.go:1404-1418 checkShards.
int check_shards(const size_t *lens, int n, bool nilok) {
    const size_t size = shard_size(lens, n);
    if (size == 0) return fail(RSX_ERR_SHARD_NO_DATA, "no shard data");
    for (int i = 0; i < n; ++i)
        if (lens[i] != size && (lens[i] != 0 || !nilok))
            return fail(RSX_ERR_SHARD_SIZE, "shard sizes do not match");
    return RSX_OK;
}

// Classifies one pointer as device (HBM / managed) or host memory.
MemKind classify(const void *p) {
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return MemKind::Host;
    }
    if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
        attr.type == hipMemoryTypeUnified)
        return MemKind::Device;
    return MemKind::Host;
}

bool DeviceRanges::is_device(const void *p) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    if (last_ < ranges_.size() && a >= ranges_[last_].first && a < ranges_[last_].second) return true;
    for (size_t i = 0; i < ranges_.size(); ++i)
        if (a >= ranges_[i].first && a < ranges_[i].second) {
            last_ = i;
            return true;
        }
    if (classify(p) != MemKind::Device) return false;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void *>(p)) == hipSuccess && size > 0) {
        const uintptr_t b = reinterpret_cast<uintptr_t>(base);
        ranges_.emplace_back(b, b + size);
        last_ = ranges_.size() - 1;
    } else {
        (void)hipGetLastError();
    }
    return true;
}

// All pointers of one call must share a memory class.
int classify_all(const std::vector<const void *> &ptrs, MemKind &kind) {
    bool have = false;
    DeviceRanges ranges;
    for (const void *p : ptrs) {
        if (!p) continue;
        const MemKind k = ranges.is_device(p) ? MemKind::Device : MemKind::Host;
        if (!have) {
            kind = k;
            have = true;
        } else if (k != kind) {
            return fail(RSX_ERR_MIXED_MEMORY, "host and device shards mixed in one call");
        }
    }
    if (!have) kind = MemKind::Host;
    return RSX_OK;
}

// Runs out = rows * in (or out ^= rows * in) in passes of at most kMaxOut
// outputs (pass_outputs).  `tables` is [nin][nrows][5].
int run_matmul(rsx_codec *c, const uint32_t *tables, const std::vector<const uint8_t *> &in,
               const std::vector<uint8_t *> &out, size_t len, int mode, uint32_t *flag, uint32_t flag_value) {
    const int nin = static_cast<int>(in.size());
    const int nrows = static_cast<int>(out.size());
    if (nin > kMaxIn) return fail(RSX_ERR_INVALID_INPUT, "too many inputs for one pass");
    for (int o0 = 0; o0 < nrows; o0 += pass_outputs(nrows - o0, nin)) {
        Gf8Pass p{};
        for (int j = 0; j < nin; ++j) p.in[j] = in[j];
        p.nout = pass_outputs(nrows - o0, nin);
        for (int i = 0; i < p.nout; ++i) p.out[i] = out[o0 + i];
        p.tables = tables;
        p.len = len;
        p.k = nin;
        p.mtot = nrows;
        p.o0 = o0;
        p.mode = mode;
        p.flag = flag;
        p.flag_value = flag_value;
        RSX_HIP(launch_gf8_pass(p, c->stream));
    }
    return RSX_OK;
}

int verify_flag_begin(rsx_codec *c, uint32_t **dev_flag, uint32_t *epoch) {
    if (!c->host_flag) {
        if (hipHostMalloc(reinterpret_cast<void **>(&c->host_flag), sizeof(uint32_t)) != hipSuccess ||
            hipHostGetDevicePointer(reinterpret_cast<void **>(&c->dev_flag), c->host_flag, 0) != hipSuccess) {
            if (c->host_flag) (void)hipHostFree(c->host_flag);
            c->host_flag = c->dev_flag = nullptr;
            (void)hipGetLastError();
            return fail(RSX_ERR_NOMEM, "pinned Verify flag");
        }
        *static_cast<volatile uint32_t *>(c->host_flag) = 0;
        c->verify_epoch = 0;
    }
    if (c->verify_epoch == 0xFFFFFFFFu) {  // wrapped; the previous call has synchronised
        *static_cast<volatile uint32_t *>(c->host_flag) = 0;
        c->verify_epoch = 0;
    }
    *epoch = ++c->verify_epoch;
    *dev_flag = c->dev_flag;
    return RSX_OK;
}

bool verify_flag_mismatch(const rsx_codec *c, uint32_t epoch) {
    return *static_cast<volatile const uint32_t *>(c->host_flag) == epoch;
}

// Uploads coefficient tables for `rows` (nrows x nin) into a [nin][nrows][5]
// device buffer and leaves the same tables in `host`.
int upload_tables(const std::vector<const uint8_t *> &rows, int nin, DeviceBuffer &dst,
                  std::vector<uint32_t> &host, hipStream_t stream) {
    const int nrows = static_cast<int>(rows.size());
    host.assign(static_cast<size_t>(nin) * nrows * kTableDwords, 0);
    for (int j = 0; j < nin; ++j)
        for (int i = 0; i < nrows; ++i)
            coefficient_table(rows[i][j], &host[(static_cast<size_t>(j) * nrows + i) * kTableDwords]);
    int err = dst.reserve(std::max<size_t>(host.size() * 4, 4));
    if (err) return err;
    RSX_HIP(hipMemcpyAsync(dst.ptr, host.data(), host.size() * 4, hipMemcpyHostToDevice, stream));
    RSX_HIP(hipStreamSynchronize(stream));
    return RSX_OK;
}

int stage_slots(rsx_codec *c, size_t len, size_t nslots, std::vector<uint8_t *> &dev) {
    const size_t pitch = align_up(len, kPitchAlign);
    int err = c->staging.reserve(std::max<size_t>(pitch * nslots, 1));
    if (err) return err;
    dev.resize(nslots);
    for (size_t i = 0; i < nslots; ++i) dev[i] = c->staging.as<uint8_t>() + pitch * i;
    return RSX_OK;
}

// Stages host shards into HBM: returns device addresses per index in `idx`.
// Caller holds c->mu.
int stage_in(rsx_codec *c, size_t len, const std::vector<int> &load, const std::vector<int> &slots,
             uint8_t *const *shards, std::vector<uint8_t *> &dev, size_t nslots) {
    const size_t pitch = align_up(len, kPitchAlign);
    int err = c->staging.reserve(pitch * nslots);
    if (err) return err;
    uint8_t *base = c->staging.as<uint8_t>();
    dev.assign(slots.size(), nullptr);
    for (size_t s = 0; s < slots.size(); ++s) dev[s] = base + pitch * slots[s];
    for (size_t s = 0; s < load.size(); ++s)
        if (load[s] >= 0)
            RSX_HIP(hipMemcpyAsync(dev[s], shards[load[s]], len, hipMemcpyHostToDevice, c->stream));
    return RSX_OK;
}

}  // namespace rsx

using namespace rsx;

extern "C" {

void rsx_options_default(rsx_options *o) {
    std::memset(o, 0, sizeof *o);
    o->leopard = RSX_LEOPARD_AS_NEEDED;
    o->matrix = RSX_MATRIX_DEFAULT;
    o->inversion_cache = 1;
    o->device = 0;
}

const char *rsx_last_error(void) { return g_last_error.c_str(); }

int rsx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

// reedsolomon.go:425-601 New.
int rsx_new(int k, int m, const rsx_options *opt, rsx_codec **out) {
    *out = nullptr;
    rsx_options o;
    if (opt) o = *opt;
    else rsx_options_default(&o);
    const long total = static_cast<long>(k) + m;

    int engine = 0;
    // reedsolomon.go:434-440 engine routing.
    if ((o.leopard == RSX_LEOPARD_GF16 && m > 0) || total > 256) engine = RSX_LEOPARD_GF16;
    else if (o.leopard == RSX_LEOPARD_GF8 && m > 0) engine = RSX_LEOPARD_GF8;

    if (engine) {
        // leopard.go:51-69 / leopard8.go:42-66 newFF16 / newFF8.
        if (k <= 0 || m <= 0) return fail(RSX_ERR_INV_SHARD_NUM, "invalid shard counts");
        if (total > 65536) return fail(RSX_ERR_MAX_SHARD_NUM, "more than 65536 shards");
        if (!leopard_shape_ok(engine == RSX_LEOPARD_GF8 ? 8 : 16, k, m))
            return fail(RSX_ERR_MAX_SHARD_NUM,
                        "Leopard needs k + ceilPow2(m) <= field order (the reference indexes past its tables)");
    } else {
        if (total > 256) return fail(RSX_ERR_MAX_SHARD_NUM, "more than 256 shards");
        if (k <= 0 || m < 0) return fail(RSX_ERR_INV_SHARD_NUM, "invalid shard counts");
    }

    int ndev = rsx_device_count();
    if (ndev <= 0) return fail(RSX_ERR_DEVICE, "no HIP device available");
    if (o.device < 0 || o.device >= ndev) return fail(RSX_ERR_DEVICE, "invalid device ordinal");
    RSX_HIP(hipSetDevice(o.device));

    std::unique_ptr<rsx_codec> c(new rsx_codec);
    c->k = k;
    c->m = m;
    c->total = static_cast<int>(total);
    c->engine = engine;
    c->device = o.device;
    c->inversion_cache = o.inversion_cache != 0;
    RSX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;

    if (engine) {
        int err = leopard_init(c.get());
        if (err) return err;
    } else if (m > 0) {
        int err = build_coding_matrix(k, m, o.matrix, o.fast_one_parity != 0, o.custom_matrix,
                                      o.custom_rows, o.custom_cols, c->matrix);
        if (err) return fail(err, "coding matrix construction failed");
        std::vector<const uint8_t *> rows(m);
        for (int i = 0; i < m; ++i) rows[i] = c->matrix.row(k + i);
        err = upload_tables(rows, k, c->parity_tables, c->parity_tables_host, c->stream);
        if (err) return err;
        c->noalloc_work.assign(static_cast<size_t>(2 * k + m) * k, 0);
    }
    *out = c.release();
    return RSX_OK;
}

void rsx_free(rsx_codec *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->decode_cache.clear();
    c->leo.reset();
    if (c->host_flag) (void)hipHostFree(c->host_flag);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int rsx_shard_size_multiple(const rsx_codec *c) { return c->engine ? 64 : 1; }
int rsx_data_shards(const rsx_codec *c) { return c->k; }
int rsx_parity_shards(const rsx_codec *c) { return c->m; }
int rsx_total_shards(const rsx_codec *c) { return c->total; }
int rsx_engine(const rsx_codec *c) { return c->engine; }

int rsx_coding_matrix(const rsx_codec *c, uint8_t *out) {
    if (c->engine) return fail(RSX_ERR_NOT_SUPPORTED, "Leopard engines have no coding matrix");
    if (c->m == 0) {
        std::memset(out, 0, static_cast<size_t>(c->k) * c->k);
        for (int i = 0; i < c->k; ++i) out[i * c->k + i] = 1;
        return RSX_OK;
    }
    std::memcpy(out, c->matrix.v.data(), c->matrix.v.size());
    return RSX_OK;
}

int rsx_set_stream(rsx_codec *c, void *hip_stream) {
    (void)hipSetDevice(c->device);
    RSX_HIP(hipStreamSynchronize(c->stream));
    if (c->own_stream) RSX_HIP(hipStreamDestroy(c->stream));
    c->own_stream = hip_stream == nullptr;
    if (hip_stream) c->stream = static_cast<hipStream_t>(hip_stream);
    else RSX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    return RSX_OK;
}

void *rsx_get_stream(rsx_codec *c) { return c->stream; }

int rsx_synchronize(rsx_codec *c) {
    (void)hipSetDevice(c->device);
    RSX_HIP(hipStreamSynchronize(c->stream));
    return RSX_OK;
}

int rsx_decode_cache_size(const rsx_codec *c) {
    return static_cast<int>(c->decode_cache.size());
}

// reedsolomon.go:686-702 Encode.
int rsx_encode(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards) {
    RSX_TRACE("rsx_encode");
    if (c->engine) return leopard_encode(c, shards, lens, nshards);
    if (nshards != c->total) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    int err = check_shards(lens, nshards, false);
    if (err) return err;
    if (c->m == 0) return RSX_OK;
    (void)hipSetDevice(c->device);
    const size_t len = lens[0];
    std::vector<const void *> ptrs(shards, shards + nshards);
    MemKind kind;
    if ((err = classify_all(ptrs, kind))) return err;
    const uint32_t *tab = c->parity_tables.as<uint32_t>();
    if (kind == MemKind::Device) {
        std::vector<const uint8_t *> in(shards, shards + c->k);
        std::vector<uint8_t *> out(shards + c->k, shards + c->total);
        return run_matmul(c, tab, in, out, len, kGf8Clear);
    }
    std::lock_guard<std::mutex> lock(c->mu);
    std::vector<int> load(c->total, -1), slots(c->total);
    for (int i = 0; i < c->total; ++i) {
        slots[i] = i;
        if (i < c->k) load[i] = i;
    }
    std::vector<uint8_t *> dev;
    if ((err = stage_in(c, len, load, slots, shards, dev, c->total))) return err;
    std::vector<const uint8_t *> in(dev.begin(), dev.begin() + c->k);
    std::vector<uint8_t *> out(dev.begin() + c->k, dev.end());
    if ((err = run_matmul(c, tab, in, out, len, kGf8Clear))) return err;
    for (int i = 0; i < c->m; ++i)
        RSX_HIP(hipMemcpyAsync(shards[c->k + i], out[i], len, hipMemcpyDeviceToHost, c->stream));
    RSX_HIP(hipStreamSynchronize(c->stream));
    return RSX_OK;
}

// reedsolomon.go:708-757 EncodeIdx: parity ^= column idx of the parity rows * data.
int rsx_encode_idx(rsx_codec *c, const uint8_t *data, size_t len, int idx, uint8_t *const *parity,
                   const size_t *parity_lens, int nparity) {
    RSX_TRACE("rsx_encode_idx");
    if (c->engine) return fail(RSX_ERR_NOT_SUPPORTED, "operation not supported");
    if (nparity != c->m) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    if (nparity == 0) return RSX_OK;
    if (idx < 0 || idx >= c->k) return fail(RSX_ERR_INV_SHARD_NUM, "shard index out of range");
    int err = check_shards(parity_lens, nparity, false);
    if (err) return err;
    if (parity_lens[0] != len) return fail(RSX_ERR_SHARD_SIZE, "shard sizes do not match");
    (void)hipSetDevice(c->device);
    std::vector<const void *> ptrs(parity, parity + nparity);
    ptrs.push_back(data);
    MemKind kind;
    if ((err = classify_all(ptrs, kind))) return err;
    // Tables are [k][m][5]: the block of input idx is the column needed here.
    const uint32_t *tab = c->parity_tables.as<uint32_t>() + static_cast<size_t>(idx) * c->m * kTableDwords;
    if (kind == MemKind::Device) {
        std::vector<const uint8_t *> in{data};
        std::vector<uint8_t *> out(parity, parity + nparity);
        return run_matmul(c, tab, in, out, len, kGf8Accumulate);
    }
    std::lock_guard<std::mutex> lock(c->mu);
    std::vector<uint8_t *> all(parity, parity + nparity);
    all.push_back(const_cast<uint8_t *>(data));
    std::vector<int> load(nparity + 1), slots(nparity + 1);
    for (int i = 0; i <= nparity; ++i) load[i] = slots[i] = i;
    std::vector<uint8_t *> dev;
    if ((err = stage_in(c, len, load, slots, all.data(), dev, nparity + 1))) return err;
    std::vector<const uint8_t *> in{dev[nparity]};
    std::vector<uint8_t *> out(dev.begin(), dev.begin() + nparity);
    if ((err = run_matmul(c, tab, in, out, len, kGf8Accumulate))) return err;
    for (int i = 0; i < nparity; ++i)
        RSX_HIP(hipMemcpyAsync(parity[i], out[i], len, hipMemcpyDeviceToHost, c->stream));
    RSX_HIP(hipStreamSynchronize(c->stream));
    return RSX_OK;
}

// reedsolomon.go:857-871 Verify + 1373-1387 checkSomeShards.
int rsx_verify(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards, int *ok) {
    RSX_TRACE("rsx_verify");
    *ok = 0;
    if (c->engine) return leopard_verify(c, shards, lens, nshards, ok);
    if (nshards != c->total) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    int err = check_shards(lens, nshards, false);
    if (err) return err;
    if (c->m == 0) {
        *ok = 1;
        return RSX_OK;
    }
    (void)hipSetDevice(c->device);
    const size_t len = lens[0];
    std::vector<const void *> ptrs(shards, shards + nshards);
    MemKind kind;
    if ((err = classify_all(ptrs, kind))) return err;
    std::lock_guard<std::mutex> lock(c->mu);
    std::vector<uint8_t *> dev(shards, shards + nshards);
    if (kind == MemKind::Host) {
        std::vector<int> load(c->total), slots(c->total);
        for (int i = 0; i < c->total; ++i) load[i] = slots[i] = i;
        if ((err = stage_in(c, len, load, slots, shards, dev, c->total))) return err;
    }
    // The parity is recomputed in registers and compared with the stored
    // parity in the same pass (reedsolomon.go:1373-1387 re-encodes into
    // temporaries and compares); a mismatch stores this call's epoch straight
    // into the mapped host flag, so there is neither a reset nor a copy back.
    uint32_t *flag = nullptr, epoch = 0;
    if ((err = verify_flag_begin(c, &flag, &epoch))) return err;
    std::vector<const uint8_t *> in(dev.begin(), dev.begin() + c->k);
    std::vector<uint8_t *> out(dev.begin() + c->k, dev.end());
    if ((err = run_matmul(c, c->parity_tables.as<uint32_t>(), in, out, len, kGf8Compare, flag, epoch))) return err;
    RSX_HIP(hipStreamSynchronize(c->stream));
    *ok = !verify_flag_mismatch(c, epoch);
    return RSX_OK;
}

}  // extern "C"

namespace rsx {

// Finds or builds the decode plan for one erasure pattern.  The key is the
// full presence mask plus the requested outputs, which determines both the
// inverse (the reference keys its inversion tree by the invalid-row prefix,
// inversion_tree.go:41-91) and the output rows.  With WithInversionCache
// (false) (options.go:152-157) every call inverts afresh and nothing is
// kept beyond the call's own plan.  Caller holds c->mu.
int decode_plan(rsx_codec *c, const size_t *lens, bool data_only, const uint8_t *required,
                std::shared_ptr<const DecodePlan> *plan) {
    const int k = c->k, total = c->total;
    std::string key;
    if (c->inversion_cache) {
        key.assign(1 + 2 * ((total + 7) / 8), '\0');
        key[0] = static_cast<char>(data_only ? 1 : 0) | static_cast<char>(required ? 2 : 0);
        for (int i = 0; i < total; ++i) {
            if (lens[i]) key[1 + i / 8] |= static_cast<char>(1 << (i % 8));
            if (required && required[i]) key[1 + (total + 7) / 8 + i / 8] |= static_cast<char>(1 << (i % 8));
        }
        auto it = c->decode_cache.find(key);
        if (it != c->decode_cache.end()) {
            *plan = it->second;
            return RSX_OK;
        }
        if (c->decode_cache.size() >= kDecodeCacheCap) {
            RSX_HIP(hipStreamSynchronize(c->stream));
            c->decode_cache.clear();
            ++c->decode_epoch;
        }
    }
    auto p = std::make_shared<DecodePlan>();
    // reedsolomon.go:1718-1730: the first k present rows.
    for (int r = 0; r < total && static_cast<int>(p->inputs.size()) < k; ++r)
        if (lens[r]) p->inputs.push_back(r);
    Matrix sub(k, k), inv;
    for (int r = 0; r < k; ++r) std::memcpy(sub.row(r), c->matrix.row(p->inputs[r]), k);
    if (!invert(sub, inv)) return fail(RSX_ERR_SINGULAR, "matrix is singular");
    std::vector<std::vector<uint8_t>> rows;
    for (int i = 0; i < k; ++i)
        if (!lens[i] && (!required || required[i])) {
            rows.emplace_back(inv.row(i), inv.row(i) + k);
            p->outputs.push_back(i);
        }
    if (!data_only) {
        const Gf256 &f = gf();
        for (int i = k; i < total; ++i)
            if (!lens[i] && (!required || required[i])) {
                // reedsolomon.go:620-631 multiplyRowWithMatrix(parity row, inverse).
                std::vector<uint8_t> r(k, 0);
                const uint8_t *prow = c->matrix.row(i);
                for (int j = 0; j < k; ++j)
                    for (int t = 0; t < k; ++t) r[j] ^= f.mul[prow[t]][inv.at(t, j)];
                rows.push_back(std::move(r));
                p->outputs.push_back(i);
            }
    }
    std::vector<const uint8_t *> rp;
    for (auto &r : rows) rp.push_back(r.data());
    if (!rp.empty()) {
        int err = upload_tables(rp, k, p->tables, p->host_tables, c->stream);
        if (err) return err;
    }
    if (c->inversion_cache) c->decode_cache.emplace(std::move(key), p);
    else c->uncached_plan = p;  // frees the previous call's plan (hipFree waits for the device)
    *plan = std::move(p);
    return RSX_OK;
}

// reedsolomon.go:1670-1781 reconstruct.
int reconstruct_classic(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
                        bool data_only, const uint8_t *required, int nrequired, uint8_t *filled) {
    const int k = c->k, total = c->total;
    if (nshards != total || (required && nrequired < k))
        return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    int err = check_shards(lens, nshards, true);
    if (err) return err;
    const size_t len = shard_size(lens, nshards);
    int present = 0, data_present = 0, missing_required = 0;
    for (int i = 0; i < total; ++i) {
        if (lens[i]) {
            ++present;
            if (i < k) ++data_present;
        } else if (required && required[i]) {
            ++missing_required;
        }
    }
    if (present == total || (data_only && data_present == k) || (required && missing_required == 0))
        return RSX_OK;
    if (present < k) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    (void)hipSetDevice(c->device);

    std::lock_guard<std::mutex> lock(c->mu);
    std::shared_ptr<const DecodePlan> plan;
    if ((err = decode_plan(c, lens, data_only, required, &plan))) return err;
    const int nout = static_cast<int>(plan->outputs.size());
    std::vector<const void *> ptrs;
    for (int i : plan->inputs) ptrs.push_back(shards[i]);
    for (int i : plan->outputs) {
        if (!shards[i]) return fail(RSX_ERR_INVALID_INPUT, "no output buffer for an absent shard");
        ptrs.push_back(shards[i]);
    }
    MemKind kind;
    if ((err = classify_all(ptrs, kind))) return err;
    const uint32_t *tab = plan->tables.as<uint32_t>();
    if (kind == MemKind::Device) {
        std::vector<const uint8_t *> in;
        for (int i : plan->inputs) in.push_back(shards[i]);
        std::vector<uint8_t *> out;
        for (int i : plan->outputs) out.push_back(shards[i]);
        if ((err = run_matmul(c, tab, in, out, len, kGf8Clear))) return err;
    } else {
        std::vector<int> load(k + nout, -1), slots(k + nout);
        for (int s = 0; s < k + nout; ++s) slots[s] = s;
        for (int s = 0; s < k; ++s) load[s] = plan->inputs[s];
        std::vector<uint8_t *> dev;
        if ((err = stage_in(c, len, load, slots, shards, dev, k + nout))) return err;
        std::vector<const uint8_t *> in(dev.begin(), dev.begin() + k);
        std::vector<uint8_t *> out(dev.begin() + k, dev.end());
        if ((err = run_matmul(c, tab, in, out, len, kGf8Clear))) return err;
        for (int s = 0; s < nout; ++s)
            RSX_HIP(hipMemcpyAsync(shards[plan->outputs[s]], out[s], len, hipMemcpyDeviceToHost,
                                   c->stream));
        RSX_HIP(hipStreamSynchronize(c->stream));
    }
    for (int i : plan->outputs) filled[i] = 1;
    return RSX_OK;
}

}  // namespace rsx

extern "C" {

int rsx_reconstruct(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
                    uint8_t *filled) {
    RSX_TRACE("rsx_reconstruct");
    std::memset(filled, 0, static_cast<size_t>(std::max(nshards, 0)));
    if (c->engine) return leopard_reconstruct(c, shards, lens, nshards, true, filled);
    return reconstruct_classic(c, shards, lens, nshards, false, nullptr, 0, filled);
}

int rsx_reconstruct_data(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
                         uint8_t *filled) {
    RSX_TRACE("rsx_reconstruct_data");
    std::memset(filled, 0, static_cast<size_t>(std::max(nshards, 0)));
    if (c->engine) return leopard_reconstruct(c, shards, lens, nshards, false, filled);
    return reconstruct_classic(c, shards, lens, nshards, true, nullptr, 0, filled);
}

// reedsolomon.go:1655-1660 ReconstructSome; leopard.go:378-383.
int rsx_reconstruct_some(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
                         const uint8_t *required, int nrequired, uint8_t *filled) {
    RSX_TRACE("rsx_reconstruct_some");
    std::memset(filled, 0, static_cast<size_t>(std::max(nshards, 0)));
    if (c->engine)
        return leopard_reconstruct(c, shards, lens, nshards, nrequired == c->total, filled);
    if ((nrequired != c->k && nrequired != c->total) || !required)
        return fail(RSX_ERR_INVALID_INPUT, "invalid input");
    // A k-entry mask leaves the parity shards not required (the reference
    // indexes required[i] for every missing shard and panics past k).
    std::vector<uint8_t> req(static_cast<size_t>(c->total), 0);
    std::copy(required, required + nrequired, req.begin());
    return reconstruct_classic(c, shards, lens, nshards, nrequired == c->k, req.data(), c->total,
                               filled);
}

}  // extern "C"

// leopard.cpp — the Leopard GF(2^16) / GF(2^8) engines on the GPU.
//
// Encode (leopard.go:169-266, leopard8.go:139-243): per column chunk, the
// IFFT of each group of m data rows (zero-padded) is XORed into an m-row
// accumulator, then an FFT truncated to the parity count produces the
// parity rows.  Reconstruct (leopard.go:421-647, leopard8.go:393-634): rows
// are scaled by the error-locator logs, IFFT over n rows, formal derivative,
// FFT truncated to m+k rows, and the erased rows are scaled back.
//
// The work arrays live in HBM.  Columns are processed in chunks sized so the
// work set stays within the GPU's last-level (Infinity) cache.  Error
// locators are computed on the host once per erasure pattern and kept in
// HBM, keyed by the full erasure mask and mode.
#include "leopard.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>

#include "batch.hpp"
#include "kernels.hpp"
#include "leo_kernels.hpp"
#include "leopard_field.hpp"

namespace rsx {

// Field tables resident on one device (shared by every codec on it).
struct LeoDeviceTables {
    DeviceBuffer tables, skew, walsh;
    LeoDev dev;
};

static std::mutex g_leo_mu;
static std::map<std::pair<int, int>, std::unique_ptr<LeoDeviceTables>> g_leo_tables;

static int device_tables(int device, int bits, hipStream_t stream, const LeoDev **out) {
    std::lock_guard<std::mutex> lock(g_leo_mu);
    auto key = std::make_pair(device, bits);
    auto it = g_leo_tables.find(key);
    if (it == g_leo_tables.end()) {
        const LeoField &f = leo_field(bits);
        auto t = std::make_unique<LeoDeviceTables>();
        std::vector<uint32_t> mt = f.mul_tables();
        int err = t->tables.reserve(mt.size() * 4);
        if (err) return err;
        err = t->skew.reserve(f.skew.size() * 2 + 64);
        if (err) return err;
        err = t->walsh.reserve(f.walsh.size() * 2);
        if (err) return err;
        RSX_HIP(hipMemcpyAsync(t->tables.ptr, mt.data(), mt.size() * 4, hipMemcpyHostToDevice, stream));
        RSX_HIP(hipMemcpyAsync(t->skew.ptr, f.skew.data(), f.skew.size() * 2, hipMemcpyHostToDevice, stream));
        RSX_HIP(hipMemcpyAsync(t->walsh.ptr, f.walsh.data(), f.walsh.size() * 2, hipMemcpyHostToDevice, stream));
        RSX_HIP(hipStreamSynchronize(stream));
        t->dev.bits = bits;
        t->dev.modulus = f.modulus;
        t->dev.tables = t->tables.as<uint32_t>();
        t->dev.skew = t->skew.as<uint16_t>();
        t->dev.walsh = t->walsh.as<uint16_t>();
        it = g_leo_tables.emplace(key, std::move(t)).first;
    }
    *out = &it->second->dev;
    return RSX_OK;
}

// Error locators of one erasure pattern.  A GF(2^16) pattern first met by a
// per-stripe reconstruct gets them from the device (leo_error_locators_device,
// straight into d_blob); the host copy, which batched plans copy into their
// own uploads, is computed when a plan first asks for it.
struct LeoPattern {
    std::vector<uint16_t> errlocs;  // n entries, or empty until a host copy is needed
    int input_count = 0;
    // What a per-stripe reconstruct reads, uploaded on its first use in one
    // allocation and one copy: the error locators, the work rows of the
    // erased shards, their reveal logs, row -> output index, FFT pruning bits.
    bool outs_ready = false;
    int nout = 0;
    uint64_t pass3_blocks = 0;  // three-pass reconstruct: row blocks with an output row
    std::vector<uint8_t> host_blob;  // source of the copy (kept while it may run)
    DeviceBuffer d_blob;
    const uint16_t *d_errlocs = nullptr, *d_logs = nullptr;
    const int *d_rows = nullptr, *d_outidx = nullptr;
    const uint32_t *d_need = nullptr;
};

struct LeopardState {
    int bits = 16;
    bool fused = true;  // LDS-resident kernels where the rows fit
    const LeoField *field = nullptr;
    const LeoDev *dev = nullptr;
    DeviceBuffer work, work2, ptrs, rows, stage, scratch, flag, cmp;
    std::vector<uint8_t *> h_ptrs;  // what `ptrs` holds on the device
    int enc_rows = 0;               // `rows` holds 0 .. enc_rows-1 (encode's output rows)
    // Shared so a batch planner that looked a pattern up under the codec's
    // lock keeps it alive if another thread flushes the map afterwards.
    std::unordered_map<std::string, std::shared_ptr<LeoPattern>> patterns;
};

void LeopardDeleter::operator()(LeopardState *s) const { delete s; }

bool leopard8_shape_ok(int k, int m) { return leopard_shape_ok(8, k, m); }

int leo_device_tables(int device, int bits, hipStream_t stream, const LeoDev **out) {
    return device_tables(device, bits, stream, out);
}

// Largest fftSkew index the encoder's grouped IFFT reads (leopard.go:764-873
// on skewLUT = fftSkew[m-1+g*m:]).
static long max_encode_skew_index(int k, int p) {
    const int m = ceil_pow2(p);
    long mx = -1;
    for (int g = 0; g * m < k; ++g) {
        const int mtrunc = std::min(m, k - g * m);
        const long off = m - 1 + static_cast<long>(g) * m;
        int dist = 1, dist4 = 4;
        while (dist4 <= m) {
            const long r_last = static_cast<long>((mtrunc - 1) / dist4) * dist4;
            mx = std::max(mx, off + r_last + 3L * dist);
            dist = dist4;
            dist4 <<= 2;
        }
        if (dist < m) mx = std::max(mx, off + dist);
    }
    return mx;
}

// Shapes whose tables the reference indexes in range: errLocs has `order`
// entries and row m+k-1 is written, fftSkew has `modulus` entries.
bool leopard_shape_ok(int bits, int k, int m) {
    const long order = 1L << bits;
    if (static_cast<long>(k) + ceil_pow2(m) > order) return false;
    return max_encode_skew_index(k, m) < order - 1;
}

int leopard_init(rsx_codec *c) {
    std::unique_ptr<LeopardState, LeopardDeleter> s(new LeopardState);
    s->bits = c->engine == RSX_LEOPARD_GF8 ? 8 : 16;
    s->field = &leo_field(s->bits);
    // RSX_LEOPARD_LAYERED=1 keeps every codec created afterwards on the
    // layer-per-launch kernels (both schedules give the same bytes).
    const char *layered = std::getenv("RSX_LEOPARD_LAYERED");
    s->fused = !(layered && layered[0] == '1');
    int err = device_tables(c->device, s->bits, c->stream, &s->dev);
    if (err) return err;
    c->leo = std::move(s);
    return RSX_OK;
}

namespace {

size_t shard_size_of(const size_t *lens, int n) {
    for (int i = 0; i < n; ++i)
        if (lens[i]) return lens[i];
    return 0;
}

int check(const size_t *lens, int n, bool nilok) {
    const size_t size = shard_size_of(lens, n);
    if (size == 0) return fail(RSX_ERR_SHARD_NO_DATA, "no shard data");
    for (int i = 0; i < n; ++i)
        if (lens[i] != size && (lens[i] != 0 || !nilok))
            return fail(RSX_ERR_SHARD_SIZE, "shard sizes do not match");
    return RSX_OK;
}

// Column chunk (multiple of 64 bytes) so that `rows` rows stay within budget.
size_t chunk_for(size_t size, int rows, size_t budget) {
    size_t c = budget / static_cast<size_t>(rows);
    c -= c % 64;
    c = std::max<size_t>(c, 64);
    return std::min(c, size);
}

constexpr size_t kEncodeBudget = 64u << 20;
constexpr size_t kDecodeBudget = 64u << 20;
constexpr bool kQuarterTiles = true;  // GF(2^16) reconstructs with n = 4096: one quarter-chunk tile
constexpr size_t kQuarterEncodeMin = 4096;  // shard bytes: 256 quarter slices, one per CU
constexpr int kPassEncodeMin = 2048;  // m from which encodes without a tile run the three passes

// LDS budget of a one-tile encode: the m rows of a slice, and with several
// groups their XOR accumulator in registers (at most the derivative's
// m * slice <= 64 KiB).
size_t encode_lds_budget(int m, int groups) {
    return groups > 1 ? std::min(kLeoFusedEncodeLds, kLeoFusedDecodeLds / m * (m + 1)) : kLeoFusedEncodeLds;
}

bool all_aligned(const std::vector<uint8_t *> &p) {
    uintptr_t bits = 0;
    for (uint8_t *q : p) bits |= reinterpret_cast<uintptr_t>(q);
    return (bits & 15) == 0;
}

}  // namespace

std::vector<uint32_t> leo_fft_need(int n, const std::vector<int> &out_rows) {
    std::vector<uint32_t> v;
    auto level = [&](int gsize) {  // one bit per group of gsize rows
        const size_t off = v.size();
        v.resize(off + (n / gsize + 31) / 32, 0u);
        for (int r : out_rows) {
            const int g = r / gsize;
            v[off + g / 32] |= 1u << (g % 32);
        }
    };
    int dist4 = n, dist = n >> 2;
    while (dist != 0) {  // the FFT's radix-4 steps (leopard.go:696-735)
        level(4 * dist);
        dist4 = dist;
        dist >>= 2;
    }
    if (dist4 == 2) level(2);  // its final radix-2 layer
    return v;
}

namespace {

// Row pointer table of a call in s.ptrs; skipped when the device copy already
// holds this table (repeated calls on the same buffers).  The copy is ordered
// on the codec's stream after every kernel that read the previous table.
// Uploads the call's device pointer table, ordered on the codec stream.  A
// copy from pageable host memory has consumed its source when the call
// returns; the host copy is kept only to skip re-uploading an unchanged
// table (repeated calls on the same buffers).
int upload_ptrs(rsx_codec *c, std::vector<uint8_t *> &&p) {
    LeopardState &s = *c->leo;
    const size_t bytes = std::max<size_t>(p.size(), 1) * sizeof(void *);
    if (p == s.h_ptrs && bytes <= s.ptrs.bytes) return RSX_OK;
    int err = s.ptrs.reserve(bytes);
    if (err) return err;
    RSX_HIP(hipMemcpyAsync(s.ptrs.ptr, p.data(), p.size() * sizeof(void *), hipMemcpyHostToDevice, c->stream));
    s.h_ptrs = std::move(p);
    return RSX_OK;
}

// IFFT of one encoder group into `w` (m rows), leopard.go:764-873.
int encode_group(rsx_codec *c, const LeopardState &s, const uint8_t *const *dptrs, size_t col_off,
                 int mtrunc, int m, const LeoWork &w, int skew_off) {
    const LeoDev &f = *s.dev;
    RSX_HIP(leo_load_rows(f, dptrs, nullptr, col_off, m, w, c->stream));
    const uint16_t *skew = f.skew + skew_off;
    int dist = 1, dist4 = 4;
    while (dist4 <= m) {
        RSX_HIP(leo_step4(f, true, w, dist, mtrunc, skew, 0, c->stream));
        dist = dist4;
        dist4 <<= 2;
    }
    if (dist < m) RSX_HIP(leo_ifft2_last(f, w, dist, s.field->skew[skew_off + dist], c->stream));
    return RSX_OK;
}

// In-place FFT truncated to mtrunc outputs, leopard.go:696-735.
int fft(rsx_codec *c, const LeopardState &s, const LeoWork &w, int mtrunc, int m) {
    const LeoDev &f = *s.dev;
    int dist4 = m, dist = m >> 2;
    while (dist != 0) {
        RSX_HIP(leo_step4(f, false, w, dist, mtrunc, f.skew, -1, c->stream));
        dist4 = dist;
        dist >>= 2;
    }
    if (dist4 == 2) RSX_HIP(leo_fft2_last(f, w, mtrunc, c->stream));
    return RSX_OK;
}

// Column slice of the one-LDS-tile encode of `size`-byte shards, or 0 when
// the shape runs one launch per layer.
int encode_tile_slice(const rsx_codec *c, size_t size) {
    const LeopardState &s = *c->leo;
    const int m = ceil_pow2(c->m);
    const int groups = (c->k + m - 1) / m;
    if (!s.fused) return 0;
    // GF(2^16) m >= 2048 below 4 KiB shards: too few tiles to fill the device;
    // the three passes run 1.2-1.5x faster there (4096+4096 x 1 KiB 0.086 ->
    // 0.056 ms), the tiles win from 16 KiB.
    if (s.bits == 16 && m >= 2048 && size < kQuarterEncodeMin) return 0;
    const int slice = leo_fused_slice(s.bits, m, 1, m, size, encode_lds_budget(m, groups));
    // GF(2^16) m = 4096 with several groups: quarter-chunk slices once the
    // shard gives every CU one (4 KiB; about even with the three passes, which
    // take m = 8192 at 1.3-1.5x the quarter tile and everything below 4 KiB).
    if (!slice && s.bits == 16 && kQuarterTiles && m <= 4096 && size >= kQuarterEncodeMin)
        return leo_quarter_encode_slice(m, groups);
    return slice;
}

// Encodes without a tile run the three tiled passes (GF(2^16), m >= 2048).
bool encode_three_pass(const rsx_codec *c) {
    const LeopardState &s = *c->leo;
    return s.fused && s.bits == 16 && ceil_pow2(c->m) >= kPassEncodeMin;
}

// Encode with device pointers (16-byte aligned) for data (k) and parity (p).
// With `flag` (one-tile and three-pass shapes) the parity rows are compared
// with `parity` instead of written, and a mismatch stores flag_value to *flag.
int encode_device(rsx_codec *c, const std::vector<uint8_t *> &data, const std::vector<uint8_t *> &parity,
                  size_t size, uint32_t *flag = nullptr, uint32_t flag_value = 0) {
    LeopardState &s = *c->leo;
    const int k = c->k, p = c->m;
    const int m = ceil_pow2(p);
    const int mtrunc = std::min(k, m);
    const int last_count = k % m;
    const int groups = (k + m - 1) / m;
    // Group-major pointer table: group g occupies entries [g*m, g*m+m), null past its rows.
    std::vector<uint8_t *> tab(static_cast<size_t>(groups) * m + p, nullptr);
    for (int j = 0; j < k; ++j) tab[j] = data[j];
    for (int i = 0; i < p; ++i) tab[static_cast<size_t>(groups) * m + i] = parity[i];
    int err = upload_ptrs(c, std::move(tab));
    if (err) return err;
    if (s.enc_rows != p) {  // the layer-per-launch path writes rows 0 .. p-1
        std::vector<int> rows(p);
        for (int i = 0; i < p; ++i) rows[i] = i;
        if ((err = s.rows.reserve(p * sizeof(int)))) return err;
        RSX_HIP(hipMemcpyAsync(s.rows.ptr, rows.data(), p * sizeof(int), hipMemcpyHostToDevice, c->stream));
        s.enc_rows = p;
    }

    const uint8_t *const *dptr = reinterpret_cast<const uint8_t *const *>(s.ptrs.ptr);
    uint8_t *const *pptr = reinterpret_cast<uint8_t *const *>(s.ptrs.as<uint8_t *>() + static_cast<size_t>(groups) * m);
    const int slice = encode_tile_slice(c, size);
    int eC = 0, eP = 0, eA = 0;
    const bool passes = !slice && encode_three_pass(c) && leo_multipass_shape(m, &eC, &eP, &eA);
    if (flag && !slice && !passes) return fail(RSX_ERR_INVALID_INPUT, "compare-mode encode needs a tiled schedule");
    if (slice) {
        const LeoEncodeArgs a{dptr, pptr, size, k, p, m, groups, slice, flag, flag_value};
        RSX_HIP(leo_encode_fused(*s.dev, a, c->stream));
        return RSX_OK;
    }
    // No tile: three tiled passes (per group: IFFT layers < C, then >= C and
    // the XOR into the accumulator; after the last group the FFT); 1.1-2.7x
    // the layered schedule.
    if (passes) {
        const size_t cw = leo_multipass_chunk(m, size);
        if ((err = s.work.reserve(cw * m))) return err;
        if ((err = s.work2.reserve(cw * m))) return err;
        for (size_t off = 0; off < size; off += cw) {
            LeoEncPassArgs a{};
            a.parity = pptr;
            a.work = s.work.as<uint8_t>();
            a.work2 = s.work2.as<uint8_t>();
            a.col0 = off;
            a.slices = static_cast<int>(std::min(cw, size - off) / 64);
            a.m = m;
            a.p = p;
            a.C = eC;
            a.P = eP;
            a.A = eA;
            a.flag = flag;
            a.flag_value = flag_value;
            for (int g = 0; g < groups; ++g) {
                a.rows = dptr + static_cast<size_t>(g) * m;
                a.mtrunc = std::min(m, k - g * m);
                a.skew_off = m - 1 + g * m;
                a.first = g == 0;
                a.last = g == groups - 1;
                RSX_HIP(leo_encode_multipass(*s.dev, a, 1, c->stream));
                RSX_HIP(leo_encode_multipass(*s.dev, a, 2, c->stream));
            }
            RSX_HIP(leo_encode_multipass(*s.dev, a, 3, c->stream));
        }
        return RSX_OK;
    }
    const size_t chunk = chunk_for(size, 2 * m, kEncodeBudget);
    if ((err = s.work.reserve(chunk * 2 * m))) return err;
    for (size_t off = 0; off < size; off += chunk) {
        const size_t cols = std::min(chunk, size - off);
        LeoWork acc{s.work.as<uint8_t>(), chunk, cols};
        LeoWork tmp{s.work.as<uint8_t>() + chunk * m, chunk, cols};
        if ((err = encode_group(c, s, dptr, off, mtrunc, m, acc, m - 1))) return err;
        if (m < k) {
            int g = 1;
            for (int i = m; i + m <= k; i += m, ++g) {
                if ((err = encode_group(c, s, dptr + static_cast<size_t>(g) * m, off, m, m, tmp, m - 1 + g * m)))
                    return err;
                RSX_HIP(leo_xor_rows(acc, tmp, m, c->stream));
            }
            if (last_count) {
                if ((err = encode_group(c, s, dptr + static_cast<size_t>(g) * m, off, last_count, m, tmp,
                                        m - 1 + g * m)))
                    return err;
                RSX_HIP(leo_xor_rows(acc, tmp, m, c->stream));
            }
        }
        if ((err = fft(c, s, acc, p, m))) return err;
        RSX_HIP(leo_store_rows(*s.dev, acc, s.rows.as<int>(), nullptr, pptr, off, p, c->stream));
    }
    return RSX_OK;
}

// Caller holds c->mu.  host_errlocs: the caller reads pat->errlocs.
int pattern_for(rsx_codec *c, const uint8_t *erased, bool recover_all, std::shared_ptr<LeoPattern> *out,
                bool host_errlocs = true) {
    LeopardState &s = *c->leo;
    const int k = c->k, p = c->m, total = c->total;
    std::string key(1 + (total + 7) / 8, '\0');
    key[0] = recover_all ? 1 : 0;
    for (int i = 0; i < total; ++i)
        if (erased[i]) key[1 + i / 8] |= static_cast<char>(1 << (i % 8));
    auto it = s.patterns.find(key);
    if (it == s.patterns.end()) {
        if (s.patterns.size() >= 4096) {
            RSX_HIP(hipStreamSynchronize(c->stream));
            s.patterns.clear();
        }
        auto pat = std::make_shared<LeoPattern>();
        std::vector<uint8_t> ind;
        pat->input_count = leo_erasure_indicator(k, p, erased, ind);
        it = s.patterns.emplace(std::move(key), std::move(pat)).first;
    }
    LeoPattern &pat = *it->second;
    // Filled once, under c->mu, and never changed after: readers that got
    // the pattern earlier never see it change.
    if (host_errlocs && pat.errlocs.empty()) pat.errlocs = leo_error_locators(*s.field, k, p, erased, nullptr);
    *out = it->second;
    return RSX_OK;
}

// Reconstruct with aligned device pointers; shards[i] is the buffer of shard i
// (input when present, output when erased and requested).
int reconstruct_device(rsx_codec *c, const std::vector<uint8_t *> &shards, const uint8_t *erased,
                       bool recover_all, size_t size, uint8_t *filled) {
    LeopardState &s = *c->leo;
    const int k = c->k, p = c->m, total = c->total;
    const int m = ceil_pow2(p), n = ceil_pow2(m + k);
    std::shared_ptr<LeoPattern> pat;
    // GF(2^16) error locators of a new pattern come from the device below.
    int err = pattern_for(c, erased, recover_all, &pat, s.bits != 16);
    if (err) return err;
    // Row table for the scaled load: parity rows, zero rows, data rows, zero rows.
    std::vector<uint8_t *> tab(static_cast<size_t>(n), nullptr);
    for (int i = 0; i < p; ++i) tab[i] = erased[k + i] ? nullptr : shards[k + i];
    for (int i = 0; i < k; ++i) tab[m + i] = erased[i] ? nullptr : shards[i];
    const int end = recover_all ? total : k;
    for (int i = 0; i < end; ++i) {
        if (!erased[i]) continue;
        tab.push_back(shards[i]);
        filled[i] = 1;
    }
    if (!pat->outs_ready) {
        // Output rows, reveal logs (leopard.go:640-646) and row -> output
        // index depend only on the pattern: uploaded once.
        std::vector<int> rows, out_index(static_cast<size_t>(n), -1);
        for (int i = 0; i < end; ++i) {
            if (!erased[i]) continue;
            const int pos = i >= k ? i - k : i + m;
            out_index[pos] = static_cast<int>(rows.size());
            rows.push_back(pos);
        }
        const int no = static_cast<int>(rows.size());
        // Without a host copy, the error locators and reveal logs are
        // computed into the blob by the device from the erasure indicator
        // (one workgroup, ~0.35 ms of host transforms saved at n = 8192).
        const bool on_device = pat->errlocs.empty();
        std::vector<uint8_t> ind;
        if (on_device) leo_erasure_indicator(k, p, erased, ind);
        const std::vector<uint32_t> need = leo_fft_need(n, rows);
        // One buffer: errlocs (n), logs (no), rows (no), out_index (n), need, indicator.
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t o_err = 0, o_logs = up(static_cast<size_t>(n) * 2), o_rows = o_logs + up(rows.size() * 2),
                     o_idx = o_rows + up(rows.size() * 4), o_need = o_idx + up(out_index.size() * 4),
                     o_ind = o_need + up(std::max<size_t>(need.size(), 1) * 4), total_bytes = o_ind + up(ind.size());
        pat->host_blob.assign(total_bytes, 0);
        uint8_t *h = pat->host_blob.data();
        if (!on_device) {
            std::memcpy(h + o_err, pat->errlocs.data(), static_cast<size_t>(n) * 2);
            uint16_t *hl = reinterpret_cast<uint16_t *>(h + o_logs);
            for (int q = 0; q < no; ++q) hl[q] = static_cast<uint16_t>(s.field->modulus - pat->errlocs[rows[q]]);
        }
        if (no) std::memcpy(h + o_rows, rows.data(), rows.size() * 4);
        std::memcpy(h + o_idx, out_index.data(), out_index.size() * 4);
        if (!need.empty()) std::memcpy(h + o_need, need.data(), need.size() * 4);
        if (!ind.empty()) std::memcpy(h + o_ind, ind.data(), ind.size());
        if ((err = pat->d_blob.reserve(total_bytes))) return err;
        RSX_HIP(hipMemcpyAsync(pat->d_blob.ptr, h, total_bytes, hipMemcpyHostToDevice, c->stream));
        uint8_t *d = pat->d_blob.as<uint8_t>();
        if (on_device)
            RSX_HIP(leo_error_locators_device(*s.dev, d + o_ind, static_cast<int>(ind.size()), n,
                                              reinterpret_cast<const int *>(d + o_rows), no,
                                              reinterpret_cast<uint16_t *>(d + o_err),
                                              reinterpret_cast<uint16_t *>(d + o_logs), c->stream));
        pat->d_errlocs = reinterpret_cast<const uint16_t *>(d + o_err);
        pat->d_logs = reinterpret_cast<const uint16_t *>(d + o_logs);
        pat->d_rows = reinterpret_cast<const int *>(d + o_rows);
        pat->d_outidx = reinterpret_cast<const int *>(d + o_idx);
        pat->d_need = reinterpret_cast<const uint32_t *>(d + o_need);
        pat->nout = no;
        int pC = 0, pP = 0, pA = 0;
        if (leo_multipass_shape(n, &pC, &pP, &pA)) pat->pass3_blocks = leo_pass3_blocks(pP, rows);
        pat->outs_ready = true;
    }
    const int nout = pat->nout;
    if ((err = upload_ptrs(c, std::move(tab)))) return err;
    const int ifft_trunc = s.bits == 16 ? pat->input_count : m + k;
    const uint8_t *const *lptr = reinterpret_cast<const uint8_t *const *>(s.ptrs.ptr);
    uint8_t *const *optr = reinterpret_cast<uint8_t *const *>(s.ptrs.as<uint8_t *>() + n);
    // n rows in LDS; the formal derivative also holds n * slice bytes in registers.
    int slice = s.fused ? leo_fused_slice(s.bits, n, 1, n, size, std::min(kLeoFusedEncodeLds,
                                                                      kLeoFusedDecodeLds / n * (n + 1)))
                        : 0;
    // n = 4096 GF(2^16) transforms: one tile of quarter-chunk slices.
    if (!slice && s.fused && s.bits == 16 && kQuarterTiles) slice = leo_quarter_slice(n);
    if (slice) {
        const LeoDecodeArgs a{lptr, pat->d_errlocs, pat->d_rows, pat->d_logs, optr,
                              size, n, nout, ifft_trunc, m + k, slice, pat->d_need};
        if (nout) RSX_HIP(leo_decode_fused(*s.dev, a, c->stream));
        return RSX_OK;
    }
    int mC = 0, mP = 0, mA = 0;
    if (s.fused && s.bits == 16 && leo_multipass_shape(n, &mC, &mP, &mA)) {
        if (!nout) return RSX_OK;
        const size_t cw = leo_multipass_chunk(n, size);
        if ((err = s.work.reserve(cw * n))) return err;
        if ((err = s.work2.reserve(cw * n))) return err;
        for (size_t off = 0; off < size; off += cw) {
            const size_t cols = std::min(cw, size - off);
            const LeoMultipassArgs a{lptr, pat->d_errlocs, pat->d_outidx, pat->d_logs,
                                     optr, s.work.as<uint8_t>(), s.work2.as<uint8_t>(), off,
                                     static_cast<int>(cols / 64), n, ifft_trunc, m + k, mC, mP, mA,
                                     pat->d_need, pat->pass3_blocks};
            RSX_HIP(leo_decode_multipass(*s.dev, a, c->stream));
        }
        return RSX_OK;
    }
    const size_t chunk = chunk_for(size, 2 * n, kDecodeBudget);
    if ((err = s.work.reserve(chunk * n))) return err;
    if ((err = s.work2.reserve(chunk * n))) return err;
    const LeoDev &f = *s.dev;
    for (size_t off = 0; off < size; off += chunk) {
        const size_t cols = std::min(chunk, size - off);
        LeoWork w{s.work.as<uint8_t>(), chunk, cols};
        LeoWork w2{s.work2.as<uint8_t>(), chunk, cols};
        RSX_HIP(leo_load_rows(f, lptr, pat->d_errlocs, off, n, w, c->stream));
        // leopard.go:650-693 ifftDITDecoder
        int dist = 1, dist4 = 4;
        while (dist4 <= n) {
            RSX_HIP(leo_step4(f, true, w, dist, ifft_trunc, f.skew, -1, c->stream));
            dist = dist4;
            dist4 <<= 2;
        }
        if (dist < n) RSX_HIP(leo_ifft2_last(f, w, dist, s.field->skew[dist - 1], c->stream));
        RSX_HIP(leo_derivative(w, w2, n, c->stream));
        if ((err = fft(c, s, w2, m + k, n))) return err;
        if (nout)
            RSX_HIP(leo_store_rows(f, w2, pat->d_rows, pat->d_logs, optr, off, nout, c->stream));
    }
    return RSX_OK;
}

// Resolves the device addresses of a call: device shards in place when
// aligned, otherwise staged in HBM.  `load[i]` = shard i must be copied in,
// `save[i]` = shard i must be copied back after the work.
struct Resolved {
    std::vector<uint8_t *> dev;
    bool staged = false;
    MemKind kind = MemKind::Device;
};

int resolve(rsx_codec *c, uint8_t *const *shards, int total, size_t size, const std::vector<bool> &use,
            const std::vector<bool> &load, Resolved &r) {
    LeopardState &s = *c->leo;
    bool have = false;
    DeviceRanges ranges;  // one runtime query per allocation, not per shard
    for (int i = 0; i < total; ++i) {
        if (!use[i] || !shards[i]) continue;
        const MemKind kk = ranges.is_device(shards[i]) ? MemKind::Device : MemKind::Host;
        if (!have) {
            r.kind = kk;
            have = true;
        } else if (kk != r.kind) {
            return fail(RSX_ERR_MIXED_MEMORY, "host and device shards mixed in one call");
        }
    }
    r.dev.assign(total, nullptr);
    for (int i = 0; i < total; ++i)
        if (use[i]) r.dev[i] = shards[i];
    std::vector<uint8_t *> used;
    for (int i = 0; i < total; ++i)
        if (use[i]) used.push_back(shards[i]);
    r.staged = r.kind == MemKind::Host || !all_aligned(used);
    if (!r.staged) return RSX_OK;
    int err = s.stage.reserve(size * total);
    if (err) return err;
    const hipMemcpyKind dir = r.kind == MemKind::Host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    for (int i = 0; i < total; ++i) {
        if (!use[i]) continue;
        r.dev[i] = s.stage.as<uint8_t>() + size * i;
        if (load[i]) RSX_HIP(hipMemcpyAsync(r.dev[i], shards[i], size, dir, c->stream));
    }
    return RSX_OK;
}

int unstage(rsx_codec *c, uint8_t *const *shards, const Resolved &r, const std::vector<bool> &save,
            size_t size) {
    if (!r.staged) return RSX_OK;
    const hipMemcpyKind dir = r.kind == MemKind::Host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    for (size_t i = 0; i < save.size(); ++i)
        if (save[i]) RSX_HIP(hipMemcpyAsync(shards[i], r.dev[i], size, dir, c->stream));
    return RSX_OK;
}

}  // namespace

// leopard.go:139-150 Encode / leopard8.go:120-130.
int leopard_encode(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards) {
    if (nshards != c->total) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    int err = check(lens, nshards, false);
    if (err) return err;
    const size_t size = shard_size_of(lens, nshards);
    if (size % 64) return fail(RSX_ERR_INVALID_SHARD_SIZE, "shard size must be a multiple of 64");
    (void)hipSetDevice(c->device);
    std::lock_guard<std::mutex> lock(c->mu);
    const int total = c->total, k = c->k;
    std::vector<bool> use(total, true), load(total, false), save(total, false);
    for (int i = 0; i < total; ++i) (i < k ? load : save)[i] = true;
    Resolved r;
    if ((err = resolve(c, shards, total, size, use, load, r))) return err;
    std::vector<uint8_t *> data(r.dev.begin(), r.dev.begin() + k), parity(r.dev.begin() + k, r.dev.end());
    if ((err = encode_device(c, data, parity, size))) return err;
    if ((err = unstage(c, shards, r, save, size))) return err;
    // Device-resident calls are ordered on the codec's stream (rsx_synchronize
    // waits); host-memory calls return with the results in place.
    if (r.staged) RSX_HIP(hipStreamSynchronize(c->stream));
    return RSX_OK;
}

// leopard.go:393-419 Verify: re-encode into temporaries and compare.
int leopard_verify(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards, int *ok) {
    *ok = 0;
    if (nshards != c->total) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    int err = check(lens, nshards, false);
    if (err) return err;
    const size_t size = shard_size_of(lens, nshards);
    if (size % 64) return fail(RSX_ERR_INVALID_SHARD_SIZE, "shard size must be a multiple of 64");
    (void)hipSetDevice(c->device);
    std::lock_guard<std::mutex> lock(c->mu);
    LeopardState &s = *c->leo;
    const int total = c->total, k = c->k, p = c->m;
    std::vector<bool> use(total, true), load(total, true), save(total, false);
    Resolved r;
    if ((err = resolve(c, shards, total, size, use, load, r))) return err;
    if (encode_tile_slice(c, size) || encode_three_pass(c)) {
        // The parity rows are recomputed in LDS and compared with the stored
        // ones by the encode kernel itself (the one tile, or the encoder's
        // pass 3): no temporaries, no compare launch.
        uint32_t *flag = nullptr, epoch = 0;
        if ((err = verify_flag_begin(c, &flag, &epoch))) return err;
        std::vector<uint8_t *> data(r.dev.begin(), r.dev.begin() + k), parity(r.dev.begin() + k, r.dev.end());
        if ((err = encode_device(c, data, parity, size, flag, epoch))) return err;
        RSX_HIP(hipStreamSynchronize(c->stream));
        *ok = !verify_flag_mismatch(c, epoch);
        return RSX_OK;
    }
    if ((err = s.scratch.reserve(size * p))) return err;
    if ((err = s.flag.reserve(sizeof(int)))) return err;
    std::vector<uint8_t *> data(r.dev.begin(), r.dev.begin() + k), tmp(p);
    for (int i = 0; i < p; ++i) tmp[i] = s.scratch.as<uint8_t>() + size * i;
    if ((err = encode_device(c, data, tmp, size))) return err;
    std::vector<CompareEntry> cmp(p);
    for (int i = 0; i < p; ++i) cmp[i] = {tmp[i], r.dev[k + i], size, 0, 0};
    if ((err = s.cmp.reserve(p * sizeof(CompareEntry)))) return err;
    RSX_HIP(hipMemcpyAsync(s.cmp.ptr, cmp.data(), p * sizeof(CompareEntry), hipMemcpyHostToDevice, c->stream));
    RSX_HIP(launch_fill_u32(s.flag.as<uint32_t>(), 1, 0u, c->stream));
    RSX_HIP(launch_batch_compare(s.cmp.as<CompareEntry>(), p, s.flag.as<uint32_t>(), c->stream));
    uint32_t mismatch = 0;
    RSX_HIP(hipMemcpyAsync(&mismatch, s.flag.ptr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    RSX_HIP(hipStreamSynchronize(c->stream));
    *ok = mismatch == 0;
    return RSX_OK;
}

// leopard.go:421-580 / leopard8.go:393-634 reconstruct.
int leopard_reconstruct(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
                        bool recover_all, uint8_t *filled) {
    const int k = c->k, total = c->total;
    if (nshards != total) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    int err = check(lens, nshards, true);
    if (err) return err;
    int present = 0, data_present = 0;
    for (int i = 0; i < total; ++i)
        if (lens[i]) {
            ++present;
            if (i < k) ++data_present;
        }
    if (present == total || (!recover_all && data_present == k)) return RSX_OK;
    if (present < k) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    const size_t size = shard_size_of(lens, nshards);
    if (size % 64) return fail(RSX_ERR_INVALID_SHARD_SIZE, "shard size must be a multiple of 64");
    (void)hipSetDevice(c->device);
    std::lock_guard<std::mutex> lock(c->mu);
    std::vector<uint8_t> erased(total);
    const int end = recover_all ? total : k;
    std::vector<bool> use(total, false), load(total, false), save(total, false);
    for (int i = 0; i < total; ++i) {
        erased[i] = lens[i] == 0;
        if (!erased[i]) use[i] = load[i] = true;
        else if (i < end) {
            if (!shards[i]) return fail(RSX_ERR_INVALID_INPUT, "no output buffer for an absent shard");
            use[i] = save[i] = true;
        }
    }
    Resolved r;
    if ((err = resolve(c, shards, total, size, use, load, r))) return err;
    err = reconstruct_device(c, r.dev, erased.data(), recover_all, size, filled);
    if (!err) err = unstage(c, shards, r, save, size);
    if (!err && r.staged) {
        const hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) err = hip_fail(e, "hipStreamSynchronize(c->stream)");
    }
    // filled marks only shards written by a call that succeeded.
    if (err) std::memset(filled, 0, static_cast<size_t>(total));
    return err;
}

namespace {

// Shard size handed to the slice chooser for batched items: large, so the
// slice is sized for full workgroups rather than for slice count.
constexpr size_t kBatchedSize = size_t{1} << 40;

size_t aux_append(std::vector<uint8_t> &aux, const void *src, size_t bytes) {
    const size_t off = (aux.size() + 15) / 16 * 16;
    aux.resize(off + bytes);
    std::memcpy(aux.data() + off, src, bytes);
    return off;
}

}  // namespace

int leopard_plan_item(rsx_codec *c, int op, uint8_t *const *shards, const size_t *lens, int nshards,
                      std::vector<const uint8_t *> &pool, std::vector<uint8_t> &aux, LeoItemPlan &out,
                      uint8_t *filled, DeviceRanges &ranges) {
    const int k = c->k, p = c->m, total = c->total;
    LeopardState &s = *c->leo;
    out = LeoItemPlan{};
    out.bits = s.bits;
    out.kind = s.bits == 16 ? 1 : 0;
    out.dev = s.dev;
    if (nshards != total) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    const bool rec = op == RSX_OP_RECONSTRUCT || op == RSX_OP_RECONSTRUCT_DATA;
    int err = check(lens, nshards, rec);
    if (err) return err;
    const size_t size = shard_size_of(lens, nshards);
    const int m = ceil_pow2(p), n = ceil_pow2(m + k);
    std::vector<uint8_t> erased(total, 0);
    int present = 0, data_present = 0;
    for (int i = 0; i < total; ++i) {
        erased[i] = lens[i] == 0;
        if (!erased[i]) {
            ++present;
            if (i < k) ++data_present;
        }
    }
    if (rec) {
        // leopard.go:421-440: nothing to do, or too few shards.
        if (present == total || (op == RSX_OP_RECONSTRUCT_DATA && data_present == k)) {
            out.empty = true;
            return RSX_OK;
        }
        if (present < k) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    }
    if (size % 64) return fail(RSX_ERR_INVALID_SHARD_SIZE, "shard size must be a multiple of 64");
    const bool recover_all = op != RSX_OP_RECONSTRUCT_DATA;
    const int end = recover_all ? total : k;
    // Device memory, 16-byte aligned, a buffer for every output.
    uintptr_t bits = 0;
    for (int i = 0; i < total; ++i) {
        const bool used = !erased[i] || (rec && i < end) || op == RSX_OP_ENCODE;
        if (!used) continue;
        if (!shards[i]) return fail(RSX_ERR_INVALID_INPUT, "no buffer for a shard of a batched item");
        if (!ranges.is_device(shards[i]))
            return fail(RSX_ERR_MIXED_MEMORY, "batched items need shards in device memory");
        bits |= reinterpret_cast<uintptr_t>(shards[i]);
    }
    if (bits & 15) return fail(RSX_ERR_NOT_SUPPORTED, "batched Leopard items need 16-byte aligned shards");
    const LeoDev &f = *s.dev;
    if (!rec) {
        const int groups = (k + m - 1) / m;
        // Other items fill the device: size the slice for full workgroups.
        int slice = leo_fused_slice(s.bits, m, 1, m, kBatchedSize, encode_lds_budget(m, groups));
        // Quarter-chunk tiles for m = 4096 (the batch kernel's 256 threads hold
        // the groups' XOR); m = 8192 takes the three passes, as per call.
        if (!slice && s.bits == 16 && kQuarterTiles && m <= 4096 &&
            (slice = leo_quarter_encode_slice(m, groups)) != 0)
            out.kind = 2;
        int eC = 0, eP = 0, eA = 0;
        const bool passes = !slice && s.bits == 16 && s.fused && leo_multipass_shape(m, &eC, &eP, &eA);
        if (!slice && !passes) return fail(RSX_ERR_NOT_SUPPORTED, "Leopard shape too tall for the batched kernel");
        out.ptr_off = pool.size();
        for (int g = 0; g < groups; ++g)
            for (int j = 0; j < m; ++j) {
                const int r = g * m + j;
                pool.push_back(r < k ? shards[r] : nullptr);
            }
        out.parity_pool = pool.size();
        for (int i = 0; i < p; ++i) pool.push_back(shards[k + i]);
        if (passes) {
            out.kind = 4;
            out.k = k;
            out.groups = groups;
            out.emp.slices = static_cast<int>(size / 64);
            out.emp.m = m;
            out.emp.p = p;
            out.emp.C = eC;
            out.emp.P = eP;
            out.emp.A = eA;
            out.work_bytes = 2 * static_cast<size_t>(m) * size;
            out.slices = static_cast<uint32_t>(size / 64);
            return RSX_OK;
        }
        out.task.op = 0;
        out.task.enc = LeoEncodeArgs{nullptr, nullptr, size, k, p, m, groups, slice};
        // GF(2^8) with several groups: a second tile-shaped LDS region holds
        // the groups' XOR (the batch kernel's pipelined encode).
        out.lds = static_cast<size_t>(slice) * (m + 1) * (s.bits == 8 && groups > 1 ? 2 : 1);
        out.regs = groups > 1 ? static_cast<size_t>(m) * slice : 0;
        out.slices = static_cast<uint32_t>((size + slice - 1) / slice);
        (void)f;
        return RSX_OK;
    }
    int slice = leo_fused_slice(s.bits, n, 1, n, kBatchedSize, std::min(kLeoFusedEncodeLds,
                                                                         kLeoFusedDecodeLds / n * (n + 1)));
    if (!slice && s.bits == 16 && kQuarterTiles && (slice = leo_quarter_slice(n)) != 0) out.kind = 2;
    int mC = 0, mP = 0, mA = 0;
    const bool passes = !slice && s.bits == 16 && s.fused && leo_multipass_shape(n, &mC, &mP, &mA);
    if (!slice && !passes) return fail(RSX_ERR_NOT_SUPPORTED, "Leopard shape too tall for the batched kernel");
    std::shared_ptr<LeoPattern> pat;
    {
        std::lock_guard<std::mutex> lock(c->mu);
        if ((err = pattern_for(c, erased.data(), recover_all, &pat))) return err;
    }
    out.ptr_off = pool.size();
    std::vector<const uint8_t *> tab(static_cast<size_t>(n), nullptr);
    for (int i = 0; i < p; ++i) tab[i] = erased[k + i] ? nullptr : shards[k + i];
    for (int i = 0; i < k; ++i) tab[m + i] = erased[i] ? nullptr : shards[i];
    std::vector<int> rows;
    std::vector<uint16_t> logs;
    for (int i = 0; i < end; ++i) {
        if (!erased[i]) continue;
        const int pos = i >= k ? i - k : i + m;
        rows.push_back(pos);
        logs.push_back(static_cast<uint16_t>(s.field->modulus - pat->errlocs[pos]));
        tab.push_back(shards[i]);
        if (filled) filled[i] = 1;
    }
    pool.insert(pool.end(), tab.begin(), tab.end());
    out.errl_off = aux_append(aux, pat->errlocs.data(), pat->errlocs.size() * sizeof(uint16_t));
    out.rows_off = aux_append(aux, rows.data(), rows.size() * sizeof(int));
    out.logs_off = aux_append(aux, logs.data(), logs.size() * sizeof(uint16_t));
    const std::vector<uint32_t> need = leo_fft_need(n, rows);
    out.need_off = aux_append(aux, need.data(), need.size() * sizeof(uint32_t));
    const int nout = static_cast<int>(rows.size());
    if (passes) {
        // Three passes over the whole column in the batch's work arrays.
        std::vector<int> out_index(static_cast<size_t>(n), -1);
        for (int q = 0; q < nout; ++q) out_index[rows[q]] = q;
        out.outidx_off = aux_append(aux, out_index.data(), out_index.size() * sizeof(int));
        out.kind = 3;
        out.mp = LeoMultipassArgs{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                                  static_cast<int>(size / 64), n, pat->input_count, m + k, mC, mP, mA, nullptr,
                                  leo_pass3_blocks(mP, rows)};
        out.work_bytes = 2 * static_cast<size_t>(n) * size;
        out.slices = static_cast<uint32_t>(size / 64);
        out.empty = nout == 0;
        return RSX_OK;
    }
    out.task.op = 1;
    out.task.dec = LeoDecodeArgs{nullptr, nullptr, nullptr, nullptr, nullptr, size, n, nout,
                                 s.bits == 16 ? pat->input_count : m + k, m + k, slice};
    out.lds = static_cast<size_t>(slice) * (n + 1);
    out.regs = static_cast<size_t>(n) * slice;
    out.slices = static_cast<uint32_t>((size + slice - 1) / slice);
    out.empty = nout == 0;
    return RSX_OK;
}

}  // namespace rsx

// update.cpp — Encoder.Update and Extensions.DecodeIdx of the classic engine
// (include/rsx.h) on the same GF(2^8) pass kernels as Encode.
//
// Update: every changed old data shard becomes old ^ new (a pass with the
// coefficient 1, accumulating), then parity ^= M[parity rows][changed] *
// (old ^ new) in one accumulating pass.  DecodeIdx: the rows of the inverse of
// the first k expected shards (times the parity rows for parity outputs),
// restricted to the inputs given in this call, accumulated into dst.
// Shards in host memory are staged through HBM and copied back.
#include <algorithm>
#include <cstring>

#include "codec.hpp"
#include "kernels.hpp"
#include "trace.hpp"

using namespace rsx;

namespace {

int one_table(rsx_codec *c, const uint32_t **out) {
    if (!c->one_table.ptr) {
        const uint8_t one = 1;
        std::vector<const uint8_t *> rows{&one};
        std::vector<uint32_t> host;
        int err = upload_tables(rows, 1, c->one_table, host, c->stream);
        if (err) return err;
    }
    *out = c->one_table.as<uint32_t>();
    return RSX_OK;
}

// dst[i] ^= src[i] for every pair (device pointers).
int xor_into(rsx_codec *c, const std::vector<const uint8_t *> &src, const std::vector<uint8_t *> &dst,
             const std::vector<size_t> &len) {
    const uint32_t *one = nullptr;
    int err = one_table(c, &one);
    if (err) return err;
    for (size_t i = 0; i < src.size(); ++i) {
        if (!len[i]) continue;
        if ((err = run_matmul(c, one, {src[i]}, {dst[i]}, len[i], kGf8Accumulate))) return err;
    }
    return RSX_OK;
}

// Runs `body` on device addresses: device shards as they are, host shards
// staged (all of `ptrs` copied in; `back` copied out afterwards).
template <class Body>
int on_device(rsx_codec *c, const std::vector<uint8_t *> &ptrs, const std::vector<bool> &back, size_t len,
              Body body) {
    std::vector<const void *> all(ptrs.begin(), ptrs.end());
    MemKind kind;
    int err = classify_all(all, kind);
    if (err) return err;
    if (kind == MemKind::Device) return body(ptrs);
    std::vector<uint8_t *> dev;
    if ((err = stage_slots(c, len, ptrs.size(), dev))) return err;
    for (size_t i = 0; i < ptrs.size(); ++i)
        RSX_HIP(hipMemcpyAsync(dev[i], ptrs[i], len, hipMemcpyHostToDevice, c->stream));
    if ((err = body(dev))) return err;
    for (size_t i = 0; i < ptrs.size(); ++i)
        if (back[i]) RSX_HIP(hipMemcpyAsync(ptrs[i], dev[i], len, hipMemcpyDeviceToHost, c->stream));
    RSX_HIP(hipStreamSynchronize(c->stream));
    return RSX_OK;
}

}  // namespace

extern "C" {

// reedsolomon.go:762-797 Update + 799-818 updateParityShards.
int rsx_update(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards, uint8_t *const *new_data,
               const size_t *new_lens, int nnew) {
    RSX_TRACE("rsx_update");
    if (c->engine) return fail(RSX_ERR_NOT_SUPPORTED, "operation not supported");
    const int k = c->k, m = c->m;
    if (nshards != c->total) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    if (nnew != k) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    int err = check_shards(lens, nshards, true);
    if (err) return err;
    if ((err = check_shards(new_lens, nnew, true))) return err;
    for (int i = 0; i < k; ++i)
        if (new_data[i] && !shards[i]) return fail(RSX_ERR_INVALID_INPUT, "invalid input");
    for (int i = k; i < c->total; ++i)
        if (!shards[i]) return fail(RSX_ERR_INVALID_INPUT, "invalid input");
    const size_t size = shard_size(lens, nshards);
    // The reference slices past shorter buffers here; mismatched lengths are
    // rejected instead.
    for (int i = 0; i < k; ++i)
        if (new_data[i] && (new_lens[i] != size || lens[i] != size))
            return fail(RSX_ERR_SHARD_SIZE, "shard sizes do not match");
    for (int i = k; i < c->total; ++i)
        if (lens[i] != size) return fail(RSX_ERR_SHARD_SIZE, "shard sizes do not match");
    std::vector<int> cs;
    for (int i = 0; i < k; ++i)
        if (new_data[i]) cs.push_back(i);
    if (cs.empty()) return RSX_OK;
    (void)hipSetDevice(c->device);
    std::lock_guard<std::mutex> lock(c->mu);
    const int nc = static_cast<int>(cs.size());
    // Buffers: changed old shards, their new data, parity.
    std::vector<uint8_t *> ptrs;
    std::vector<bool> back;
    for (int i : cs) {
        ptrs.push_back(shards[i]);
        back.push_back(true);
    }
    for (int i : cs) {
        ptrs.push_back(new_data[i]);
        back.push_back(false);
    }
    for (int i = k; i < c->total; ++i) {
        ptrs.push_back(shards[i]);
        back.push_back(true);
    }
    // Rows: parity row r restricted to the changed columns.
    std::vector<std::vector<uint8_t>> rows(static_cast<size_t>(m), std::vector<uint8_t>(static_cast<size_t>(nc)));
    std::vector<const uint8_t *> rp;
    for (int r = 0; r < m; ++r) {
        for (int j = 0; j < nc; ++j) rows[r][j] = c->matrix.at(k + r, cs[j]);
        rp.push_back(rows[r].data());
    }
    std::vector<uint32_t> host;
    if (m && (err = upload_tables(rp, nc, c->aux_tables, host, c->stream))) return err;
    return on_device(c, ptrs, back, size, [&](const std::vector<uint8_t *> &d) {
        std::vector<const uint8_t *> src(d.begin() + nc, d.begin() + 2 * nc);
        std::vector<uint8_t *> old(d.begin(), d.begin() + nc);
        int e = xor_into(c, src, old, std::vector<size_t>(static_cast<size_t>(nc), size));
        if (e || m == 0) return e;
        std::vector<const uint8_t *> in(old.begin(), old.end());
        std::vector<uint8_t *> out(d.begin() + 2 * nc, d.end());
        return run_matmul(c, c->aux_tables.as<uint32_t>(), in, out, size, kGf8Accumulate);
    });
}

// reedsolomon.go:1464-1617 DecodeIdx.
int rsx_decode_idx(rsx_codec *c, uint8_t *const *dst, const size_t *dst_lens, int ndst, const uint8_t *expect,
                   int nexpect, uint8_t *const *input, const size_t *input_lens, int ninput) {
    RSX_TRACE("rsx_decode_idx");
    if (c->engine) return fail(RSX_ERR_NOT_SUPPORTED, "operation not supported");
    const int k = c->k, total = c->total;
    (void)hipSetDevice(c->device);
    if (!expect) {
        // Merge mode: dst[i] ^= input[i]; the reference stops at the first
        // bad pair after merging the pairs before it.
        if (ndst != ninput) return fail(RSX_ERR_INVALID_INPUT, "dst and input must have same length for merging");
        int stop_err = RSX_OK;
        std::vector<uint8_t *> ptrs;
        std::vector<bool> back;
        std::vector<size_t> lens;
        for (int i = 0; i < ndst; ++i) {
            if (!input[i]) continue;
            if (!dst[i]) {
                stop_err = fail(RSX_ERR_INVALID_INPUT, "input provided but dst is nil");
                break;
            }
            if (dst_lens[i] != input_lens[i]) {
                stop_err = fail(RSX_ERR_INVALID_SHARD_SIZE, "dst and input sizes differ");
                break;
            }
            ptrs.push_back(dst[i]);
            ptrs.push_back(input[i]);
            back.push_back(true);
            back.push_back(false);
            lens.push_back(input_lens[i]);
        }
        if (!ptrs.empty()) {
            std::lock_guard<std::mutex> lock(c->mu);
            std::vector<const void *> all(ptrs.begin(), ptrs.end());
            MemKind kind;
            int err = classify_all(all, kind);
            if (err) return err;
            const size_t np = lens.size();
            if (kind == MemKind::Device) {
                std::vector<const uint8_t *> src;
                std::vector<uint8_t *> d;
                for (size_t q = 0; q < np; ++q) {
                    d.push_back(ptrs[2 * q]);
                    src.push_back(ptrs[2 * q + 1]);
                }
                if ((err = xor_into(c, src, d, lens))) return err;
            } else {
                // Pairs may differ in length: stage each pair on its own.
                for (size_t q = 0; q < np; ++q) {
                    if (!lens[q]) continue;
                    std::vector<uint8_t *> pair{ptrs[2 * q], ptrs[2 * q + 1]};
                    err = on_device(c, pair, {true, false}, lens[q], [&](const std::vector<uint8_t *> &dv) {
                        return xor_into(c, {dv[1]}, {dv[0]}, {lens[q]});
                    });
                    if (err) return err;
                }
            }
            if (kind == MemKind::Device) RSX_HIP(hipStreamSynchronize(c->stream));
        }
        return stop_err;
    }
    if (nexpect != total) return fail(RSX_ERR_INVALID_INPUT, "expectInput length does not match TotalShards");
    if (ndst != total) return fail(RSX_ERR_INVALID_INPUT, "dst length does not match TotalShards");
    if (ninput != total) return fail(RSX_ERR_INVALID_INPUT, "input length does not match TotalShards");
    for (int i = 0; i < total; ++i)
        if (input[i] && !expect[i]) return fail(RSX_ERR_INVALID_INPUT, "unexpected input");
    for (int i = 0; i < total; ++i)
        if (expect[i] && dst[i]) return fail(RSX_ERR_INVALID_INPUT, "dst should be nil for an expected input");
    std::vector<int> valid;
    for (int i = 0; i < total; ++i)
        if (expect[i]) valid.push_back(i);
    if (static_cast<int>(valid.size()) < k) return fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    // getDecodeMatrix (reedsolomon.go:634-673): inverse of the first k valid rows.
    Matrix sub(k, k), inv;
    for (int r = 0; r < k; ++r) std::memcpy(sub.row(r), c->matrix.row(valid[r]), k);
    if (!invert(sub, inv)) return fail(RSX_ERR_SINGULAR, "matrix is singular");
    size_t size = 0;
    bool have = false;
    for (int i = 0; i < total; ++i) {
        if (!input[i]) continue;
        if (!have) {
            size = input_lens[i];
            have = size != 0;
        } else if (input_lens[i] != size) {
            return fail(RSX_ERR_INVALID_SHARD_SIZE, "input sizes differ");
        }
    }
    for (int i = 0; i < total; ++i)
        if (dst[i] && dst_lens[i] != size) return fail(RSX_ERR_INVALID_SHARD_SIZE, "dst size differs from input");
    // Rows of the outputs; only columns of the first k valid shards are non-zero.
    const Gf256 &f = gf();
    std::vector<int> outs;
    std::vector<std::vector<uint8_t>> rows;
    for (int d = 0; d < total; ++d) {
        if (!dst[d]) continue;
        std::vector<uint8_t> row(static_cast<size_t>(k), 0);
        if (d < k) {
            std::memcpy(row.data(), inv.row(d), k);
        } else {
            const uint8_t *prow = c->matrix.row(d);
            for (int col = 0; col < k; ++col) {
                uint8_t v = 0;
                for (int i = 0; i < k; ++i) v ^= f.mul[prow[i]][inv.at(i, col)];
                row[col] = v;
            }
        }
        rows.push_back(std::move(row));
        outs.push_back(d);
    }
    if (outs.empty() || size == 0) return RSX_OK;
    std::vector<int> ins, cols;
    for (int col = 0; col < k; ++col)
        if (input[valid[col]]) {
            ins.push_back(valid[col]);
            cols.push_back(col);
        }
    if (ins.empty()) return RSX_OK;  // inputs past the first k valid ones add zero
    // An empty input beside sized ones: the reference indexes past it (a
    // panic there); here it is ErrInvalidShardSize.
    for (int i : ins)
        if (input_lens[i] != size) return fail(RSX_ERR_INVALID_SHARD_SIZE, "input size differs");
    const int nin = static_cast<int>(ins.size()), nout = static_cast<int>(outs.size());
    std::vector<std::vector<uint8_t>> adj(static_cast<size_t>(nout), std::vector<uint8_t>(static_cast<size_t>(nin)));
    std::vector<const uint8_t *> rp;
    for (int o = 0; o < nout; ++o) {
        for (int j = 0; j < nin; ++j) adj[o][j] = rows[o][cols[j]];
        rp.push_back(adj[o].data());
    }
    std::lock_guard<std::mutex> lock(c->mu);
    std::vector<uint32_t> host;
    int err = upload_tables(rp, nin, c->aux_tables, host, c->stream);
    if (err) return err;
    std::vector<uint8_t *> ptrs;
    std::vector<bool> back;
    for (int d : outs) {
        ptrs.push_back(dst[d]);
        back.push_back(true);
    }
    for (int i : ins) {
        ptrs.push_back(input[i]);
        back.push_back(false);
    }
    err = on_device(c, ptrs, back, size, [&](const std::vector<uint8_t *> &d) {
        std::vector<uint8_t *> out(d.begin(), d.begin() + nout);
        std::vector<const uint8_t *> in(d.begin() + nout, d.end());
        return run_matmul(c, c->aux_tables.as<uint32_t>(), in, out, size, kGf8Accumulate);
    });
    return err;
}

}  // extern "C"

// sanitize_main.cpp — drives the library's host code, built against the
// host HIP stand-in (hip_host.cpp, kernels_host.cpp), under AddressSanitizer
// + UndefinedBehaviorSanitizer or ThreadSanitizer:
//   - per-call Encode / Verify / Reconstruct / ReconstructData /
//     ReconstructSome / Update on device and host shards, classic and Leopard;
//   - batch plans of mixed classic and Leopard items, re-planned with graphs
//     on and off, autotuned and fixed tunings, and with the tile -> task
//     table replaced by the binary search;
//   - the request loop with four submitting threads, a polling thread, and a
//     thread that pauses, resumes, cancels and reads the counters meanwhile,
//     Leopard requests on the worker thread, and a trace with a rejected
//     request;
//   - the stream encoder over in-memory callbacks, sequential and concurrent.
// GF(2^8) results are checked (the host launchers compute them); Leopard
// items are checked for status only.  Exit status 0 = every check passed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../../../include/rsx.h"

namespace {

std::atomic<int> g_failures{0};

#define CHECK(cond)                                                                       \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            std::fprintf(stderr, "%s:%d: check failed: %s (%s)\n", __FILE__, __LINE__, #cond, \
                         rsx_last_error());                                               \
            ++g_failures;                                                                 \
        }                                                                                 \
    } while (0)

uint8_t *dev_alloc(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, n) != hipSuccess) std::abort();
    return static_cast<uint8_t *>(p);
}

rsx_codec *make_codec(int k, int m, int leopard) {
    rsx_options o;
    rsx_options_default(&o);
    o.leopard = leopard;
    rsx_codec *c = nullptr;
    CHECK(rsx_new(k, m, &o, &c) == RSX_OK);
    return c;
}

// One stripe in "device" memory with its encoded reference bytes.
struct Stripe {
    rsx_codec *c = nullptr;
    int k = 0, m = 0;
    size_t len = 0;
    bool classic = true;
    std::vector<uint8_t *> base;  // allocations (shards may start at an offset)
    std::vector<uint8_t *> sh;
    std::vector<size_t> lens;
    std::vector<std::vector<uint8_t>> ref;

    void fill_from_ref() {
        for (int i = 0; i < k + m; ++i) std::memcpy(sh[i], ref[i].data(), len);
        lens.assign(k + m, len);
    }
    bool equals_ref(int upto) const {
        for (int i = 0; i < upto; ++i)
            if (std::memcmp(sh[i], ref[i].data(), len) != 0) return false;
        return true;
    }
    void release() {
        for (uint8_t *p : base) (void)hipFree(p);
    }
};

Stripe make_stripe(rsx_codec *c, size_t len, std::mt19937 &rng, size_t offset = 0) {
    Stripe s;
    s.c = c;
    s.k = rsx_data_shards(c);
    s.m = rsx_parity_shards(c);
    s.len = len;
    s.classic = rsx_engine(c) == 0;
    const int n = s.k + s.m;
    for (int i = 0; i < n; ++i) {
        s.base.push_back(dev_alloc(len + offset));
        s.sh.push_back(s.base.back() + offset);
    }
    s.lens.assign(n, len);
    for (int i = 0; i < s.k; ++i)
        for (size_t b = 0; b < len; ++b) s.sh[i][b] = static_cast<uint8_t>(rng());
    for (int i = s.k; i < n; ++i) std::memset(s.sh[i], 0, len);
    CHECK(rsx_encode(c, s.sh.data(), s.lens.data(), n) == RSX_OK);
    CHECK(rsx_synchronize(c) == RSX_OK);
    s.ref.resize(n);
    for (int i = 0; i < n; ++i) s.ref[i].assign(s.sh[i], s.sh[i] + len);
    return s;
}

// ---- per-call paths ----
void per_call(std::mt19937 &rng) {
    const int shapes[][2] = {{10, 4}, {6, 3}, {17, 3}, {20, 12}, {1, 1}, {5, 0}};
    for (const auto &sh : shapes) {
        rsx_codec *c = make_codec(sh[0], sh[1], RSX_LEOPARD_AS_NEEDED);
        for (size_t len : {size_t{1}, size_t{100}, size_t{4096}, size_t{5003}}) {
            Stripe s = make_stripe(c, len, rng);
            const int n = s.k + s.m;
            int ok = 0;
            CHECK(rsx_verify(c, s.sh.data(), s.lens.data(), n, &ok) == RSX_OK && ok == 1);
            if (s.m > 0) {
                s.sh[0][len / 2] ^= 0x5A;
                CHECK(rsx_verify(c, s.sh.data(), s.lens.data(), n, &ok) == RSX_OK && ok == 0);
                s.sh[0][len / 2] ^= 0x5A;
                // erase up to m shards, rebuild all, then only the data
                std::vector<uint8_t> filled(n);
                for (int data_only = 0; data_only < 2; ++data_only) {
                    std::vector<int> idx(n);
                    for (int i = 0; i < n; ++i) idx[i] = i;
                    std::shuffle(idx.begin(), idx.end(), rng);
                    const int e = 1 + static_cast<int>(rng() % s.m);
                    for (int i = 0; i < e; ++i) {
                        std::memset(s.sh[idx[i]], 0xEE, len);
                        s.lens[idx[i]] = 0;
                    }
                    const int rc = data_only ? rsx_reconstruct_data(c, s.sh.data(), s.lens.data(), n, filled.data())
                                             : rsx_reconstruct(c, s.sh.data(), s.lens.data(), n, filled.data());
                    CHECK(rc == RSX_OK);
                    CHECK(rsx_synchronize(c) == RSX_OK);
                    CHECK(s.equals_ref(data_only ? s.k : n));
                    s.fill_from_ref();
                }
                // ReconstructSome with a k-entry mask
                std::vector<uint8_t> req(s.k, 0);
                req[0] = 1;
                s.lens[0] = 0;
                std::memset(s.sh[0], 0, len);
                CHECK(rsx_reconstruct_some(c, s.sh.data(), s.lens.data(), n, req.data(), s.k, filled.data()) ==
                      RSX_OK);
                CHECK(rsx_synchronize(c) == RSX_OK);
                CHECK(s.equals_ref(1));
                s.fill_from_ref();
                // too many erasures (all of them: no shard data)
                for (int i = 0; i <= s.m; ++i) s.lens[i] = 0;
                CHECK(rsx_reconstruct(c, s.sh.data(), s.lens.data(), n, filled.data()) ==
                      (n > s.m + 1 ? RSX_ERR_TOO_FEW_SHARDS : RSX_ERR_SHARD_NO_DATA));
                s.fill_from_ref();
                // Update: new data for shard 0, parity updated in place
                std::vector<uint8_t *> upd(s.sh);
                std::vector<size_t> ulens(s.lens);
                for (int i = 1; i < s.k; ++i) {
                    upd[i] = nullptr;
                    ulens[i] = 0;
                }
                uint8_t *nd = dev_alloc(len);
                for (size_t b = 0; b < len; ++b) nd[b] = static_cast<uint8_t>(rng());
                std::vector<uint8_t *> newd(s.k, nullptr);
                std::vector<size_t> nlens(s.k, 0);
                newd[0] = nd;
                nlens[0] = len;
                CHECK(rsx_update(c, upd.data(), ulens.data(), n, newd.data(), nlens.data(), s.k) == RSX_OK);
                CHECK(rsx_synchronize(c) == RSX_OK);
                std::memcpy(s.sh[0], nd, len);
                CHECK(rsx_verify(c, s.sh.data(), s.lens.data(), n, &ok) == RSX_OK && ok == 1);
                (void)hipFree(nd);
            }
            // host-memory shards: staged through the device buffers
            std::vector<std::vector<uint8_t>> host(n);
            std::vector<uint8_t *> hp(n);
            for (int i = 0; i < n; ++i) {
                host[i] = s.ref[i];
                if (i >= s.k) std::fill(host[i].begin(), host[i].end(), 0);
                hp[i] = host[i].data();
            }
            CHECK(rsx_encode(c, hp.data(), s.lens.data(), n) == RSX_OK);
            for (int i = s.k; i < n; ++i) CHECK(host[i] == s.ref[i]);
            s.release();
        }
        rsx_free(c);
    }
    // Leopard codecs: planners and argument checks only
    const int leo[][3] = {{32, 16, RSX_LEOPARD_GF8}, {100, 50, RSX_LEOPARD_GF16}, {300, 40, RSX_LEOPARD_AS_NEEDED}};
    for (const auto &l : leo) {
        rsx_codec *c = make_codec(l[0], l[1], l[2]);
        Stripe s = make_stripe(c, 4096, rng);
        const int n = s.k + s.m;
        std::vector<uint8_t> filled(n);
        int ok = 0;
        CHECK(rsx_verify(c, s.sh.data(), s.lens.data(), n, &ok) == RSX_OK);
        for (int i = 0; i < s.m; i += 2) s.lens[i * 2 % n] = 0;
        CHECK(rsx_reconstruct(c, s.sh.data(), s.lens.data(), n, filled.data()) == RSX_OK);
        s.fill_from_ref();
        s.lens[1] = 0;
        CHECK(rsx_reconstruct_data(c, s.sh.data(), s.lens.data(), n, filled.data()) == RSX_OK);
        CHECK(rsx_synchronize(c) == RSX_OK);
        s.release();
        rsx_free(c);
    }
}

// ---- batch plans ----
struct Item {
    Stripe *s;
    int op;
    bool expect_ok;
    std::vector<uint8_t *> sh;
    std::vector<size_t> lens;
};

void batch_plans(std::mt19937 &rng) {
    std::vector<rsx_codec *> codecs = {make_codec(10, 4, 0), make_codec(6, 3, 0),  make_codec(17, 3, 0),
                                       make_codec(20, 12, 0), make_codec(4, 2, 0), make_codec(1, 1, 0)};
    rsx_codec *leo8 = make_codec(32, 16, RSX_LEOPARD_GF8);
    rsx_codec *leo16 = make_codec(100, 50, RSX_LEOPARD_GF16);
    const size_t sizes[] = {100, 1000, 4096, 5000, 16 * 1024 + 7};
    std::vector<Stripe> stripes;
    stripes.reserve(200);
    for (int i = 0; i < 120; ++i) stripes.push_back(make_stripe(codecs[rng() % codecs.size()], sizes[rng() % 5], rng));
    for (int i = 0; i < 6; ++i) stripes.push_back(make_stripe(i % 2 ? leo16 : leo8, 4096, rng));
    rsx_batch *b = nullptr;
    CHECK(rsx_batch_create(0, &b) == RSX_OK);
    for (int round = 0; round < 6; ++round) {
        // round 0-1: graph on / off; 2: autotune; 3: binary search; 4-5: fixed tunings
        rsx_batch_set_graph(b, round != 1);
        if (round == 2) rsx_batch_set_tuning(b, 0, 0);
        if (round >= 4) rsx_batch_set_tuning(b, round == 4 ? 4096 : 65536, round == 4 ? 2 : 16);
        if (round == 3) setenv("RSX_BATCH_TILE_TABLE_MAX", "1", 1);
        std::vector<Item> items;
        for (Stripe &s : stripes) {
            if (rng() % 4 == 0) continue;  // plans of different sizes
            s.fill_from_ref();
            Item it{&s, RSX_OP_ENCODE, true, s.sh, s.lens};
            const int n = s.k + s.m;
            const unsigned r = rng() % 10;
            if (r < 4 || s.m == 0) {
                for (int i = s.k; i < n; ++i) std::memset(s.sh[i], 0, s.len);
            } else if (r < 6) {
                it.op = RSX_OP_VERIFY;
                if (rng() % 2) {
                    s.sh[rng() % n][rng() % s.len] ^= 1;
                    it.expect_ok = false;
                }
            } else {
                it.op = r < 8 ? RSX_OP_RECONSTRUCT : RSX_OP_RECONSTRUCT_DATA;
                const int e = 1 + static_cast<int>(rng() % s.m);
                for (int q = 0; q < e; ++q) {
                    const int i = static_cast<int>(rng() % n);
                    std::memset(s.sh[i], 0xEE, s.len);
                    it.lens[i] = 0;
                }
            }
            items.push_back(std::move(it));
        }
        std::vector<rsx_batch_item> bi;
        size_t nflags = 0;
        for (Item &it : items) {
            bi.push_back({it.s->c, it.op, it.s->k + it.s->m, it.sh.data(), it.lens.data()});
            nflags += it.sh.size();
        }
        std::vector<int> status(bi.size());
        std::vector<uint8_t> filled(nflags);
        CHECK(rsx_batch_plan(b, bi.data(), static_cast<int>(bi.size()), status.data(), filled.data()) == RSX_OK);
        CHECK(rsx_batch_launch(b) == RSX_OK);
        CHECK(rsx_batch_synchronize(b) == RSX_OK);
        for (size_t i = 0; i < items.size(); ++i) {
            const Item &it = items[i];
            const Stripe &s = *it.s;
            if (!s.classic) {
                CHECK(status[i] == RSX_OK || status[i] == RSX_ERR_NOT_SUPPORTED);
                continue;
            }
            CHECK(status[i] == RSX_OK);
            if (it.op == RSX_OP_VERIFY) {
                int ok = -1;
                CHECK(rsx_batch_verify_result(b, static_cast<int>(i), &ok) == RSX_OK);
                CHECK((ok == 1) == it.expect_ok);
            } else {
                CHECK(s.equals_ref(it.op == RSX_OP_RECONSTRUCT_DATA ? s.k : s.k + s.m));
            }
        }
        if (round == 3) unsetenv("RSX_BATCH_TILE_TABLE_MAX");
    }
    rsx_batch_free(b);
    for (Stripe &s : stripes) s.release();
    for (rsx_codec *c : codecs) rsx_free(c);
    rsx_free(leo8);
    rsx_free(leo16);
}

// ---- request loop ----
struct Pending {
    Stripe *s;
    int op;
    uint64_t ticket;
    bool expect_ok;
};

void request_loop(std::mt19937 &rng0) {
    std::vector<rsx_codec *> codecs = {make_codec(10, 4, 0), make_codec(6, 3, 0), make_codec(4, 2, 0)};
    rsx_codec *leo8 = make_codec(32, 16, RSX_LEOPARD_GF8);
    rsx_server *srv = nullptr;
    CHECK(rsx_server_create(0, size_t{2} << 20, 16, &srv) == RSX_OK);
    constexpr int kThreads = 4, kPerThread = 60;
    std::vector<std::vector<Stripe>> stripes(kThreads);
    for (int t = 0; t < kThreads; ++t) {
        stripes[t].reserve(kPerThread);
        for (int i = 0; i < kPerThread; ++i) {
            if (i % 15 == 14)  // Leopard, shards off 16-byte alignment: the worker's per-call path
                stripes[t].push_back(make_stripe(leo8, 4096, rng0, 8));
            else
                stripes[t].push_back(make_stripe(codecs[(t + i) % 3], 1000 + 37 * i, rng0));
        }
    }
    std::mutex mu;
    std::vector<uint64_t> cancellable;
    std::atomic<bool> running{true};
    std::atomic<int> cancelled{0};
    auto submitter = [&](int t) {
        std::mt19937 rng(1000 + t);
        std::vector<Pending> mine;
        std::vector<std::vector<uint8_t *>> keep_sh(kPerThread);
        std::vector<std::vector<size_t>> keep_lens(kPerThread);
        for (int i = 0; i < kPerThread; ++i) {
            Stripe &s = stripes[t][i];
            const int n = s.k + s.m;
            keep_sh[i] = s.sh;
            keep_lens[i] = s.lens;
            int op = RSX_OP_ENCODE;
            bool expect_ok = true;
            const unsigned r = rng() % 3;
            if (r == 0) {
                for (int q = s.k; q < n; ++q) std::memset(keep_sh[i][q], 0, s.len);
            } else if (r == 1) {
                op = RSX_OP_VERIFY;
                if (rng() % 2) {
                    keep_sh[i][rng() % n][rng() % s.len] ^= 1;
                    expect_ok = false;
                }
            } else {
                op = RSX_OP_RECONSTRUCT;
                const int q = static_cast<int>(rng() % n);
                std::memset(keep_sh[i][q], 0xEE, s.len);
                keep_lens[i][q] = 0;
            }
            rsx_request req{s.c, op, n, keep_sh[i].data(), keep_lens[i].data()};
            uint64_t ticket = 0;
            CHECK(rsx_server_submit(srv, &req, &ticket) == RSX_OK);
            mine.push_back({&s, op, ticket, expect_ok});
            if (i % 7 == 3) {
                std::lock_guard<std::mutex> lock(mu);
                cancellable.push_back(ticket);
            }
        }
        for (size_t i = 0; i < mine.size(); ++i) {
            const Pending &p = mine[i];
            int status = -1, ok = 0;
            uint64_t lat = 0;
            if (i % 2) {  // poll, then wait if not done yet
                int done = 0;
                CHECK(rsx_server_poll(srv, p.ticket, &done, &status, &ok, &lat) == RSX_OK);
                if (!done) CHECK(rsx_server_wait(srv, p.ticket, &status, &ok, &lat) == RSX_OK);
            } else {
                CHECK(rsx_server_wait(srv, p.ticket, &status, &ok, &lat) == RSX_OK);
            }
            const Stripe &s = *p.s;
            if (status == RSX_ERR_CANCELLED) {
                ++cancelled;
                continue;
            }
            CHECK(status == RSX_OK);
            if (!s.classic) continue;
            if (p.op == RSX_OP_VERIFY) CHECK((ok == 1) == p.expect_ok);
            else CHECK(s.equals_ref(s.k + s.m));
        }
    };
    auto controller = [&] {
        std::mt19937 rng(77);
        int cancels = 0;
        while (running.load()) {
            uint64_t a, b2, c2, d, e;
            CHECK(rsx_server_stats(srv, &a, &b2, &c2, &d, &e) == RSX_OK);
            CHECK(rsx_server_admission(srv, &a, &b2, &c2) == RSX_OK);
            if (rng() % 4 == 0) {
                CHECK(rsx_server_pause(srv, 1) == RSX_OK);
                std::this_thread::sleep_for(std::chrono::microseconds(200));
                uint64_t t = 0;
                {
                    std::lock_guard<std::mutex> lock(mu);
                    if (!cancellable.empty()) {
                        t = cancellable.back();
                        cancellable.pop_back();
                    }
                }
                int did = 0;
                // the submitter may already have reported the ticket
                if (t && rsx_server_cancel(srv, t, &did) == RSX_OK) cancels += did;
                CHECK(rsx_server_pause(srv, 0) == RSX_OK);
            }
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
        (void)cancels;
    };
    std::thread ctl(controller);
    std::vector<std::thread> subs;
    for (int t = 0; t < kThreads; ++t) subs.emplace_back(submitter, t);
    for (auto &th : subs) th.join();
    running = false;
    ctl.join();
    CHECK(rsx_server_drain(srv) == RSX_OK);
    // a trace with a request larger than the budget
    {
        std::mt19937 rng(5);
        Stripe big = make_stripe(codecs[0], size_t{1} << 18, rng);  // 14 x 256 KiB > 2 MiB
        std::vector<Stripe> small;
        for (int i = 0; i < 5; ++i) small.push_back(make_stripe(codecs[1], 3000, rng));
        std::vector<rsx_request> reqs;
        for (int i = 0; i < 5; ++i) {
            for (int q = small[i].k; q < small[i].k + small[i].m; ++q) std::memset(small[i].sh[q], 0, small[i].len);
            reqs.push_back({small[i].c, RSX_OP_ENCODE, small[i].k + small[i].m, small[i].sh.data(),
                            small[i].lens.data()});
        }
        reqs.insert(reqs.begin() + 2, rsx_request{big.c, RSX_OP_ENCODE, 14, big.sh.data(), big.lens.data()});
        std::vector<uint64_t> arr = {0, 1000, 2000, 3000, 4000, 5000};
        std::vector<int> st(6), ok(6);
        std::vector<uint64_t> lat(6);
        uint64_t span = 0;
        CHECK(rsx_server_run_trace(srv, reqs.data(), arr.data(), 6, st.data(), ok.data(), lat.data(), &span) ==
              RSX_OK);
        for (int i = 0; i < 6; ++i) CHECK(st[i] == (i == 2 ? RSX_ERR_OVER_BUDGET : RSX_OK));
        for (Stripe &s : small) CHECK(s.equals_ref(s.k + s.m));
        for (Stripe &s : small) s.release();
        big.release();
    }
    rsx_server_free(srv);
    for (auto &v : stripes)
        for (Stripe &s : v) s.release();
    for (rsx_codec *c : codecs) rsx_free(c);
    rsx_free(leo8);
    std::printf("request loop: %d requests cancelled\n", cancelled.load());
}

// ---- stream encoder ----
struct MemStreams {
    std::vector<std::vector<uint8_t>> in, out;
    std::vector<size_t> pos;
};

long long mem_read(void *ctx, int shard, uint8_t *buf, size_t n) {
    MemStreams *m = static_cast<MemStreams *>(ctx);
    const size_t left = m->in[shard].size() - m->pos[shard];
    const size_t c = std::min(left, n);
    std::memcpy(buf, m->in[shard].data() + m->pos[shard], c);
    m->pos[shard] += c;
    return static_cast<long long>(c);
}

long long mem_write(void *ctx, int shard, const uint8_t *buf, size_t n) {
    MemStreams *m = static_cast<MemStreams *>(ctx);
    m->out[shard].insert(m->out[shard].end(), buf, buf + n);
    return static_cast<long long>(n);
}

void streams(std::mt19937 &rng) {
    rsx_options o;
    rsx_options_default(&o);
    rsx_stream *s = nullptr;
    CHECK(rsx_stream_create(4, 2, &o, 4096, &s) == RSX_OK);
    MemStreams m;
    m.in.resize(6);
    m.out.resize(6);
    m.pos.assign(6, 0);
    for (int i = 0; i < 4; ++i) {
        m.in[i].resize(10000);
        for (auto &x : m.in[i]) x = static_cast<uint8_t>(rng());
    }
    int es = -1;
    CHECK(rsx_stream_encode(s, mem_read, &m, mem_write, &m, &es) == RSX_OK);
    // parity written to out[0..1]; verify all six
    MemStreams v;
    v.in = {m.in[0], m.in[1], m.in[2], m.in[3], m.out[0], m.out[1]};
    v.pos.assign(6, 0);
    int ok = 0;
    CHECK(rsx_stream_verify(s, mem_read, &v, &ok, &es) == RSX_OK && ok == 1);
    // rebuild data shard 1 from the others
    MemStreams r;
    r.in = v.in;
    r.in[1].clear();
    r.pos.assign(6, 0);
    r.out.resize(6);
    uint8_t valid[6] = {1, 0, 1, 1, 1, 1}, fill[6] = {0, 1, 0, 0, 0, 0};
    CHECK(rsx_stream_reconstruct(s, valid, fill, mem_read, &r, mem_write, &r, &es) == RSX_OK);
    CHECK(r.out[1] == m.in[1]);
    // concurrent reads and writes: callbacks for all shards at once
    CHECK(rsx_stream_set_concurrency(s, 1, 1) == RSX_OK);
    MemStreams c;
    c.in = m.in;
    c.out.resize(6);
    c.pos.assign(6, 0);
    CHECK(rsx_stream_encode(s, mem_read, &c, mem_write, &c, &es) == RSX_OK);
    CHECK(c.out[0] == m.out[0] && c.out[1] == m.out[1]);
    MemStreams cv;
    cv.in = v.in;
    cv.pos.assign(6, 0);
    CHECK(rsx_stream_verify(s, mem_read, &cv, &ok, &es) == RSX_OK && ok == 1);
    rsx_stream_free(s);
}

}  // namespace

int main(int argc, char **argv) {
    std::mt19937 rng(argc > 1 ? static_cast<unsigned>(std::atoi(argv[1])) : 1234u);
    per_call(rng);
    std::printf("per-call paths done (%d failures)\n", g_failures.load());
    batch_plans(rng);
    std::printf("batch plans done (%d failures)\n", g_failures.load());
    request_loop(rng);
    std::printf("request loop done (%d failures)\n", g_failures.load());
    streams(rng);
    std::printf("streams done (%d failures)\n", g_failures.load());
    return g_failures.load() ? 1 : 0;
}

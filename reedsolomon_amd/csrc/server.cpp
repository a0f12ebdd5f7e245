// server.cpp — request loop with continuous batching (include/rsx.h,
// "Serving").
//
// Submit copies the request and appends it to a FIFO; a request larger than
// the HBM budget, or one that would take the queued bytes past the queue
// limit, is rejected there.  A queued request can be cancelled until the
// scheduler admits it.  The scheduler thread loops: (1) retire launch slots
// whose completion event has fired, (2) when a slot is idle, admit queued
// requests in order while the shard bytes in flight stay within the HBM
// budget (unless paused) and hand them to that slot.  Each slot has a
// planner thread that plans its admitted requests (classic and Leopard) into
// the slot's batch and launches it on the slot's stream, one batch at a time,
// while earlier batches run and the scheduler keeps retiring; Leopard requests the batch
// kernel does not take (unaligned shards, rows too tall for one LDS tile) go
// to a worker thread that runs the per-stripe call.  Slots have their own
// streams, so batches overlap on the device.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

#include "batch.hpp"
#include "codec.hpp"
#include "trace.hpp"

using Clock = std::chrono::steady_clock;

namespace {

constexpr int kSlots = 4;

struct Req {
    rsx_request r{};
    std::vector<uint8_t *> shards;
    std::vector<size_t> lens;
    size_t bytes = 0;
    Clock::time_point arrival;
    bool leopard = false;
    bool admitted = false;  // taken off the queue by the scheduler
    bool done = false;
    int status = 0;
    int verify_ok = 0;
    uint64_t latency_ns = 0;
};

// Idle: free for the scheduler; Planning: its planner thread owns it;
// Busy: launched, waiting for its completion event.
enum class SlotState { Idle, Planning, Busy };

struct Slot {
    rsx_batch *b = nullptr;
    hipEvent_t ev = nullptr;     // after the batch's kernels (completion, timing)
    hipEvent_t ev0 = nullptr;    // before them (timing)
    std::vector<uint64_t> admitted;                 // handed to the planner
    std::vector<std::pair<uint64_t, int>> tickets;  // (ticket, item index in the plan)
    SlotState state = SlotState::Idle;
    std::condition_variable cv;  // planner wake-up
    std::thread planner;
};

}  // namespace

struct rsx_server {
    int device = 0;
    size_t budget = 0;
    int max_batch = 4096;
    mutable std::mutex mu;
    std::condition_variable cv;       // scheduler wake-up
    std::condition_variable done_cv;  // completions
    std::condition_variable work_cv;  // worker wake-up
    std::unordered_map<uint64_t, Req> reqs;
    std::deque<uint64_t> queue, work;
    uint64_t next = 1;
    size_t inflight = 0, max_inflight = 0;
    size_t queued_bytes = 0, queue_limit = 0;  // 0: no limit
    size_t outstanding = 0;  // submitted and not completed
    uint64_t rejected = 0, cancelled = 0;
    bool paused = false;
    uint64_t batches = 0, batched_requests = 0;
    uint64_t plan_ns = 0, device_ns = 0;  // host planning + launch; device time of the batches
    Slot slots[kSlots];
    bool stop = false, stop_planners = false;
    // Slots being planned now, and the cap: one batch is planned at a time
    // (while earlier batches run).  Two planners at once measured slower
    // under Poisson arrivals (2.44 vs 2.36 TB/s saturated, but p50 6.5 vs
    // 0.63 ms at 0.7 load): concurrent HIP launches and copies from several
    // threads contend in the runtime.
    int planning = 0, max_planning = 1;
    // Scheduler poll interval while batches run: 5 and 25 µs measured the
    // same (2.35-2.44 TB/s saturated, p99 0.89-1.03 ms), 100 µs worse.
    int poll_us = 20;
    std::thread sched, worker;
};

namespace {

// Caller holds s->mu and notifies done_cv / cv once it has completed
// everything at hand (one wake-up per retired batch, not per request).
void complete(rsx_server *s, uint64_t t, int status, int verify_ok) {
    Req &q = s->reqs.at(t);
    q.done = true;
    q.status = status;
    q.verify_ok = verify_ok;
    q.latency_ns = static_cast<uint64_t>(
        std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - q.arrival).count());
    if (q.admitted) s->inflight -= q.bytes;
    --s->outstanding;
}

void notify_completed(rsx_server *s) {
    s->done_cv.notify_all();
    s->cv.notify_one();  // budget freed: the scheduler may admit more
}

void worker_loop(rsx_server *s) {
    (void)hipSetDevice(s->device);
    roctxNameOsThread("rsx worker");
    std::unique_lock<std::mutex> lock(s->mu);
    for (;;) {
        s->work_cv.wait(lock, [&] { return s->stop || !s->work.empty(); });
        if (s->work.empty()) {
            if (s->stop) return;
            continue;
        }
        const uint64_t t = s->work.front();
        s->work.pop_front();
        Req &q = s->reqs.at(t);
        rsx_codec *c = q.r.codec;
        std::vector<uint8_t *> sh = q.shards;
        std::vector<size_t> ln = q.lens;
        const int op = q.r.op, n = q.r.nshards;
        lock.unlock();
        int rc = RSX_OK, ok = 0;
        std::vector<uint8_t> filled(static_cast<size_t>(std::max(n, 1)));
        switch (op) {
        case RSX_OP_ENCODE: rc = rsx_encode(c, sh.data(), ln.data(), n); break;
        case RSX_OP_VERIFY: rc = rsx_verify(c, sh.data(), ln.data(), n, &ok); break;
        case RSX_OP_RECONSTRUCT: rc = rsx_reconstruct(c, sh.data(), ln.data(), n, filled.data()); break;
        case RSX_OP_RECONSTRUCT_DATA:
            rc = rsx_reconstruct_data(c, sh.data(), ln.data(), n, filled.data());
            break;
        default: rc = rsx::fail(RSX_ERR_INVALID_INPUT, "unknown request op");
        }
        if (rc == RSX_OK) rc = rsx_synchronize(c);
        lock.lock();
        complete(s, t, rc, ok);
        notify_completed(s);
    }
}

// Plans and launches the requests the scheduler handed to `sl`.
void planner_loop(rsx_server *s, Slot *sl) {
    (void)hipSetDevice(s->device);
    roctxNameOsThread("rsx planner");
    std::unique_lock<std::mutex> lock(s->mu);
    for (;;) {
        sl->cv.wait(lock, [&] { return s->stop_planners || sl->state == SlotState::Planning; });
        if (sl->state != SlotState::Planning) return;
        const std::vector<uint64_t> admitted = std::move(sl->admitted);
        sl->admitted.clear();
        // Requests stay in s->reqs (node-stable) until reported, and an
        // admitted request is neither cancelled nor reported before it
        // completes, so the pointers taken here stay valid while unlocked.
        std::vector<rsx_batch_item> items(admitted.size());
        size_t nflags = 0;
        for (size_t i = 0; i < admitted.size(); ++i) {
            const Req &q = s->reqs.at(admitted[i]);
            items[i] = {q.r.codec, q.r.op, q.r.nshards, q.shards.data(), q.lens.data()};
            nflags += static_cast<size_t>(std::max(q.r.nshards, 0));
        }
        std::vector<int> status(admitted.size());
        std::vector<uint8_t> filled(std::max<size_t>(nflags, 1));
        lock.unlock();
        RSX_TRACE("rsx_server batch");
        const Clock::time_point p0 = Clock::now();
        int rc = rsx_batch_plan(sl->b, items.data(), static_cast<int>(items.size()), status.data(), filled.data());
        hipStream_t bs = static_cast<hipStream_t>(rsx_batch_stream(sl->b));
        if (rc == RSX_OK) {
            const hipError_t e = hipEventRecord(sl->ev0, bs);
            if (e != hipSuccess) rc = rsx::hip_fail(e, "hipEventRecord");
        }
        if (rc == RSX_OK) rc = rsx_batch_launch(sl->b);
        if (rc == RSX_OK) {
            const hipError_t e = hipEventRecord(sl->ev, bs);
            if (e != hipSuccess) rc = rsx::hip_fail(e, "hipEventRecord");
        }
        const uint64_t plan_ns = static_cast<uint64_t>(
            std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - p0).count());
        lock.lock();
        s->plan_ns += plan_ns;
        ++s->batches;
        s->batched_requests += admitted.size();
        bool failed = false;
        for (size_t i = 0; i < admitted.size(); ++i) {
            if (rc == RSX_OK && status[i] == RSX_ERR_NOT_SUPPORTED && s->reqs.at(admitted[i]).leopard) {
                // Leopard shapes / alignments the batch kernel does not
                // take run through the per-stripe call on the worker.
                s->work.push_back(admitted[i]);
                s->work_cv.notify_one();
            } else if (rc != RSX_OK || status[i] != RSX_OK) {
                complete(s, admitted[i], rc != RSX_OK ? rc : status[i], 0);
                failed = true;
            } else {
                sl->tickets.emplace_back(admitted[i], static_cast<int>(i));
            }
        }
        sl->state = sl->tickets.empty() ? SlotState::Idle : SlotState::Busy;
        --s->planning;
        if (failed) notify_completed(s);
        s->cv.notify_one();  // the scheduler may retire or refill this slot
    }
}

void scheduler_loop(rsx_server *s) {
    (void)hipSetDevice(s->device);
    roctxNameOsThread("rsx scheduler");
    std::unique_lock<std::mutex> lock(s->mu);
    for (;;) {
        // (1) retire finished slots
        for (Slot &sl : s->slots) {
            if (sl.state != SlotState::Busy || hipEventQuery(sl.ev) != hipSuccess) continue;
            float ms = 0;
            if (hipEventElapsedTime(&ms, sl.ev0, sl.ev) == hipSuccess) s->device_ns += static_cast<uint64_t>(ms * 1e6);
            else (void)hipGetLastError();
            for (const auto &ti : sl.tickets) {
                int ok = 0;
                if (s->reqs.at(ti.first).r.op == RSX_OP_VERIFY) rsx_batch_verify_result(sl.b, ti.second, &ok);
                complete(s, ti.first, RSX_OK, ok);
            }
            sl.tickets.clear();
            sl.state = SlotState::Idle;
            notify_completed(s);
        }
        // (2) admit into an idle slot and hand it to its planner
        Slot *free_slot = nullptr;
        for (Slot &sl : s->slots)
            if (sl.state == SlotState::Idle && s->planning < s->max_planning) {
                free_slot = &sl;
                break;
            }
        std::vector<uint64_t> admitted;
        size_t batch_bytes = 0;
        while (free_slot && !s->paused && !s->queue.empty() && static_cast<int>(admitted.size()) < s->max_batch) {
            const uint64_t t = s->queue.front();
            Req &q = s->reqs.at(t);
            if (s->inflight > 0 && s->inflight + q.bytes > s->budget) break;  // FIFO: wait for room
            // A batch takes at most an eighth of the budget, so under load the
            // next batch is planned while earlier ones run.
            if (!admitted.empty() && batch_bytes + q.bytes > s->budget / 8) break;
            batch_bytes += q.bytes;
            s->queue.pop_front();
            q.admitted = true;
            s->queued_bytes -= q.bytes;
            s->inflight += q.bytes;
            s->max_inflight = std::max(s->max_inflight, s->inflight);
            admitted.push_back(t);
        }
        if (!admitted.empty()) {
            free_slot->admitted = std::move(admitted);
            free_slot->state = SlotState::Planning;
            ++s->planning;
            free_slot->cv.notify_one();
            continue;  // look for more work right away
        }
        bool busy = false, idle = true;
        for (Slot &sl : s->slots) {
            busy |= sl.state == SlotState::Busy;
            idle &= sl.state == SlotState::Idle;
        }
        if (s->stop && s->queue.empty() && idle) return;
        if (busy) {
            lock.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(s->poll_us));
            lock.lock();
        } else {
            // planners notify cv when they finish; submits when they queue
            s->cv.wait_for(lock, std::chrono::milliseconds(1));
        }
    }
}

int submit_at(rsx_server *s, const rsx_request *req, Clock::time_point arrival, uint64_t *ticket) {
    if (!req->codec || req->nshards != req->codec->total)
        return rsx::fail(RSX_ERR_TOO_FEW_SHARDS, "too few shards given");
    Req q;
    q.r = *req;
    q.shards.assign(req->shards, req->shards + req->nshards);
    q.lens.assign(req->lens, req->lens + req->nshards);
    size_t size = 0;
    for (size_t l : q.lens) size = std::max(size, l);
    q.bytes = size * static_cast<size_t>(req->nshards);
    q.arrival = arrival;
    q.leopard = req->codec->engine != 0;
    std::lock_guard<std::mutex> lock(s->mu);
    if (q.bytes > s->budget) {
        ++s->rejected;
        return rsx::fail(RSX_ERR_OVER_BUDGET, "request larger than the server's HBM budget");
    }
    if (s->queue_limit && s->queued_bytes + q.bytes > s->queue_limit) {
        ++s->rejected;
        return rsx::fail(RSX_ERR_OVER_BUDGET, "server queue full");
    }
    s->queued_bytes += q.bytes;
    const uint64_t t = s->next++;
    s->reqs.emplace(t, std::move(q));
    s->queue.push_back(t);
    ++s->outstanding;
    *ticket = t;
    s->cv.notify_one();
    return RSX_OK;
}

}  // namespace

extern "C" {

int rsx_server_create(int device, size_t hbm_budget, int max_batch, rsx_server **out) {
    *out = nullptr;
    auto s = std::make_unique<rsx_server>();
    s->device = device;
    s->budget = hbm_budget ? hbm_budget : (size_t{4} << 30);
    s->max_batch = max_batch > 0 ? max_batch : 4096;
    auto release_slots = [&] {
        for (Slot &sl : s->slots) {
            if (sl.ev) (void)hipEventDestroy(sl.ev);
            if (sl.ev0) (void)hipEventDestroy(sl.ev0);
            rsx_batch_free(sl.b);
        }
    };
    for (Slot &sl : s->slots) {
        int rc = rsx_batch_create(device, &sl.b);
        if (rc) {
            release_slots();
            return rc;
        }
        rsx_batch_set_graph(sl.b, 0);
        rsx_batch_set_tuning(sl.b, 16 << 10, 8);
        // Plans of up to max_batch requests fit these without regrowing.
        rc = rsx::batch_reserve(sl.b, size_t{16} << 20, s->max_batch);
        if (rc) {
            release_slots();
            return rc;
        }
        // Batches already run side by side on the slots' streams: classes
        // in order on each (no fork / join events per launch; 2.36 vs 2.26
        // TB/s saturated, host plan 3.3 vs 4.1 ms per 4096 requests).
        rsx::batch_set_class_streams(sl.b, false);
        if (hipEventCreate(&sl.ev) != hipSuccess || hipEventCreate(&sl.ev0) != hipSuccess) {
            (void)hipGetLastError();
            release_slots();
            return rsx::fail(RSX_ERR_DEVICE, "event creation failed");
        }
    }
    rsx_server *p = s.release();
    for (Slot &sl : p->slots) sl.planner = std::thread(planner_loop, p, &sl);
    p->sched = std::thread(scheduler_loop, p);
    p->worker = std::thread(worker_loop, p);
    *out = p;
    return RSX_OK;
}

int rsx_server_drain(rsx_server *s) {
    std::unique_lock<std::mutex> lock(s->mu);
    // A paused server admits nothing, so queued requests would never finish.
    s->done_cv.wait(lock, [&] { return s->outstanding == 0 || (s->paused && !s->queue.empty()); });
    if (s->outstanding) return rsx::fail(RSX_ERR_INVALID_INPUT, "server paused with queued requests");
    return RSX_OK;
}

void rsx_server_free(rsx_server *s) {
    if (!s) return;
    rsx_server_pause(s, 0);
    rsx_server_drain(s);
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    s->work_cv.notify_all();
    s->sched.join();
    s->worker.join();
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop_planners = true;  // the scheduler has returned: every slot is idle
    }
    for (Slot &sl : s->slots) {
        sl.cv.notify_all();
        sl.planner.join();
    }
    for (Slot &sl : s->slots) {
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.ev0) (void)hipEventDestroy(sl.ev0);
        rsx_batch_free(sl.b);
    }
    delete s;
}

int rsx_server_submit(rsx_server *s, const rsx_request *req, uint64_t *ticket) {
    return submit_at(s, req, Clock::now(), ticket);
}

int rsx_server_cancel(rsx_server *s, uint64_t ticket, int *cancelled) {
    *cancelled = 0;
    std::lock_guard<std::mutex> lock(s->mu);
    const auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return rsx::fail(RSX_ERR_INVALID_INPUT, "unknown ticket");
    Req &q = it->second;
    if (q.done || q.admitted) return RSX_OK;  // finished, or its batch is already planned
    s->queue.erase(std::find(s->queue.begin(), s->queue.end(), ticket));
    s->queued_bytes -= q.bytes;
    ++s->cancelled;
    complete(s, ticket, RSX_ERR_CANCELLED, 0);
    *cancelled = 1;
    notify_completed(s);
    return RSX_OK;
}

int rsx_server_pause(rsx_server *s, int paused) {
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->paused = paused != 0;
    }
    s->cv.notify_all();
    s->done_cv.notify_all();  // a drain in progress re-checks its condition
    return RSX_OK;
}

int rsx_server_set_queue_limit(rsx_server *s, size_t max_queued_bytes) {
    std::lock_guard<std::mutex> lock(s->mu);
    s->queue_limit = max_queued_bytes;
    return RSX_OK;
}

int rsx_server_admission(const rsx_server *s, uint64_t *rejected, uint64_t *cancelled, uint64_t *queued_bytes) {
    std::lock_guard<std::mutex> lock(s->mu);
    *rejected = s->rejected;
    *cancelled = s->cancelled;
    *queued_bytes = s->queued_bytes;
    return RSX_OK;
}

int rsx_server_poll(rsx_server *s, uint64_t ticket, int *done, int *status, int *verify_ok,
                    uint64_t *latency_ns) {
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return rsx::fail(RSX_ERR_INVALID_INPUT, "unknown ticket");
    *done = it->second.done;
    if (it->second.done) {
        *status = it->second.status;
        *verify_ok = it->second.verify_ok;
        *latency_ns = it->second.latency_ns;
        s->reqs.erase(it);  // a ticket is reported done once
    }
    return RSX_OK;
}

int rsx_server_wait(rsx_server *s, uint64_t ticket, int *status, int *verify_ok, uint64_t *latency_ns) {
    std::unique_lock<std::mutex> lock(s->mu);
    if (s->reqs.find(ticket) == s->reqs.end()) return rsx::fail(RSX_ERR_INVALID_INPUT, "unknown ticket");
    // Look the ticket up again after every wake-up: submits may rehash the
    // map while this thread waits, and another thread may report it first.
    s->done_cv.wait(lock, [&] {
        const auto f = s->reqs.find(ticket);
        return f == s->reqs.end() || f->second.done;
    });
    const auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return rsx::fail(RSX_ERR_INVALID_INPUT, "ticket reported by another call");
    *status = it->second.status;
    *verify_ok = it->second.verify_ok;
    *latency_ns = it->second.latency_ns;
    s->reqs.erase(it);
    return RSX_OK;
}

int rsx_server_run_trace(rsx_server *s, const rsx_request *reqs, const uint64_t *arrival_ns, int n,
                         int *status, int *verify_ok, uint64_t *latency_ns, uint64_t *makespan_ns) {
    std::vector<uint64_t> tickets(static_cast<size_t>(std::max(n, 0)));
    const Clock::time_point t0 = Clock::now();
    for (int i = 0; i < n; ++i) {
        const Clock::time_point at = t0 + std::chrono::nanoseconds(arrival_ns[i]);
        while (Clock::now() < at) {
            const auto left = at - Clock::now();
            if (left > std::chrono::microseconds(200)) std::this_thread::sleep_for(left / 2);
        }
        // A rejected request keeps its error as its status and no ticket.
        status[i] = submit_at(s, &reqs[i], at, &tickets[i]);
        if (status[i] != RSX_OK) tickets[i] = 0;
    }
    int rc = rsx_server_drain(s);
    if (rc) return rc;
    Clock::time_point last = t0;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        for (int i = 0; i < n; ++i) {
            if (!tickets[i]) {
                verify_ok[i] = 0;
                latency_ns[i] = 0;
                continue;
            }
            const auto it = s->reqs.find(tickets[i]);
            if (it == s->reqs.end()) {  // already reported by a concurrent poll / wait
                status[i] = RSX_ERR_INVALID_INPUT;
                verify_ok[i] = 0;
                latency_ns[i] = 0;
                rc = rsx::fail(RSX_ERR_INVALID_INPUT, "ticket reported by another call");
                continue;
            }
            const Req &q = it->second;
            status[i] = q.status;
            verify_ok[i] = q.verify_ok;
            latency_ns[i] = q.latency_ns;
            last = std::max(last, q.arrival + std::chrono::nanoseconds(q.latency_ns));
            s->reqs.erase(it);
        }
    }
    *makespan_ns = static_cast<uint64_t>(std::chrono::duration_cast<std::chrono::nanoseconds>(last - t0).count());
    return rc;
}

int rsx_server_stats(const rsx_server *s, uint64_t *batches, uint64_t *requests, uint64_t *max_inflight_bytes,
                     uint64_t *plan_ns, uint64_t *device_ns) {
    std::lock_guard<std::mutex> lock(s->mu);
    *batches = s->batches;
    *requests = s->batched_requests;
    *max_inflight_bytes = s->max_inflight;
    if (plan_ns) *plan_ns = s->plan_ns;
    if (device_ns) *device_ns = s->device_ns;
    return RSX_OK;
}

}  // extern "C"

// lqg_serve.cpp -- serving engine: single-query requests, continuously
// batched over several batch launches in flight on separate streams.
//
// Requests live in a fixed table of slots (query in pinned host memory, ef,
// k, results).  Whenever a lane (stream + device buffers + launch scratch) is
// idle and requests are pending, the dispatcher thread gathers up to
// max_batch of them into the lane's pinned staging buffer and enqueues
// copy-in and the search launch on the lane's stream.  The kernels write each
// request's results straight into the lane's host-coherent result buffers
// and then set the request's flag, so the dispatcher hands every request
// back as soon as its own walk ends, not when its batch's slowest request
// does; the batch's completion event frees the lane.  A request arriving
// while launches run is admitted to the next free lane without waiting for
// the others; with several lanes busy, several batches occupy the device at
// once.  Each query is walked by one wavefront independently of the others,
// so a request's results do not depend on its batch.
#include <hip/hip_runtime.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "lqg_index.h"

namespace {

using lqg::fail;

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

enum SlotState : int { kFree = 0, kPending = 1, kRunning = 2, kDone = 3, kFailed = 4 };

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    lqg::Scratch scratch;
    lqg::DevBuf<float> d_q;
    lqg::DevBuf<uint32_t> d_ef, d_k;
    float* h_q = nullptr;      // pinned staging: max_batch x dim
    uint32_t* h_ek = nullptr;  // pinned: ef[max_batch] | k[max_batch]
    // host-coherent, written by the kernels: results and per-request flags
    uint32_t* h_ids = nullptr;  // max_batch x kmax
    float* h_dist = nullptr;    // max_batch x kmax
    uint32_t* h_flag = nullptr; // max_batch
    uint32_t* g_ids = nullptr;  // the same three as device addresses
    float* g_dist = nullptr;
    uint32_t* g_flag = nullptr;
    std::vector<uint32_t> slots;
    std::vector<uint8_t> out;   // request i of the batch already handed back
    size_t n_out = 0;
    bool busy = false;
    bool failed = false;
};

}  // namespace

struct lqg_server {
    lqg_index* h = nullptr;
    lqg_serve_config cfg{};
    uint32_t dim = 0;
    int device = 0;
    float* q_store = nullptr;  // pinned: slots x dim
    std::vector<uint32_t> ef, k, gen, ids;
    std::vector<float> dist;
    std::vector<int64_t> t_done;
    std::unique_ptr<std::atomic<int>[]> state;
    std::vector<uint32_t> free_list;
    std::deque<uint32_t> pending;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<Lane> lanes;
    std::thread worker;
    bool stop = false;
    lqg_serve_stats st{};
};

namespace {

void free_lane(Lane& l) {
    if (l.stream) (void)hipStreamSynchronize(l.stream);
    l.scratch.release();
    l.d_q.release();
    l.d_ef.release();
    l.d_k.release();
    if (l.h_q) (void)hipHostFree(l.h_q);
    if (l.h_ek) (void)hipHostFree(l.h_ek);
    if (l.h_ids) (void)hipHostFree(l.h_ids);
    if (l.h_dist) (void)hipHostFree(l.h_dist);
    if (l.h_flag) (void)hipHostFree(l.h_flag);
    if (l.done) (void)hipEventDestroy(l.done);
    if (l.stream) (void)hipStreamDestroy(l.stream);
    l = Lane();
}

void destroy(lqg_server* s) {
    for (Lane& l : s->lanes) free_lane(l);
    if (s->q_store) (void)hipHostFree(s->q_store);
    delete s;
}

// A lane's stream.  HIP serves a process's ordinary streams from a small pool
// of hardware queues (4 here), and launches of two streams that share a queue
// run one after the other, so lanes created that way did not overlap (on the
// 768-d workload at 20k requests/s, p50 grew from 9 ms with 1 lane to 19 ms
// with 4).  A stream created with a CU mask gets a hardware queue of its own;
// the mask here holds every CU, so a lane still runs on the whole device.
hipError_t lane_stream(hipStream_t* stream) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        cus > 0) {
        std::vector<uint32_t> mask((static_cast<size_t>(cus) + 31) / 32, 0xFFFFFFFFu);
        if (hipExtStreamCreateWithCUMask(stream, static_cast<uint32_t>(mask.size()),
                                         mask.data()) == hipSuccess)
            return hipSuccess;
        (void)hipGetLastError();
    }
    return hipStreamCreateWithFlags(stream, hipStreamNonBlocking);
}

// Enqueues one batch on an idle lane (dispatcher thread, no lock held: the
// batch's slots are in state kRunning, which nobody else touches).
void launch(lqg_server* s, Lane& l, const std::vector<uint32_t>& batch) {
    const uint32_t b = static_cast<uint32_t>(batch.size()), d = s->dim, mb = s->cfg.max_batch;
    for (uint32_t i = 0; i < b; ++i) {
        const uint32_t slot = batch[i];
        memcpy(l.h_q + static_cast<size_t>(i) * d, s->q_store + static_cast<size_t>(slot) * d,
               4ull * d);
        l.h_ek[i] = s->ef[slot];
        l.h_ek[mb + i] = s->k[slot];
    }
    l.slots = batch;
    l.out.assign(b, 0);
    l.n_out = 0;
    // the lane's previous launch has completed, so nothing writes these now
    memset(l.h_flag, 0, 4ull * b);
    l.scratch.done = l.g_flag;
    l.failed = false;
    bool ok = hipMemcpyAsync(l.d_q.p, l.h_q, 4ull * b * d, hipMemcpyHostToDevice, l.stream) ==
                  hipSuccess &&
              hipMemcpyAsync(l.d_ef.p, l.h_ek, 4ull * b, hipMemcpyHostToDevice, l.stream) ==
                  hipSuccess &&
              hipMemcpyAsync(l.d_k.p, l.h_ek + mb, 4ull * b, hipMemcpyHostToDevice, l.stream) ==
                  hipSuccess;
    ok = ok && lqg::run_batch(s->h, l.d_q.p, b, l.d_ef.p, l.d_k.p, s->cfg.ef_max, s->cfg.kmax,
                              s->cfg.order, l.g_ids, l.g_dist, nullptr, l.stream, false,
                              &l.scratch) == LQG_OK;
    ok = ok && hipEventRecord(l.done, l.stream) == hipSuccess;
    l.failed = !ok;
    l.busy = true;
}

// Request i of lane l's batch, whose results are in the lane's buffers, is
// done (or failed): copy them to its slot.  Caller holds s->mu.
void hand_back(lqg_server* s, Lane& l, size_t i, bool bad, int64_t t) {
    const uint32_t slot = l.slots[i], km = s->cfg.kmax;
    if (!bad) {
        memcpy(&s->ids[static_cast<size_t>(slot) * km], l.h_ids + i * km, 4ull * km);
        memcpy(&s->dist[static_cast<size_t>(slot) * km], l.h_dist + i * km, 4ull * km);
    }
    s->t_done[slot] = t;
    s->state[slot].store(bad ? kFailed : kDone, std::memory_order_release);
    l.out[i] = 1;
    ++l.n_out;
    ++s->st.completed;
}

// Hands back the requests of a running lane whose flags the kernels have set
// (their results were written before the flag; kFlagFailed marks a request
// that could not finish, a FastScan visited set that overflowed).  True if
// any.
bool deliver(lqg_server* s, Lane& l, std::vector<uint32_t>& ready) {
    if (l.failed || l.n_out == l.slots.size()) return false;
    ready.clear();
    for (size_t i = 0; i < l.slots.size(); ++i)
        if (!l.out[i] && __atomic_load_n(&l.h_flag[i], __ATOMIC_ACQUIRE) != 0u)
            ready.push_back(static_cast<uint32_t>(i));
    if (ready.empty()) return false;
    const int64_t t = now_ns();
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (uint32_t i : ready)
            hand_back(s, l, i, __atomic_load_n(&l.h_flag[i], __ATOMIC_ACQUIRE) != lqg::kFlagDone,
                      t);
    }
    s->cv_done.notify_all();
    return true;
}

// The lane's launch has completed: hands back what is left and frees the
// lane.  Every request whose flag is set is handed back by its flag (done or
// failed on its own); a request without a flag failed with its launch.
void retire(lqg_server* s, Lane& l) {
    // a launch that failed to enqueue may still have copies in flight that
    // read the lane's staging buffers
    if (l.failed) (void)hipStreamSynchronize(l.stream);
    const int64_t t = now_ns();
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (size_t i = 0; i < l.slots.size(); ++i) {
            if (l.out[i]) continue;
            const uint32_t f = __atomic_load_n(&l.h_flag[i], __ATOMIC_ACQUIRE);
            hand_back(s, l, i, f != lqg::kFlagDone, t);
        }
    }
    l.slots.clear();
    l.busy = false;
    s->cv_done.notify_all();
}

void dispatcher(lqg_server* s) {
    (void)hipSetDevice(s->device);
    std::vector<uint32_t> batch, ready;
    for (;;) {
        bool progressed = false;
        uint32_t busy = 0;
        for (Lane& l : s->lanes) {
            if (!l.busy) continue;
            if (deliver(s, l, ready)) progressed = true;
            if (l.failed || hipEventQuery(l.done) == hipSuccess) {
                retire(s, l);
                progressed = true;
            } else {
                ++busy;
            }
        }
        Lane* idle = nullptr;
        for (Lane& l : s->lanes)
            if (!l.busy) {
                idle = &l;
                break;
            }
        batch.clear();
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->stop && s->pending.empty() && busy == 0) break;
            if (idle && !s->pending.empty()) {
                while (!s->pending.empty() && batch.size() < s->cfg.max_batch) {
                    const uint32_t slot = s->pending.front();
                    s->pending.pop_front();
                    s->state[slot].store(kRunning, std::memory_order_relaxed);
                    batch.push_back(slot);
                }
                ++s->st.batches;
                s->st.max_batch_seen =
                    std::max(s->st.max_batch_seen, static_cast<uint32_t>(batch.size()));
                s->st.max_lanes_busy = std::max(s->st.max_lanes_busy, busy + 1);
            } else if (!progressed && busy == 0) {
                s->cv_work.wait(lk, [s] { return s->stop || !s->pending.empty(); });
                continue;
            }
        }
        if (!batch.empty()) {
            launch(s, *idle, batch);
            continue;
        }
        if (!progressed) std::this_thread::yield();
    }
}

// capacity: entries results / dists hold; below the request's k the call
// fails (LQG_ERR_ARG) and the ticket stays valid.
lqg_status poll_impl(lqg_server* s, uint64_t ticket, uint32_t* results, float* dists,
                     uint32_t capacity, bool block, int64_t* t_done) {
    const uint32_t slot = static_cast<uint32_t>(ticket & 0xFFFFFFFFu);
    const uint32_t g = static_cast<uint32_t>(ticket >> 32);
    std::unique_lock<std::mutex> lk(s->mu);
    if (slot >= s->cfg.slots || s->gen[slot] != g || s->state[slot].load() == kFree)
        return fail(LQG_ERR_ARG, "unknown or released ticket");
    if ((results || dists) && capacity < s->k[slot])
        return fail(LQG_ERR_ARG, "result buffer smaller than the request's k");
    for (;;) {
        const int st = s->state[slot].load(std::memory_order_acquire);
        if (st == kDone || st == kFailed) {
            if (st == kDone) {
                const uint32_t km = s->cfg.kmax, kq = s->k[slot];
                if (results) memcpy(results, &s->ids[static_cast<size_t>(slot) * km], 4ull * kq);
                if (dists) memcpy(dists, &s->dist[static_cast<size_t>(slot) * km], 4ull * kq);
            }
            if (t_done) *t_done = s->t_done[slot];
            s->state[slot].store(kFree, std::memory_order_relaxed);
            s->free_list.push_back(slot);
            return st == kDone ? LQG_OK
                               : fail(LQG_ERR_DEVICE, "request failed: its launch failed, or "
                                                      "its FastScan visited set overflowed");
        }
        if (!block) return LQG_ERR_BUSY;
        s->cv_done.wait(lk);
    }
}

lqg_status submit_impl(lqg_server* s, const float* query, uint32_t ef, uint32_t k,
                       uint64_t* ticket) {
    if (ef == 0 || ef > s->cfg.ef_max) return fail(LQG_ERR_ARG, "ef must be in [1, ef_max]");
    if (k == 0 || k > s->cfg.kmax) return fail(LQG_ERR_ARG, "k must be in [1, kmax]");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->stop) return fail(LQG_ERR_STATE, "server is stopping");
    if (s->free_list.empty()) return LQG_ERR_BUSY;
    const uint32_t slot = s->free_list.back();
    s->free_list.pop_back();
    ++s->gen[slot];
    memcpy(s->q_store + static_cast<size_t>(slot) * s->dim, query, 4ull * s->dim);
    s->ef[slot] = ef;
    s->k[slot] = k;
    s->state[slot].store(kPending, std::memory_order_relaxed);
    s->pending.push_back(slot);
    ++s->st.submitted;
    *ticket = (static_cast<uint64_t>(s->gen[slot]) << 32) | slot;
    lk.unlock();
    s->cv_work.notify_one();
    return LQG_OK;
}

// Submits requests [0, n) of a replay under one lock, as many as there are
// free slots; returns how many were taken (their tickets are written).
uint64_t submit_due(lqg_server* s, const float* queries, const uint32_t* ef, const uint32_t* k,
                    uint64_t n, uint64_t* tickets) {
    uint64_t m = 0;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return 0;
        for (; m < n && !s->free_list.empty(); ++m) {
            const uint32_t slot = s->free_list.back();
            s->free_list.pop_back();
            ++s->gen[slot];
            memcpy(s->q_store + static_cast<size_t>(slot) * s->dim, queries + m * s->dim,
                   4ull * s->dim);
            s->ef[slot] = ef[m];
            s->k[slot] = k[m];
            s->state[slot].store(kPending, std::memory_order_relaxed);
            s->pending.push_back(slot);
            ++s->st.submitted;
            tickets[m] = (static_cast<uint64_t>(s->gen[slot]) << 32) | slot;
        }
    }
    if (m) s->cv_work.notify_one();
    return m;
}

}  // namespace

extern "C" {

lqg_status lqg_serve_start(lqg_index* h, const lqg_serve_config* cfg, lqg_server** out) {
    if (!h || !cfg || !out) return fail(LQG_ERR_ARG, "null argument");
    *out = nullptr;
    if (cfg->lanes < 1 || cfg->lanes > 16) return fail(LQG_ERR_ARG, "lanes must be in [1, 16]");
    if (cfg->max_batch < 1 || cfg->max_batch > 65536)
        return fail(LQG_ERR_ARG, "max_batch must be in [1, 65536]");
    if (cfg->slots < cfg->max_batch) return fail(LQG_ERR_ARG, "slots must be >= max_batch");
    if (cfg->kmax == 0) return fail(LQG_ERR_ARG, "kmax must be positive");
    if (cfg->ef_max == 0 || lqg::ef_class(cfg->ef_max) == 0)
        return fail(LQG_ERR_ARG, "ef_max must be in [1, 1024]");
    if (lqg::abi_enum(cfg->order) != LQG_ORDER_REFERENCE &&
        lqg::abi_enum(cfg->order) != LQG_ORDER_NEAREST)
        return fail(LQG_ERR_ARG, "unknown order");
    lqg_status st = lqg::ensure_device(h);
    if (st != LQG_OK) return st;
    lqg_server* s = new (std::nothrow) lqg_server();
    if (!s) return fail(LQG_ERR_NOMEM, "allocation failed");
    s->h = h;
    s->cfg = *cfg;
    s->dim = static_cast<uint32_t>(h->dim);
    (void)hipGetDevice(&s->device);
    const size_t slots = cfg->slots, km = cfg->kmax, mb = cfg->max_batch, d = s->dim;
    try {
        s->ef.assign(slots, 0);
        s->k.assign(slots, 0);
        s->gen.assign(slots, 0);
        s->ids.assign(slots * km, 0);
        s->dist.assign(slots * km, 0.0f);
        s->t_done.assign(slots, 0);
        s->state.reset(new std::atomic<int>[slots]);
        s->free_list.reserve(slots);
        s->lanes.resize(cfg->lanes);
    } catch (...) {
        delete s;
        return fail(LQG_ERR_NOMEM, "host allocation failed");
    }
    for (size_t i = 0; i < slots; ++i) {
        s->state[i].store(kFree);
        s->free_list.push_back(static_cast<uint32_t>(slots - 1 - i));
    }
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&s->q_store), 4 * slots * d);
    for (Lane& l : s->lanes) {
        if (e != hipSuccess) break;
        e = lane_stream(&l.stream);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&l.done, hipEventDisableTiming);
        if (e == hipSuccess) e = l.d_q.reserve(mb * d);
        if (e == hipSuccess) e = l.d_ef.reserve(mb);
        if (e == hipSuccess) e = l.d_k.reserve(mb);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&l.h_q), 4 * mb * d);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&l.h_ek), 8 * mb);
        // written by the kernels and read by this thread while they run
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&l.h_ids), 4 * mb * km, hipHostMallocCoherent);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&l.h_dist), 4 * mb * km, hipHostMallocCoherent);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&l.h_flag), 4 * mb, hipHostMallocCoherent);
        if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&l.g_ids), l.h_ids, 0);
        if (e == hipSuccess)
            e = hipHostGetDevicePointer(reinterpret_cast<void**>(&l.g_dist), l.h_dist, 0);
        if (e == hipSuccess)
            e = hipHostGetDevicePointer(reinterpret_cast<void**>(&l.g_flag), l.h_flag, 0);
        // every lane's launch scratch is allocated here, for its largest
        // batch: a hipMalloc / hipFree in the dispatcher can wait for the
        // device, and with it hold back every lane's hand-back
        if (e == hipSuccess) e = l.scratch.counter.reserve(2 * lqg::kQueueWords + 1);
    }
    if (e != hipSuccess) {
        destroy(s);
        return lqg::hip_fail(e, "serving buffers");
    }
    for (Lane& l : s->lanes) {
        st = lqg::h2_prepare_scratch(h, cfg->max_batch, cfg->ef_max, cfg->kmax, &l.scratch);
        if (st != LQG_OK) {
            destroy(s);
            return st;
        }
    }
    try {
        s->worker = std::thread(dispatcher, s);
    } catch (...) {
        destroy(s);
        return fail(LQG_ERR_NOMEM, "cannot start the dispatcher thread");
    }
    *out = s;
    return LQG_OK;
}

lqg_status lqg_serve_submit(lqg_server* s, const float* query, uint32_t ef, uint32_t k,
                            uint64_t* ticket) {
    if (!s || !query || !ticket) return fail(LQG_ERR_ARG, "null argument");
    for (uint32_t i = 0; i < s->dim; ++i)
        if (!std::isfinite(query[i])) return fail(LQG_ERR_ARG, "query contains a non-finite value");
    return submit_impl(s, query, ef, k, ticket);
}

lqg_status lqg_serve_poll(lqg_server* s, uint64_t ticket, uint32_t* results, float* dists,
                          uint32_t capacity) {
    if (!s) return fail(LQG_ERR_ARG, "null server");
    return poll_impl(s, ticket, results, dists, capacity, false, nullptr);
}

lqg_status lqg_serve_wait(lqg_server* s, uint64_t ticket, uint32_t* results, float* dists,
                          uint32_t capacity) {
    if (!s) return fail(LQG_ERR_ARG, "null server");
    return poll_impl(s, ticket, results, dists, capacity, true, nullptr);
}

lqg_status lqg_serve_get_stats(lqg_server* s, lqg_serve_stats* out) {
    if (!s || !out) return fail(LQG_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->st;
    return LQG_OK;
}

lqg_status lqg_serve_stop(lqg_server* s) {
    if (!s) return LQG_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    destroy(s);
    return LQG_OK;
}

lqg_status lqg_serve_replay(lqg_server* s, const float* queries, uint64_t nq, const uint32_t* ef,
                            const uint32_t* k, const uint64_t* arrival_ns, uint64_t* latency_ns,
                            uint32_t* results) {
    if (!s || !queries || !ef || !k || !arrival_ns || !latency_ns)
        return fail(LQG_ERR_ARG, "null argument");
    if (nq == 0) return fail(LQG_ERR_ARG, "nq must be positive");
    for (uint64_t i = 0; i < nq; ++i) {
        if (ef[i] == 0 || ef[i] > s->cfg.ef_max || k[i] == 0 || k[i] > s->cfg.kmax)
            return fail(LQG_ERR_ARG, "ef / k out of the server's range");
        if (i > 0 && arrival_ns[i] < arrival_ns[i - 1])
            return fail(LQG_ERR_ARG, "arrival times must be non-decreasing");
    }
    for (uint64_t i = 0; i < nq * s->dim; ++i)
        if (!std::isfinite(queries[i])) return fail(LQG_ERR_ARG, "query contains a non-finite value");
    const uint32_t km = s->cfg.kmax;
    std::vector<uint64_t> tickets(nq);
    std::unique_ptr<std::atomic<int>[]> ready(new std::atomic<int>[nq]);
    for (uint64_t i = 0; i < nq; ++i) ready[i].store(0);
    std::atomic<int> failed{0};
    const int64_t t0 = now_ns();
    // collector: waits for the tickets in submission order, watching the
    // slot's state without the lock until the request is done
    std::thread collector([&] {
        std::vector<uint32_t> buf(km);
        for (uint64_t i = 0; i < nq; ++i) {
            while (!ready[i].load(std::memory_order_acquire)) std::this_thread::yield();
            if (ready[i].load() < 0) {
                failed.store(1);
                latency_ns[i] = 0;
                continue;
            }
            const uint32_t slot = static_cast<uint32_t>(tickets[i] & 0xFFFFFFFFu);
            for (;;) {
                const int st = s->state[slot].load(std::memory_order_acquire);
                if (st == kDone || st == kFailed) break;
                std::this_thread::yield();
            }
            int64_t td = 0;
            uint32_t* out = results ? results + i * km : buf.data();
            if (poll_impl(s, tickets[i], out, nullptr, km, true, &td) != LQG_OK) failed.store(1);
            if (results)
                for (uint32_t j = k[i]; j < km; ++j) out[j] = 0xFFFFFFFFu;
            const int64_t sched = t0 + static_cast<int64_t>(arrival_ns[i]);
            latency_ns[i] = static_cast<uint64_t>(td > sched ? td - sched : 0);
        }
    });
    // submitter: every request due by now goes in under one lock (requests
    // behind schedule are submitted together), until the slots run out
    for (uint64_t i = 0; i < nq;) {
        const int64_t now = now_ns();
        uint64_t j = i;
        while (j < nq && t0 + static_cast<int64_t>(arrival_ns[j]) <= now) ++j;
        if (j == i) continue;  // the next request is not due yet
        const uint64_t m =
            submit_due(s, queries + i * s->dim, ef + i, k + i, j - i, tickets.data() + i);
        for (uint64_t t = i; t < i + m; ++t) ready[t].store(1, std::memory_order_release);
        if (m == 0) {
            bool stopping;
            {
                std::lock_guard<std::mutex> g(s->mu);
                stopping = s->stop;
            }
            if (stopping) {
                for (uint64_t t = i; t < nq; ++t) ready[t].store(-1, std::memory_order_release);
                break;
            }
            std::this_thread::yield();
        }
        i += m;
    }
    collector.join();
    if (failed.load()) return fail(LQG_ERR_DEVICE, "some replayed requests failed");
    return LQG_OK;
}

}  // extern "C"

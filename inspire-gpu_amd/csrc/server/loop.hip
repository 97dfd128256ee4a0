// Continuous-batching serving loop (capi.h ipir_loop_*). No counterpart in
// the reference, whose server is a blocking, externally serialised call
// (src/gpu_protocol.h:67, src/capi.h:12-14).
//
// The context's max_batch slots are split into partitions, each with its
// own HIP stream, scratch lane, pinned staging and completion thread. A
// dispatcher thread admits waiting requests into the first free partition
// (up to its slot count) as soon as one frees up, so a new batch starts
// while earlier batches are still on the device and their kernels overlap
// (e.g. one batch's latency-bound Horner chains beside the next batch's
// HBM-bound database pass). Each request completes independently and
// records its submit-to-completion latency. Results are bit-identical to
// answering the request alone: the pipeline is exact and per-slot.
//
// While a loop runs, the context is marked loop-owned (loop_active, set and
// cleared under the context's mutex); the other entry points that use the
// slots return "busy" instead of blocking.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>
#include <vector>

#include "capi.h"
#include "inspire_debug.h"
#include "server_ctx.h"

using namespace inspire;

namespace {
using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t ticket = 0;
    std::vector<uint64_t> q;  // flat query
    uint64_t* out = nullptr;
    Clock::time_point t0;
    double lat_us = 0;
    bool done = false;
    bool failed = false;
};

struct Partition {
    size_t slot0 = 0, nslots = 0;
    int lane = 0;  // scratch lane (and status word) of the context this partition uses
    hipStream_t stream = nullptr;
    // device timestamps of the batch in flight (its first and last command),
    // and [start, end) in ms after the loop's reference event of the last
    // completed batch
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last0 = -1, last1 = -1;
    std::vector<Req*> batch;  // admitted requests; partition busy when non-empty
    bool launched = false;    // batch's device work fully enqueued
    std::thread completer;
    std::condition_variable cv;  // wakes the completer
};
}  // namespace

struct ipir_loop {
    ipir_server* srv = nullptr;
    GpuServerCtx* ctx = nullptr;
    std::mutex m;
    std::condition_variable cv_work;  // dispatcher: requests or free partitions
    std::condition_variable cv_done;  // waiters
    std::deque<Req*> pending;
    std::unordered_map<uint64_t, Req*> reqs;
    std::vector<Partition*> parts;
    std::thread dispatcher;
    bool stop = false;
    bool hold = false;  // completers keep finished partitions busy (inspection)
    hipEvent_t ev_ref = nullptr;  // time origin of the batches' device intervals
    uint64_t next_ticket = 1;
    ipir_loop_stats st{};

    size_t busy_parts() const {
        size_t n = 0;
        for (auto* p : parts) n += !p->batch.empty();
        return n;
    }

    void dispatch_loop() {
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            Partition* free_p = nullptr;
            cv_work.wait(lk, [&] {
                if (stop) return true;
                if (pending.empty()) return false;
                for (auto* p : parts)
                    if (p->batch.empty()) { free_p = p; return true; }
                return false;
            });
            if (stop) return;
            const size_t n = pending.size() < free_p->nslots ? pending.size() : free_p->nslots;
            if (busy_parts() > 0) st.admitted_while_busy++;
            for (size_t i = 0; i < n; i++) {
                free_p->batch.push_back(pending.front());
                pending.pop_front();
            }
            const size_t busy = busy_parts();
            if (busy > st.max_busy_partitions) st.max_busy_partitions = busy;
            if (n > st.max_batch_size) st.max_batch_size = n;
            st.batches++;
            for (size_t i = 0; i < n; i++)
                std::memcpy(ctx->h_q + (free_p->slot0 + i) * ctx->qlen, free_p->batch[i]->q.data(), ctx->qlen * 8);
            lk.unlock();
            bool ok = true;
            try {
                const size_t s0 = free_p->slot0;
                hip_check(hipEventRecord(free_p->ev0, free_p->stream), "loop event");
                hip_check(hipMemcpyAsync(ctx->d_q + s0 * ctx->qlen, ctx->h_q + s0 * ctx->qlen, n * ctx->qlen * 8,
                                         hipMemcpyHostToDevice, free_p->stream), "loop h2d");
                ctx->enqueue(ctx->d_q + s0 * ctx->qlen, n, ctx->d_resp + s0 * ctx->resp_u64s, free_p->stream, s0,
                             free_p->lane, true);
                hip_check(hipMemcpyAsync(ctx->h_resp + s0 * ctx->resp_u64s, ctx->d_resp + s0 * ctx->resp_u64s,
                                         n * ctx->resp_u64s * 8, hipMemcpyDeviceToHost, free_p->stream), "loop d2h");
                hip_check(hipEventRecord(free_p->ev1, free_p->stream), "loop event");
            } catch (...) {
                ok = false;
            }
            lk.lock();
            if (!ok)
                for (Req* r : free_p->batch) r->failed = true;
            free_p->launched = true;
            free_p->cv.notify_one();
        }
    }

    void complete_loop(Partition* p) {
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            p->cv.wait(lk, [&] { return stop || p->launched; });
            if (!p->launched) return;  // stopping with nothing in flight
            lk.unlock();
            bool ok = hipStreamSynchronize(p->stream) == hipSuccess;
            if (ok) {
                try {
                    ctx->check_status(p->stream, p->lane);
                } catch (...) {
                    ok = false;
                }
            }
            float t0 = -1, t1 = -1;
            if (ok && hipEventElapsedTime(&t0, ev_ref, p->ev0) == hipSuccess &&
                hipEventElapsedTime(&t1, ev_ref, p->ev1) == hipSuccess) {
            } else {
                t0 = t1 = -1;
            }
            lk.lock();
            if (t0 >= 0) {
                // on the device beside another partition's batch: the two
                // batches' [first command, last command] intervals intersect
                bool beside = false;
                for (auto* o : parts)
                    if (o != p && o->last1 > t0 && o->last0 < t1) beside = true;
                if (beside) st.device_overlaps++;
                p->last0 = t0;
                p->last1 = t1;
            }
            p->cv.wait(lk, [&] { return !hold; });
            const auto now = Clock::now();
            for (size_t i = 0; i < p->batch.size(); i++) {
                Req* r = p->batch[i];
                if (ok && !r->failed)
                    std::memcpy(r->out, ctx->h_resp + (p->slot0 + i) * ctx->resp_u64s, ctx->resp_u64s * 8);
                else
                    r->failed = true;
                r->lat_us = std::chrono::duration<double, std::micro>(now - r->t0).count();
                r->done = true;
                st.completed++;
                st.failed += r->failed;
            }
            p->batch.clear();
            p->launched = false;
            cv_done.notify_all();
            cv_work.notify_one();
        }
    }
};

extern "C" {

ipir_loop* ipir_loop_start(ipir_server* srv, size_t max_inflight) {
    if (!srv || max_inflight > (size_t)GpuServerCtx::kLanes) return nullptr;
    GpuServerCtx* ctx = srv->ctx;
    const size_t B = ctx->cfg.max_batch;
    size_t per = 0, nparts = 0;
    if (max_inflight == 0) {
        // full batched passes in every partition: 32-query partitions (the
        // whole batch if smaller), as many as the slots and lanes allow
        per = B < ipx::kMfmaMaxQueries ? B : ipx::kMfmaMaxQueries;
        nparts = B / per;
        if (nparts > (size_t)GpuServerCtx::kLanes) nparts = GpuServerCtx::kLanes;
    } else {
        nparts = max_inflight > B ? B : max_inflight;
        per = B / nparts;
    }
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ctx->loop_active) return nullptr;  // one loop per server
        ctx->loop_active = true;
    }
    // status words left by earlier asynchronous calls are not this loop's
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemset(ctx->d_status, 0, GpuServerCtx::kLanes * sizeof(uint32_t)) != hipSuccess) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->loop_active = false;
        return nullptr;
    }
    auto* L = new ipir_loop;
    L->srv = srv;
    L->ctx = ctx;
    for (size_t i = 0; i < nparts; i++) {
        auto* p = new Partition;
        p->slot0 = i * per;
        p->lane = (int)i;
        p->nslots = per;
        if (hipStreamCreate(&p->stream) != hipSuccess || hipEventCreate(&p->ev0) != hipSuccess ||
            hipEventCreate(&p->ev1) != hipSuccess) {
            if (p->ev0) (void)hipEventDestroy(p->ev0);
            if (p->stream) (void)hipStreamDestroy(p->stream);
            delete p;
            break;
        }
        L->parts.push_back(p);
    }
    if (!L->parts.empty() &&
        (hipEventCreate(&L->ev_ref) != hipSuccess || hipEventRecord(L->ev_ref, L->parts[0]->stream) != hipSuccess)) {
        for (auto* p : L->parts) {
            (void)hipEventDestroy(p->ev0);
            (void)hipEventDestroy(p->ev1);
            (void)hipStreamDestroy(p->stream);
            delete p;
        }
        L->parts.clear();
    }
    if (L->parts.empty()) {
        if (L->ev_ref) (void)hipEventDestroy(L->ev_ref);
        delete L;
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->loop_active = false;
        return nullptr;
    }
    L->st.partitions = L->parts.size();
    L->st.slots_per_partition = per;
    {
        // work the context's other entry points left on their streams comes first
        std::lock_guard<std::mutex> lk(ctx->mu);
        for (auto* p : L->parts) {
            try {
                ctx->order_before(p->stream);
            } catch (...) {
            }
        }
    }
    for (auto* p : L->parts) p->completer = std::thread([L, p] { L->complete_loop(p); });
    L->dispatcher = std::thread([L] { L->dispatch_loop(); });
    return L;
}

uint64_t ipir_loop_submit(ipir_loop* L, const ipir_query_view* v, uint64_t* out_resp) {
    if (!L || !v || !out_resp || !v->lwe_b_limb0 || !v->lwe_b_limb1 || !v->ksk5_b || !v->kskneg1_b || !v->rgsw)
        return 0;
    auto* r = new Req;
    const size_t R = L->ctx->pp.db_rows, b = (size_t)D_EFF * 2 * N;
    r->q.resize(L->ctx->qlen);
    std::memcpy(r->q.data(), v->lwe_b_limb0, R * 8);
    std::memcpy(r->q.data() + R, v->lwe_b_limb1, R * 8);
    std::memcpy(r->q.data() + 2 * R, v->ksk5_b, b * 8);
    std::memcpy(r->q.data() + 2 * R + b, v->kskneg1_b, b * 8);
    std::memcpy(r->q.data() + 2 * R + 2 * b, v->rgsw, 4 * b * 8);
    r->out = out_resp;
    std::lock_guard<std::mutex> lk(L->m);
    if (L->stop) {
        delete r;
        return 0;
    }
    r->ticket = L->next_ticket++;
    r->t0 = Clock::now();
    L->reqs[r->ticket] = r;
    L->pending.push_back(r);
    L->st.submitted++;
    L->cv_work.notify_one();
    return r->ticket;
}

int ipir_loop_wait(ipir_loop* L, uint64_t ticket, double* latency_us) {
    if (!L) return 1;
    std::unique_lock<std::mutex> lk(L->m);
    auto it = L->reqs.find(ticket);
    if (it == L->reqs.end()) return 1;
    Req* r = it->second;
    L->cv_done.wait(lk, [&] { return r->done; });
    if (latency_us) *latency_us = r->lat_us;
    const int rc = r->failed ? 4 : 0;
    L->reqs.erase(it);
    delete r;
    return rc;
}

int ipir_loop_get_stats(ipir_loop* L, ipir_loop_stats* out) {
    if (!L || !out) return 1;
    std::lock_guard<std::mutex> lk(L->m);
    *out = L->st;
    out->pending = L->pending.size();
    out->busy_partitions = L->busy_parts();
    return 0;
}

int ipir_dbg_loop_hold(ipir_loop* L, int on) {
    if (!L || (on != 0 && on != 1)) return 1;
    std::lock_guard<std::mutex> lk(L->m);
    L->hold = on != 0;
    if (!L->hold)
        for (auto* p : L->parts) p->cv.notify_all();
    return 0;
}

void ipir_loop_stop(ipir_loop* L) {
    if (!L) return;
    {
        // let admitted work finish, then stop
        std::unique_lock<std::mutex> lk(L->m);
        L->hold = false;
        for (auto* p : L->parts) p->cv.notify_all();
        L->cv_done.wait(lk, [&] {
            if (!L->pending.empty()) return false;
            for (auto* p : L->parts)
                if (!p->batch.empty()) return false;
            return true;
        });
        L->stop = true;
    }
    L->cv_work.notify_all();
    for (auto* p : L->parts) p->cv.notify_all();
    L->dispatcher.join();
    GpuServerCtx* ctx = L->ctx;
    for (auto* p : L->parts) {
        p->completer.join();
        (void)hipStreamSynchronize(p->stream);  // every completer synchronised already
        (void)hipEventDestroy(p->ev0);
        (void)hipEventDestroy(p->ev1);
        (void)hipStreamDestroy(p->stream);
        delete p;
    }
    (void)hipEventDestroy(L->ev_ref);
    for (auto& kv : L->reqs) delete kv.second;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->order_stream_set = false;  // the loop's work has finished and its streams are gone
        ctx->loop_active = false;
    }
    delete L;
}

}  // extern "C"

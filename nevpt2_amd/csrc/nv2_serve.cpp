// Request loop for NEVPT2 evaluations: continuous batching on one GPU.
//
// A server owns a fixed number of slots (host threads).  Requests arrive at any time and are
// admitted first-in first-out as soon as a slot is free and their device memory fits the
// HBM budget; admitted requests run concurrently, each on its own evaluation context (own
// streams, rocBLAS handles and stage-2 worker threads), so the latency-bound solver chains of
// one request overlap with the kernels of the others.  Contexts are pooled by problem shape:
// a finished request leaves its context (RDM, integral and intermediate buffers) resident for
// the next request of the same shape; idle contexts are freed when the budget is needed.
//
// Two request kinds: host-side inputs (uploaded into a pooled or new context on admission),
// and a caller-owned context whose inputs are already resident in HBM.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/nevpt2_amd.h"

namespace nv2 {

void set_error(const char *fmt, ...);   // nv2_capi.cpp: thread-local message for nv2_last_error

typedef std::chrono::steady_clock Clock;

struct ServeReq {
    int64_t id = 0;
    nv2_request r;                 // host-data request (ctx == nullptr)
    nv2_ctx *ctx = nullptr;        // resident-context request
    int32_t prec = 0;
    double thr = 1e-6;
    double *out = nullptr;
    int64_t bytes = 0;             // device memory a new context for it would take
    Clock::time_point t_submit;
};
struct ServeDone {
    int status = 0;
    double latency_ms = 0.0, queue_ms = 0.0;
    char err[256] = "";
};
struct PoolCtx {
    nv2_ctx *c;
    int64_t key[6];
    int64_t bytes;
};

}  // namespace nv2

using namespace nv2;

struct nv2_server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServeReq> queue;
    std::map<int64_t, ServeDone> done;
    std::vector<PoolCtx> idle;     // pooled contexts not in use
    std::vector<std::thread> slots;
    int64_t budget = 0, used = 0;  // device bytes of all pooled and running contexts
    int64_t next_id = 1;
    int running = 0;
    bool stop = false;
};

int64_t nv2_ctx_bytes(int64_t ncore, int64_t nact, int64_t nvirt, int64_t metat, int64_t nchovec) {
    const double n = (double)nact, nc = (double)ncore, nv = (double)nvirt, M = (double)metat;
    const double norb = nc + n + nv, npair = norb * (norb + 1) / 2;
    const double n4 = n * n * n * n, n6 = n4 * n * n;
    double b = 8.0 * n4 * n4;                                  // E4 (FP64)
    b += 8.0 * (n4 * (n + 1) * (n + 2) * (n + 3) / 24.0) * M;  // packed ro4
    b += 8.0 * ((double)nchovec * npair + npair);              // Cholesky vectors, one-electron integrals
    b += 8.0 * M * (6 * n6 + 20 * n4) + 8.0 * (8 * n6 + 8 * n4);   // RDMs, Koopmans and A matrices, bamat scratch
    b += 8.0 * ((nc * nv) * (nc * nv) + 4 * nv * nv * n * n + 3 * (double)nchovec * (norb * n + nc * nv));  // integrals
    b += 8.0 * 16.0 * n6;                                      // stage-2 metric / Koopmans problems of order n^3
    return (int64_t)(1.10 * b) + ((int64_t)64 << 20);
}

static bool same_shape(const PoolCtx &p, const nv2_request &r) {
    return p.key[0] == r.ncore && p.key[1] == r.nact && p.key[2] == r.nvirt && p.key[3] == r.metat &&
           p.key[4] == r.nele && p.key[5] == r.nchovec;
}

static int run_host_request(nv2_server *s, ServeReq &q, PoolCtx &pc) {
    const nv2_request &r = q.r;
    if (!pc.c) {
        pc.c = nv2_ctx_create(r.ncore, r.nact, r.nvirt, r.metat, r.nele, r.nchovec);
        if (!pc.c) return -1;
        const int64_t k[6] = {r.ncore, r.nact, r.nvirt, r.metat, r.nele, r.nchovec};
        memcpy(pc.key, k, sizeof(k));
        pc.bytes = q.bytes;
    }
    int rc = nv2_ctx_set_integrals(pc.c, r.cholesky, r.oneint);
    if (!rc) rc = nv2_ctx_set_ro4(pc.c, r.ro4);
    if (!rc && r.metat > 1 && r.trdm3) rc = nv2_ctx_set_trdm(pc.c, r.trdm3, r.trdm2, r.trdm1);
    if (!rc) rc = nv2_ctx_run(pc.c, r.prec, r.thr_lindep, r.out);
    if (!rc && r.heff && r.metat > 1) {
        rc = nv2_ctx_get(pc.c, "Heff-SC", r.heff, r.metat * r.metat);
        if (!rc) rc = nv2_ctx_get(pc.c, "Heff-PC", r.heff + r.metat * r.metat, r.metat * r.metat);
    }
    (void)s;
    return rc;
}

// admission of the queue head: a slot is free (the caller is one) and, for a host-data
// request, a pooled context of its shape is idle or its bytes fit once idle contexts are freed
static bool try_admit(nv2_server *s, ServeReq &q, PoolCtx &pc, std::vector<nv2_ctx *> &to_free) {
    ServeReq &h = s->queue.front();
    if (h.ctx) {
        q = h;
        s->queue.pop_front();
        return true;
    }
    for (size_t k = 0; k < s->idle.size(); k++)
        if (same_shape(s->idle[k], h.r)) {
            pc = s->idle[k];
            s->idle.erase(s->idle.begin() + k);
            q = h;
            s->queue.pop_front();
            return true;
        }
    // free idle contexts (oldest first) until the new one fits
    while (s->used + h.bytes > s->budget && !s->idle.empty()) {
        to_free.push_back(s->idle.front().c);
        s->used -= s->idle.front().bytes;
        s->idle.erase(s->idle.begin());
    }
    if (s->used + h.bytes > s->budget) return false;
    s->used += h.bytes;
    pc = PoolCtx{nullptr, {0, 0, 0, 0, 0, 0}, h.bytes};
    q = h;
    s->queue.pop_front();
    return true;
}

static void slot_main(nv2_server *s) {
    for (;;) {
        ServeReq q;
        PoolCtx pc{nullptr, {0, 0, 0, 0, 0, 0}, 0};
        std::vector<nv2_ctx *> to_free;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (!to_free.empty()) {
                    // contexts evicted for a request that still did not fit: their memory is already
                    // off the books, so free it now rather than when (or if) this slot admits one
                    lk.unlock();
                    for (nv2_ctx *c : to_free) nv2_ctx_destroy(c);
                    to_free.clear();
                    lk.lock();
                    continue;
                }
                if (s->stop) return;
                if (!s->queue.empty()) {
                    ServeReq &h = s->queue.front();
                    if (!h.ctx && h.bytes > s->budget) {       // can never fit: fail it now
                        ServeDone d;
                        d.status = -2;
                        snprintf(d.err, sizeof(d.err), "request needs %lld bytes, above the HBM budget %lld",
                                 (long long)h.bytes, (long long)s->budget);
                        s->done[h.id] = d;
                        s->queue.pop_front();
                        s->cv_done.notify_all();
                        continue;
                    }
                    if (try_admit(s, q, pc, to_free)) break;
                    if (!to_free.empty()) continue;     // free what was evicted first, then look again
                }
                s->cv_work.wait(lk);
            }
            s->running++;
        }
        const Clock::time_point t_start = Clock::now();
        int rc;
        ServeDone d;
        // nothing may escape a slot thread: an exception there would end the process (std::terminate)
        try {
            for (nv2_ctx *c : to_free) nv2_ctx_destroy(c);
            if (q.ctx) rc = nv2_ctx_run(q.ctx, q.prec, q.thr, q.out);
            else rc = run_host_request(s, q, pc);
            if (rc) nv2_last_error(d.err, sizeof(d.err));
        } catch (const std::exception &x) {
            rc = -1;
            snprintf(d.err, sizeof(d.err), "request failed with an exception: %s", x.what());
        } catch (...) {
            rc = -1;
            snprintf(d.err, sizeof(d.err), "request failed with an unknown exception");
        }
        d.status = rc;
        const Clock::time_point t_end = Clock::now();
        d.latency_ms = std::chrono::duration<double, std::milli>(t_end - q.t_submit).count();
        d.queue_ms = std::chrono::duration<double, std::milli>(t_start - q.t_submit).count();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (!q.ctx) {
                if (pc.c && rc == 0) s->idle.push_back(pc);     // keep it resident for the next request
                else {
                    if (pc.c) nv2_ctx_destroy(pc.c);
                    s->used -= pc.bytes;
                }
            }
            s->done[q.id] = d;
            s->running--;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

nv2_server *nv2_server_create(int32_t slots, int64_t hbm_budget_bytes) {
    if (slots < 1 || slots > 64) { set_error("nv2_server_create: slots=%d outside 1..64", slots); return nullptr; }
    if (hbm_budget_bytes <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) { set_error("hipMemGetInfo failed"); return nullptr; }
        hbm_budget_bytes = (int64_t)(0.8 * (double)fr);
    }
    nv2_server *s = new nv2_server();
    s->budget = hbm_budget_bytes;
    for (int k = 0; k < slots; k++) s->slots.emplace_back(slot_main, s);
    return s;
}

static int64_t enqueue(nv2_server *s, ServeReq &q) {
    std::lock_guard<std::mutex> lk(s->mu);
    q.id = s->next_id++;
    q.t_submit = Clock::now();
    s->queue.push_back(q);
    s->cv_work.notify_all();
    return q.id;
}

int64_t nv2_server_submit(nv2_server *s, const nv2_request *r) {
    if (!s || !r) { set_error("nv2_server_submit: null argument"); return -2; }
    if (r->nact < 1 || r->nact > 16 || r->metat < 1 || !r->cholesky || !r->oneint || !r->ro4 || !r->out) {
        set_error("nv2_server_submit: invalid request");
        return -2;
    }
    ServeReq q;
    q.r = *r;
    q.bytes = nv2_ctx_bytes(r->ncore, r->nact, r->nvirt, r->metat, r->nchovec);
    return enqueue(s, q);
}

int64_t nv2_server_submit_ctx(nv2_server *s, nv2_ctx *ctx, int32_t prec, double thr_lindep, double *out) {
    if (!s || !ctx || !out) { set_error("nv2_server_submit_ctx: null argument"); return -2; }
    ServeReq q;
    memset(&q.r, 0, sizeof(q.r));
    q.ctx = ctx;
    q.prec = prec;
    q.thr = thr_lindep;
    q.out = out;
    return enqueue(s, q);
}

int nv2_server_wait(nv2_server *s, int64_t id, double *latency_ms, double *queue_ms) {
    if (!s) return -2;
    std::unique_lock<std::mutex> lk(s->mu);
    if (id <= 0 || id >= s->next_id) { set_error("nv2_server_wait: unknown request %lld", (long long)id); return -2; }
    s->cv_done.wait(lk, [&] { return s->done.count(id) > 0; });
    ServeDone d = s->done[id];
    s->done.erase(id);
    if (latency_ms) *latency_ms = d.latency_ms;
    if (queue_ms) *queue_ms = d.queue_ms;
    if (d.status) set_error("request %lld: %s", (long long)id, d.err);
    return d.status;
}

void nv2_server_destroy(nv2_server *s) {
    if (!s) return;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return s->queue.empty() && s->running == 0; });   // drain
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->slots) t.join();
    for (auto &p : s->idle) nv2_ctx_destroy(p.c);
    delete s;
}

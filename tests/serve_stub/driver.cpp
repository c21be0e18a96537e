// Request-loop driver on stub contexts: many requests of mixed shapes from several submitting
// threads, resident-context requests beside host-data ones, a tight HBM budget that forces
// contexts out of the pool, and requests that fail or throw.  Prints "OK" when every result is
// the one its inputs define and every pooled context is freed.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/nevpt2_amd.h"

extern std::atomic<int> g_live, g_created;

struct Req {
    int64_t nact, metat;
    std::vector<double> chol, one, ro4, t3, out, heff;
    nv2_request r;
    int64_t id = 0;
};

static void fill(Req &q, int64_t nact, int64_t metat, double seed) {
    q.nact = nact; q.metat = metat;
    q.chol.assign(8, seed); q.one.assign(8, 0.5); q.ro4.assign(8, 0.25); q.t3.assign(8, 0.125);
    q.out.assign(32 * metat, 0.0); q.heff.assign(2 * metat * metat, 0.0);
    memset(&q.r, 0, sizeof(q.r));
    q.r.ncore = 3; q.r.nact = nact; q.r.nvirt = 5; q.r.metat = metat; q.r.nele = nact; q.r.nchovec = 24;
    q.r.cholesky = q.chol.data(); q.r.oneint = q.one.data(); q.r.ro4 = q.ro4.data();
    if (metat > 1) { q.r.trdm3 = q.t3.data(); q.r.trdm2 = q.t3.data(); q.r.trdm1 = q.t3.data(); q.r.heff = q.heff.data(); }
    q.r.thr_lindep = 1e-6;
    q.r.out = q.out.data();
}

int main() {
    int bad = 0;
    for (int round = 0; round < 3; round++) {
        const int64_t one = nv2_ctx_bytes(3, 6, 5, 1, 24);
        // round 0: room for everything; round 1: a budget of about three contexts; round 2: one slot
        const int slots = round == 2 ? 1 : 4;
        const int64_t budget = round == 1 ? 3 * one + one / 2 : 0;
        nv2_server *s = nv2_server_create(slots, budget);
        const int NT = 4, PER = 24;
        std::vector<std::vector<Req>> reqs(NT, std::vector<Req>(PER));
        std::vector<std::thread> th;
        for (int t = 0; t < NT; t++)
            th.emplace_back([&, t]() {
                for (int k = 0; k < PER; k++) {
                    Req &q = reqs[t][k];
                    const int64_t shapes[5] = {4, 5, 6, 11, 13};
                    fill(q, shapes[(t + k) % 5], 1 + (k % 2), 0.1 * t + 0.01 * k);
                    q.id = nv2_server_submit(s, &q.r);
                }
                for (int k = 0; k < PER; k++) {
                    Req &q = reqs[t][k];
                    double lat = 0, qms = 0;
                    const int rc = nv2_server_wait(s, q.id, &lat, &qms);
                    const bool want_fail = q.nact == 11 || q.nact == 13;
                    if (want_fail != (rc != 0)) { fprintf(stderr, "request %lld rc %d\n", (long long)q.id, rc); bad++; }
                    if (rc) continue;
                    const double seed = 0.1 * t + 0.01 * k + 0.5 + 0.25 + (q.metat > 1 ? 0.125 : 0.0);
                    for (int64_t j = 0; j < 32 * q.metat; j++)
                        if (std::fabs(q.out[j] - (seed + 0.001 * j + q.nact)) > 1e-12) { bad++; break; }
                    if (!(lat >= qms && qms >= 0)) bad++;
                }
            });
        for (auto &x : th) x.join();
        nv2_server_destroy(s);
        if (g_live.load() != 0) { fprintf(stderr, "%d contexts left\n", g_live.load()); bad++; }
    }
    printf(bad ? "FAIL %d\n" : "OK (%d contexts created)\n", bad ? bad : g_created.load());
    return bad ? 1 : 0;
}

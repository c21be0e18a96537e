// CPU stand-ins for the evaluation contexts, to run the request loop (nevpt2_amd/csrc/nv2_serve.cpp)
// under ThreadSanitizer / AddressSanitizer without a GPU.  Each stub context "runs" for a time set by
// its shape and writes a result that depends only on its inputs, so every request's result can be
// checked alone and in the mix.  A context with nact == 13 throws from nv2_ctx_run and one with
// nact == 11 returns an error, to exercise the loop's failure paths.
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <thread>

#include <hip/hip_runtime_api.h>

#include "../../include/nevpt2_amd.h"

static thread_local char g_err[1024];
std::atomic<int> g_live{0}, g_created{0};

namespace nv2 {
void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
}  // namespace nv2

extern "C" {
hipError_t hipMemGetInfo(size_t *fr, size_t *tot) {
    *fr = (size_t)1 << 34;
    *tot = (size_t)1 << 35;
    return hipSuccess;
}

struct nv2_ctx {
    int64_t nc, n, nv, M, nele, L;
    double seed = 0.0;
    std::atomic<int> busy{0};
};

int nv2_last_error(char *buf, int64_t len) {
    snprintf(buf, (size_t)len, "%s", g_err);
    return 0;
}

nv2_ctx *nv2_ctx_create(int64_t ncore, int64_t nact, int64_t nvirt, int64_t metat, int64_t nele, int64_t nchovec) {
    nv2_ctx *c = new nv2_ctx;
    c->nc = ncore; c->n = nact; c->nv = nvirt; c->M = metat; c->nele = nele; c->L = nchovec;
    g_live++;
    g_created++;
    return c;
}
int nv2_ctx_set_integrals(nv2_ctx *c, const double *chol, const double *one) {
    c->seed = chol[0] + one[0];
    return 0;
}
int nv2_ctx_set_ro4(nv2_ctx *c, const double *ro4) {
    c->seed += ro4[0];
    return 0;
}
int nv2_ctx_set_trdm(nv2_ctx *c, const double *t3, const double *, const double *) {
    c->seed += t3[0];
    return 0;
}
int nv2_ctx_run(nv2_ctx *c, int32_t, double, double *out) {
    if (c->busy.fetch_add(1) != 0) { nv2::set_error("context used by two requests at once"); return -9; }
    std::this_thread::sleep_for(std::chrono::milliseconds(2 + 3 * (c->n % 4)));
    c->busy--;
    if (c->n == 13) throw std::bad_alloc();
    if (c->n == 11) { nv2::set_error("stub failure"); return -1; }
    for (int64_t k = 0; k < 32 * c->M; k++) out[k] = c->seed + 0.001 * k + c->n;
    return 0;
}
int nv2_ctx_get(nv2_ctx *c, const char *, double *host, int64_t count) {
    for (int64_t k = 0; k < count; k++) host[k] = c->seed - k;
    return 0;
}
void nv2_ctx_destroy(nv2_ctx *c) {
    g_live--;
    delete c;
}
}

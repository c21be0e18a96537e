// Resident evaluation context: one NEVPT2 problem (dimensions fixed at creation)
// whose RDMs, integrals and every intermediate stay in HBM.  nv2_ctx_run is one
// full evaluation -- partial traces of the packed 4-RDM, every Koopmans matrix
// (koopro4QD_driver, koopro4QD.F90:270-424) and the SC/PC class energies of every
// state (qdnevpt, qdnevpt.F:858-993) -- with per-phase device timings.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nv2_common.h"
#include "nv2_rdm.h"
#include "nv2_stage1.h"
#include "nv2_stage2.h"
#include "../../include/nevpt2_amd.h"

namespace nv2 {

struct CtxBuf {
    std::string name;
    double *p = nullptr;
    size_t count = 0;
};

}  // namespace nv2

using namespace nv2;

struct nv2_ctx {
    int nc, n, nv, M, nele, L;
    i64 norb, np, nw;
    hipStream_t st = nullptr;
    rocblas_handle rb = nullptr;
    std::vector<CtxBuf> bufs;
    void *E4 = nullptr;
    size_t e4_bytes = 0;
    ContractPlan *plan = nullptr;
    Stage2 *s2 = nullptr;
    hipEvent_t ev[10];
    double ms[NV2_NPHASE];
    bool has_trdm = false;
    bool no4rdm = false;                    // .NO4RDM: no A-D matrices, no (+1)'/(-1)' classes
    bool natorb = false;                    // .NATORB: first-order density matrices and their eigenvalues
    std::vector<double> rho1, rho1_mo, natocc;   // per state: norb^2, norb^2, norb
    std::vector<double> qd;                 // per state s: 8 classes x (SC e2mp(j,s), PC e2en(j,s)), 16*M
    std::vector<double> heff_sc, heff_pc;   // second-order Heff, H[j + M*i] = e2mp(j,i) / e2en(j,i)
    std::vector<double> pcm;                // per state, per class: (min metric eigenvalue, kept, Cholesky path)
    std::vector<double> pct;                // per state: (K' reduction ms, its order, its batch, its form)
    double *find(const char *nm) {
        for (auto &b : bufs) if (b.name == nm) return b.p;
        return nullptr;
    }
    CtxBuf *entry(const char *nm) {
        for (auto &b : bufs) if (b.name == nm) return &b;
        return nullptr;
    }
};

static int ctx_alloc(nv2_ctx *c, const char *nm, size_t count) {
    CtxBuf b;
    b.name = nm;
    b.count = count;
    if (hipMalloc(&b.p, sizeof(double) * (count ? count : 1)) != hipSuccess) {
        set_error("hipMalloc(%s, %zu doubles) failed", nm, count);
        return -1;
    }
    c->bufs.push_back(b);
    return 0;
}

#define TRYC(x) do { int _r = (x); if (_r) { nv2_ctx_destroy(c); return nullptr; } } while (0)

extern "C" {

void nv2_ctx_destroy(nv2_ctx *c) {
    if (!c) return;
    if (c->st) (void)hipStreamSynchronize(c->st);
    stage2_destroy(c->s2);
    contract_plan_destroy(c->plan);
    for (auto &b : c->bufs) (void)hipFree(b.p);
    if (c->E4) (void)hipFree(c->E4);
    for (int k = 0; k < 10; k++) if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
    if (c->rb) rocblas_destroy_handle(c->rb);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

nv2_ctx *nv2_ctx_create(int64_t ncore, int64_t nact, int64_t nvirt, int64_t metat, int64_t nele, int64_t nchovec) {
    if (nact < 2 || nact > 16 || metat < 1 || ncore < 0 || nvirt < 1 || nele < 4 || nchovec < 1) {
        set_error("nv2_ctx_create: invalid dimensions (nact=%lld ncore=%lld nvirt=%lld metat=%lld nele=%lld L=%lld)",
                  (long long)nact, (long long)ncore, (long long)nvirt, (long long)metat, (long long)nele,
                  (long long)nchovec);
        return nullptr;
    }
    nv2_ctx *c = new nv2_ctx();
    for (int k = 0; k < 10; k++) c->ev[k] = nullptr;
    c->nc = (int)ncore; c->n = (int)nact; c->nv = (int)nvirt; c->M = (int)metat; c->nele = (int)nele;
    c->L = (int)nchovec;
    c->norb = ncore + nact + nvirt;
    c->np = c->norb * (c->norb + 1) / 2;
    const i64 n = nact, M = metat;
    c->nw = n * (n + 1) * (n + 2) * (n + 3) / 24;
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) { set_error("stream"); delete c; return nullptr; }
    if (rocblas_create_handle(&c->rb) != rocblas_status_success) { set_error("rocblas handle"); nv2_ctx_destroy(c); return nullptr; }
    for (int k = 0; k < 10; k++)
        if (hipEventCreate(&c->ev[k]) != hipSuccess) { set_error("event"); nv2_ctx_destroy(c); return nullptr; }
    const i64 n2 = n * n, n4 = n2 * n2, n6 = n4 * n2;
    TRYC(ctx_alloc(c, "cholesky", (size_t)nchovec * c->np));
    TRYC(ctx_alloc(c, "oneint", c->np));
    TRYC(ctx_alloc(c, "ro4", (size_t)c->nw * n4 * M));
    TRYC(ctx_alloc(c, "3-RDM", M * n6)); TRYC(ctx_alloc(c, "2-RDM", M * n4)); TRYC(ctx_alloc(c, "1-RDM", M * n2));
    TRYC(ctx_alloc(c, "aj", c->np)); TRYC(ctx_alloc(c, "atwo", n4));
    TRYC(ctx_alloc(c, "AJ", n2)); TRYC(ctx_alloc(c, "Fhalf", n2)); TRYC(ctx_alloc(c, "Fone", n2));
    const char *m2[] = {"1-IPKO", "1-EAKO", "gk0pf", "dmatIP", "dmatEA"};
    const char *m4[] = {"2-HRDM", "2-IPKO", "2-EAKO", "gk0pa", "gk0pd", "bmatIP", "bmatEA", "cmatIP", "cmatEA"};
    for (auto nm : m2) TRYC(ctx_alloc(c, nm, M * n2));
    for (auto nm : m4) TRYC(ctx_alloc(c, nm, M * n4));
    TRYC(ctx_alloc(c, "amatIP", M * n6)); TRYC(ctx_alloc(c, "amatEA", M * n6));
    for (const char *nm : {"C1", "C2", "C3", "C4"}) TRYC(ctx_alloc(c, nm, M * n6));
    const i64 P = M * (M - 1) / 2;
    if (P) {   // transition RDMs of the state pairs (nevpt.h5 1/2/3-TRDM)
        TRYC(ctx_alloc(c, "1-TRDM", P * n2)); TRYC(ctx_alloc(c, "2-TRDM", P * n4)); TRYC(ctx_alloc(c, "3-TRDM", P * n6));
    }
    TRYC(ctx_alloc(c, "bamat_work", (size_t)bamat_lo_workspace((int)n)));
    TRYC(ctx_alloc(c, "koop0p_work", (size_t)koop0p_workspace((int)n)));
    TRYC(ctx_alloc(c, "e4_d3t", n6));
    c->e4_bytes = sizeof(double) * (size_t)(n4 * n4);
    if (hipMalloc(&c->E4, c->e4_bytes) != hipSuccess) { set_error("hipMalloc(E4) failed"); nv2_ctx_destroy(c); return nullptr; }
    c->plan = contract_plan_create((int)n, (int)M);
    if (!c->plan) { set_error("contract plan"); nv2_ctx_destroy(c); return nullptr; }
    S2Inputs in{c->nc, c->n, c->nv, c->L, c->M, c->find("cholesky"), c->find("oneint"), c->find("1-RDM"),
                c->find("2-RDM"), c->find("3-RDM"), c->find("1-IPKO"), c->find("1-EAKO"), c->find("2-HRDM"),
                c->find("2-IPKO"), c->find("2-EAKO"), c->find("gk0pa"), c->find("gk0pd"), c->find("gk0pf"),
                c->find("amatIP"), c->find("bmatIP"), c->find("cmatIP"), c->find("dmatIP"), c->find("amatEA"),
                c->find("bmatEA"), c->find("cmatEA"), c->find("dmatEA")};
    in.trdm1 = c->find("1-TRDM"); in.trdm2 = c->find("2-TRDM"); in.trdm3 = c->find("3-TRDM");
    c->s2 = stage2_create(in, c->rb, c->st);
    if (!c->s2) { set_error("stage2_create failed"); nv2_ctx_destroy(c); return nullptr; }
    return c;
}

int nv2_ctx_set_integrals(nv2_ctx *c, const double *cholesky, const double *oneint) {
    if (!c) return -2;
    NV2_HIP(hipMemcpyAsync(c->find("cholesky"), cholesky, sizeof(double) * (size_t)c->L * c->np, hipMemcpyHostToDevice, c->st));
    NV2_HIP(hipMemcpyAsync(c->find("oneint"), oneint, sizeof(double) * c->np, hipMemcpyHostToDevice, c->st));
    NV2_HIP(hipStreamSynchronize(c->st));
    return 0;
}

int nv2_ctx_set_trdm(nv2_ctx *c, const double *trdm3, const double *trdm2, const double *trdm1) {
    if (!c) return -2;
    if (c->M < 2) { set_error("nv2_ctx_set_trdm: transition RDMs need metat >= 2"); return -2; }
    const char *nm[3] = {"3-TRDM", "2-TRDM", "1-TRDM"};
    const double *src[3] = {trdm3, trdm2, trdm1};
    for (int k = 0; k < 3; k++) {
        CtxBuf *b = c->entry(nm[k]);
        NV2_HIP(hipMemcpyAsync(b->p, src[k], sizeof(double) * b->count, hipMemcpyHostToDevice, c->st));
    }
    NV2_HIP(hipStreamSynchronize(c->st));
    c->has_trdm = true;
    return 0;
}

int nv2_ctx_set_ci(nv2_ctx *c, int64_t nalpha, int64_t nbeta, const double *ci) {
    if (!c || !ci) { set_error("nv2_ctx_set_ci: null argument"); return -2; }
    if (nalpha < 0 || nbeta < 0 || nalpha > c->n || nbeta > c->n || nalpha + nbeta != c->nele) {
        set_error("nv2_ctx_set_ci: nalpha=%lld nbeta=%lld do not fit CAS(%d,%d)", (long long)nalpha, (long long)nbeta,
                  c->nele, c->n);
        return -2;
    }
    const i64 ndet = ci_nstrings(c->n, (int)nalpha) * ci_nstrings(c->n, (int)nbeta);
    double *dC = nullptr;
    NV2_HIP(hipMalloc(&dC, sizeof(double) * ndet * c->M));
    int rc = hipMemcpyAsync(dC, ci, sizeof(double) * ndet * c->M, hipMemcpyHostToDevice, c->st) == hipSuccess ? 0 : -1;
    if (rc) set_error("nv2_ctx_set_ci: upload of the CI vectors failed");
    const bool qd = c->M > 1;
    if (!rc)
        rc = ci_fill_rdms(c->n, (int)nalpha, (int)nbeta, c->M, dC, c->rb, c->st, c->find("ro4"),
                          qd ? c->find("3-TRDM") : nullptr, qd ? c->find("2-TRDM") : nullptr,
                          qd ? c->find("1-TRDM") : nullptr);
    (void)hipStreamSynchronize(c->st);
    (void)hipFree(dC);
    if (!rc && qd) c->has_trdm = true;
    return rc;
}

int nv2_ctx_set_ro4(nv2_ctx *c, const double *ro4) {
    if (!c) return -2;
    CtxBuf *b = c->entry("ro4");
    NV2_HIP(hipMemcpyAsync(b->p, ro4, sizeof(double) * b->count, hipMemcpyHostToDevice, c->st));
    NV2_HIP(hipStreamSynchronize(c->st));
    return 0;
}

int nv2_ctx_run(nv2_ctx *c, int32_t prec, double thr_lindep, double *out) {
    if (!c) return -2;
    if (prec != NV2_PREC_FP64 && prec != NV2_PREC_FP32 && prec != NV2_PREC_BF16X3) { set_error("nv2_ctx_run: unsupported precision %d", prec); return -2; }
    hipStream_t st = c->st;
    const int n = c->n, M = c->M;
    Rdm R{n, M, c->nw, c->find("3-RDM"), c->find("2-RDM"), c->find("1-RDM"), c->find("ro4")};
    NV2_HIP(hipEventRecord(c->ev[0], st));
    // phase 0: partial traces (bro3/bro2/bro1)
    if (launch_bro(R, c->nele, c->find("3-RDM"), c->find("2-RDM"), c->find("1-RDM"), st)) return -1;
    NV2_HIP(hipEventRecord(c->ev[1], st));
    // the (+-1)' metric of the first state depends on the 3-RDM only: its solve overlaps stage 1
    if (stage2_pm1p_begin(c->s2, 0, thr_lindep)) return -1;
    // phase 1 (first part): Dyall h, active ERIs and the active-space f that bamat needs
    if (launch_aj_atwo(n, c->nc, (int)c->norb, c->L, c->find("cholesky"), c->find("oneint"), c->find("aj"),
                       c->find("atwo"), st)) return -1;
    Stage1Scratch w; w.AJ = c->find("AJ"); w.Fhalf = c->find("Fhalf"); w.Fone = c->find("Fone");
    w.h = c->rb; w.k0p = c->find("koop0p_work");
    Stage1Out so;
    so.ipko1 = c->find("1-IPKO"); so.eako1 = c->find("1-EAKO"); so.hrdm = c->find("2-HRDM");
    so.ipko2 = c->find("2-IPKO"); so.eako2 = c->find("2-EAKO"); so.gk0pa = c->find("gk0pa");
    so.gk0pd = c->find("gk0pd"); so.gk0pf = c->find("gk0pf"); so.bmat = c->find("bmatIP"); so.btmat = c->find("bmatEA");
    so.cmat = c->find("cmatIP"); so.ctmat = c->find("cmatEA"); so.dmat = c->find("dmatIP"); so.dtmat = c->find("dmatEA");
    if (c->no4rdm) so.bmat = nullptr;     // B, C and D serve only the (+1)'/(-1)' classes (koopro4QD.F90:357-424)
    if (launch_koop_prep(n, c->nc, c->find("aj"), c->find("atwo"), w, st)) return -1;
    // phases 2-3: E4 build and the four 4-RDM contractions of bamat, per state
    float t_build = 0.f, t_contr = 0.f;
    for (int s = 0; s < M && !c->no4rdm; s++) {
        hipEvent_t a = c->ev[7], b = c->ev[8], d = c->ev[9];
        NV2_HIP(hipEventRecord(a, st));
        if (launch_build_e4(R, s, c->E4, prec, c->find("e4_d3t"), st)) return -1;
        NV2_HIP(hipEventRecord(b, st));
        if (launch_e4_contractions(c->plan, c->E4, prec, c->find("atwo"), s, c->find("C1"), c->find("C2"),
                                   c->find("C3"), c->find("C4"), st)) return -1;
        NV2_HIP(hipEventRecord(d, st));
        NV2_HIP(hipEventSynchronize(d));
        float x = 0.f, y = 0.f;
        NV2_HIP(hipEventElapsedTime(&x, a, b));
        NV2_HIP(hipEventElapsedTime(&y, b, d));
        t_build += x; t_contr += y;
    }
    float t01 = 0.f;
    NV2_HIP(hipEventSynchronize(c->ev[1]));   // already complete unless no E4 phase synchronised above
    NV2_HIP(hipEventElapsedTime(&t01, c->ev[0], c->ev[1]));
    NV2_HIP(hipEventRecord(c->ev[4], st));
    // phase 4: lower-order part of bamat and final A-matrices
    for (int s = 0; s < M && !c->no4rdm; s++)
        if (launch_bamat_lo(c->rb, R, s, c->find("atwo"), c->find("Fhalf"), c->find("bamat_work"), c->find("C1"),
                            c->find("C2"), c->find("C3"), c->find("C4"), c->find("amatIP"), c->find("amatEA"), st))
            return -1;
    NV2_HIP(hipEventRecord(c->ev[2], st));
    // the (+-1)' chain of the first state needs amat only: it runs on while the rest of stage 1 is built
    if (!c->no4rdm && stage2_pm1p_kready(c->s2, 0)) return -1;
    // phase 1 (second part): Koopmans matrices of rank <= 3 and B/C/D
    if (launch_koop_rest(R, c->find("atwo"), w, so, st)) return -1;
    NV2_HIP(hipEventRecord(c->ev[5], st));
    // phase 5: stage 2 for every state, with the QD couplings when transition RDMs are present
    const bool qd = c->has_trdm && M > 1;
    c->qd.assign(qd ? (size_t)16 * M * M : 0, 0.0);
    c->pcm.assign((size_t)24 * M, 0.0);
    c->pct.assign((size_t)4 * M, 0.0);
    const size_t no = (size_t)c->norb;
    c->rho1.assign(c->natorb ? no * no * M : 0, 0.0);
    c->rho1_mo.assign(c->natorb ? no * no * M : 0, 0.0);
    c->natocc.assign(c->natorb ? no * M : 0, 0.0);
    for (int s = 0; s < M; s++) {
        S2Stats sd{};
        int rc = stage2_state(c->s2, s, thr_lindep, prec, out + 32 * s, &sd, qd ? c->qd.data() + 16 * M * s : nullptr);
        if (rc) return rc;
        if (c->natorb &&
            stage2_rho1_get(c->s2, c->rho1.data() + no * no * s, c->rho1_mo.data() + no * no * s, c->natocc.data() + no * s))
            return -1;
        for (int k = 0; k < 8; k++) {
            c->pcm[24 * s + 3 * k] = sd.min_metric_eig[k];
            c->pcm[24 * s + 3 * k + 1] = sd.nnew[k];
            c->pcm[24 * s + 3 * k + 2] = sd.chol[k];
        }
        c->pct[4 * s] = sd.kred_ms;
        c->pct[4 * s + 1] = sd.kred_N;
        c->pct[4 * s + 2] = sd.kred_nb;
        c->pct[4 * s + 3] = sd.kred_form;
    }
    c->heff_sc.assign((size_t)M * M, 0.0);
    c->heff_pc.assign((size_t)M * M, 0.0);
    for (int i = 0; i < M; i++) {   // column i: perturbers of state i (qdnevpt.F e2mp(j,i), e2en(j,i))
        for (int k = 0; k < 8; k++) {
            c->heff_sc[i + M * i] += out[32 * i + 4 * k + 1];
            c->heff_pc[i + M * i] += out[32 * i + 4 * k + 3];
            if (!qd) continue;
            for (int j = 0; j < M; j++) {
                if (j == i) continue;
                c->heff_sc[j + M * i] += c->qd[16 * M * i + k * 2 * M + j];
                c->heff_pc[j + M * i] += c->qd[16 * M * i + k * 2 * M + M + j];
            }
        }
    }
    NV2_HIP(hipEventRecord(c->ev[6], st));
    NV2_HIP(hipEventSynchronize(c->ev[6]));
    float t42 = 0.f, t25 = 0.f, t56 = 0.f;
    NV2_HIP(hipEventElapsedTime(&t42, c->ev[4], c->ev[2]));
    NV2_HIP(hipEventElapsedTime(&t25, c->ev[2], c->ev[5]));
    NV2_HIP(hipEventElapsedTime(&t56, c->ev[5], c->ev[6]));
    c->ms[0] = t01; c->ms[1] = t25; c->ms[2] = t_build; c->ms[3] = t_contr; c->ms[4] = t42; c->ms[5] = t56;
    return 0;
}

int nv2_ctx_set_option(nv2_ctx *c, const char *name, int64_t value) {
    if (!c || !name) { set_error("nv2_ctx_set_option: null argument"); return -2; }
    if (!strcmp(name, "no4rdm")) {
        c->no4rdm = value != 0;
        stage2_set_no4rdm(c->s2, c->no4rdm);
        return 0;
    }
    if (!strcmp(name, "natorb")) {
        c->natorb = value != 0;
        stage2_set_rho1(c->s2, c->natorb);
        return 0;
    }
    set_error("nv2_ctx_set_option: unknown option '%s'", name);
    return -2;
}

int nv2_ctx_set_frozen(nv2_ctx *c, int64_t nfrozen, const int64_t *orbitals) {
    if (!c || nfrozen < 0 || (nfrozen > 0 && !orbitals)) { set_error("nv2_ctx_set_frozen: invalid argument"); return -2; }
    std::vector<int> idx;
    for (int64_t k = 0; k < nfrozen; k++) {
        if (orbitals[k] == 0) continue;            // the reference ignores zero entries (qdnevpt.F:735)
        idx.push_back((int)(orbitals[k] - 1));
    }
    return stage2_set_frozen(c->s2, (int)idx.size(), idx.data());
}

int nv2_ctx_timings(nv2_ctx *c, double *ms) {
    if (!c || !ms) return -2;
    for (int k = 0; k < NV2_NPHASE; k++) ms[k] = c->ms[k];
    return 0;
}

int nv2_ctx_get(nv2_ctx *c, const char *name, double *host, int64_t count) {
    if (!c || !name) return -2;
    if (!strcmp(name, "PC-graphs")) {        // hipGraph launches of the batched eigensolver call sites
        if (count != 3) { set_error("nv2_ctx_get(PC-graphs): count %lld != 3", (long long)count); return -2; }
        if (!c->s2) { for (int k = 0; k < 3; k++) host[k] = 0.0; return 0; }
        stage2_graph_launches(c->s2, host);
        return 0;
    }
    const std::vector<double> *hv = !strcmp(name, "Heff-SC") ? &c->heff_sc : !strcmp(name, "Heff-PC") ? &c->heff_pc
                                  : !strcmp(name, "QD") ? &c->qd : !strcmp(name, "PC-metric") ? &c->pcm
                                  : !strcmp(name, "PC-timings") ? &c->pct : !strcmp(name, "rho1") ? &c->rho1
                                  : !strcmp(name, "rho1-mo") ? &c->rho1_mo : !strcmp(name, "natocc") ? &c->natocc
                                  : nullptr;
    if (hv) {
        if ((size_t)count != hv->size()) {
            set_error("nv2_ctx_get(%s): count %lld != %zu (run the context first)", name, (long long)count, hv->size());
            return -2;
        }
        for (size_t k = 0; k < hv->size(); k++) host[k] = (*hv)[k];
        return 0;
    }
    CtxBuf *b = c->entry(name);
    if (!b) { set_error("nv2_ctx_get: unknown array '%s'", name); return -2; }
    if ((size_t)count != b->count) { set_error("nv2_ctx_get(%s): count %lld != %zu", name, (long long)count, b->count); return -2; }
    NV2_HIP(hipStreamSynchronize(c->st));
    NV2_HIP(hipMemcpy(host, b->p, sizeof(double) * b->count, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"

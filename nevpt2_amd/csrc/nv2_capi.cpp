// C ABI of the NEVPT2 hot-path library (declarations in include/nevpt2_amd.h).
//
// The per-routine entry points take host arrays in the reference layouts,
// stage them to HBM, run the gfx950 kernels and copy the results back; they
// are the drop-in replacements for the Fortran routines they name.
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "nv2_common.h"
#include "nv2_stage1.h"
#include "nv2_stage2.h"
#include "nv2_syevb.h"
#include "../../include/nevpt2_amd.h"

namespace nv2 {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

// device buffer owned for the duration of one call
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    ~DBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t b) {
        bytes = b;
        if (b == 0) return 0;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) { set_error("hipMalloc(%zu bytes): %s", b, hipGetErrorString(e)); p = nullptr; return -1; }
        return 0;
    }
    int up(const void *h, size_t b) {
        if (alloc(b)) return -1;
        if (b && h) {
            hipError_t e = hipMemcpy(p, h, b, hipMemcpyHostToDevice);
            if (e != hipSuccess) { set_error("hipMemcpy H2D: %s", hipGetErrorString(e)); return -1; }
        }
        return 0;
    }
    int down(void *h) const {
        if (!bytes || !h) return 0;
        hipError_t e = hipMemcpy(h, p, bytes, hipMemcpyDeviceToHost);
        if (e != hipSuccess) { set_error("hipMemcpy D2H: %s", hipGetErrorString(e)); return -1; }
        return 0;
    }
    double *d() const { return (double *)p; }
};

static int check_dims(i64 nact, i64 metat) {
    if (nact < 1 || nact > 16) { set_error("nact=%lld outside 1..16", (long long)nact); return -2; }
    if (metat < 1) { set_error("metat=%lld must be >= 1", (long long)metat); return -2; }
    return 0;
}

static inline i64 nwords(i64 n) { return n * (n + 1) * (n + 2) * (n + 3) / 24; }
static inline i64 p(i64 n, int k) { i64 r = 1; while (k--) r *= n; return r; }

}  // namespace nv2

using namespace nv2;

#define TRY(x) do { int _r = (x); if (_r) return _r; } while (0)

extern "C" {

int nv2_last_error(char *buf, int64_t len) {
    if (!buf || len <= 0) return -2;
    strncpy(buf, g_err, (size_t)len - 1);
    buf[len - 1] = 0;
    return 0;
}

int nv2_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int nv2_bro(int64_t nact, int64_t metat, int64_t nele, const double *ro4, double *daaa, double *taa, double *dal) {
    TRY(check_dims(nact, metat));
    if (nele < 4) { set_error("nv2_bro needs nele >= 4 (got %lld)", (long long)nele); return -2; }
    const i64 n = nact, M = metat;
    DBuf dro4, d3, d2, d1;
    TRY(dro4.up(ro4, sizeof(double) * nwords(n) * p(n, 4) * M));
    TRY(d3.alloc(sizeof(double) * M * p(n, 6)));
    TRY(d2.alloc(sizeof(double) * M * p(n, 4)));
    TRY(d1.alloc(sizeof(double) * M * p(n, 2)));
    Rdm R{(int)n, (int)M, nwords(n), nullptr, nullptr, nullptr, dro4.d()};
    TRY(launch_bro(R, (int)nele, d3.d(), d2.d(), d1.d(), 0));
    NV2_HIP(hipDeviceSynchronize());
    TRY(d3.down(daaa)); TRY(d2.down(taa)); TRY(d1.down(dal));
    return 0;
}

int nv2_get_aj_atwo(int64_t ncore, int64_t nact, int64_t norb, int64_t nchovec, const double *cholesky,
                    const double *oneint, double *aj, double *atwo) {
    TRY(check_dims(nact, 1));
    const i64 np = norb * (norb + 1) / 2;
    DBuf dB, d1, daj, datwo;
    TRY(dB.up(cholesky, sizeof(double) * nchovec * np));
    TRY(d1.up(oneint, sizeof(double) * np));
    TRY(daj.alloc(sizeof(double) * np));
    TRY(datwo.alloc(sizeof(double) * p(nact, 4)));
    TRY(launch_aj_atwo((int)nact, (int)ncore, (int)norb, (int)nchovec, dB.d(), d1.d(), daj.d(), datwo.d(), 0));
    NV2_HIP(hipDeviceSynchronize());
    TRY(daj.down(aj)); TRY(datwo.down(atwo));
    return 0;
}

// all Koopmans-type matrices that need at most the 3-RDM: koop_matrices_ctl + bbmat/bcmat/bdmat
int nv2_koop_matrices(int64_t metat, int64_t nact, int64_t ncore, int64_t norb, const double *atwo,
                      const double *daaa, const double *taa, const double *dal, const double *aj, double *ipko1,
                      double *eako1, double *hrdm, double *ipko2, double *eako2, double *gk0pa, double *gk0pd,
                      double *gk0pf, double *bmat, double *btmat, double *cmat, double *ctmat, double *dmat,
                      double *dtmat) {
    TRY(check_dims(nact, metat));
    const i64 n = nact, M = metat, np = norb * (norb + 1) / 2;
    DBuf dat, d3, d2, d1, daj, AJ, Fh, F1;
    TRY(dat.up(atwo, sizeof(double) * p(n, 4)));
    TRY(d3.up(daaa, sizeof(double) * M * p(n, 6)));
    TRY(d2.up(taa, sizeof(double) * M * p(n, 4)));
    TRY(d1.up(dal, sizeof(double) * M * p(n, 2)));
    TRY(daj.up(aj, sizeof(double) * np));
    TRY(AJ.alloc(sizeof(double) * n * n)); TRY(Fh.alloc(sizeof(double) * n * n)); TRY(F1.alloc(sizeof(double) * n * n));
    const size_t s2 = sizeof(double) * M * n * n, s4 = sizeof(double) * M * p(n, 4);
    DBuf o[14];
    size_t sz[14] = {s2, s2, s4, s4, s4, s4, s4, s2, s4, s4, s4, s4, s2, s2};
    for (int k = 0; k < 14; k++) TRY(o[k].alloc(sz[k]));
    DBuf k0p;
    TRY(k0p.alloc(sizeof(double) * koop0p_workspace((int)n)));
    Stage1Scratch w; w.AJ = AJ.d(); w.Fhalf = Fh.d(); w.Fone = F1.d(); w.k0p = k0p.d();
    if (rocblas_create_handle(&w.h) != rocblas_status_success) { set_error("rocblas_create_handle failed"); return -1; }
    Stage1Out so;
    so.ipko1 = o[0].d(); so.eako1 = o[1].d(); so.hrdm = o[2].d(); so.ipko2 = o[3].d(); so.eako2 = o[4].d();
    so.gk0pa = o[5].d(); so.gk0pd = o[6].d(); so.gk0pf = o[7].d();
    so.bmat = o[8].d(); so.btmat = o[9].d(); so.cmat = o[10].d(); so.ctmat = o[11].d();
    so.dmat = o[12].d(); so.dtmat = o[13].d();
    Rdm R{(int)n, (int)M, nwords(n), d3.d(), d2.d(), d1.d(), nullptr};
    int rc = launch_koop_small(R, (int)ncore, daj.d(), dat.d(), w, so, 0);
    if (!rc) { hipError_t e = hipDeviceSynchronize(); if (e != hipSuccess) { set_error("koop: %s", hipGetErrorString(e)); rc = -1; } }
    rocblas_destroy_handle(w.h);
    if (rc) return rc;
    double *hs[14] = {ipko1, eako1, hrdm, ipko2, eako2, gk0pa, gk0pd, gk0pf, bmat, btmat, cmat, ctmat, dmat, dtmat};
    for (int k = 0; k < 14; k++) TRY(o[k].down(hs[k]));
    return 0;
}

int nv2_bamat(int64_t nact, int64_t ncore, int64_t norb, int64_t metat, const double *atwo, const double *ro4,
              const double *daaa, const double *taa, const double *dal, const double *aj, double *amat,
              double *atmat, int32_t prec) {
    TRY(check_dims(nact, metat));
    if (prec != NV2_PREC_FP64 && prec != NV2_PREC_FP32 && prec != NV2_PREC_BF16X3) { set_error("nv2_bamat: unsupported precision %d", prec); return -2; }
    const i64 n = nact, M = metat, np = norb * (norb + 1) / 2;
    DBuf dat, dro4, d3, d2, d1, daj, Fh, E4, C[4], dam, dat6, wk, d3t;
    TRY(dat.up(atwo, sizeof(double) * p(n, 4)));
    TRY(dro4.up(ro4, sizeof(double) * nwords(n) * p(n, 4) * M));
    TRY(d3.up(daaa, sizeof(double) * M * p(n, 6)));
    TRY(d2.up(taa, sizeof(double) * M * p(n, 4)));
    TRY(d1.up(dal, sizeof(double) * M * p(n, 2)));
    TRY(daj.up(aj, sizeof(double) * np));
    TRY(Fh.alloc(sizeof(double) * n * n));
    TRY(E4.alloc((prec == NV2_PREC_FP64 ? sizeof(double) : sizeof(float)) * p(n, 8)));
    for (int k = 0; k < 4; k++) TRY(C[k].alloc(sizeof(double) * M * p(n, 6)));
    TRY(dam.alloc(sizeof(double) * M * p(n, 6)));
    TRY(dat6.alloc(sizeof(double) * M * p(n, 6)));
    TRY(wk.alloc(sizeof(double) * bamat_lo_workspace((int)n)));
    TRY(d3t.alloc(sizeof(double) * p(n, 6)));
    Rdm R{(int)n, (int)M, nwords(n), d3.d(), d2.d(), d1.d(), dro4.d()};
    TRY(launch_make_f((int)n, (int)ncore, daj.d(), dat.d(), 0.5, Fh.d(), 0));
    ContractPlan *plan = contract_plan_create((int)n, (int)M);
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) { contract_plan_destroy(plan); set_error("rocblas_create_handle failed"); return -1; }
    int rc = plan ? 0 : -1;
    for (int s = 0; s < M && !rc; s++) {
        rc = launch_build_e4(R, s, E4.p, prec, d3t.d(), 0);
        if (!rc) rc = launch_e4_contractions(plan, E4.p, prec, dat.d(), s, C[0].d(), C[1].d(), C[2].d(), C[3].d(), 0);
        if (!rc) rc = launch_bamat_lo(h, R, s, dat.d(), Fh.d(), wk.d(), C[0].d(), C[1].d(), C[2].d(), C[3].d(), dam.d(),
                                      dat6.d(), 0);
    }
    if (!rc) { hipError_t e = hipDeviceSynchronize(); if (e != hipSuccess) { set_error("bamat: %s", hipGetErrorString(e)); rc = -1; } }
    rocblas_destroy_handle(h);
    contract_plan_destroy(plan);
    if (rc) return rc;
    TRY(dam.down(amat)); TRY(dat6.down(atmat));
    return 0;
}

int nv2_qdnevpt_state(int64_t ncore, int64_t nact, int64_t nvirt, int64_t metat, int64_t nchovec,
                      const double *cholesky, const double *oneint, const double *dal, const double *taa,
                      const double *daaa, const double *ipko1, const double *eako1, const double *hrdm,
                      const double *ipko2, const double *eako2, const double *gk0pa, const double *gk0pd,
                      const double *gk0pf, const double *amatIP, const double *bmatIP, const double *cmatIP,
                      const double *dmatIP, const double *amatEA, const double *bmatEA, const double *cmatEA,
                      const double *dmatEA, int64_t istate, double thr_lindep, double *out) {
    TRY(check_dims(nact, metat));
    if (istate < 0 || istate >= metat) { set_error("istate %lld out of range", (long long)istate); return -2; }
    const i64 n = nact, M = metat, norb = ncore + nact + nvirt, np = norb * (norb + 1) / 2;
    const size_t s2 = sizeof(double) * M * n * n, s4 = sizeof(double) * M * p(n, 4), s6 = sizeof(double) * M * p(n, 6);
    DBuf dB, d1e, dd1, dd2, dd3, k[16];
    TRY(dB.up(cholesky, sizeof(double) * nchovec * np));
    TRY(d1e.up(oneint, sizeof(double) * np));
    TRY(dd1.up(dal, s2)); TRY(dd2.up(taa, s4)); TRY(dd3.up(daaa, s6));
    const double *hk[16] = {ipko1, eako1, hrdm, ipko2, eako2, gk0pa, gk0pd, gk0pf,
                            amatIP, bmatIP, cmatIP, dmatIP, amatEA, bmatEA, cmatEA, dmatEA};
    size_t sz[16] = {s2, s2, s4, s4, s4, s4, s4, s2, s6, s4, s4, s2, s6, s4, s4, s2};
    for (int q = 0; q < 16; q++) if (hk[q]) TRY(k[q].up(hk[q], sz[q]));
    S2Inputs in{(int)ncore, (int)n, (int)nvirt, (int)nchovec, (int)M, dB.d(), d1e.d(), dd1.d(), dd2.d(), dd3.d(),
                k[0].d(), k[1].d(), k[2].d(), k[3].d(), k[4].d(), k[5].d(), k[6].d(), k[7].d(),
                amatIP ? k[8].d() : nullptr, k[9].d(), k[10].d(), k[11].d(),
                amatEA ? k[12].d() : nullptr, k[13].d(), k[14].d(), k[15].d()};
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) { set_error("rocblas_create_handle failed"); return -1; }
    Stage2 *st2 = stage2_create(in, h, 0);
    int rc = st2 ? stage2_state(st2, (int)istate, thr_lindep, 0, out, nullptr) : -1;
    stage2_destroy(st2);
    rocblas_destroy_handle(h);
    return rc;
}

int nv2_syev_batched(int64_t nprob, const int64_t *N, double *const *A, double *const *w) {
    if (nprob < 0 || (nprob > 0 && (!N || !A || !w))) { set_error("nv2_syev_batched: bad arguments"); return -2; }
    for (i64 q = 0; q < nprob; q++)
        if (N[q] < 1 || N[q] > SYEVB_NMAX) { set_error("nv2_syev_batched: order %lld outside 1..%d", (long long)N[q], SYEVB_NMAX); return -2; }
    std::vector<DBuf> dA(nprob), dw(nprob), dwk(nprob);
    std::vector<SyevbProb> P(nprob);
    for (i64 q = 0; q < nprob; q++) {
        const i64 n = N[q];
        TRY(dA[q].up(A[q], sizeof(double) * n * n));
        TRY(dw[q].alloc(sizeof(double) * n));
        TRY(dwk[q].alloc(sizeof(double) * syevb_workspace((int)n)));
        P[q] = SyevbProb{dA[q].d(), dw[q].d(), dwk[q].d(), (int)n};
    }
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) { set_error("rocblas_create_handle failed"); return -1; }
    int rc = rocblas_set_stream(h, 0) == rocblas_status_success ? 0 : -1;
    if (!rc) rc = syevb(h, 0, (int)nprob, P.data());
    if (!rc) { hipError_t e = hipDeviceSynchronize(); if (e != hipSuccess) { set_error("syevb: %s", hipGetErrorString(e)); rc = -1; } }
    rocblas_destroy_handle(h);
    if (rc) return rc;
    for (i64 q = 0; q < nprob; q++) { TRY(dA[q].down(A[q])); TRY(dw[q].down(w[q])); }
    return 0;
}

}  // extern "C"

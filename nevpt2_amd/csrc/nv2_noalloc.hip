// Koopmans matrices of rank <= 3 and B/C/D without any device allocation (nv2_koop_matrices_noalloc).
//
// For callers that cannot allocate (code under HIP graph capture, tight serving loops): inputs and
// outputs are device arrays the caller owns, the work is enqueued on the caller's stream, and the
// only device memory beyond them -- the three active-space Dyall f matrices (aj block, f with
// alpha = 1/2 and alpha = 1), 3 n^2 doubles -- lives in the file-scope array `scratch`, one copy per
// GPU, sized for nact <= 16.  The kernels are those of nv2_koop_matrices' direct form (one wavefront
// per output element), which need no rocBLAS workspace.  Every call on a GPU shares that GPU's
// scratch: calls that overlap on one GPU (two streams, two host threads) race.  Each call rewrites
// all of the scratch it reads before reading it, so nothing has to start at zero.
#include <hip/hip_runtime.h>

#include <mutex>

#include "nv2_common.h"
#include "nv2_stage1.h"
#include "../../include/nevpt2_amd.h"

namespace nv2 {

#define NOALLOC_MAXN 16
static __device__ __align__(256) double scratch[3 * NOALLOC_MAXN * NOALLOC_MAXN];

// Device address of `scratch` on the current GPU, looked up once per device with that device current.
// In host code the symbol's name is only a host-side stand-in; kernels get this address.
static int scratch_address(double **out) {
    static std::mutex mu;
    static double *addr[64] = {};
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= 64) {
        set_error("nv2_koop_matrices_noalloc: hipGetDevice: %s", hipGetErrorString(e));
        return -1;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("nv2_koop_matrices_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return -1;
        }
        addr[dev] = (double *)p;
    }
    *out = addr[dev];
    return 0;
}

}  // namespace nv2

using namespace nv2;

extern "C" int nv2_koop_matrices_noalloc(int64_t metat, int64_t nact, int64_t ncore, int64_t norb,
                                         const double *atwo, const double *daaa, const double *taa,
                                         const double *dal, const double *aj, double *ipko1, double *eako1,
                                         double *hrdm, double *ipko2, double *eako2, double *gk0pa, double *gk0pd,
                                         double *gk0pf, double *bmat, double *btmat, double *cmat, double *ctmat,
                                         double *dmat, double *dtmat, void *stream) {
    if (nact < 1 || nact > NOALLOC_MAXN) {
        set_error("nv2_koop_matrices_noalloc: nact=%lld outside 1..%d (the static scratch holds 3*%d^2 doubles)",
                  (long long)nact, NOALLOC_MAXN, NOALLOC_MAXN);
        return -2;
    }
    if (metat < 1 || ncore < 0 || norb < ncore + nact) {
        set_error("nv2_koop_matrices_noalloc: metat=%lld ncore=%lld norb=%lld do not fit nact=%lld",
                  (long long)metat, (long long)ncore, (long long)norb, (long long)nact);
        return -2;
    }
    const void *in[5] = {atwo, daaa, taa, dal, aj};
    void *out[14] = {ipko1, eako1, hrdm, ipko2, eako2, gk0pa, gk0pd, gk0pf, bmat, btmat, cmat, ctmat, dmat, dtmat};
    for (const void *q : in)
        if (!q) { set_error("nv2_koop_matrices_noalloc: null input array"); return -2; }
    for (void *q : out)
        if (!q) { set_error("nv2_koop_matrices_noalloc: null output array"); return -2; }
    double *ptr = nullptr;
    if (scratch_address(&ptr)) return -1;
    const int n = (int)nact;
    hipStream_t st = (hipStream_t)stream;
    Stage1Scratch w;                 // no rocBLAS handle: the direct kernels, no workspace
    w.AJ = ptr;
    w.Fhalf = ptr + n * n;
    w.Fone = ptr + 2 * n * n;
    Stage1Out so;
    so.ipko1 = ipko1; so.eako1 = eako1; so.hrdm = hrdm; so.ipko2 = ipko2; so.eako2 = eako2;
    so.gk0pa = gk0pa; so.gk0pd = gk0pd; so.gk0pf = gk0pf;
    so.bmat = bmat; so.btmat = btmat; so.cmat = cmat; so.ctmat = ctmat; so.dmat = dmat; so.dtmat = dtmat;
    const i64 nw = (i64)n * (n + 1) * (n + 2) * (n + 3) / 24;
    Rdm R{n, (int)metat, nw, const_cast<double *>(daaa), const_cast<double *>(taa), const_cast<double *>(dal),
          nullptr};
    return launch_koop_small(R, (int)ncore, aj, atwo, w, so, st);
}

// Internal stage-1 interfaces (device pointers, reference layouts).
#pragma once
#include <rocblas/rocblas.h>

#include "nv2_common.h"

namespace nv2 {

struct Stage1Scratch {
    double *AJ = nullptr;      // active aj block, n x n
    double *Fhalf = nullptr;   // f with alpha = 1/2
    double *Fone = nullptr;    // f with alpha = 1
    rocblas_handle h = nullptr;   // with k0p: koopman0pE's matrices as GEMMs (else one wavefront per element)
    double *k0p = nullptr;        // koop0p_workspace(n) doubles
};
i64 koop0p_workspace(int n);

struct Stage1Out {            // every array: state index fastest, reference shapes
    double *ipko1 = nullptr, *eako1 = nullptr;                         // (M,n,n)
    double *hrdm = nullptr, *ipko2 = nullptr, *eako2 = nullptr;        // (M,n^4)
    double *gk0pa = nullptr, *gk0pd = nullptr, *gk0pf = nullptr;       // (M,n^4),(M,n^4),(M,n,n)
    double *amat = nullptr, *atmat = nullptr;                          // (M,n^6)
    double *bmat = nullptr, *btmat = nullptr, *cmat = nullptr, *ctmat = nullptr;   // (M,n^4)
    double *dmat = nullptr, *dtmat = nullptr;                          // (M,n,n)
};

int launch_bro(const Rdm &R, int nele, double *daaa, double *taa, double *dal, hipStream_t st);
// bro2 and bro1 only, from R.d3 with R.M states (or state pairs)
int launch_bro21(const Rdm &R, int nele, double *taa, double *dal, hipStream_t st);
int launch_aj_atwo(int n, int ncore, int norb, int L, const double *B, const double *oneint, double *aj,
                   double *atwo, hipStream_t st);
int launch_make_f(int n, int ncore, const double *aj, const double *atwo, double alpha, double *F, hipStream_t st);
// launch_koop_small in two parts: the active-space Dyall f (AJ, Fhalf, Fone), which bamat needs,
// and the Koopmans matrices of rank <= 3 and B/C/D, which the (+-1)' chain does not
int launch_koop_prep(int n, int ncore, const double *aj, const double *atwo, Stage1Scratch &w, hipStream_t st);
int launch_koop_rest(const Rdm &R, const double *atwo, Stage1Scratch &w, Stage1Out &o, hipStream_t st);
int launch_koop_small(const Rdm &R, int ncore, const double *aj, const double *atwo, Stage1Scratch &w,
                      Stage1Out &o, hipStream_t st);
// d3t: n^6 doubles of scratch (the 3-RDM of state s transposed for the delta terms)
int launch_build_e4(const Rdm &R, int s, void *E4, int prec, double *d3t, hipStream_t st);
// launches each contraction pass is split into (NV2_PASS_CHUNKS)
int stage1_chunks();
// lower-order part of bamat for state s as GEMMs + one combine kernel (nv2_bamat_lo.hip);
// work holds bamat_lo_workspace(n) doubles
i64 bamat_lo_workspace(int n);
int launch_bamat_lo(rocblas_handle h, const Rdm &R, int s, const double *atwo, const double *Fhalf, double *work,
                    const double *C1, const double *C2, const double *C3, const double *C4, double *amat,
                    double *atmat, hipStream_t st);

// 4-RDM contractions of bamat for one state s; C1..C4 in amat's layout (see nv2_stage1.hip)
struct ContractPlan;
ContractPlan *contract_plan_create(int n, int M);
void contract_plan_destroy(ContractPlan *p);
int launch_e4_contractions(ContractPlan *p, const void *E4, int prec, const double *atwo, int s, double *C1,
                           double *C2, double *C3, double *C4, hipStream_t st);

}  // namespace nv2

// Internal stage-2 interfaces: canonicalisation, Cholesky integral assembly and
// the eight perturber classes (SC and PC) for one state, all on the device.
#pragma once
#include <rocblas/rocblas.h>
#include "nv2_common.h"

namespace nv2 {

struct S2Inputs {             // device pointers, reference layouts (state index fastest)
    int nc, n, nv, L, M;
    const double *B;          // cholesky(L, norb(norb+1)/2)
    const double *oneint;     // norb(norb+1)/2
    const double *dal, *taa, *daaa;
    const double *ipko1, *eako1, *hrdm, *ipko2, *eako2, *gk0pa, *gk0pd, *gk0pf;
    const double *amatIP, *bmatIP, *cmatIP, *dmatIP, *amatEA, *bmatEA, *cmatEA, *dmatEA;  // may be null (no 4-RDM)
    const double *trdm1 = nullptr, *trdm2 = nullptr, *trdm3 = nullptr;   // (P, n^k), P = M(M-1)/2; null: no QD
};

struct S2Stats {              // diagnostics for threshold-sensitive seeds
    double min_metric_eig[8];    // on the Cholesky path: thr, a lower bound
    int nnew[8];
    int chol[8];                 // 1: the metric was factored (every eigenvalue at or above thr)
    double kred_ms;              // device time of the K' reduction of the (+-1)' problems (0: none)
    int kred_N, kred_nb;         // its order and batch
    int kred_form;               // 1: tridiagonal form, 2: band form (half-bandwidth 64)
    double eps_min_virt, eps_max_core;
};

struct Stage2;
Stage2 *stage2_create(const S2Inputs &in, rocblas_handle h, hipStream_t st);
void stage2_destroy(Stage2 *p);
// hipGraph launches so far of the batched eigensolver call sites: worker-0 metrics, worker-0 K', Fock blocks
void stage2_graph_launches(Stage2 *p, double *out3);
// out[k*4 + {0,1,2,3}] = (norm_sc, e_sc, norm_pc, e_pc) for class k in the reference order
// (0),(+1),(-1),(+2),(-2),(+1)',(-1)',(0)'.  With qd_out (16*M doubles) and transition RDMs
// present, the QD couplings of state s: qd_out[k*2M + j] = SC e2mp(j,s), qd_out[k*2M + M + j] =
// PC e2en(j,s) of class k (zero for j == s).
int stage2_state(Stage2 *p, int s, double thr, int prec, double *out, S2Stats *stats, double *qd_out = nullptr);
// starts the (+-1)' metric solve of state s on worker 1 ahead of stage2_state (it needs the 3-RDM
// only); stage2_state(s) then waits for it instead of starting it
// Without the 4-RDM terms (the reference's .NO4RDM) the (+1)' and (-1)' classes and their QD
// couplings are not formed; their table entries stay zero (qdnevpt.F:949-991).
void stage2_set_no4rdm(Stage2 *p, bool on);
// Frozen core orbitals (the reference's .FROZEN, qdnevpt.F:733-745): canonical core orbitals, 0-based in
// ascending orbital energy, that no class uses as a hole.  They stay in the Fock operator; their orbital
// energy is replaced by -1e30 after canonicalisation, so every term with a hole in one has a denominator
// of ~1e30 and vanishes to rounding.  Returns -2 for an index outside the core.
int stage2_set_frozen(Stage2 *p, int nfrozen, const int *core_index);
// .NATORB (qdnevpt.F:1240-1467): with it on, stage2_state also forms the state's density matrix to first
// order -- core block 2, active block the 1-RDM, and the core-virtual, active-virtual and core-active
// blocks <Psi_0|E_pq|Psi_1> of the PC classes (0)', (-1)' and (+1)' -- in its canonical core and virtual
// orbitals; stage2_rho1_get copies the last state's matrix (norb x norb), the same matrix back in the
// input orbitals, and its eigenvalues in ascending order (any of the three pointers may be null).
void stage2_set_rho1(Stage2 *p, bool on);
int stage2_rho1_get(Stage2 *p, double *can, double *mo, double *occ);
int stage2_pm1p_begin(Stage2 *p, int s, double thr);
// Fills the chain's Koopmans matrix from amat of state s (bamat done) and releases the chain;
// stage2_state then skips that fill.  No-op without a chain started for state s.
int stage2_pm1p_kready(Stage2 *p, int s);

}  // namespace nv2

/*
 * nevpt2_amd -- C ABI of the MI355X (gfx950) NEVPT2 RDM-contraction hot path.
 *
 * Conventions (identical to the Fortran reference, so a Fortran caller passes
 * its arrays unchanged through ISO_C_BINDING):
 *   - every array is column-major with the state index fastest, e.g.
 *     daaa(metat,n,n,n,n,n,n), taa(metat,n,n,n,n), dal(metat,n,n);
 *   - the packed 4-RDM is ro4(nwords,n,n,n,n,metat), nwords = n(n+1)(n+2)(n+3)/24
 *     (koopro4QD.F90:263, indices_norm.F90:53);
 *   - one-electron quantities (oneint, aj) are packed over the norb triangle,
 *     pair (p,q) at max(p,q)*(max(p,q)-1)/2+min(p,q) (1-based, koop_matrices.F90:180);
 *   - Cholesky vectors cholesky(nchovec, norb*(norb+1)/2), pair column as above
 *     (ijkl_utils.F90:257, 676-677);
 *   - all dimensions are int64 (the reference builds with -fdefault-integer-8,
 *     cmake/FortranFlags.cmake:17-21).
 * Every entry point returns 0 on success, -1 on a HIP error, -2 on invalid
 * arguments; nv2_last_error() returns the message (the reference stops with a
 * message instead, e.g. qdnevpt.F:521-526).
 */
#ifndef NEVPT2_AMD_H
#define NEVPT2_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* arithmetic of the 4-RDM contractions (bamat) and PC S-builds */
#define NV2_PREC_FP64 0   /* drop-in form, matches the reference */
#define NV2_PREC_FP32 1   /* FP32 storage and arithmetic of the 4-RDM contractions */
#define NV2_PREC_BF16X3 2 /* FP32 storage; split-BF16x3 products on the bf16 matrix cores, FP32 accumulation */

int nv2_last_error(char *buf, int64_t len);
int nv2_device_count(void);

/* Partial traces of the packed 4-RDM: daaa = Tr4 ro4/(N-3), taa = Tr3 daaa/(N-2),
 * dal = Tr2 taa/(N-1).  Replaces bro3/bro2/bro1 (koopro4QD.F90:4881, 4975, 5028). */
int nv2_bro(int64_t nact, int64_t metat, int64_t nele, const double *ro4, double *daaa, double *taa, double *dal);

/* Dyall core-dressed one-electron matrix aj and active ERIs atwo(n,n,n,n) from
 * Cholesky vectors.  Replaces get_aj_atwo (koopro4QD.F90:3900). */
int nv2_get_aj_atwo(int64_t ncore, int64_t nact, int64_t norb, int64_t nchovec, const double *cholesky,
                    const double *oneint, double *aj, double *atwo);

/* Koopmans matrices of rank <= 3: 1-IPKO/1-EAKO (koopE, koop_matrices.F90:165),
 * 2-HRDM (koop_matrices.F90:78-88), 2-IPKO/2-EAKO (koop2E, :253),
 * gk0pa/gk0pd/gk0pf (koopman0pE, :389) and the B/C/D matrices of the (+-1)'
 * classes (bbmat/bcmat/bdmat, koopro4QD.F90:4231, 4301, 4399).  Together with
 * nv2_bamat this replaces koop_matrices_ctl (koop_matrices.F90:16) and the
 * A-D block of koopro4QD_driver (koopro4QD.F90:352-424).  Shapes:
 * (metat,n,n) for ipko1, eako1, gk0pf, dmat, dtmat; (metat,n,n,n,n) otherwise. */
int nv2_koop_matrices(int64_t metat, int64_t nact, int64_t ncore, int64_t norb, const double *atwo,
                      const double *daaa, const double *taa, const double *dal, const double *aj, double *ipko1,
                      double *eako1, double *hrdm, double *ipko2, double *eako2, double *gk0pa, double *gk0pd,
                      double *gk0pf, double *bmat, double *btmat, double *cmat, double *ctmat, double *dmat,
                      double *dtmat);

/* Allocation-free form of nv2_koop_matrices.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Every array argument is DEVICE memory the
 * caller owns, in the layouts of nv2_koop_matrices; the work is enqueued on
 * `stream` (a hipStream_t, NULL for the default stream) and the call returns
 * without waiting.  It calls no hipMalloc/hipFree and no allocating library
 * routine: the only scratch, the active-space Dyall f matrices (3 nact^2
 * doubles), lives in a static device array sized for nact <= 16; larger nact
 * returns -2.  All calls on one GPU share that array, so calls that overlap on
 * one GPU race. */
int nv2_koop_matrices_noalloc(int64_t metat, int64_t nact, int64_t ncore, int64_t norb, const double *atwo,
                              const double *daaa, const double *taa, const double *dal, const double *aj,
                              double *ipko1, double *eako1, double *hrdm, double *ipko2, double *eako2,
                              double *gk0pa, double *gk0pd, double *gk0pf, double *bmat, double *btmat,
                              double *cmat, double *ctmat, double *dmat, double *dtmat, void *stream);

/* A-matrices of the (+-1)' classes, amat/atmat(metat,n,n,n,n,n,n) (amatIP/amatEA).
 * Replaces bamat (koopro4QD.F90:3974), C1 branch. */
int nv2_bamat(int64_t nact, int64_t ncore, int64_t norb, int64_t metat, const double *atwo, const double *ro4,
              const double *daaa, const double *taa, const double *dal, const double *aj, double *amat,
              double *atmat, int32_t prec);

/* Second-order energies of one state: the canonical core/virtual orbitals, the
 * Cholesky transform and the eight perturber classes in SC and PC form.
 * Replaces the per-state body of qdnevpt (qdnevpt.F:369-784, 858-993) with the
 * class drivers v0, v1k, vm1k, v2mod, vm2mod, v0pmod, vmupE, vpupE
 * (qdnevpt.F:1661-6445).  Koopmans-matrix arguments are the stage-1 outputs
 * (nevpt.h5 datasets of the same names); the eight A-D matrices may be NULL,
 * which skips the (+-1)' classes as .NO4RDM does.  out[4*k + {0,1,2,3}] =
 * (norm_sc, e_sc, norm_pc, e_pc) of class k in the order (0),(+1),(-1),(+2),
 * (-2),(+1)',(-1)',(0)' -- the columns of the reference's result table
 * (qdnevpt.F:1019-1045).  istate is 0-based. */
int nv2_qdnevpt_state(int64_t ncore, int64_t nact, int64_t nvirt, int64_t metat, int64_t nchovec,
                      const double *cholesky, const double *oneint, const double *dal, const double *taa,
                      const double *daaa, const double *ipko1, const double *eako1, const double *hrdm,
                      const double *ipko2, const double *eako2, const double *gk0pa, const double *gk0pd,
                      const double *gk0pf, const double *amatIP, const double *bmatIP, const double *cmatIP,
                      const double *dmatIP, const double *amatEA, const double *bmatEA, const double *cmatEA,
                      const double *dmatEA, int64_t istate, double thr_lindep, double *out);

/* Resident evaluation context (no counterpart routine in the reference: it
 * replaces the koopro4QD.x -> nevpt.h5 -> QDnevpt2.x file pipeline,
 * koopro4QD_main.F90 / qdnevpt_main.F90, with device-resident hand-over).
 * Dimensions are fixed at creation; the packed 4-RDM and the integrals are
 * uploaded once and stay in HBM.  nv2_ctx_run is one full evaluation: partial
 * traces, every Koopmans matrix, and the SC/PC class energies of every state,
 * out[32*s + 4*k + j] as in nv2_qdnevpt_state.  nv2_ctx_get copies any
 * resident array back by its nevpt.h5 dataset name ("amatIP", "3-RDM", ...). */
#define NV2_NPHASE 6   /* bro, koop<=3-RDM, E4 build, E4 contractions, bamat rest, stage 2 */
typedef struct nv2_ctx nv2_ctx;
nv2_ctx *nv2_ctx_create(int64_t ncore, int64_t nact, int64_t nvirt, int64_t metat, int64_t nele, int64_t nchovec);
int nv2_ctx_set_integrals(nv2_ctx *ctx, const double *cholesky, const double *oneint);
int nv2_ctx_set_ro4(nv2_ctx *ctx, const double *ro4);
/* Spin-free RDMs from CI vectors on the GPU, written into the context: the
 * packed 4-RDM ro4 of every state and, for metat >= 2, the transition 3-/2-/1-RDMs
 * of the state pairs (as nv2_ctx_set_trdm takes them, bra = state I > J = ket).
 * Replaces rofour and ro3off/ro2off/ro1off (koopro4QD.F90:921-1570, 1978-2411,
 * 2620-2860, 2936-3014) with their drivers state_specific_densities_driver /
 * transition_densities_driver (koopro4QD.F90:504-917), which read the CI vector
 * from file04 instead.  ci holds metat vectors C[I_alpha][I_beta] (beta string
 * fastest) over all strings of nalpha / nbeta electrons in nact orbitals, each
 * string list in lexicographic order of the ascending occupied-orbital tuple;
 * a truncated expansion is the same array with zeros.  nalpha + nbeta must
 * equal the context's nele; nact <= 16.  Takes device memory of its own while it
 * runs (the CI vectors, the spin blocks and a GEMM panel of at most 2 GiB). */
int nv2_ctx_set_ci(nv2_ctx *ctx, int64_t nalpha, int64_t nbeta, const double *ci);
/* Transition RDMs of the state pairs for QD-NEVPT2 (metat >= 2): the nevpt.h5
 * datasets 3-TRDM (P,n^6), 2-TRDM (P,n^4), 1-TRDM (P,n,n), P = metat(metat-1)/2,
 * pair (I>J) at (I-1)(I-2)/2+J (koopro4QD.F90:433-477, read at qdnevpt.F:792-852).
 * Once set, nv2_ctx_run also forms the off-diagonal class terms that every class
 * driver adds to e2mp(j,i)/e2en(j,i) (e.g. qdnevpt.F:1936-1943, 2056-2070) and
 * nv2_ctx_get returns "Heff-SC" and "Heff-PC" (metat x metat, element (j,i) at
 * j + metat*i, second order without E0) and "QD" (per state i: 8 classes x
 * (SC row e2mp(.,i), PC row e2en(.,i)), 16*metat doubles per state).
 * "PC-metric" (after a run, any metat): per state and class, 24 doubles per state,
 * (smallest metric eigenvalue, eigenvalues kept, 1 if the metric was factored by
 * Cholesky because none falls below thr_lindep -- the smallest is then reported
 * as thr_lindep, a lower bound -- else 0), to flag threshold-sensitive inputs.
 * "PC-timings" (after a run): per state 4 doubles, the device time in ms of the
 * reduction of the (+-1)' Koopmans matrices K' (HIP events on its stream), their
 * order, their number and the form reduced to (1 tridiagonal, 2 band of
 * half-bandwidth 64); zeros when the state has none. */
int nv2_ctx_set_trdm(nv2_ctx *ctx, const double *trdm3, const double *trdm2, const double *trdm1);
int nv2_ctx_run(nv2_ctx *ctx, int32_t prec, double thr_lindep, double *out);
/* Run options of a context.  "no4rdm" (value != 0) is the reference's .NO4RDM
 * keyword (nevpt2_cfg no_4rdm_terms): the A-D matrices are not built
 * (koopro4QD.F90:357-424) and the (+1)' and (-1)' classes and their QD couplings
 * are not formed (qdnevpt.F:949-991): their table entries are zero and neither
 * "Heff-SC" nor "Heff-PC" contains them.
 * "natorb" (value != 0) is the reference's .NATORB keyword (compute_rho1st and
 * compute_nooccn, qdnevpt.F:1240-1467): each run also forms, per state, the density
 * matrix to first order that the reference diagonalises -- core block 2, active
 * block the 1-RDM, and the core-virtual, active-virtual and core-active blocks
 * <Psi_0|E_pq|Psi_1> accumulated by the PC classes (0)', (-1)' and (+1)'
 * (qdnevpt.F:4823-4988, 5326-5424, 5113-5203; zero without the 4-RDM classes),
 * virtual block 0.  nv2_ctx_get then returns "rho1" (norb x norb per state, in the
 * state's canonical core and virtual orbitals, as the reference prints it),
 * "rho1-mo" (the same matrix in the input orbitals, U rho U^T with U the
 * canonicalising rotation) and "natocc" (norb per state: its eigenvalues in
 * ascending order, the natural occupation numbers).  Unknown names return -2. */
int nv2_ctx_set_option(nv2_ctx *ctx, const char *name, int64_t value);
/* Frozen core orbitals: the reference's .FROZEN list (qdnevpt.F:733-745,
 * input_reader_sections.F90:110-114), 1-based orbital numbers of canonical core
 * orbitals (ascending orbital energy); zero entries are ignored.  No class uses
 * them as holes, in SC and PC form and in the QD couplings; they stay occupied
 * in the Fock operator.  nfrozen = 0 clears the list.  -2 for an orbital outside
 * the core. */
int nv2_ctx_set_frozen(nv2_ctx *ctx, int64_t nfrozen, const int64_t *orbitals);
int nv2_ctx_timings(nv2_ctx *ctx, double *ms);
int nv2_ctx_get(nv2_ctx *ctx, const char *name, double *host, int64_t count);
void nv2_ctx_destroy(nv2_ctx *ctx);

/* One spin-free k-RDM (k = 1..4) or transition k-RDM from CI vectors on the GPU:
 * out[p1 + n*p2 + ... + n^(2k-1)*qk] = <bra| sum over spins a+_{p1}..a+_{pk} a_{qk}..a_{q1} |ket>
 * (creator p_i paired with annihilator q_i, the convention of the reference's
 * RDM arrays, e.g. eexx_functions.F90:125-133).  bra/ket in the nv2_ctx_set_ci
 * layout (one vector each); ket NULL means ket = bra.  The host-side form of the
 * builders behind nv2_ctx_set_ci (koopro4QD.F90:921-3014). */
int nv2_ci_rdm(int64_t nact, int64_t nalpha, int64_t nbeta, int32_t k, const double *bra, const double *ket,
               double *out);

/* CI vectors in the reference's file04 form (the record read at
 * koopro4QD.F90:147-176) converted to the layout nv2_ctx_set_ci takes.
 * Determinant J (1..ncf) is the closed-shell reference (nocc = NOCA orbitals,
 * core included, doubly occupied) with the excitations trou(nd(J)+k) ->
 * part(nd(J)+k), k = 1..ne(J): spin-orbitals 1..norb and norb+1..2norb, the
 * pairs a+_part a_trou applied in listed order (the phase convention of
 * Jdifgen, koopro4QD.F90:3763-3886), particle 2*norb+1 removing the electron
 * (an ion, :188-204).  Replaces giveocc2 (:3123-3167) and the determinant
 * bookkeeping of the RDM builders.  c holds C(J, state) (J fastest); ci
 * receives metat vectors C[I_alpha][I_beta] over the active strings (alpha =
 * spin-orbitals 1..norb), and nalpha/nbeta their electron counts.  ci = NULL
 * only returns nalpha/nbeta (to size ci: metat * C(nact,nalpha) *
 * C(nact,nbeta) doubles).  Host code only, no GPU.  -2 when a determinant
 * changes a core or virtual occupation or the electron counts differ. */
int nv2_ci_from_file04(int64_t norb, int64_t ncore, int64_t nact, int64_t nocc, int64_t metat, int64_t ncf,
                       const int16_t *ne, const int64_t *nd, const int16_t *trou, const int16_t *part,
                       const double *c, double *ci, int64_t *nalpha, int64_t *nbeta);

/* QD-NEVPT2 effective Hamiltonian (qdnevpt.F:1046-1204 with hermitian_heff,
 * :8089-8184): adds e0 to the diagonal of the second-order heff2 (metat x metat,
 * element (j,i) at j + metat*i), takes the right eigenvectors of the
 * non-Hermitian matrix in ascending eigenvalue order, and returns the
 * hermitised Heff~ = C' diag(w) C'^T, C' = V (V^T V)^{-1/2}, and its
 * eigenvalues (the final QD energies).  Degenerate eigenvalues (states of
 * equal energy) get linearly independent eigenvectors; Heff~ then depends on the
 * basis chosen inside the degenerate eigenspace, as with dgeev's, its eigenvalues
 * do not.  Returns -3 when the spectrum is not real or the eigenvectors are
 * linearly dependent (a defective matrix).  metat <= 16. */
int nv2_qd_heff(int64_t metat, const double *heff2, const double *e0, double *heff, double *energies);

/* Eigen-decomposition of nprob real symmetric matrices of mixed orders
 * (1 <= N[q] <= 1024) in one batched solve: A[q] (order N[q], column-major,
 * upper triangle read) is overwritten by its orthonormal eigenvectors and w[q]
 * receives the eigenvalues in ascending order, as dsyev('V','U') returns them
 * (up to the sign of each eigenvector and the basis inside a degenerate
 * eigenspace).  Replaces the per-class metric and Koopmans-matrix dsyev calls of
 * the PC classes (qdnevpt.F:1989 and 2023 for (+1), the same pair in every class
 * driver, thresholding at :4744-4769) and the core and virtual Fock dsyev calls
 * (:519, 671), which stage 2 now issues as batches.  Host arrays. */
int nv2_syev_batched(int64_t nprob, const int64_t *N, double *const *A, double *const *w);

/* Request loop (no counterpart in the reference, whose drivers evaluate one
 * problem per process run: koopro4QD_main.F90:71, qdnevpt_main.F90:75).  A
 * server runs `slots` evaluations concurrently on one GPU; requests are admitted
 * first-in first-out when a slot is free and their device memory fits the HBM
 * budget (<= 0: 80% of the free memory), so requests arriving while others run
 * overlap with them on the device.  Contexts are kept resident per problem shape
 * and reused by later requests of that shape.  A host-data request points at
 * arrays in the nv2_ctx_set_* layouts, which must stay valid until its wait
 * returns; out (32*metat) and, for metat > 1 with transition RDMs, heff
 * (Heff-SC then Heff-PC, 2*metat^2) are filled on completion.
 * nv2_server_submit_ctx runs an existing context whose inputs are resident
 * (the caller keeps it, and must not run it elsewhere meanwhile).  Submit
 * returns a request id (> 0) or a negative status; nv2_server_wait blocks until
 * the request is done and returns its status, with the submit-to-completion
 * latency and the time spent queued.  Destroy drains the queue first. */
typedef struct nv2_request {
    int64_t ncore, nact, nvirt, metat, nele, nchovec;
    const double *cholesky, *oneint, *ro4;
    const double *trdm3, *trdm2, *trdm1;   /* NULL for one state or no QD couplings */
    int32_t prec;
    double thr_lindep;
    double *out;
    double *heff;                           /* NULL when not wanted */
} nv2_request;
typedef struct nv2_server nv2_server;
nv2_server *nv2_server_create(int32_t slots, int64_t hbm_budget_bytes);
int64_t nv2_server_submit(nv2_server *srv, const nv2_request *req);
int64_t nv2_server_submit_ctx(nv2_server *srv, nv2_ctx *ctx, int32_t prec, double thr_lindep, double *out);
int nv2_server_wait(nv2_server *srv, int64_t id, double *latency_ms, double *queue_ms);
void nv2_server_destroy(nv2_server *srv);
/* device bytes one evaluation context of this shape holds (the admission estimate) */
int64_t nv2_ctx_bytes(int64_t ncore, int64_t nact, int64_t nvirt, int64_t metat, int64_t nchovec);

#ifdef __cplusplus
}
#endif
#endif

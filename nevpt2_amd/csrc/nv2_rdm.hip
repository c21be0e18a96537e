// Spin-free reduced density matrices from CI vectors, on the GPU: the packed 4-RDM of every
// state and the transition 3-/2-/1-RDMs of the state pairs, written straight into a resident
// evaluation context (or returned to the host for one k-RDM).
//
// Replaces the reference's determinant-pair builders rofour (koopro4QD.F90:921-1570 with
// case0..case40 and accum) and ro3off/ro2off/ro1off (koopro4QD.F90:1978-2411, 2620-2860,
// 2936-3014), as driven by state_specific_densities_driver and transition_densities_driver
// (koopro4QD.F90:504-917).  Those walk every pair of determinants within four excitations of each
// other; here the same quantities come from string algebra that maps onto dense GEMMs:
//
//   CI vector C[Ia, Ib] over the alpha and beta occupation strings (each list in lexicographic
//   order of the ascending occupied-orbital tuple, beta index fastest).  For a alpha and b = k-a
//   beta annihilators, with Qa, Qb ascending orbital sets and Ka, Kb the strings left behind,
//     W[(Ka,Kb), (Qa,Qb)] = <Ka Kb| a_{Qb} a_{Qa} |C> = s(Ka,Qa) s(Kb,Qb) C[Ka+Qa, Kb+Qb],
//     s(K,Q) = (-1)^(sum over q in Q of the number of orbitals of K below q);
//   D_ab = W_bra^T W_ket (rocBLAS dsyrk when bra = ket, dgemm otherwise), summed over row chunks;
//   Gamma_k(p1..pk; q1..qk) = sum over the 2^k spin patterns (which positions i carry alpha spin;
//     creator p_i and annihilator q_i share it) of
//     sgn(p_alpha) sgn(p_beta) sgn(q_alpha) sgn(q_beta) D_ab[(sort p_alpha, sort p_beta), (sort q_alpha, sort q_beta)],
//   zero when a same-spin tuple repeats an orbital; sgn is the parity of the sort.
// The reordering sign that brings each spin pattern into alpha-then-beta order is the same for
// the creators and the annihilators, so it cancels.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "nv2_common.h"
#include "nv2_rdm.h"
#include "nv2_stage1.h"

namespace nv2 {

#define CI_NMAX 16

static i64 binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    i64 r = 1;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

// parity of a_{q_m} .. a_{q_1} |K + Q> -> |K> for Q ascending: each q passes the orbitals of K below it
__device__ inline int ann_parity(uint32_t K, uint32_t Q) {
    int p = 0;
    while (Q) {
        const int q = __ffs(Q) - 1;
        Q &= Q - 1;
        p += __popc(K & ((1u << q) - 1u));
    }
    return p;
}

// mask and sort parity of an orbital tuple; false when it repeats an orbital
__device__ inline bool tuple_key(const int *v, int cnt, uint32_t &mask, int &par) {
    mask = 0;
    par = 0;
    for (int i = 0; i < cnt; i++) {
        const uint32_t bit = 1u << v[i];
        if (mask & bit) return false;
        par += __popc(mask & ~((bit << 1) - 1u));    // earlier entries above v[i]
        mask |= bit;
    }
    return true;
}

struct DBlocks {
    const double *D[5];   // D_ab for a = 0..k alpha annihilators (nullptr: the split does not exist)
    int nqb[5];           // beta sets of size k-a (row or column = rank_alpha * nqb + rank_beta)
    i64 ld[5];            // leading dimension (= number of (Qa,Qb) sets)
    int sym;              // only the upper triangle is stored (bra = ket)
};

__device__ double spinfree_elem(const DBlocks &B, const int *rank, const int *p, const int *q, int k) {
    double acc = 0.0;
    for (int S = 0; S < (1 << k); S++) {
        const int a = __popc(S);
        const double *D = B.D[a];
        if (!D) continue;
        int pa[4], pb[4], qa[4], qb[4], na = 0, nb = 0;
        for (int i = 0; i < k; i++) {
            if ((S >> i) & 1) { pa[na] = p[i]; qa[na] = q[i]; na++; }
            else { pb[nb] = p[i]; qb[nb] = q[i]; nb++; }
        }
        uint32_t mpa, mpb, mqa, mqb;
        int s1, s2, s3, s4;
        if (!tuple_key(pa, na, mpa, s1) || !tuple_key(pb, nb, mpb, s2) || !tuple_key(qa, na, mqa, s3) ||
            !tuple_key(qb, nb, mqb, s4))
            continue;
        i64 r = (i64)rank[mpa] * B.nqb[a] + rank[mpb], c = (i64)rank[mqa] * B.nqb[a] + rank[mqb];
        if (B.sym && r > c) { const i64 t = r; r = c; c = t; }
        const double v = D[r + B.ld[a] * c];
        acc += ((s1 + s2 + s3 + s4) & 1) ? -v : v;
    }
    return acc;
}

// W rows rl = 0..rc-1 of the chunk starting at global row r0 = ka * nKb + kb, column (qa, qb)
__global__ void k_ci_w(const double *C, i64 nbs, const int *rank, const uint32_t *Ka, const uint32_t *Kb, int nKb,
                       const uint32_t *Qa, const uint32_t *Qb, int nQb, i64 r0, int rc, double *W) {
    const int rl = blockIdx.x * blockDim.x + threadIdx.x;
    if (rl >= rc) return;
    const int col = blockIdx.y;
    const i64 r = r0 + rl;
    const uint32_t ka = Ka[r / nKb], kb = Kb[r % nKb];
    const uint32_t qa = Qa[col / nQb], qb = Qb[col % nQb];
    double v = 0.0;
    if (!(ka & qa) && !(kb & qb)) {
        v = C[(i64)rank[ka | qa] * nbs + rank[kb | qb]];
        if ((ann_parity(ka, qa) + ann_parity(kb, qb)) & 1) v = -v;
    }
    W[rl + (i64)rc * col] = v;
}

// packed 4-RDM of one state: ro4[w, q1, q2, q3, q4] = Gamma_4(p(w); q), p(w) the descending creator
// tuple of packed word w (lindice order, indices_norm.F90:41-44)
__global__ void k_ci_ro4(DBlocks B, const int *rank, const int *wtup, i64 nw, int n, double *ro4) {
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nw * n * n * n * n) return;
    const i64 w = e % nw;
    i64 r = e / nw;
    int q[4];
    for (int i = 0; i < 4; i++) { q[i] = (int)(r % n); r /= n; }
    const int t = wtup[w];
    const int p[4] = {t & 255, (t >> 8) & 255, (t >> 16) & 255, (t >> 24) & 255};
    ro4[e] = spinfree_elem(B, rank, p, q, 4);
}

// full k-RDM, indices (p1..pk, q1..qk) with p1 fastest, element e at out[off + s0 * e]
__global__ void k_ci_full(DBlocks B, const int *rank, int n, int k, i64 tot, i64 s0, i64 off, double *out) {
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= tot) return;
    int p[4], q[4];
    i64 r = e;
    for (int i = 0; i < k; i++) { p[i] = (int)(r % n); r /= n; }
    for (int i = 0; i < k; i++) { q[i] = (int)(r % n); r /= n; }
    out[off + s0 * e] = spinfree_elem(B, rank, p, q, k);
}

static inline unsigned nblk(i64 tot, int bs) { return (unsigned)((tot + bs - 1) / bs); }

i64 ci_nstrings(int n, int k) { return binom(n, k); }

// String tables of one CAS: every orbital subset of n orbitals as a bit mask, grouped by size in
// lexicographic order, and the rank of every mask within its size.
struct CiPlan {
    int n = 0, na = 0, nb = 0;
    i64 nas = 0, nbs = 0;                 // alpha and beta strings
    i64 off[CI_NMAX + 2];                 // first mask of each size in `masks`
    int *d_rank = nullptr;
    uint32_t *d_masks = nullptr;
    int *d_wtup = nullptr;
    i64 nw = 0;
    ~CiPlan() {
        if (d_rank) (void)hipFree(d_rank);
        if (d_masks) (void)hipFree(d_masks);
        if (d_wtup) (void)hipFree(d_wtup);
    }
    const uint32_t *list(int size) const { return d_masks + off[size]; }
};

static int ci_plan(CiPlan &P, int n, int na, int nb, hipStream_t st) {
    P.n = n; P.na = na; P.nb = nb;
    P.nas = binom(n, na); P.nbs = binom(n, nb);
    std::vector<uint32_t> masks;
    std::vector<int> rank((size_t)1 << n, 0);
    for (int k = 0; k <= n; k++) {
        P.off[k] = (i64)masks.size();
        std::vector<int> c(k);
        for (int i = 0; i < k; i++) c[i] = i;
        int r = 0;
        for (;;) {                            // combinations in lexicographic order
            uint32_t m = 0;
            for (int i = 0; i < k; i++) m |= 1u << c[i];
            masks.push_back(m);
            rank[m] = r++;
            int i = k - 1;
            while (i >= 0 && c[i] == n - k + i) i--;
            if (i < 0) break;
            c[i]++;
            for (int j = i + 1; j < k; j++) c[j] = c[j - 1] + 1;
        }
    }
    P.off[n + 1] = (i64)masks.size();
    P.nw = (i64)n * (n + 1) * (n + 2) * (n + 3) / 24;
    std::vector<int> wtup((size_t)P.nw);
    for (int i1 = 1; i1 <= n; i1++)
        for (int i2 = 1; i2 <= i1; i2++)
            for (int i3 = 1; i3 <= i2; i3++)
                for (int i4 = 1; i4 <= i3; i4++) {
                    const i64 w = (i64)(i1 - 1) * i1 * (i1 + 1) * (i1 + 2) / 24 + (i2 - 1) * i2 * (i2 + 1) / 6 +
                                  (i3 - 1) * i3 / 2 + i4 - 1;
                    wtup[w] = (i1 - 1) | ((i2 - 1) << 8) | ((i3 - 1) << 16) | ((i4 - 1) << 24);
                }
    NV2_HIP(hipMalloc(&P.d_rank, sizeof(int) * rank.size()));
    NV2_HIP(hipMalloc(&P.d_masks, sizeof(uint32_t) * masks.size()));
    NV2_HIP(hipMalloc(&P.d_wtup, sizeof(int) * wtup.size()));
    NV2_HIP(hipMemcpyAsync(P.d_rank, rank.data(), sizeof(int) * rank.size(), hipMemcpyHostToDevice, st));
    NV2_HIP(hipMemcpyAsync(P.d_masks, masks.data(), sizeof(uint32_t) * masks.size(), hipMemcpyHostToDevice, st));
    NV2_HIP(hipMemcpyAsync(P.d_wtup, wtup.data(), sizeof(int) * wtup.size(), hipMemcpyHostToDevice, st));
    NV2_HIP(hipStreamSynchronize(st));       // the host vectors go out of scope
    return 0;
}

// Work buffers for one k: the D_ab blocks and the W chunk(s)
struct CiWork {
    double *D = nullptr, *W = nullptr, *W2 = nullptr;
    i64 wcap = 0;
    ~CiWork() {
        if (D) (void)hipFree(D);
        if (W) (void)hipFree(W);
        if (W2) (void)hipFree(W2);
    }
};

static const i64 CI_WCAP = (i64)1 << 28;     // W chunk: at most 2 GiB of FP64

static bool split_ok(const CiPlan &P, int k, int a) {
    const int b = k - a;
    return a <= P.na && b <= P.nb && a <= P.n && b <= P.n;
}

static int ci_work(const CiPlan &P, int k, bool two, CiWork &w) {
    i64 dtot = 0, wneed = 0;
    for (int a = 0; a <= k; a++) {
        if (!split_ok(P, k, a)) continue;
        const i64 nQ = binom(P.n, a) * binom(P.n, k - a);
        const i64 R = binom(P.n, P.na - a) * binom(P.n, P.nb - (k - a));
        dtot += nQ * nQ;
        wneed = std::max(wneed, R * nQ);
    }
    w.wcap = std::min(wneed, CI_WCAP);
    NV2_HIP(hipMalloc(&w.D, sizeof(double) * (dtot ? dtot : 1)));
    NV2_HIP(hipMalloc(&w.W, sizeof(double) * (w.wcap ? w.wcap : 1)));
    if (two) NV2_HIP(hipMalloc(&w.W2, sizeof(double) * (w.wcap ? w.wcap : 1)));
    return 0;
}

#define RBC(call) do { if ((call) != rocblas_status_success) { set_error("rocBLAS failed: %s", #call); return -1; } } while (0)

// D_ab for every spin split of k annihilators between bra and ket (device CI vectors)
static int ci_blocks(const CiPlan &P, int k, const double *Cb, const double *Ck, rocblas_handle h, hipStream_t st,
                     CiWork &w, DBlocks &B) {
    const bool sym = Ck == Cb;
    B.sym = sym ? 1 : 0;
    double *D = w.D;
    for (int a = 0; a <= k; a++) {
        B.D[a] = nullptr;
        B.nqb[a] = 0;
        B.ld[a] = 0;
        if (!split_ok(P, k, a)) continue;
        const int b = k - a;
        const i64 nQa = binom(P.n, a), nQb = binom(P.n, b), nQ = nQa * nQb;
        const i64 nKa = binom(P.n, P.na - a), nKb = binom(P.n, P.nb - b), R = nKa * nKb;
        if (nQ > 65535) { set_error("ci_blocks: %lld annihilator sets above the grid limit", (long long)nQ); return -2; }
        const i64 rc_max = std::max<i64>(1, std::min<i64>(R, w.wcap / nQ));
        const double one = 1.0, zero = 0.0;
        for (i64 r0 = 0; r0 < R; r0 += rc_max) {
            const int rc = (int)std::min<i64>(rc_max, R - r0);
            const dim3 grid(nblk(rc, 256), (unsigned)nQ);
            hipLaunchKernelGGL(k_ci_w, grid, dim3(256), 0, st, Cb, P.nbs, P.d_rank, P.list(P.na - a), P.list(P.nb - b),
                               (int)nKb, P.list(a), P.list(b), (int)nQb, r0, rc, w.W);
            NV2_LAUNCH_CHECK();
            const double *beta = r0 == 0 ? &zero : &one;
            if (sym) {
                RBC(rocblas_dsyrk(h, rocblas_fill_upper, rocblas_operation_transpose, (int)nQ, rc, &one, w.W, rc, beta,
                                  D, (int)nQ));
            } else {
                hipLaunchKernelGGL(k_ci_w, grid, dim3(256), 0, st, Ck, P.nbs, P.d_rank, P.list(P.na - a),
                                   P.list(P.nb - b), (int)nKb, P.list(a), P.list(b), (int)nQb, r0, rc, w.W2);
                NV2_LAUNCH_CHECK();
                RBC(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, (int)nQ, (int)nQ, rc, &one,
                                  w.W, rc, w.W2, rc, beta, D, (int)nQ));
            }
        }
        B.D[a] = D;
        B.nqb[a] = (int)nQb;
        B.ld[a] = nQ;
        D += nQ * nQ;
    }
    return 0;
}

static int check_dims(i64 nact, i64 nalpha, i64 nbeta) {
    if (nact < 1 || nact > CI_NMAX || nalpha < 0 || nbeta < 0 || nalpha > nact || nbeta > nact) {
        set_error("CI RDMs: unsupported space nact=%lld nalpha=%lld nbeta=%lld (nact <= %d)", (long long)nact,
                  (long long)nalpha, (long long)nbeta, CI_NMAX);
        return -2;
    }
    return 0;
}

struct StreamHandle {           // stream + rocBLAS handle of a standalone call
    hipStream_t st = nullptr;
    rocblas_handle h = nullptr;
    ~StreamHandle() {
        if (h) rocblas_destroy_handle(h);
        if (st) (void)hipStreamDestroy(st);
    }
};

// ro4 of every state and, for nroots >= 2, the transition 3-/2-/1-RDMs of the pairs (i > j) at
// (i-1)i/2 + j (0-based), bra = state i and ket = state j; CI vectors already on the device
int ci_fill_rdms(int n, int na, int nb, int nroots, const double *dC, rocblas_handle h, hipStream_t st, double *ro4,
                 double *trdm3, double *trdm2, double *trdm1) {
    if (check_dims(n, na, nb)) return -2;
    RBC(rocblas_set_stream(h, st));
    CiPlan P;
    if (ci_plan(P, n, na, nb, st)) return -1;
    const i64 ndet = P.nas * P.nbs, n4 = (i64)n * n * n * n;
    {
        CiWork w;
        if (ci_work(P, 4, false, w)) return -1;
        for (int s = 0; s < nroots; s++) {
            DBlocks B;
            const double *Cs = dC + s * ndet;
            if (ci_blocks(P, 4, Cs, Cs, h, st, w, B)) return -1;
            hipLaunchKernelGGL(k_ci_ro4, dim3(nblk(P.nw * n4, 256)), dim3(256), 0, st, B, P.d_rank, P.d_wtup, P.nw, n,
                               ro4 + s * P.nw * n4);
            NV2_LAUNCH_CHECK();
        }
        NV2_HIP(hipStreamSynchronize(st));    // the work buffers are freed on scope exit
    }
    const i64 npair = (i64)nroots * (nroots - 1) / 2;
    if (npair && trdm3) {
        CiWork w;
        if (ci_work(P, 3, true, w)) return -1;
        const i64 n6 = n4 * n * n;
        i64 ij = 0;
        for (int i = 0; i < nroots; i++)
            for (int j = 0; j < i; j++, ij++) {
                DBlocks B;
                if (ci_blocks(P, 3, dC + i * ndet, dC + j * ndet, h, st, w, B)) return -1;
                hipLaunchKernelGGL(k_ci_full, dim3(nblk(n6, 256)), dim3(256), 0, st, B, P.d_rank, n, 3, n6, npair, ij,
                                   trdm3);
                NV2_LAUNCH_CHECK();
            }
        Rdm R{n, (int)npair, 0, trdm3, trdm2, trdm1, nullptr};
        if (launch_bro21(R, na + nb, trdm2, trdm1, st)) return -1;
        NV2_HIP(hipStreamSynchronize(st));
    }
    return 0;
}

}  // namespace nv2

using namespace nv2;

extern "C" {

int nv2_ci_rdm(int64_t nact, int64_t nalpha, int64_t nbeta, int32_t k, const double *bra, const double *ket,
               double *out) {
    if (check_dims(nact, nalpha, nbeta)) return -2;
    if (k < 1 || k > 4 || !bra || !out) { set_error("nv2_ci_rdm: k=%d outside 1..4 or null argument", k); return -2; }
    StreamHandle sh;
    NV2_HIP(hipStreamCreateWithFlags(&sh.st, hipStreamNonBlocking));
    if (rocblas_create_handle(&sh.h) != rocblas_status_success) { set_error("rocblas handle"); return -1; }
    RBC(rocblas_set_stream(sh.h, sh.st));
    CiPlan P;
    if (ci_plan(P, (int)nact, (int)nalpha, (int)nbeta, sh.st)) return -1;
    const i64 ndet = P.nas * P.nbs;
    i64 tot = 1;
    for (int i = 0; i < 2 * k; i++) tot *= nact;
    double *dC = nullptr, *dOut = nullptr;
    const bool two = ket && ket != bra;
    struct Free { double **a, **b; ~Free() { if (*a) (void)hipFree(*a); if (*b) (void)hipFree(*b); } } fr{&dC, &dOut};
    NV2_HIP(hipMalloc(&dC, sizeof(double) * ndet * (two ? 2 : 1)));
    NV2_HIP(hipMalloc(&dOut, sizeof(double) * tot));
    NV2_HIP(hipMemcpyAsync(dC, bra, sizeof(double) * ndet, hipMemcpyHostToDevice, sh.st));
    if (two) NV2_HIP(hipMemcpyAsync(dC + ndet, ket, sizeof(double) * ndet, hipMemcpyHostToDevice, sh.st));
    CiWork w;
    if (ci_work(P, k, two, w)) return -1;
    DBlocks B;
    if (ci_blocks(P, k, dC, two ? dC + ndet : dC, sh.h, sh.st, w, B)) return -1;
    hipLaunchKernelGGL(k_ci_full, dim3(nblk(tot, 256)), dim3(256), 0, sh.st, B, P.d_rank, (int)nact, (int)k, tot,
                       (i64)1, (i64)0, dOut);
    NV2_LAUNCH_CHECK();
    NV2_HIP(hipMemcpyAsync(out, dOut, sizeof(double) * tot, hipMemcpyDeviceToHost, sh.st));
    NV2_HIP(hipStreamSynchronize(sh.st));
    return 0;
}

}  // extern "C"

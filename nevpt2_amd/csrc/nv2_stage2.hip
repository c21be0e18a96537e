// Stage 2 of the NEVPT2 hot path on gfx950: canonical core/virtual orbitals,
// Cholesky integral assembly, and the eight perturber classes in SC and PC
// form for one state (qdnevpt.F:369-784, 858-993, 1661-6445, 7599-8086).
//
// Plain library GEMMs (the Cholesky transforms and integral blocks, the PC
// metric transforms and the n^9 S-builds of the (+-1)' classes) go to rocBLAS;
// the metric/Koopmans eigen-solves go to rocSOLVER syevd with the reference's
// triangle choices.  Every denominator-weighted energy sum is a hand-written
// kernel that reads the resident integral blocks once and reduces per block;
// the block partials are summed on the device with compensated (Neumaier)
// FP64 accumulation.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <utility>
#include <string>
#include <thread>
#include <vector>

#include "nv2_common.h"
#include "nv2_stage2.h"
#include "nv2_band.h"
#include "nv2_chol.h"
#include "nv2_trd.h"
#include "nv2_syevb.h"
#include "../../include/nevpt2_amd.h"

namespace nv2 {

#define RB(call)                                                                          \
    do {                                                                                  \
        rocblas_status _s = (call);                                                       \
        if (_s != rocblas_status_success) {                                               \
            set_error("%s:%d %s: rocblas status %d", __FILE__, __LINE__, #call, (int)_s); \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)
#define TRYS(x) do { int _r = (x); if (_r) return _r; } while (0)

static inline unsigned nb(i64 tot, int bs) { return (unsigned)((tot + bs - 1) / bs); }

// ------------------------------------------------------------------ reductions
__device__ __forceinline__ double wsum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
// Wavefront reductions by DPP lane permutes (VALU latency, no LDS round trip): xor 1 and 2 by quad
// permutes, then the half-row and row mirrors, then rows 0 + 1 and 2 + 3 by row_bcast15 and the halves by
// row_bcast31; lane 63 ends with the total, which readlane returns to every lane (a wave-uniform value).
// Lanes outside a broadcast's row mask receive 0, the identity of both reductions below.
template <int CTRL, int RM>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long x) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)x, CTRL, RM, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(x >> 32), CTRL, RM, 0xf, false);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
__device__ __forceinline__ unsigned long long umax64(unsigned long long a, unsigned long long b) { return a > b ? a : b; }
__device__ __forceinline__ unsigned long long wmax_u64(unsigned long long x) {   // every lane active
    x = umax64(x, dpp_u64<0xb1, 0xf>(x));       // quad_perm [1,0,3,2]
    x = umax64(x, dpp_u64<0x4e, 0xf>(x));       // quad_perm [2,3,0,1]
    x = umax64(x, dpp_u64<0x141, 0xf>(x));      // row_half_mirror
    x = umax64(x, dpp_u64<0x140, 0xf>(x));      // row_mirror
    x = umax64(x, dpp_u64<0x142, 0xa>(x));      // row_bcast15
    x = umax64(x, dpp_u64<0x143, 0xc>(x));      // row_bcast31
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_f64(double v) {
    return __longlong_as_double((long long)dpp_u64<CTRL, RM>((unsigned long long)__double_as_longlong(v)));
}
__device__ __forceinline__ double wsum_dpp(double v) {                    // every lane active
    v += dpp_f64<0xb1, 0xf>(v);
    v += dpp_f64<0x4e, 0xf>(v);
    v += dpp_f64<0x141, 0xf>(v);
    v += dpp_f64<0x140, 0xf>(v);
    v += dpp_f64<0x142, 0xa>(v);
    v += dpp_f64<0x143, 0xc>(v);
    const unsigned long long x = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), 63);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// block sum of four values; thread 0 writes part[4*block + k]
__device__ void block_store4(double a, double b, double c, double d, double *part) {
    __shared__ double sh[4][16];
    a = wsum(a); b = wsum(b); c = wsum(c); d = wsum(d);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if (lane == 0) { sh[0][wid] = a; sh[1][wid] = b; sh[2][wid] = c; sh[3][wid] = d; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const i64 bid = (i64)blockIdx.x + (i64)gridDim.x * blockIdx.y;
        for (int k = 0; k < 4; k++) {
            double s = 0.0;
            for (int w = 0; w < nw; w++) s += sh[k][w];
            part[4 * bid + k] = s;
        }
    }
}
__device__ double block_sum(double v, double *sh16) {
    v = wsum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) sh16[wid] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += sh16[w];
    return s;
}
// Neumaier-compensated sum of the block partials; out[k] += sum_b part[4b+k]
__global__ void k_reduce_parts(const double *part, i64 nparts, double *out) {
    const int k = blockIdx.x;
    double s = 0.0, c = 0.0;
    for (i64 i = threadIdx.x; i < nparts; i += blockDim.x) {
        double x = part[4 * i + k], t = s + x;
        c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __shared__ double ss[256], cc[256];
    ss[threadIdx.x] = s; cc[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = 0.0, Cc = 0.0;
        for (int t = 0; t < (int)blockDim.x; t++) {
            double x = ss[t], u = S + x;
            Cc += (fabs(S) >= fabs(x)) ? (S - u) + x : (x - u) + S;
            S = u;
            Cc += cc[t];
        }
        out[k] += S + Cc;
    }
}

// ------------------------------------------------------------------ small utility kernels
__global__ void k_slice(const double *src, int M, int s, i64 cnt, double *dst) {
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) dst[i] = src[s + (i64)M * i];
}
// out[l + L*(p + np*q)] = B[l, tri(p0+p, q0+q)]
__global__ void k_unpack(const double *B, int L, int p0, int np, int q0, int nq, double *out) {
    i64 tot = (i64)L * np * nq;
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tot) return;
    int l = i % L;
    i64 r = i / L;
    int p = r % np, q = r / np;
    out[i] = B[l + (i64)L * tri(p0 + p, q0 + q)];
}
__global__ void k_unpack_h(const double *oneint, int norb, double *h) {
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (i64)norb * norb) return;
    int p = i % norb, q = i / norb;
    h[i] = oneint[tri(p, q)];
}
// Jcore[l] = sum_k B[l,tri(k,k)] ; Jact[l] = sum_ab B[l,tri(a,b)] da(b,a)
__global__ void k_jvec(const double *B, int L, int nc, int n, const double *da, double *Jc, double *Ja) {
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    double jc = 0.0, ja = 0.0;
    for (int k = 0; k < nc; k++) jc += B[l + (i64)L * tri(k, k)];
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) ja += B[l + (i64)L * tri(nc + a, nc + b)] * da[b + n * a];
    if (Jc) Jc[l] = jc;
    Ja[l] = ja;
}
// F[p,q] += J . B[:, tri(p0+p, p0+q)] (one wavefront per element)
__global__ void k_coul(const double *B, int L, int p0, int np, const double *J, double *F) {
    i64 e = (i64)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (e >= (i64)np * np) return;
    int p = e % np, q = e / np, lane = threadIdx.x & 63;
    const double *col = B + (i64)L * tri(p0 + p, p0 + q);
    double s = 0.0;
    for (int l = lane; l < L; l += 64) s += col[l] * J[l];
    s = wsum(s);
    if (lane == 0) F[e] += s;
}
__global__ void k_jdiag(const double *Tcc, int L, int nc, double *Jc) {
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    double s = 0.0;
    for (int j = 0; j < nc; j++) s += Tcc[l + (i64)L * (j + (i64)nc * j)];
    Jc[l] = s;
}
// caux[:,j] = U[:, first+j] / sqrt(w[first+j])
__global__ void k_caux(const double *U, const double *w, int N, int first, int nnew, double *caux) {
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (i64)N * nnew) return;
    int r = i % N, j = i / N;
    caux[i] = U[r + (i64)N * (first + j)] / sqrt(w[first + j]);
}
// symmetric matrix from the upper triangle (row <= col), as dsymm('L','U') reads it
__global__ void k_sym_upper(const double *K, int N, double *Ks) {
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (i64)N * N) return;
    int r = i % N, c = i / N;
    Ks[i] = (r <= c) ? K[i] : K[c + (i64)N * r];
}
__global__ void k_fill_diag_minus(const double *da, int n, double two, double *S) {   // S = two*I - da (or da)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n) return;
    int r = i % n, c = i / n;
    S[i] = (two != 0.0) ? ((r == c ? two : 0.0) - da[i]) : da[i];
}
// out[c + cols*r] = in[r + rows*c]
__global__ void k_transpose(const double *in, int rows, int cols, double *out) {
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (i64)rows * cols) return;
    int r = i % rows, c = i / rows;
    out[c + (i64)cols * r] = in[i];
}
__global__ void k_scale(const double *src, i64 cnt, double f, double *dst) {
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) dst[i] = f * src[i];
}

// ------------------------------------------------------------------ stage-2 accessors on per-state slices
struct Sl {           // da(n,n), taa(n^4), ro3(n^6) column-major
    int n;
    const double *da, *taa, *ro3;
};
__device__ __forceinline__ double DA(const Sl &S, int p, int q) { return S.da[p + S.n * q]; }
__device__ __forceinline__ double TA(const Sl &S, int p, int q, int r, int t) {
    const i64 n = S.n;
    return S.taa[p + n * (q + n * (r + n * t))];
}
__device__ __forceinline__ double R3(const Sl &S, int p, int q, int r, int t, int u, int v) {
    const i64 n = S.n;
    return S.ro3[p + n * (q + n * (r + n * (t + n * (u + n * v))))];
}
// qdnevpt.F:6462-6469
__device__ __forceinline__ double s_ee2(const Sl &S, int a, int b, int c, int d) {
    double v = TA(S, b, d, a, c);
    if (b == c) v += DA(S, d, a);
    return v;
}
// qdnevpt.F:6490-6497
__device__ __forceinline__ double s_ee2tr(const Sl &S, int a, int b, int c, int d) {
    double v = -s_ee2(S, a, b, d, c);
    if (c == d) v += 2.0 * DA(S, b, a);
    return v;
}
// qdnevpt.F:6450-6460
__device__ __forceinline__ double s_eee2(const Sl &S, int a, int b, int c, int d, int e, int f) {
    double v = R3(S, b, d, f, a, c, e);
    if (e == d) v += TA(S, b, f, a, c);
    if (e == b) v += TA(S, f, d, a, c);
    if (b == c) v += TA(S, d, f, a, e);
    if (b == c && e == d) v += DA(S, f, a);
    return v;
}
// qdnevpt.F:6480-6488
__device__ __forceinline__ double s_eee2t(const Sl &S, int a, int b, int c, int d, int e, int f) {
    double v = -s_eee2(S, a, b, d, c, e, f);
    if (c == d) v += 2.0 * s_ee2(S, a, b, e, f);
    return v;
}

// qdnevpt.F:6499-6506
__device__ __forceinline__ double s_ee2tl(const Sl &S, int a, int b, int c, int d) {
    double v = -s_ee2(S, b, a, c, d);
    if (a == b) v += 2.0 * DA(S, d, c);
    return v;
}
// ro2toff (qdnevpt.F:6508-6518): two-hole density of a state pair, without constant terms
__device__ __forceinline__ double s_ro2toff(const Sl &S, int ap, int bp, int a, int b) {
    double v = TA(S, ap, bp, a, b);
    if (ap == a) v -= 2.0 * DA(S, bp, b);
    if (bp == b) v -= 2.0 * DA(S, ap, a);
    if (ap == b) v += DA(S, bp, a);
    if (bp == a) v += DA(S, ap, b);
    return v;
}

// ------------------------------------------------------------------ QD-NEVPT2 couplings
// For state i (self) every class also sums, per other state j, the reference's
// off-diagonal terms e2mp(j,i) (SC) and e2en(j,i) (PC).  They reuse i's
// integrals, Koopmans matrices and denominators; the RDM-dependent factors come
// from the transition RDMs of the pair, prepared per j by k_qd_fill into the
// arrays a..e (their meaning is documented at each kernel).  M = 0 disables.
#define QM 8
struct QDArgs {
    int M, self;
    double *qpart;                 // per block: QM SC sums then QM PC sums
    const double *a[QM], *b[QM], *c[QM], *d[QM], *e[QM];
};
__device__ __forceinline__ bool qd_on(const QDArgs &q, int j) { return j < q.M && j != q.self; }

// block sums of sc[QM], pc[QM]; thread 0 writes qpart[2*QM*block + k]
__device__ void block_storeQ(const double *sc, const double *pc, const QDArgs &q) {
    __shared__ double sh[2 * QM][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < QM; k++) {
        double x = wsum(sc[k]), y = wsum(pc[k]);
        if (lane == 0) { sh[k][wid] = x; sh[QM + k][wid] = y; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const i64 bid = (i64)blockIdx.x + (i64)gridDim.x * blockIdx.y;
        for (int k = 0; k < 2 * QM; k++) {
            double s = 0.0;
            for (int w = 0; w < nw; w++) s += sh[k][w];
            q.qpart[2 * QM * bid + k] = s;
        }
    }
}
// Neumaier sum of the per-block QD partials: out[k] += sum_b part[2*QM*b + k]
__global__ void k_reduce_q(const double *part, i64 nparts, double *out) {
    const int k = blockIdx.x;
    double s = 0.0, c = 0.0;
    for (i64 i = threadIdx.x; i < nparts; i += blockDim.x) {
        double x = part[2 * QM * i + k], t = s + x;
        c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __shared__ double ss[256], cc[256];
    ss[threadIdx.x] = s; cc[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = 0.0, Cc = 0.0;
        for (int t = 0; t < (int)blockDim.x; t++) {
            double x = ss[t], u = S + x;
            Cc += (fabs(S) >= fabs(x)) ? (S - u) + x : (x - u) + S;
            S = u;
            Cc += cc[t];
        }
        out[k] += S + Cc;
    }
}

// Per-pair arrays of the off-diagonal terms.  S holds the transition RDMs of the
// pair (i,j) as stored by the reference (pair I>J, koopro4QD.F90:2969-2974);
// lt = (j < i) selects the reference's index orientation for that case.
enum QdKind {
    QK_V1_RO,     // (+1) SC  ro(a,b) at a+n*b                 qdnevpt.F:1912-1918
    QK_V1_R,      // (+1) PC  R(ap,a) at ap+n*a                qdnevpt.F:3128-3138
    QK_VM1_RO,    // (-1) SC                                   qdnevpt.F:2227-2234
    QK_VM1_R,     // (-1) PC                                   qdnevpt.F:2445-2452
    QK_P2_X,      // (+2) SC  X(C,D,c,d) at C+n(D+n(c+n d))    qdnevpt.F:3401-3414
    QK_P2_XS,     // (+2) SC, diagonal core pair
    QK_M2_X,      // (-2) SC                                   qdnevpt.F:3806-3819
    QK_M2_XS,
    QK_P2_Y,      // (+2) PC  Y[(A,B),(a,b)] at (A*n+B)+n^2(a*n+b)   qdnevpt.F:2859-2870
    QK_M2_Y,      // (-2) PC                                         qdnevpt.F:2613-2624
    QK_P2P_Y,     // (+2) PC diagonal pairs, upper-pair index        qdnevpt.F:2990-3003
    QK_M2P_Y,     // (-2) PC diagonal pairs                          qdnevpt.F:2737-2750
    QK_0P_X1,     // (0)' SC  ee2(b',a',a,b) at A+n(B+n(a+n b)) (same order for j<i and j>i) qdnevpt.F:4531
    QK_0P_X2,     // (0)' SC  ee2tr(b',b,a',a) + d(a,b) daoff(a',b')                    qdnevpt.F:4533-4536
    QK_0P_DS,     // (0)' SC  daoff(b,a)+daoff(a,b) at a+n*b                            qdnevpt.F:4569
    QK_0P_V,      // (0)' PC  V[ab,a'b'] (ab = a*n+b rows)                              qdnevpt.F:4862-4880
    QK_0P_V2,
    QK_0P_RO,     // (0)' PC  ro(ab)                                                    qdnevpt.F:4935-4945
    QK_P1P_R1,    // (+1)' SC -daoff(a',a) / -daoff(a,a') at a'+n*a                      qdnevpt.F:5602-5611
    QK_M1P_R1,    // (-1)' SC                                                           qdnevpt.F:6077-6088
    QK_P1P_QSC,   // (+1)' SC ee2tr(c',a',b',a)+ee2tl(a,b',a',c') at a'b'c' + n^3 a      qdnevpt.F:5642-5644
    QK_M1P_QSC,   // (-1)' SC ee2(c',a',b',a)+ee2(a,b',a',c')                          qdnevpt.F:6098-6100
    QK_P1P_MSC,   // (+1)' SC eee2t pattern, as [abc, a'b'c'] (transposed M)           qdnevpt.F:5661-5668
    QK_M1P_MSC,   // (-1)' SC eee2 pattern                                             qdnevpt.F:6120-6131
    QK_P1P_MPC,   // (+1)' PC eee2t pattern, [abc, a'b'c']                             qdnevpt.F:5082-5091
    QK_M1P_MPC,   // (-1)' PC eee2 pattern                                             qdnevpt.F:5295-5304
    QK_P1P_QPC,   // (+1)' PC ee2tl(a,b',a',c') (j<i) / ee2tr(c',a',b',a) (j>i)         qdnevpt.F:5084-5097
    QK_M1P_QPC    // (-1)' PC ee2(a,b',a',c') / ee2(c',a',b',a)                         qdnevpt.F:5297-5310
};
__global__ void k_qd_fill(Sl S, int kind, int lt, const int2 *up, double *out, i64 total) {
    const int n = S.n, nn = n * n;
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    double v = 0.0;
    switch (kind) {
    case QK_V1_RO: { int a = e % n, b = e / n; v = lt ? -DA(S, b, a) : -DA(S, a, b); break; }
    case QK_V1_R: { int ap = e % n, a = e / n; v = lt ? -DA(S, a, ap) : -DA(S, ap, a); break; }
    case QK_VM1_RO: { int a = e % n, b = e / n; v = lt ? DA(S, a, b) : DA(S, b, a); break; }
    case QK_VM1_R: { int ap = e % n, a = e / n; v = lt ? DA(S, ap, a) : DA(S, a, ap); break; }
    case QK_P2_X: case QK_P2_XS: {
        int C = e % n, D = (e / n) % n, c = (e / nn) % n, d = e / (nn * n);
        if (lt) { v = s_ro2toff(S, C, D, c, d); if (kind == QK_P2_XS) v += s_ro2toff(S, D, C, c, d); }
        else { v = s_ro2toff(S, c, d, C, D); if (kind == QK_P2_XS) v += s_ro2toff(S, c, d, D, C); }
        break;
    }
    case QK_M2_X: case QK_M2_XS: {
        int C = e % n, D = (e / n) % n, c = (e / nn) % n, d = e / (nn * n);
        if (lt) { v = TA(S, c, d, C, D); if (kind == QK_M2_XS) v += TA(S, c, d, D, C); }
        else { v = TA(S, C, D, c, d); if (kind == QK_M2_XS) v += TA(S, D, C, c, d); }
        break;
    }
    case QK_P2_Y: case QK_M2_Y: {
        int row = e % nn, col = e / nn;
        int A = row / n, B = row % n, a = col / n, b = col % n;
        if (kind == QK_P2_Y) v = lt ? s_ro2toff(S, b, a, B, A) : s_ro2toff(S, B, A, b, a);
        else v = lt ? TA(S, B, A, b, a) : TA(S, b, a, B, A);
        break;
    }
    case QK_P2P_Y: case QK_M2P_Y: {
        const int m = n * (n + 1) / 2;
        int row = e % m, col = e / m;
        int A = up[row].x, B = up[row].y, a = up[col].x, b = up[col].y;
        if (kind == QK_P2P_Y)
            v = lt ? 0.5 * (s_ro2toff(S, b, a, B, A) + s_ro2toff(S, b, a, A, B))
                   : 0.5 * (s_ro2toff(S, B, A, b, a) + s_ro2toff(S, A, B, b, a));
        else
            v = lt ? 0.5 * (TA(S, B, A, b, a) + TA(S, A, B, b, a)) : 0.5 * (TA(S, b, a, A, B) + TA(S, b, a, B, A));
        break;
    }
    case QK_0P_X1: { int A = e % n, B = (e / n) % n, a = (e / nn) % n, b = e / (nn * n); v = s_ee2(S, B, A, a, b); break; }
    case QK_0P_X2: {
        int A = e % n, B = (e / n) % n, a = (e / nn) % n, b = e / (nn * n);
        v = s_ee2tr(S, B, b, A, a) + (a == b ? DA(S, A, B) : 0.0);
        break;
    }
    case QK_0P_DS: { int a = e % n, b = e / n; v = DA(S, a, b) + DA(S, b, a); break; }
    case QK_0P_V: case QK_0P_V2: {
        int row = e % nn, col = e / nn;
        int a = row / n, b = row % n, A = col / n, B = col % n;
        if (kind == QK_0P_V) v = lt ? s_ee2(S, b, a, A, B) : s_ee2(S, B, A, a, b);
        else {
            v = lt ? -TA(S, a, B, A, b) : -TA(S, A, b, a, B);
            if (A == a) v += 2.0 * (lt ? DA(S, B, b) : DA(S, b, B));
        }
        break;
    }
    case QK_0P_RO: { int a = e / n, b = e % n; v = lt ? DA(S, b, a) : DA(S, a, b); break; }
    case QK_P1P_R1: { int A = e % n, a = e / n; v = lt ? -DA(S, A, a) : -DA(S, a, A); break; }
    case QK_M1P_R1: { int A = e % n, a = e / n; v = lt ? DA(S, a, A) : DA(S, A, a); break; }
    case QK_P1P_QSC: case QK_M1P_QSC: case QK_P1P_QPC: case QK_M1P_QPC: {
        const i64 N = (i64)nn * n;
        i64 row = e % N;
        int a = e / N;
        int A = row / nn, B = (row / n) % n, C = row % n;
        const bool plus = kind == QK_P1P_QSC || kind == QK_P1P_QPC;
        const double t1 = plus ? s_ee2tr(S, C, A, B, a) : s_ee2(S, C, A, B, a);     // X2(c',a',b',a)
        const double t2 = plus ? s_ee2tl(S, a, B, A, C) : s_ee2(S, a, B, A, C);     // (a,b',a',c')
        if (kind == QK_P1P_QSC || kind == QK_M1P_QSC) v = t1 + t2;
        else v = lt ? t2 : t1;
        break;
    }
    case QK_P1P_MSC: case QK_M1P_MSC: case QK_P1P_MPC: case QK_M1P_MPC: {
        const i64 N = (i64)nn * n;
        i64 row = e % N, col = e / N;
        int a = row / nn, b = (row / n) % n, c = row % n;
        int A = col / nn, B = (col / n) % n, C = col % n;
        const bool plus = kind == QK_P1P_MSC || kind == QK_P1P_MPC;
        // SC: X3(c',a',b',b,a,c) for j<i, X3(c,a,b,b',a',c') for j>i; PC the other way round
        const bool first = (kind == QK_P1P_MSC || kind == QK_M1P_MSC) ? (lt != 0) : (lt == 0);
        if (first) v = plus ? s_eee2t(S, C, A, B, b, a, c) : s_eee2(S, C, A, B, b, a, c);
        else v = plus ? s_eee2t(S, c, a, b, B, A, C) : s_eee2(S, c, a, b, B, A, C);
        break;
    }
    default: break;
    }
    out[e] = v;
}

// ------------------------------------------------------------------ class (0) (qdnevpt.F:1661-1773)
// E = -sum dint^2 / (e_r + e_s - e_i - e_j) over core pairs i <= j and virtual pairs r <= s, from the pair
// blocks B_p(r, s) = (i r | j s) of the core pairs p = (i <= j) (pairs_c order),
// nv x nv each at Bb + p nv^2: traint(r,s) = (j r | s i) = B_p(s, r), (j s | r i) = B_p(r, s).  Workgroup
// (tile pair t, pair p) loads the tiles (r0.., s0..) and (s0.., r0..) of B_p (s0 >= r0) and sums r <= s.
#define V0T 32
__global__ __launch_bounds__(256) void k_v0b(const double *Bb, const int2 *pairs, int nv, const double *ec,
                                             const double *ev, double *part) {
    __shared__ double ta[V0T][V0T + 1], tb[V0T][V0T + 1];
    const int nt = (nv + V0T - 1) / V0T;
    int t = blockIdx.x, rt = 0;                      // tile pair t -> (rt <= st)
    while (t >= nt - rt) { t -= nt - rt; rt++; }
    const int st = rt + t, r0 = V0T * rt, s0 = V0T * st;
    const int2 pr = pairs[blockIdx.y];
    const double *B = Bb + (i64)blockIdx.y * nv * nv;
    for (int e = threadIdx.x; e < V0T * V0T; e += 256) {
        const int a = e % V0T, b = e / V0T;
        ta[a][b] = r0 + a < nv && s0 + b < nv ? B[(r0 + a) + (i64)nv * (s0 + b)] : 0.0;   // B(r, s)
        tb[b][a] = s0 + a < nv && r0 + b < nv ? B[(s0 + a) + (i64)nv * (r0 + b)] : 0.0;   // tb[i][j] = B(s0 + j, r0 + i)
    }
    __syncthreads();
    const double dij = ec[pr.x] + ec[pr.y];
    const bool zij = pr.x == pr.y;
    double E = 0.0, P = 0.0;
    for (int e = threadIdx.x; e < V0T * V0T; e += 256) {
        const int a = e % V0T, b = e / V0T, r = r0 + a, s = s0 + b;
        if (r >= nv || s >= nv || s < r) continue;
        const double ab1 = tb[a][b], ab2 = ta[a][b];   // (j r | s i) = B(s, r), (j s | r i) = B(r, s)
        double dint;
        const bool zrs = r == s;
        if (zrs && zij) dint = ab1;
        else if (zrs != zij) dint = sqrt(2.0) * ab1;
        else dint = 2.0 * sqrt(ab1 * ab1 + ab2 * ab2 - ab1 * ab2);
        const double co = -dint / (ev[r] + ev[s] - dij);
        E += dint * co;
        P += co * co;
    }
    block_store4(E, P, E, P, part);
}

// ------------------------------------------------------------------ class (+1) (qdnevpt.F:1775-1934, 3076-3254)
template <int NM>
__global__ void k_v1(const double *I, i64 ld, const int2 *pairs, int nc, int n, int nv, const double *ec,
                     const double *ev, const double *gk, const double *ro, const double *spc, const double *w, int nnew,
                     double *part, QDArgs q) {
    // QD: q.a[j] = ro of pair (i,j) at a+n*b, q.b[j] = s(a,mu) of pair (i,j)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int2 pr = pairs[blockIdx.y];
    const int ii = pr.x, ij = pr.y;
    double Esc = 0, Psc = 0, Epc = 0, Ppc = 0;
    double qsc[QM], qpc[QM];
#pragma unroll
    for (int j = 0; j < QM; j++) { qsc[j] = 0.0; qpc[j] = 0.0; }
    if (r < nv) {
        double Tj[NM], Tk[NM];
#pragma unroll
        for (int a = 0; a < NM; a++) {
            Tj[a] = a < n ? I[(ii + (i64)nc * a) + ld * (ij + (i64)nc * r)] : 0.0;     // (a ii | r ij)
            Tk[a] = a < n ? I[(ij + (i64)nc * a) + ld * (ii + (i64)nc * r)] : 0.0;     // (a ij | r ii)
        }
        const bool same = ii == ij;
        double dn = 0.0, ek = 0.0, dnj[QM];
#pragma unroll
        for (int j = 0; j < QM; j++) dnj[j] = 0.0;
#pragma unroll
        for (int a = 0; a < NM; a++) {
            if (a >= n) break;
#pragma unroll
            for (int b = 0; b < NM; b++) {
                if (b >= n) break;
                double cint = same ? 0.5 * Tj[a] * Tj[b]
                                   : Tj[a] * Tj[b] + Tk[a] * Tk[b] - 0.5 * Tj[a] * Tk[b] - 0.5 * Tk[a] * Tj[b];
                dn += 2.0 * ro[a + n * b] * cint;
                ek += gk[a + n * b] * cint;
#pragma unroll
                for (int j = 0; j < QM; j++)
                    if (qd_on(q, j)) dnj[j] += 2.0 * q.a[j][a + n * b] * cint;
            }
        }
        if (fabs(dn) > 0.0) {
            double den = ec[ii] + ec[ij] - ev[r] - ek / dn;
            double co = dn / den;
            Esc += co;
            Psc += co / den;
#pragma unroll
            for (int j = 0; j < QM; j++) qsc[j] += dnj[j] / den;
        }
        for (int mu = 0; mu < nnew; mu++) {
            double rj = 0.0, rjp = 0.0;
#pragma unroll
            for (int a = 0; a < NM; a++) {
                if (a >= n) break;
                double s = spc[a + n * mu];
                rj += (same ? Tj[a] : Tj[a] + Tk[a]) * s;
                rjp += (Tj[a] - Tk[a]) * s;
            }
            double deno = w[mu] + ev[r] - ec[ii] - ec[ij];
            double contri = same ? rj * rj / deno : (0.5 * rj * rj + 1.5 * rjp * rjp) / deno;
            Epc -= contri;
            Ppc += contri / deno;
#pragma unroll
            for (int j = 0; j < QM; j++) {
                if (!qd_on(q, j)) continue;
                double xj = 0.0, xjp = 0.0;
                for (int a = 0; a < n; a++) {
                    double sj = q.b[j][a + n * mu];
                    xj += (same ? Tj[a] : Tj[a] + Tk[a]) * sj;
                    xjp += (Tj[a] - Tk[a]) * sj;
                }
                qpc[j] -= same ? xj * rj / deno : (0.5 * xj * rj + 1.5 * xjp * rjp) / deno;
            }
        }
    }
    block_store4(Esc, Psc, Epc, Ppc, part);
    if (q.M) block_storeQ(qsc, qpc, q);
}

// ------------------------------------------------------------------ class (-1) (qdnevpt.F:2095-2262, 2400-2561)
template <int NM>
__global__ void k_vm1(const double *I, i64 ld, const int2 *pairs, i64 npairs, int nc, int n, const double *ec,
                      const double *ev, const double *da, const double *gk, const double *spc, const double *w,
                      int nnew, double *part, QDArgs q) {
    // QD: q.a[j] = ro of pair (i,j) at a+n*b, q.b[j] = s(a,mu) of pair (i,j)
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    double Esc = 0, Psc = 0, Epc = 0, Ppc = 0;
    double qsc[QM], qpc[QM];
#pragma unroll
    for (int j = 0; j < QM; j++) { qsc[j] = 0.0; qpc[j] = 0.0; }
    if (t < npairs * nc) {
        const int i = t % nc;
        const int2 pr = pairs[t / nc];
        const int ir = pr.x, is = pr.y;
        double Tj[NM], Tk[NM];
#pragma unroll
        for (int a = 0; a < NM; a++) {
            Tj[a] = a < n ? I[(i + (i64)nc * ir) + ld * (a + (i64)n * is)] : 0.0;    // (r i | s a)
            Tk[a] = a < n ? I[(i + (i64)nc * is) + ld * (a + (i64)n * ir)] : 0.0;    // (s i | r a)
        }
        const bool same = ir == is;
        double dn = 0.0, ek = 0.0, dnj[QM];
#pragma unroll
        for (int j = 0; j < QM; j++) dnj[j] = 0.0;
#pragma unroll
        for (int a = 0; a < NM; a++) {
            if (a >= n) break;
#pragma unroll
            for (int b = 0; b < NM; b++) {
                if (b >= n) break;
                double cint = same ? 0.5 * Tj[a] * Tj[b]
                                   : Tj[a] * Tj[b] + Tk[a] * Tk[b] - 0.5 * Tj[a] * Tk[b] - 0.5 * Tk[a] * Tj[b];
                dn += 2.0 * da[a + n * b] * cint;
                ek += gk[a + n * b] * cint;
#pragma unroll
                for (int j = 0; j < QM; j++)
                    if (qd_on(q, j)) dnj[j] += 2.0 * q.a[j][a + n * b] * cint;
            }
        }
        if (fabs(dn) >= 1.0e-14) {
            double den = -ek / dn + ec[i] - ev[ir] - ev[is];
            double co = dn / den;
            Esc += co;
            Psc += co / den;
#pragma unroll
            for (int j = 0; j < QM; j++) qsc[j] += dnj[j] / den;
        }
        for (int mu = 0; mu < nnew; mu++) {
            double r1 = 0.0, r2 = 0.0;
#pragma unroll
            for (int a = 0; a < NM; a++) {
                if (a >= n) break;
                double s = spc[a + n * mu];
                r1 += (same ? Tj[a] : Tj[a] + Tk[a]) * s;
                r2 += (Tj[a] - Tk[a]) * s;
            }
            double deno = w[mu] + ev[ir] + ev[is] - ec[i];
            double contri = same ? r1 * r1 / deno : (0.5 * r1 * r1 + 1.5 * r2 * r2) / deno;
            Epc -= contri;
            Ppc += contri / deno;
#pragma unroll
            for (int j = 0; j < QM; j++) {
                if (!qd_on(q, j)) continue;
                double xj = 0.0, xjp = 0.0;
                for (int a = 0; a < n; a++) {
                    double sj = q.b[j][a + n * mu];
                    xj += (same ? Tj[a] : Tj[a] + Tk[a]) * sj;
                    xjp += (Tj[a] - Tk[a]) * sj;
                }
                qpc[j] -= same ? xj * r1 / deno : (0.5 * xj * r1 + 1.5 * xjp * r2) / deno;
            }
        }
    }
    block_store4(Esc, Psc, Epc, Ppc, part);
    if (q.M) block_storeQ(qsc, qpc, q);
}

// ------------------------------------------------------------------ classes (+2)/(-2), SC (qdnevpt.F:3327-3445, 3740-3849)
// one workgroup per external pair; T(c,d) is the pair's integral block, X the metric RDM and K the Koopmans matrix:
//   dn = sum cint*X, dk = sum cint*K, cint = T(c',d')T(c,d) [x 1/4 and symmetrised when the pair is diagonal]
__global__ void k_pm2_sc(const double *I, i64 ld, const int2 *pairs, int mode, int nc, int n, const double *X,
                         const double *K, const double *eo, double *part, QDArgs q) {
    // QD: q.a[j] = X of pair (i,j) for x != y, q.b[j] = the symmetrised X for x == y (both at C+n(D+n(c+n d)))
    // mode 0: (+2), pair (i<=j) core, T(a,b) = (i a | j b) from I_caca ; mode 1: (-2), pair (r<=s) virt,
    // T(a,b) = (s a | r b) from I_avav
    __shared__ double T[256];
    __shared__ double sh[16];
    const int2 pr = pairs[blockIdx.x];
    const int x = pr.x, y = pr.y;
    const int nn = n * n;
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        int a = e % n, b = e / n;
        T[e] = mode == 0 ? I[(x + (i64)nc * a) + ld * (y + (i64)nc * b)] : I[(a + (i64)n * y) + ld * (b + (i64)n * x)];
    }
    __syncthreads();
    const bool same = x == y;
    double dn = 0.0, dk = 0.0, dnj[QM];
#pragma unroll
    for (int j = 0; j < QM; j++) dnj[j] = 0.0;
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {      // (c',d') row
        const int C = e % n, D = e / n;
        double sx = 0.0, sk = 0.0;
        for (int f = 0; f < nn; f++) {
            const int c = f % n, d = f / n;
            const i64 o = C + (i64)n * (D + (i64)n * (c + (i64)n * d));
            double xv = X[o], kv = K[o];
            if (same) {
                xv += X[C + (i64)n * (D + (i64)n * (d + (i64)n * c))];
                kv += K[D + (i64)n * (C + (i64)n * (c + (i64)n * d))];
            }
            sx += T[f] * xv;
            sk += T[f] * kv;
        }
        const double f = same ? 0.25 : 1.0;
        dn += f * T[e] * sx;
        dk += f * T[e] * sk;
#pragma unroll
        for (int j = 0; j < QM; j++) {
            if (!qd_on(q, j)) continue;
            const double *Xj = same ? q.b[j] : q.a[j];
            double sj = 0.0;
            for (int g = 0; g < nn; g++) sj += T[g] * Xj[e + (i64)nn * g];
            dnj[j] += f * T[e] * sj;
        }
    }
    dn = block_sum(dn, sh);
    dk = block_sum(dk, sh);
#pragma unroll
    for (int j = 0; j < QM; j++)
        if (qd_on(q, j)) dnj[j] = block_sum(dnj[j], sh);
    double E = 0.0, P = 0.0, qsc[QM], qpc[QM];
#pragma unroll
    for (int j = 0; j < QM; j++) { qsc[j] = 0.0; qpc[j] = 0.0; }
    if (threadIdx.x == 0 && fabs(dn) >= 1.0e-14) {
        double den = mode == 0 ? eo[x] + eo[y] - dk / dn : -dk / dn - eo[x] - eo[y];
        double co = dn / den;
        E = co;
        P = co / den;
#pragma unroll
        for (int j = 0; j < QM; j++)
            if (qd_on(q, j)) qsc[j] = dnj[j] / den;
    }
    if (threadIdx.x != 0) { E = 0.0; P = 0.0; }
    block_store4(E, P, 0.0, 0.0, part);
    if (q.M) block_storeQ(qsc, qpc, q);
}

// The same sums for the pairs x < y as one GEMM: T (npairs x nn, pair-major: the pair's integral block
// T(a,b) at a + n b) times [X | K | X_j ...] (nn x ncol) gives U = T [X | K ...]; then dn = T_p . U_p[X],
// dk = T_p . U_p[K] and the QD forms T_p . U_p[X_j] (k_pm2_dot, one wavefront per pair).
__global__ void k_pm2_tgather(const double *I, i64 ld, const int2 *pairs, i64 npairs, int mode, int nc, int n,
                             double *T) {
    const int nn = n * n;
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs * nn) return;
    const int e = (int)(t % nn), a = e % n, b = e / n;
    const i64 pi = t / nn;
    const int2 pr = pairs[pi];
    const int x = pr.x, y = pr.y;
    T[t] = mode == 0 ? I[(x + (i64)nc * a) + ld * (y + (i64)nc * b)] : I[(a + (i64)n * y) + ld * (b + (i64)n * x)];
}
__global__ void k_pm2_dot(const double *T, const double *U, int nn, int ncol, const int2 *pairs, i64 npairs, int mode,
                          const double *eo, double *part, QDArgs q) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const i64 pi = (i64)blockIdx.x * (blockDim.x >> 6) + wv;
    double E = 0.0, P = 0.0, qsc[QM], qpc[QM];
#pragma unroll
    for (int j = 0; j < QM; j++) { qsc[j] = 0.0; qpc[j] = 0.0; }
    if (pi < npairs) {
        const double *Tp = T + pi * nn, *Up = U + pi * ncol;
        double dn = 0.0, dk = 0.0, dnj[QM];
#pragma unroll
        for (int j = 0; j < QM; j++) dnj[j] = 0.0;
        for (int e = lane; e < nn; e += 64) {
            const double te = Tp[e];
            dn += te * Up[e];
            dk += te * Up[nn + e];
            int col = 2;
#pragma unroll
            for (int j = 0; j < QM; j++)
                if (qd_on(q, j)) dnj[j] += te * Up[(i64)nn * col++ + e];
        }
        dn = wsum(dn);
        dk = wsum(dk);
#pragma unroll
        for (int j = 0; j < QM; j++)
            if (qd_on(q, j)) dnj[j] = wsum(dnj[j]);
        if (lane == 0 && fabs(dn) >= 1.0e-14) {
            const int2 pr = pairs[pi];
            const double den = mode == 0 ? eo[pr.x] + eo[pr.y] - dk / dn : -dk / dn - eo[pr.x] - eo[pr.y];
            E = dn / den;
            P = E / den;
#pragma unroll
            for (int j = 0; j < QM; j++)
                if (qd_on(q, j)) qsc[j] = dnj[j] / den;
        }
    }
    if (lane != 0) {
        E = 0.0; P = 0.0;
#pragma unroll
        for (int j = 0; j < QM; j++) qsc[j] = 0.0;
    }
    block_store4(E, P, 0.0, 0.0, part);
    if (q.M) block_storeQ(qsc, qpc, q);
}

// ------------------------------------------------------------------ classes (+2)/(-2), PC (qdnevpt.F:2563-3072)
// off-diagonal pairs: v(ab) = T(b,a) with ab = a*n+b ; contribution -(v.s_mu)^2/deno
__global__ void k_pm2_pc(const double *I, i64 ld, const int2 *pairs, i64 npairs, int mode, int nc, int n,
                         const double *spc, const double *w, int nnew, const double *eo, double *part, QDArgs q) {
    // QD: q.a[j] = s(ab,mu) of pair (i,j)
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    double E = 0.0, P = 0.0, qsc[QM], qpc[QM];
#pragma unroll
    for (int j = 0; j < QM; j++) { qsc[j] = 0.0; qpc[j] = 0.0; }
    if (t < npairs * nnew) {
        const int mu = t % nnew;
        const int2 pr = pairs[t / nnew];
        const int x = pr.x, y = pr.y;
        const int nn = n * n;
        double acc = 0.0, accj[QM];
#pragma unroll
        for (int j = 0; j < QM; j++) accj[j] = 0.0;
        for (int ab = 0; ab < nn; ab++) {
            const int a = ab / n, b = ab % n;
            // (+2): traint(b,a) = (i b | j a) ; (-2): dint1(b,a) = (r b | s a)
            double v = mode == 0 ? I[(x + (i64)nc * b) + ld * (y + (i64)nc * a)] : I[(b + (i64)n * x) + ld * (a + (i64)n * y)];
            acc += v * spc[ab + (i64)nn * mu];
#pragma unroll
            for (int j = 0; j < QM; j++)
                if (qd_on(q, j)) accj[j] += v * q.a[j][ab + (i64)nn * mu];
        }
        double deno = mode == 0 ? w[mu] - eo[x] - eo[y] : w[mu] + eo[x] + eo[y];
        double contri = acc * acc / deno;
        E = -contri;
        P = contri / deno;
#pragma unroll
        for (int j = 0; j < QM; j++)
            if (qd_on(q, j)) qpc[j] = -accj[j] * acc / deno;
    }
    block_store4(0.0, 0.0, E, P, part);
    if (q.M) block_storeQ(qsc, qpc, q);
}
// The same sums as k_pm2_pc as one GEMM: the pairs' n^2 integral blocks gathered into Vt (n^2 x npairs,
// ab fastest), R = Vt^T spc (npairs x nnew, rocBLAS), then one thread per (pair, mu) for the
// denominators; QD partners get R_j = Vt^T s_j the same way (their pointers replace q.a[j]).
__global__ void k_pm2_gather(const double *I, i64 ld, const int2 *pairs, i64 npairs, int mode, int nc, int n,
                             double *Vt) {
    const int nn = n * n;
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs * nn) return;
    const int ab = (int)(t % nn), a = ab / n, b = ab % n;
    const int2 pr = pairs[t / nn];
    const int x = pr.x, y = pr.y;
    Vt[t] = mode == 0 ? I[(x + (i64)nc * b) + ld * (y + (i64)nc * a)] : I[(b + (i64)n * x) + ld * (a + (i64)n * y)];
}
__global__ void k_pm2_denom(const double *R, const int2 *pairs, i64 npairs, int mode, const double *w, int nnew,
                            const double *eo, double *part, QDArgs q) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    double E = 0.0, P = 0.0, qsc[QM], qpc[QM];
#pragma unroll
    for (int j = 0; j < QM; j++) { qsc[j] = 0.0; qpc[j] = 0.0; }
    if (t < npairs * nnew) {
        const i64 pi = t % npairs;
        const int mu = (int)(t / npairs);
        const int2 pr = pairs[pi];
        const double acc = R[t];
        const double deno = mode == 0 ? w[mu] - eo[pr.x] - eo[pr.y] : w[mu] + eo[pr.x] + eo[pr.y];
        const double contri = acc * acc / deno;
        E = -contri;
        P = contri / deno;
#pragma unroll
        for (int j = 0; j < QM; j++)
            if (qd_on(q, j)) qpc[j] = -q.a[j][t] * acc / deno;
    }
    block_store4(0.0, 0.0, E, P, part);
    if (q.M) block_storeQ(qsc, qpc, q);
}
// diagonal pairs (i=j or r=s) with the n(n+1)/2 metric
__global__ void k_pm2p_pc(const double *I, i64 ld, int next, int mode, int nc, int n, const double *spc,
                          const double *w, int nnew, const double *eo, double *part, QDArgs q) {
    // QD: q.a[j] = s(ab,mu) of pair (i,j) over upper pairs
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    double E = 0.0, P = 0.0, qsc[QM], qpc[QM];
#pragma unroll
    for (int j = 0; j < QM; j++) { qsc[j] = 0.0; qpc[j] = 0.0; }
    const int m = n * (n + 1) / 2;
    if (t < (i64)next * nnew) {
        const int mu = t % nnew, x = t / nnew;
        double acc = 0.0, accj[QM];
#pragma unroll
        for (int j = 0; j < QM; j++) accj[j] = 0.0;
        int ab = 0;
        for (int a = 0; a < n; a++)
            for (int b = a; b < n; b++, ab++) {
                double v;
                if (mode == 0) {   // (i b | i a), doubled off the diagonal (qdnevpt.F:3018-3047)
                    v = I[(x + (i64)nc * b) + ld * (x + (i64)nc * a)];
                    if (a != b) v *= 2.0;
                } else {           // (r a | r b), doubled off the diagonal (qdnevpt.F:2760-2779)
                    v = I[(a + (i64)n * x) + ld * (b + (i64)n * x)];
                    if (a != b) v *= 2.0;
                }
                acc += v * spc[ab + (i64)m * mu];
#pragma unroll
                for (int j = 0; j < QM; j++)
                    if (qd_on(q, j)) accj[j] += v * q.a[j][ab + (i64)m * mu];
            }
        double deno = mode == 0 ? w[mu] - 2.0 * eo[x] : w[mu] + 2.0 * eo[x];
        double contri = acc * acc / deno;
        E = -contri;
        P = contri / deno;
#pragma unroll
        for (int j = 0; j < QM; j++)
            if (qd_on(q, j)) qpc[j] = -accj[j] * acc / deno;
    }
    block_store4(0.0, 0.0, E, P, part);
    if (q.M) block_storeQ(qsc, qpc, q);
}

// metric / Koopmans matrices of the (+-2) PC solves: fillcm2 and fillcm2p (qdnevpt.F:4156-4193)
__global__ void k_fillcm2(const double *X, int n, double *S) {
    const i64 nn = (i64)n * n;
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nn * nn) return;
    i64 ij = e % nn, kl = e / nn;
    int i = ij / n, j = ij % n, k = kl / n, l = kl % n;
    S[e] = X[i + n * (j + (i64)n * (k + (i64)n * l))];
}
__global__ void k_fillcm2p(const double *X, int n, const int2 *up, double *S) {
    const int m = n * (n + 1) / 2;
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (i64)m * m) return;
    int p = e % m, q = e / m;
    int i = up[p].x, j = up[p].y, k = up[q].x, l = up[q].y;
    S[e] = X[i + n * (j + (i64)n * (k + (i64)n * l))] + X[i + n * (j + (i64)n * (l + (i64)n * k))];
}

// ------------------------------------------------------------------ class (0)' (qdnevpt.F:4326-4742, 4771-4999)
// X1 = ee2(b',a',a,b); X2 = ee2tr(b',b,a',a) + d(a,b) da(b',a'); G1 = gk0pa(b',a',a,b); G2 = gk0pd(b',a',a,b)
// all stored at index A + n*(B + n*(a + n*b)) with (A,B) = (a',b')
__global__ void k_v0p_prep(Sl S, const double *ga, const double *gd, double *X1, double *X2, double *G1, double *G2) {
    const int n = S.n;
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (i64)n * n * n * n) return;
    int A = e % n, B = (e / n) % n, a = (e / n / n) % n, b = e / n / n / n;
    X1[e] = s_ee2(S, B, A, a, b);
    X2[e] = s_ee2tr(S, B, b, A, a) + (a == b ? DA(S, B, A) : 0.0);
    const i64 g = B + (i64)n * (A + (i64)n * (a + (i64)n * b));
    G1[e] = ga[g];
    G2[e] = gd[g];
}
// fillc0p (qdnevpt.F:4195-4264): 2n^2 metric S and Koopmans matrix K, index ab = a*n+b
__global__ void k_fillc0p(Sl S, const double *ga, const double *gd, double *Sm, double *Km) {
    const int n = S.n, nn = n * n, N = 2 * nn;
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (i64)N * N) return;
    int row = e % N, col = e / N;
    int br = row / nn, bc = col / nn, ab = row % nn, abp = col % nn;
    int a = ab / n, b = ab % n, ap = abp / n, bp = abp % n;
    const double gav = ga[b + (i64)n * (a + (i64)n * (ap + (i64)n * bp))];
    double s, k;
    if (br == 0 && bc == 0) { k = gav; s = 2.0 * s_ee2(S, bp, ap, a, b); }
    else if (br != bc) { k = -0.5 * gav; s = -s_ee2(S, bp, ap, a, b); }
    else {
        k = gd[b + (i64)n * (a + (i64)n * (ap + (i64)n * bp))];
        s = -TA(S, bp, a, b, ap);
        if (a == ap) s += 2.0 * DA(S, bp, b);
    }
    Sm[e] = s;
    Km[e] = k;
}
// e20pcont's V[ab,a'b'] = ee2(b',a',a,b) and V2[ab,a'b'] = -daa(a',b,a,b') + 2 d(a',a) da(b,b')
__global__ void k_v0p_vmats(Sl S, double *V, double *V2) {
    const int n = S.n, nn = n * n;
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (i64)nn * nn) return;
    int row = e % nn, col = e / nn;
    int a = row / n, b = row % n, ap = col / n, bp = col % n;
    V[e] = s_ee2(S, bp, ap, a, b);
    double v2 = -TA(S, ap, b, a, bp);
    if (ap == a) v2 += 2.0 * DA(S, b, bp);
    V2[e] = v2;
}
// The (0)' sums as GEMMs over all (i, r) pairs p = i + nc r.  DJK holds, pair-major,
// dj(ab) = (i r | a b) and dk(ab) = (i a | r b): row p = [dj | dk], with ab = a + n b (conv 0, the SC
// sums) or a n + b (conv 1, the PC sums).
__global__ void k_v0p_gather(const double *Icvaa, i64 ldj, const double *Icaav, i64 ldk, int nc, int n, i64 npairs,
                             int conv, double *DJK) {
    const int nn = n * n;
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs * 2 * nn) return;
    const int k = (int)(t % (2 * nn)), e = k % nn;
    const i64 pp = t / (2 * nn);
    const int i = (int)(pp % nc), r = (int)(pp / nc);
    const int a = conv ? e / n : e % n, b = conv ? e % n : e / n;
    DJK[t] = k < nn ? Icvaa[(i + (i64)nc * r) + ldj * (a + (i64)n * b)] : Icaav[(i + (i64)nc * a) + ldk * (b + (i64)n * r)];
}
// SC: W (2 npairs x ncw, row 2p: dj^T [X1 | X2 | G1 | G2 | A_j | B_j ...], row 2p+1: dk^T [...], ncw = nn (4 +
// 2 partners)); one wavefront per pair forms e20pcont's SC quadratic forms (qdnevpt.F:4436-4611), the
// denominator and, for QD,
// the couplings with A_j = q.a[j], B_j = q.b[j] (the partners' X1, X2) and q.c[j]
__global__ void k_v0p_scdot(const double *W, int ncw, const double *DJK, int n, i64 npairs, int nc, const double *da,
                            const double *gf, const double *heff, const double *ec, const double *ev, double *part,
                            QDArgs q) {
    const int lane = threadIdx.x & 63, nn = n * n;
    const i64 pp = (i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    double E = 0.0, P = 0.0, qsc[QM], qpc[QM];
#pragma unroll
    for (int j = 0; j < QM; j++) { qsc[j] = 0.0; qpc[j] = 0.0; }
    if (pp < npairs) {
        const double *dj = DJK + pp * 2 * nn, *dk = dj + nn;
        const double *wj = W + pp * 2 * ncw, *wk = wj + ncw;      // blocks X1 | X2 | G1 | G2 | A_j | B_j ...
        double s1 = 0.0, s2 = 0.0, dum = 0.0, dumk = 0.0, dnj[QM];
#pragma unroll
        for (int j = 0; j < QM; j++) dnj[j] = 0.0;
        const double hh = heff[pp];
        for (int e = lane; e < nn; e += 64) {
            const double dje = dj[e], dke = dk[e];
            s1 += 2.0 * wj[e] * dje - wj[e] * dke - wk[e] * dje + wk[nn + e] * dke;
            s2 += wj[2 * nn + e] * dje - 0.5 * wj[2 * nn + e] * dke - 0.5 * wk[2 * nn + e] * dje + wk[3 * nn + e] * dke;
            const int a = e % n, b = e / n;
            const double d1 = 2.0 * dje - dke;
            dum += da[b + n * a] * d1;
            dumk += gf[a + n * b] * d1 * 0.5;
            int blk = 4;
#pragma unroll
            for (int j = 0; j < QM; j++)
                if (qd_on(q, j)) {
                    const i64 oa = (i64)nn * blk + e, ob = oa + nn;
                    dnj[j] += 2.0 * wj[oa] * dje - wj[oa] * dke - wk[oa] * dje + wk[ob] * dke + hh * q.c[j][e] * d1;
                    blk += 2;
                }
        }
        s1 = wsum(s1); s2 = wsum(s2); dum = wsum(dum); dumk = wsum(dumk);
#pragma unroll
        for (int j = 0; j < QM; j++)
            if (qd_on(q, j)) dnj[j] = wsum(dnj[j]);
        if (lane == 0) {
            const int i = (int)(pp % nc), r = (int)(pp / nc);
            const double dn = s1 + 2.0 * hh * dum + 2.0 * hh * hh, dko = s2 + hh * dumk;
            if (fabs(dn) > 1.0e-14) {
                const double den = ec[i] - ev[r] - dko / dn;
                E = dn / den;
                P = E / den;
#pragma unroll
                for (int j = 0; j < QM; j++)
                    if (qd_on(q, j)) qsc[j] = dnj[j] / den;
            }
        }
    }
    if (lane != 0) {
        E = 0.0; P = 0.0;
#pragma unroll
        for (int j = 0; j < QM; j++) qsc[j] = 0.0;
    }
    block_store4(E, P, 0.0, 0.0, part);
    if (q.M) block_storeQ(qsc, qpc, q);
}
// PC (qdnevpt.F:4892-4986): CM (2nn x nnew) = [2 sm - sp ; spp - sm], so that CO = CM^T [dj | dk] is
// r_mu = sum_ab (2 dj - dk) sm + (-dj sp + dk spp), with sm = V coef[:nn], sp = V coef[nn:], spp = V2 coef[nn:]
// (the QD partners' s, s', s'' give the next nnew-column blocks of CM the same way)
__global__ void k_v0p_cm(const double *sm, const double *sp, const double *spp, int nn, int nnew, double *CM) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (i64)2 * nn * nnew) return;
    const int k = (int)(t % (2 * nn));
    const i64 mu = t / (2 * nn), o = (k % nn) + (i64)nn * mu;
    CM[t] = k < nn ? 2.0 * sm[o] - sp[o] : spp[o] - sm[o];
}
// PC: CO (nco x npairs, nco = nnew (1 + partners)); one wavefront per pair over mu for the denominators of
// the PC sum (c_mu = r_mu + 2 h ro.coef[:nn] - h ro.coef[nn:], cir = c_mu / (w_mu + e_r - e_i)) and, for
// QD, the couplings -r_j cir with r_j from the partner's block of CO, q.d[j] and q.e[j]
__global__ void k_v0p_pcden(const double *CO, int nnew, int nco, i64 npairs, int nc, const double *dc1,
                            const double *dc2, const double *w, const double *heff, const double *ec, const double *ev,
                            double *part, const double *vx, double *rcv, QDArgs q) {
    const int lane = threadIdx.x & 63;
    const i64 pp = (i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    double E = 0.0, P = 0.0, racc = 0.0, qsc[QM], qpc[QM];
#pragma unroll
    for (int j = 0; j < QM; j++) { qsc[j] = 0.0; qpc[j] = 0.0; }
    if (pp < npairs) {
        const int i = (int)(pp % nc), r = (int)(pp / nc);
        const double hh = heff[pp], sh = ev[r] - ec[i];
        const double *co = CO + pp * nco;
        for (int mu = lane; mu < nnew; mu += 64) {
            const double c = co[mu] + 2.0 * hh * dc1[mu] - hh * dc2[mu];
            const double cir = c / (w[mu] + sh);
            E -= c * cir;
            P += cir * cir;
            if (rcv) racc += cir * vx[mu];
            int blk = 1;
#pragma unroll
            for (int j = 0; j < QM; j++)
                if (qd_on(q, j)) {
                    const double rj = co[(i64)nnew * blk + mu] + 2.0 * hh * q.d[j][mu] - hh * q.e[j][mu];
                    qpc[j] -= rj * cir;
                    blk++;
                }
        }
        if (rcv) {
            racc = wsum(racc);
            if (lane == 0) rcv[pp] = -racc;
        }
    }
    block_store4(0.0, 0.0, E, P, part);
    if (q.M) block_storeQ(qsc, qpc, q);
}
// .NATORB, class (0)' (qdnevpt.F:4854-4861): vx[mu] = sum_ab da(a,b) (2 coef[ab, mu] - coef[n^2 + ab, mu]),
// ab = a*n+b: the overlap of PC function mu with E_ri Psi_0 (per unit of the external pair)
__global__ void k_rho_vaux0p(const double *da, int n, const double *coef, int ldc, int nnew, double *vx) {
    const int mu = blockIdx.x * blockDim.x + threadIdx.x;
    if (mu >= nnew) return;
    const int nn = n * n;
    double s = 0.0;
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) {
            const i64 o = a * n + b + (i64)ldc * mu;
            s += da[a + n * b] * (2.0 * coef[o] - coef[nn + o]);
        }
    vx[mu] = s;
}
// dc[mu] = sum_ab da(b,a) coef[off + ab, mu], ab = a*n+b
__global__ void k_dacoef(const double *da, int n, const double *coef, int ldc, int off, int nnew, double *dc) {
    int mu = blockIdx.x * blockDim.x + threadIdx.x;
    if (mu >= nnew) return;
    double s = 0.0;
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) s += da[b + n * a] * coef[off + a * n + b + (i64)ldc * mu];
    dc[mu] = s;
}

// ------------------------------------------------------------------ classes (+1)' / (-1)' (qdnevpt.F:5437-6445, 5001-5435)
// index order abc = a*n^2 + b*n + c (the reference's iabc)
// Spc[abc, A'B'C'] = X3(c',a',b',b,a,c) with X3 = eee2 (minus) or eee2t (plus); Kpc[abc, A'B'C'] = amat(a',b',c',a,b,c)
__global__ void k_fill_pm1p_s(Sl S, int plus, double *Spc) {
    const int n = S.n;
    const i64 N = (i64)n * n * n;
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * N) return;
    i64 row = e % N, col = e / N;
    int a = row / (n * n), b = (row / n) % n, c = row % n;
    int ap = col / (n * n), bp = (col / n) % n, cp = col % n;
    Spc[e] = plus ? s_eee2t(S, cp, ap, bp, b, a, c) : s_eee2(S, cp, ap, bp, b, a, c);
}
__global__ void k_fill_pm1p_k(int n, const double *amat, double *Kpc) {
    const i64 N = (i64)n * n * n;
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * N) return;
    i64 row = e % N, col = e / N;
    int a = row / (n * n), b = (row / n) % n, c = row % n;
    int ap = col / (n * n), bp = (col / n) % n, cp = col % n;
    Kpc[e] = amat[ap + (i64)n * (bp + (i64)n * (cp + (i64)n * (a + (i64)n * (b + (i64)n * c))))];
}
// Q2[A'B'C', a] = X2(c',a',b',a) (ee2 or ee2tr); BC[A'B'C', a] = bmat(a',b',c',a) + cmat(a,a',b',c')
__global__ void k_fill_pm1p_q(Sl S, int plus, const double *bmat, const double *cmat, double *Q2, double *BC) {
    const int n = S.n;
    const i64 N = (i64)n * n * n;
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * n) return;
    i64 row = e % N;
    int a = e / N;
    int ap = row / (n * n), bp = (row / n) % n, cp = row % n;
    Q2[e] = plus ? s_ee2tr(S, cp, ap, bp, a) : s_ee2(S, cp, ap, bp, a);
    BC[e] = bmat[ap + (i64)n * (bp + (i64)n * (cp + (i64)n * a))] + cmat[a + (i64)n * (ap + (i64)n * (bp + (i64)n * cp))];
}
// Vmat[abc, x] = dint1(a,c,b): (a c | b r) from I_aaav (minus) or (a c | b i) from I_aaca (plus)
__global__ void k_fill_pm1p_v(const double *I, int plus, int n, int nc, int next, double *V) {
    const i64 N = (i64)n * n * n, ld = (i64)n * n;
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * next) return;
    i64 row = e % N;
    int x = e / N;
    int a = row / (n * n), b = (row / n) % n, c = row % n;
    V[e] = plus ? I[(a + (i64)n * c) + ld * (x + (i64)nc * b)] : I[(a + (i64)n * c) + ld * (b + (i64)n * x)];
}
// .NATORB, classes (+1)' / (-1)': T[abc + N3 d] = ee2tl(d,b,a,c) (qdnevpt.F:5113-5130) or ee2(d,b,a,c)
// (:5326-5341), so that vaux(d, mu) = (coef^T T)[mu, d]
__global__ void k_fill_rho_t(Sl S, int plus, double *T) {
    const int n = S.n;
    const i64 N = (i64)n * n * n;
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * n) return;
    const i64 row = e % N;
    const int d = (int)(e / N);
    const int a = row / (n * n), b = (row / n) % n, c = row % n;
    T[e] = plus ? s_ee2tl(S, d, b, a, c) : s_ee2(S, d, b, a, c);
}
// The .NATORB density matrix of one state in the canonical orbitals (qdnevpt.F:1247-1296): core block
// 2, active block the 1-RDM, the first-order core-virtual, active-virtual and core-active blocks
// (zero when absent) mirrored, virtual block 0; and U = diag(C_core, 1, C_virt) for the back-transform.
__global__ void k_rho_assemble(int nc, int n, int nv, const double *da, const double *rcv, const double *rca,
                               const double *rav, double *R, double *U, const double *Cc, const double *Cv) {
    const int norb = nc + n + nv;
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (i64)norb * norb) return;
    const int p = e % norb, q = e / norb;
    auto cls = [&](int o) { return o < nc ? 0 : o < nc + n ? 1 : 2; };
    const int cp = cls(p), cq = cls(q);
    double v = 0.0, u = 0.0;
    if (cp == 0 && cq == 0) { v = p == q ? 2.0 : 0.0; u = Cc[p + (i64)nc * q]; }
    else if (cp == 1 && cq == 1) { v = da[(p - nc) + n * (q - nc)]; u = p == q ? 1.0 : 0.0; }
    else if (cp == 2 && cq == 2) { u = Cv[(p - nc - n) + (i64)nv * (q - nc - n)]; }
    else if (cp == 0 && cq == 2) v = rcv ? rcv[p + (i64)nc * (q - nc - n)] : 0.0;
    else if (cp == 2 && cq == 0) v = rcv ? rcv[q + (i64)nc * (p - nc - n)] : 0.0;
    else if (cp == 1 && cq == 2) v = rav ? rav[(p - nc) + (i64)n * (q - nc - n)] : 0.0;
    else if (cp == 2 && cq == 1) v = rav ? rav[(q - nc) + (i64)n * (p - nc - n)] : 0.0;
    else if (cp == 0 && cq == 1) v = rca ? rca[p + (i64)nc * (q - nc)] : 0.0;
    else v = rca ? rca[q + (i64)nc * (p - nc)] : 0.0;      // cp == 1, cq == 0
    R[e] = v;
    U[e] = u;
}
// SC sums per external orbital x
__global__ void k_pm1p_sc(int plus, int n, int next, const double *H, const double *W1, const double *W2,
                          const double *V, const double *Z1, const double *Z2, const double *Y1, const double *Y2,
                          const double *eo, double *part, QDArgs q) {
    // QD: q.a[j] = R1_j H (n x next), q.b[j] = Q2_j H and q.c[j] = M_j V (n^3 x next) of pair (i,j)
    __shared__ double sh[16];
    const int x = blockIdx.x;
    const i64 N = (i64)n * n * n;
    double dm = 0.0, dk = 0.0, dmj[QM];
#pragma unroll
    for (int j = 0; j < QM; j++) dmj[j] = 0.0;
    for (i64 e = threadIdx.x; e < N; e += blockDim.x) {
        const double v = V[e + N * x];
        dm += 2.0 * v * Z1[e + N * x] + Y1[e + N * x] * v;
        dk += v * Z2[e + N * x] + Y2[e + N * x] * v;
#pragma unroll
        for (int j = 0; j < QM; j++)
            if (qd_on(q, j)) dmj[j] += v * q.b[j][e + N * x] + q.c[j][e + N * x] * v;
    }
    for (int a = threadIdx.x; a < n; a += blockDim.x) {
        const double h = H[a + (i64)n * x];
        dm += h * W1[a + (i64)n * x];
        dk += h * W2[a + (i64)n * x];
#pragma unroll
        for (int j = 0; j < QM; j++)
            if (qd_on(q, j)) dmj[j] += h * q.a[j][a + (i64)n * x];
    }
    dm = block_sum(dm, sh);
    dk = block_sum(dk, sh);
#pragma unroll
    for (int j = 0; j < QM; j++)
        if (qd_on(q, j)) dmj[j] = block_sum(dmj[j], sh);
    double E = 0.0, P = 0.0, qsc[QM], qpc[QM];
#pragma unroll
    for (int j = 0; j < QM; j++) { qsc[j] = 0.0; qpc[j] = 0.0; }
    if (threadIdx.x == 0 && dm != 0.0) {
        double den = plus ? eo[x] - dk / dm : -dk / dm - eo[x];
        double co = dm / den;
        E = co;
        P = co / den;
#pragma unroll
        for (int j = 0; j < QM; j++)
            if (qd_on(q, j)) qsc[j] = dmj[j] / den;
    }
    block_store4(E, P, 0.0, 0.0, part);
    if (q.M) block_storeQ(qsc, qpc, q);
}
// (+-1)' PC sums of one external orbital x per thread: (T + s_x) u = z_x with s_x = -e_x for (+1)'
// (denominator w_mu - e_i) and +e_x for (-1)' (w_mu + e_r); E -= z.u, norm += u.u, and for QD the
// coupling -z_j.u with z_j = q.a[j] (same layout as Z)
// .NATORB (rho non-null): rho[x rsx + d rsd] = -u.vt_d for the nr columns vt_d = Vt + ldz d, the
// projections of the density vectors (qdnevpt.F:5187-5192, 5407-5412)
__global__ void k_pm1p_shift(int plus, int nnew, int next, const double *d, const double *e, const double *Z, i64 ldz,
                             const double *eo, double *rows, double *part, QDArgs q, const double *Vt, int nr,
                             double *rho, i64 rsx, i64 rsd) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    double E = 0.0, P = 0.0, qsc[QM], qpc[QM], racc[16];
#pragma unroll
    for (int j = 0; j < QM; j++) { qsc[j] = 0.0; qpc[j] = 0.0; }
#pragma unroll
    for (int k = 0; k < 16; k++) racc[k] = 0.0;
    if (x < next) {
        const double sh = plus ? -eo[x] : eo[x];
        const double *z = Z + ldz * x;
        double ez = 0.0, uu = 0.0;
        trd_shift_solve(nnew, d, e, sh, z, 1, rows + x, next, [&](int i, double u) {
            ez += z[i] * u;
            uu += u * u;
#pragma unroll
            for (int j = 0; j < QM; j++)
                if (qd_on(q, j)) qpc[j] -= q.a[j][ldz * x + i] * u;
            if (rho) {
#pragma unroll
                for (int k = 0; k < 16; k++)
                    if (k < nr) racc[k] += u * Vt[i + ldz * k];
            }
        });
        E = -ez;
        P = uu;
        if (rho) {
#pragma unroll
            for (int k = 0; k < 16; k++)
                if (k < nr) rho[x * rsx + k * rsd] = -racc[k];
        }
    }
    block_store4(0.0, 0.0, E, P, part);
    if (q.M) block_storeQ(qsc, qpc, q);
}

// The same sums on the band form K' = Q B Q^T (nv2_band.hip): one workgroup per external orbital x
// solves (B + s_x) u = z_x by Gaussian elimination with partial pivoting (k_pm1p_bandr below).  The
// active part of step j is the window of rows j..j+BAND_B and columns j..j+2*BAND_B (a pivot row from
// below the diagonal carries the fill up to 2*BAND_B columns right of it).  Global data moves in
// blocks of BSTEP rows (band rows entering the window, rows of U), loaded one block ahead into
// registers and staged through LDS, so that no step waits for memory.
// E -= z.u, norm += u.u, and for QD the coupling -z_j.u with z_j = q.a[j] (same layout as Z).
#define PBR (BAND_B + 1)
#define UW (BAND_W + 1)
#define BSTEP 8
// LDS ordering inside one wavefront of a larger workgroup (in-order LDS; no workgroup barrier)
#define WAVE_SYNC() __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory")
// one class of the launch: band rows, right-hand sides z_x at Z + N*x, orbital energies, U rows and
// solutions (N * UR and N per orbital), QD partners (q.a[j]: same layout as Z)
struct BandCls {
    const double *Br, *Z, *eo;
    double *U, *y;
    int plus, next;
    QDArgs q;
};
// Blocks 0 .. c0.next-1 solve for the orbitals of class c0, the following c1.next those of class c1.
// The window lives in registers: thread g of three wavefronts owns column slot g
// (g = 0..2*BAND_B) or, for g = BAND_W, the right-hand side, and keeps that column for the BAND_B+1
// row slots of the window in registers.  Rows have no order in the window: partial pivoting picks the
// largest |entry| of column j over all rows present (rows j..j+BAND_B), the pivot row's slot takes row
// j+BAND_B+1 after the step, and column j's slot (its owner zeroes it) becomes column j+2*BAND_B+1.  A
// step moves three small vectors through LDS (column j, the multipliers, the pivot) instead of the
// whole window.  U rows are stored in slot order: value of column c at slot c % BAND_W (the pivot
// column's own slot 0), then the eliminated right-hand side and the reciprocal pivot; the back
// substitution keeps u in a ring by slot, so it reads U rows as they are.
#define RW_T (3 * 64)
#define UR (BAND_W + 3)
__device__ __forceinline__ double rw_get(const double (&w)[PBR], int s) {
    switch (s) {
        case 0: return w[0];
        case 1: return w[1];
        case 2: return w[2];
        case 3: return w[3];
        case 4: return w[4];
        case 5: return w[5];
        case 6: return w[6];
        case 7: return w[7];
        case 8: return w[8];
        case 9: return w[9];
        case 10: return w[10];
        case 11: return w[11];
        case 12: return w[12];
        case 13: return w[13];
        case 14: return w[14];
        case 15: return w[15];
        case 16: return w[16];
        case 17: return w[17];
        case 18: return w[18];
        case 19: return w[19];
        case 20: return w[20];
        case 21: return w[21];
        case 22: return w[22];
        case 23: return w[23];
        case 24: return w[24];
        case 25: return w[25];
        case 26: return w[26];
        case 27: return w[27];
        case 28: return w[28];
        case 29: return w[29];
        case 30: return w[30];
        case 31: return w[31];
        case 32: return w[32];
        case 33: return w[33];
        case 34: return w[34];
        case 35: return w[35];
        case 36: return w[36];
        case 37: return w[37];
        case 38: return w[38];
        case 39: return w[39];
        case 40: return w[40];
        case 41: return w[41];
        case 42: return w[42];
        case 43: return w[43];
        case 44: return w[44];
        case 45: return w[45];
        case 46: return w[46];
        case 47: return w[47];
        case 48: return w[48];
        case 49: return w[49];
        case 50: return w[50];
        case 51: return w[51];
        case 52: return w[52];
        case 53: return w[53];
        case 54: return w[54];
        case 55: return w[55];
        case 56: return w[56];
        case 57: return w[57];
        case 58: return w[58];
        case 59: return w[59];
        case 60: return w[60];
        case 61: return w[61];
        case 62: return w[62];
        case 63: return w[63];
        case 64: return w[64];
    }
    return 0.0;
}
__device__ __forceinline__ void rw_set(double (&w)[PBR], int s, double v) {
    switch (s) {
        case 0: w[0] = v; break;
        case 1: w[1] = v; break;
        case 2: w[2] = v; break;
        case 3: w[3] = v; break;
        case 4: w[4] = v; break;
        case 5: w[5] = v; break;
        case 6: w[6] = v; break;
        case 7: w[7] = v; break;
        case 8: w[8] = v; break;
        case 9: w[9] = v; break;
        case 10: w[10] = v; break;
        case 11: w[11] = v; break;
        case 12: w[12] = v; break;
        case 13: w[13] = v; break;
        case 14: w[14] = v; break;
        case 15: w[15] = v; break;
        case 16: w[16] = v; break;
        case 17: w[17] = v; break;
        case 18: w[18] = v; break;
        case 19: w[19] = v; break;
        case 20: w[20] = v; break;
        case 21: w[21] = v; break;
        case 22: w[22] = v; break;
        case 23: w[23] = v; break;
        case 24: w[24] = v; break;
        case 25: w[25] = v; break;
        case 26: w[26] = v; break;
        case 27: w[27] = v; break;
        case 28: w[28] = v; break;
        case 29: w[29] = v; break;
        case 30: w[30] = v; break;
        case 31: w[31] = v; break;
        case 32: w[32] = v; break;
        case 33: w[33] = v; break;
        case 34: w[34] = v; break;
        case 35: w[35] = v; break;
        case 36: w[36] = v; break;
        case 37: w[37] = v; break;
        case 38: w[38] = v; break;
        case 39: w[39] = v; break;
        case 40: w[40] = v; break;
        case 41: w[41] = v; break;
        case 42: w[42] = v; break;
        case 43: w[43] = v; break;
        case 44: w[44] = v; break;
        case 45: w[45] = v; break;
        case 46: w[46] = v; break;
        case 47: w[47] = v; break;
        case 48: w[48] = v; break;
        case 49: w[49] = v; break;
        case 50: w[50] = v; break;
        case 51: w[51] = v; break;
        case 52: w[52] = v; break;
        case 53: w[53] = v; break;
        case 54: w[54] = v; break;
        case 55: w[55] = v; break;
        case 56: w[56] = v; break;
        case 57: w[57] = v; break;
        case 58: w[58] = v; break;
        case 59: w[59] = v; break;
        case 60: w[60] = v; break;
        case 61: w[61] = v; break;
        case 62: w[62] = v; break;
        case 63: w[63] = v; break;
        case 64: w[64] = v; break;
    }
}
__global__ __launch_bounds__(RW_T) void k_pm1p_bandr(int N, BandCls c0, BandCls c1, double *part) {
    __shared__ __align__(16) double stg[BSTEP][UR];    // band rows entering (Br order) + z; then rows of U
    __shared__ __align__(16) double colj[PBR + 1];
    __shared__ __align__(16) double ltab[PBR + 1];
    __shared__ double ys[BAND_W];
    __shared__ int sp_s;
    __shared__ double rp_s;
    const bool second = (int)blockIdx.x >= c0.next;
    const int x = second ? (int)blockIdx.x - c0.next : (int)blockIdx.x;
    const int g = threadIdx.x, lane = g & 63, wv = g >> 6;
    const int plus = second ? c1.plus : c0.plus;
    const double *Br = second ? c1.Br : c0.Br;
    const double *eo = second ? c1.eo : c0.eo;
    const QDArgs &q = c0.q;
    const double *qa[QM];
#pragma unroll
    for (int k = 0; k < QM; k++) qa[k] = second ? c1.q.a[k] : c0.q.a[k];
    const i64 ldz = N;
    const double sh = plus ? -eo[x] : eo[x];
    const double *z = (second ? c1.Z : c0.Z) + ldz * x;
    double *U = (second ? c1.U : c0.U) + (i64)x * N * UR;
    double *yx = (second ? c1.y : c0.y) + (i64)x * N;
    const bool col = g < BAND_W, zl = g == BAND_W;
    double w[PBR];
#pragma unroll
    for (int s = 0; s < PBR; s++) {          // slot s = row s; slot g = column g
        double v = 0.0;
        if (col) {
            const int d = g - s;
            if (s < N && g < N && d >= -BAND_B && d <= BAND_B) v = Br[(i64)s * BAND_W + d + BAND_B] + (d == 0 ? sh : 0.0);
        } else if (zl) {
            v = s < N ? z[s] : 0.0;
        }
        w[s] = v;
    }
    // block prefetch of the band rows entering next (Br order, plus z at UW - 1): BSTEP*UW values
    constexpr int PF = (BSTEP * UW + RW_T - 1) / RW_T;
    double pf[PF];
    auto fetch_rows = [&](int r0) {
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int e = g + RW_T * k, u = e / UW, t = e % UW, r = r0 + u;
            double v = 0.0;
            if (e < BSTEP * UW && r < N) v = t < BAND_W ? Br[(i64)r * BAND_W + t] : z[r];
            pf[k] = v;
        }
    };
    fetch_rows(PBR);
#pragma unroll 1
    for (int j0 = 0; j0 < N; j0 += BSTEP) {
        __syncthreads();                     // the previous block's stage is consumed
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int e = g + RW_T * k;
            if (e < BSTEP * UW) stg[e / UW][e % UW] = pf[k];
        }
        fetch_rows(j0 + PBR + BSTEP);
        const int jn = N - j0 < BSTEP ? N - j0 : BSTEP;
#pragma unroll 1
        for (int jj = 0; jj < jn; jj++) {
            const int j = j0 + jj, cs0 = j % BAND_W;
            if (g == cs0) {
#pragma unroll
                for (int s = 0; s < PBR; s++) colj[s] = w[s];
            }
            __syncthreads();                 // column j (and the stage) in LDS
            if (wv == 0) {
                // pivot: largest |entry| of column j over the row slots; the key is |entry| with its 7
                // lowest mantissa bits replaced by 127 - slot (ties to 2^-45 relative go to the lower slot)
                auto key = [](double v, int i) -> unsigned long long {
                    return ((unsigned long long)__double_as_longlong(fabs(v)) & ~0x7Full) | (unsigned long long)(127 - i);
                };
                const double cv = colj[lane], cv64 = colj[BAND_B];
                unsigned long long kk = key(cv, lane);
                if (lane == 0) { const unsigned long long k64 = key(cv64, BAND_B); if (k64 > kk) kk = k64; }
                kk = wmax_u64(kk);
                const int pv = 127 - (int)(kk & 0x7F);
                const double piv = colj[pv];
                const double rp = 1.0 / piv;
                ltab[lane] = cv * rp;
                if (lane == 0) { ltab[BAND_B] = cv64 * rp; sp_s = pv; rp_s = rp; }
            }
            __syncthreads();                 // multipliers and pivot slot
            const int sp = __builtin_amdgcn_readfirstlane(sp_s);
            const double u = rw_get(w, sp);  // this column's entry of the pivot row
            double *Uj = U + (i64)j * UR;
            if (col) Uj[g] = g == cs0 ? 0.0 : u;
            if (zl) Uj[BAND_W] = u;          // the eliminated right-hand side
            if (g == 0) Uj[BAND_W + 1] = rp_s;
#pragma unroll
            for (int s = 0; s < PBR - 1; s += 2) {
                const double2 l2 = *(const double2 *)&ltab[s];
                w[s] -= l2.x * u;
                w[s + 1] -= l2.y * u;
            }
            w[BAND_B] -= ltab[BAND_B] * u;
            if (g == cs0) {                  // column j leaves; its slot becomes column j + 2*BAND_B + 1
#pragma unroll
                for (int s = 0; s < PBR; s++) w[s] = 0.0;
            }
            // row j + BAND_B + 1 enters the pivot's slot: column c at slot g is c = j+1+((g-cs0-1) mod BAND_W)
            double v = 0.0;
            if (col) {
                const int t = g - cs0 - 1 < 0 ? g - cs0 - 1 + BAND_W : g - cs0 - 1;
                v = stg[jj][t] + (t == BAND_B && j + PBR < N ? sh : 0.0);
            } else if (zl) {
                v = stg[jj][BAND_W];
            }
            rw_set(w, sp, v);
        }
    }
    __syncthreads();
    // back substitution by wavefront 0: u_j = (zf_j - sum_g U[j][g] u(slot g)) * rp_j, u in a ring by slot
    double ez = 0.0, uu = 0.0, qsc[QM], qpc[QM];
#pragma unroll
    for (int k = 0; k < QM; k++) { qsc[k] = 0.0; qpc[k] = 0.0; }
    if (wv == 0) {
        for (int t = lane; t < BAND_W; t += 64) ys[t] = 0.0;
        double fa[BSTEP], fb[BSTEP], fc[BSTEP];
        auto fetch_u = [&](int jtop) {
#pragma unroll
            for (int uu2 = 0; uu2 < BSTEP; uu2++) {
                const bool in = jtop - uu2 >= 0;
                const double *row = U + (i64)(jtop - uu2) * UR;
                fa[uu2] = in ? row[lane] : 0.0;
                fb[uu2] = in ? row[lane + 64] : 0.0;
                fc[uu2] = in && lane < 3 ? row[2 * BAND_B + lane] : 0.0;
            }
        };
        fetch_u(N - 1);
        WAVE_SYNC();
#pragma unroll 1
        for (int jt = N - 1; jt >= 0; jt -= BSTEP) {
#pragma unroll
            for (int uu2 = 0; uu2 < BSTEP; uu2++) {
                stg[uu2][lane] = fa[uu2];
                stg[uu2][lane + 64] = fb[uu2];
                if (lane < 3) stg[uu2][2 * BAND_B + lane] = fc[uu2];
            }
            fetch_u(jt - BSTEP);
            WAVE_SYNC();
            const int jn = jt + 1 < BSTEP ? jt + 1 : BSTEP;
#pragma unroll 1
            for (int uu2 = 0; uu2 < jn; uu2++) {
                const int j = jt - uu2, cs = j % BAND_W;
                double sm = stg[uu2][lane] * ys[lane] + stg[uu2][lane + 64] * ys[lane + 64];
                if (lane == 0) sm += stg[uu2][2 * BAND_B] * ys[2 * BAND_B];
                sm = wsum_dpp(sm);
                if (lane == 0) {
                    const double y = (stg[uu2][BAND_W] - sm) * stg[uu2][BAND_W + 1];
                    ys[cs] = y;
                    yx[j] = y;
                }
                WAVE_SYNC();
            }
        }
    }
    __threadfence();
    __syncthreads();
    for (int i = g; i < N; i += RW_T) {
        const double uv = yx[i];
        ez += z[i] * uv;
        uu += uv * uv;
#pragma unroll
        for (int k = 0; k < QM; k++)
            if (qd_on(q, k)) qpc[k] -= qa[k][ldz * x + i] * uv;
    }
    block_store4(0.0, 0.0, -ez, uu, part);
    if (q.M) block_storeQ(qsc, qpc, q);
}

// ------------------------------------------------------------------ host side
struct Buf {
    void *p = nullptr;
    size_t cap = 0;
};

// independent PC eigen-solves run on worker streams (each with its own rocBLAS handle)
#define NW 2
struct Pm1pEarly;
struct Stage2 {
    S2Inputs in;
    bool no4 = false;              // .NO4RDM: no (+-1)' classes
    bool rho1 = false;             // .NATORB: first-order density matrix of each state
    std::vector<int> frozen;       // .FROZEN: canonical core orbitals that are never holes
    Pm1pEarly *early = nullptr;    // the (+-1)' chain when started ahead of stage2_state
    hipEvent_t evS = nullptr, evK = nullptr;   // its metric filled / its Koopmans matrix filled
    hipEvent_t evX = nullptr;                  // its PC right-hand sides built (band-form sums on worker 1)
    hipEvent_t evQt = nullptr;                 // Q^T applied to them (the held SC sums start after it)
    hipEvent_t evKr[2] = {nullptr, nullptr};   // around the K' reduction of the (+-1)' problems (timed)
    bool kr_recorded = false;
    int kr_N = 0, kr_nb = 0, kr_form = 0;
    rocblas_handle h;
    hipStream_t st;
    rocblas_handle wh[NW] = {};
    hipStream_t ws[NW] = {};
    hipEvent_t wev[NW + 1] = {};
    double *hpin = nullptr;        // pinned host copies of metric eigenvalues
    size_t hpin_cap = 0;
    int *hinfo = nullptr;          // pinned host copies of solver info words
    int *dinfo_w = nullptr;        // device info words, 2 per solve slot
    double *wE[NW] = {};           // syevd workspace per worker
    SyevbPlan *sbplan[3] = {};     // batched eigensolver call sites: worker-0 metrics, worker-0 K', Fock blocks
    SyevbPlan *plan(int k) {
        if (!sbplan[k]) sbplan[k] = syevb_plan_create();
        return sbplan[k];
    }
    hipStream_t xs = nullptr;      // second stream + handle for the "-" problem's dstedc of the n^3 metric
    rocblas_handle xh = nullptr;
    hipEvent_t xev[2] = {};
    int nc, n, nv, norb, L, M;
    std::map<std::string, Buf> bufs;
    int *dinfo = nullptr;
    double *part = nullptr;
    size_t part_cap = 0;
    double *acc = nullptr;         // 8 classes x 4
    double *accq = nullptr;        // 8 classes x 2*QM (QD off-diagonal sums)
    double *qpart = nullptr;
    size_t qpart_cap = 0;
    int2 *pairs_c = nullptr, *pairs_cs = nullptr, *pairs_v = nullptr, *pairs_vs = nullptr, *up = nullptr;
    int2 *pairs_cd = nullptr, *pairs_vd = nullptr;   // diagonal pairs (i, i), (r, r)
    i64 np_c, np_cs, np_v, np_vs;
    bool blocks_ready = false;

    // Waits for this evaluation's own streams only: other contexts on the device (concurrent
    // requests of the request loop) keep running.
    void sync_own() {
        (void)hipStreamSynchronize(st);
        for (int w = 0; w < NW; w++)
            if (ws[w]) (void)hipStreamSynchronize(ws[w]);
        if (xs) (void)hipStreamSynchronize(xs);
    }
    // Buffers are looked up and grown on the thread that drives the evaluation only; worker threads
    // use the pointers handed to them.
    double *get(const std::string &name, i64 count) {
        Buf &b = bufs[name];
        size_t bytes = sizeof(double) * (size_t)(count > 0 ? count : 1);
        if (b.cap < bytes) {
            if (b.p) { sync_own(); (void)hipFree(b.p); }
            b.p = nullptr;
            if (hipMalloc(&b.p, bytes) != hipSuccess) { b.cap = 0; return nullptr; }
            b.cap = bytes;
        }
        return (double *)b.p;
    }
};

#define GET(var, name, cnt) double *var = p->get(name, (cnt)); if (!var) { set_error("hipMalloc failed for %s", name); return -1; }

static int gemm_on(rocblas_handle h, hipStream_t st, bool tA, bool tB, int m, int nn, int k, double alpha,
                   const double *A, i64 lda, const double *B, i64 ldb, double beta, double *C, i64 ldc) {
    if (m <= 0 || nn <= 0) return 0;
    if (k <= 0) {
        if (beta == 0.0) NV2_HIP(hipMemset2DAsync(C, sizeof(double) * ldc, 0, sizeof(double) * m, nn, st));
        return 0;
    }
    RB(rocblas_dgemm(h, tA ? rocblas_operation_transpose : rocblas_operation_none,
                     tB ? rocblas_operation_transpose : rocblas_operation_none, m, nn, k, &alpha, A, (int)lda, B,
                     (int)ldb, &beta, C, (int)ldc));
    return 0;
}
__global__ void k_splitk_reduce(const double *P, int ns, int m, int nn, double alpha, double beta, double *C, i64 ldc) {
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (i64)m * nn) return;
    const i64 i = e % m, j = e / m;
    double acc = 0.0;
    for (int c = 0; c < ns; c++) acc += P[(i64)c * m * nn + e];
    const double old = beta == 0.0 ? 0.0 : beta * C[i + ldc * j];
    C[i + ldc * j] = old + alpha * acc;
}
// GEMMs with a small output and a long reduction (the Fock and heff contractions over the
// Cholesky index times an orbital index, k up to L*nc) run split over k as a strided batch,
// so enough workgroups share the reduction; the slices are summed by k_splitk_reduce.
static int gemm(Stage2 *p, bool tA, bool tB, int m, int nn, int k, double alpha, const double *A, i64 lda,
                const double *B, i64 ldb, double beta, double *C, i64 ldc) {
    if (m > 0 && nn > 0 && k >= 4096 && (i64)m * nn <= 512 * 512) {
        int ns = k / 1024;
        if (ns > 64) ns = 64;
        const int kc = k / ns;          // the last slice takes the remainder below
        const i64 mn = (i64)m * nn;
        double *P = p->get("splitk", mn * (ns + 1));
        if (!P) { set_error("hipMalloc failed for splitk"); return -1; }
        const double one = 1.0, zero = 0.0;
        const i64 sA = tA ? kc : kc * lda, sB = tB ? kc * ldb : kc;
        RB(rocblas_dgemm_strided_batched(p->h, tA ? rocblas_operation_transpose : rocblas_operation_none,
                                         tB ? rocblas_operation_transpose : rocblas_operation_none, m, nn, kc, &one, A,
                                         (int)lda, sA, B, (int)ldb, sB, &zero, P, m, mn, ns));
        int nsum = ns;
        const int rem = k - kc * ns;
        if (rem > 0) {
            TRYS(gemm_on(p->h, p->st, tA, tB, m, nn, rem, 1.0, A + ns * sA, lda, B + ns * sB, ldb, 0.0, P + ns * mn, m));
            nsum++;
        }
        hipLaunchKernelGGL(k_splitk_reduce, dim3(nb(mn, 256)), dim3(256), 0, p->st, P, nsum, m, nn, alpha, beta, C, ldc);
        NV2_LAUNCH_CHECK();
        return 0;
    }
    return gemm_on(p->h, p->st, tA, tB, m, nn, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
static int gemm_sb(Stage2 *p, int m, int nn, int k, const double *A, i64 lda, i64 sA, const double *B, i64 ldb,
                   i64 sB, double *C, i64 ldc, i64 sC, int batch) {
    if (m <= 0 || nn <= 0 || batch <= 0) return 0;
    const double one = 1.0, zero = 0.0;
    RB(rocblas_dgemm_strided_batched(p->h, rocblas_operation_none, rocblas_operation_none, m, nn, k, &one, A, (int)lda,
                                     sA, B, (int)ldb, sB, &zero, C, (int)ldc, sC, batch));
    return 0;
}
// NV2_SYEVB=0: the Fock blocks and the worker-0 classes take the per-matrix paths (rocSOLVER dsyevd,
// trd_reduce + dstedc) instead of the batched eigensolver
static bool syevb_on() {
    const char *e = getenv("NV2_SYEVB");
    return !(e && atoi(e) == 0);
}
// eigen-decomposition in place (ascending eigenvalues in w), reference triangle choice.
// NV2_EIGEN selects the rocSOLVER algorithm: "dc" (syevd, default), "jacobi" (syevj), "dcj" (syevdj)
static int eigen_algo() {
    static const int algo = [] {      // initialised once, thread-safely (several contexts run concurrently)
        const char *e = getenv("NV2_EIGEN");
        return (e && !strcmp(e, "jacobi")) ? 1 : (e && !strcmp(e, "dcj")) ? 2 : 0;
    }();
    return algo;
}
static int syevd(Stage2 *p, bool upper, int N, double *A, double *w) {
    if (N <= 0) return 0;
    GET(E, "syevd_E", N);
    const rocblas_fill uplo = upper ? rocblas_fill_upper : rocblas_fill_lower;
    const int algo = eigen_algo();
    if (algo == 1) {
        RB(rocsolver_dsyevj(p->h, rocblas_esort_ascending, rocblas_evect_original, uplo, N, A, N, 0.0, E, 100,
                            p->dinfo + 1, w, p->dinfo));
    } else if (algo == 2) {
        RB(rocsolver_dsyevdj(p->h, rocblas_evect_original, uplo, N, A, N, w, p->dinfo));
    } else {
        RB(rocsolver_dsyevd(p->h, rocblas_evect_original, uplo, N, A, N, w, E, p->dinfo));
    }
    int info = 0;
    NV2_HIP(hipMemcpyAsync(&info, p->dinfo, sizeof(int), hipMemcpyDeviceToHost, p->st));
    NV2_HIP(hipStreamSynchronize(p->st));
    if (info != 0) { set_error("syevd(N=%d) failed, info=%d", N, info); return -3; }
    return 0;
}
// the same for nb problems A_b = A + b*sA (ld N), eigenvalues ev + b*sD, on worker w without
// waiting; info words go to dinfo_w[idx + b] and are copied to hinfo[idx + b]
static int syevd_async(Stage2 *p, int w, int idx, bool upper, int N, double *A, i64 sA, double *ev, i64 sD, int nb) {
    if (N <= 0) return 0;
    double *E = p->wE[w];
    const rocblas_fill uplo = upper ? rocblas_fill_upper : rocblas_fill_lower;
    int *info = p->dinfo_w + idx;
    const int algo = eigen_algo();
    if (algo == 0 && nb > 1) {
        RB(rocsolver_dsyevd_strided_batched(p->wh[w], rocblas_evect_original, uplo, N, A, N, sA, ev, sD, E, N, info, nb));
    } else {
        for (int b = 0; b < nb; b++) {
            double *Ab = A + b * sA, *eb = ev + b * sD;
            if (algo == 1) {
                RB(rocsolver_dsyevj(p->wh[w], rocblas_esort_ascending, rocblas_evect_original, uplo, N, Ab, N, 0.0, E,
                                    100, p->dinfo_w + 62, eb, info + b));
            } else if (algo == 2) {
                RB(rocsolver_dsyevdj(p->wh[w], rocblas_evect_original, uplo, N, Ab, N, eb, info + b));
            } else {
                RB(rocsolver_dsyevd(p->wh[w], rocblas_evect_original, uplo, N, Ab, N, eb, E, info + b));
            }
        }
    }
    NV2_HIP(hipMemcpyAsync(p->hinfo + idx, info, sizeof(int) * nb, hipMemcpyDeviceToHost, p->ws[w]));
    return 0;
}
// Pads problem b of a batch (n_b x n_b used block of an m x m matrix, stride sA) to m x m
// with a diagonal value above every eigenvalue of every used block: twice the largest
// Gershgorin row sum plus one, so the padded block stays decoupled and sorts last while
// the matrix norm, and with it the solver's tolerances, stays that of the used blocks.
__global__ void k_pad_gersh(double *A, int ld, i64 sA, int nb, int n0, int n1, int m) {
    __shared__ double sh[256];
    double mx = 0.0;
    for (int b = 0; b < nb; b++) {
        const int nu = b == 0 ? n0 : n1;
        const double *Ab = A + b * sA;
        for (int i = threadIdx.x; i < nu; i += blockDim.x) {
            double r = 0.0;
            for (int j = 0; j < nu; j++) r += fabs(Ab[i + (i64)ld * j]);
            mx = fmax(mx, r);
        }
    }
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    const double pad = 2.0 * sh[0] + 1.0;
    for (int b = 0; b < nb; b++) {
        const int nu = b == 0 ? n0 : n1;
        double *Ab = A + b * sA;
        for (int i = nu + threadIdx.x; i < m; i += blockDim.x) Ab[i + (i64)ld * i] = pad;
    }
}
static double *parts(Stage2 *p, i64 nblocks) {
    size_t need = sizeof(double) * 4 * (size_t)nblocks;
    if (p->part_cap < need) {
        if (p->part) { (void)hipStreamSynchronize(p->st); (void)hipFree(p->part); }
        p->part = nullptr;
        if (hipMalloc(&p->part, need) != hipSuccess) { p->part_cap = 0; return nullptr; }
        p->part_cap = need;
    }
    return p->part;
}
static int reduce_into(Stage2 *p, i64 nblocks, int cls, i64 first = 0) {   // blocks first .. first+nblocks-1
    hipLaunchKernelGGL(k_reduce_parts, dim3(4), dim3(256), 0, p->st, p->part + 4 * first, nblocks, p->acc + 4 * cls);
    NV2_LAUNCH_CHECK();
    return 0;
}
// QD partial buffer for a launch of nblocks blocks (q.M == 0 leaves q disabled)
static int qd_prepare(Stage2 *p, QDArgs &q, i64 nblocks) {
    if (!q.M) return 0;
    size_t need = sizeof(double) * 2 * QM * (size_t)(nblocks > 0 ? nblocks : 1);
    if (p->qpart_cap < need) {
        if (p->qpart) { (void)hipStreamSynchronize(p->st); (void)hipFree(p->qpart); }
        p->qpart = nullptr;
        p->qpart_cap = 0;
        if (hipMalloc(&p->qpart, need) != hipSuccess) { set_error("hipMalloc(qpart) failed"); return -1; }
        p->qpart_cap = need;
    }
    q.qpart = p->qpart;
    return 0;
}
static int qd_reduce(Stage2 *p, const QDArgs &q, i64 nblocks, int cls, i64 first = 0) {
    if (!q.M) return 0;
    hipLaunchKernelGGL(k_reduce_q, dim3(2 * QM), dim3(256), 0, p->st, p->qpart + 2 * QM * first, nblocks,
                       p->accq + 2 * QM * cls);
    NV2_LAUNCH_CHECK();
    return 0;
}

// One or two PC metric problems of equal size N (qdnevpt.F:4744-4769 and e.g. :1974-2037),
// solved in two launches on worker stream w: the metric eigen-decompositions (one batched
// call), then, after the host has counted the kept eigenvalues of each, the transformed
// Koopmans matrices K' and their eigen-decompositions.  The second problem's K' is padded
// to the larger kept dimension with a decoupled diagonal block far above its spectrum, so
// both still go through one batched call; the padding never reaches the coefficients.
struct PcJob {
    std::string tag;
    int N = 0, nb = 1, w = 0, slot = 0;        // info words at 4*slot + {metric b, K' 2+b}
    bool trd = false;                          // K' only reduced (the (+-1)' sums solve with its T)
    bool chol = false;                         // metric factored as S = U^T U (every eigenvalue kept)
    bool band = false;                         // K' reduced to band form (nv2_band.hip) instead of T
    double *bandw = nullptr;                   // band_reduce workspace (trd jobs)
    double *trdw = nullptr;                    // K' reduction scratch; T and tau of problem b from trd_outputs(nmax)
    double *trdm = nullptr, *trde = nullptr;   // eig_trd: reduction scratch, off-diagonal copy for dstedc
    double *trdq = nullptr;                    // scratch of applying Q
    double *cholw = nullptr;                   // Cholesky scratch (trd jobs)
    double *sbw = nullptr;                     // batched-eigensolver scratch, nb problems (other jobs)
    double *S = nullptr, *K = nullptr;         // nb consecutive N x N matrices; S is overwritten
    double *w1 = nullptr, *hw1 = nullptr, *coef = nullptr, *w2 = nullptr;
    double *caux = nullptr, *Ks = nullptr, *T = nullptr, *Kp = nullptr;   // nb x N x N scratch each
    int nnew[2] = {0, 0}, nmax = 0;
    double minw[2] = {0.0, 0.0};
    i64 NN() const { return (i64)N * N; }
    double *Sb(int b) const { return S + b * NN(); }
    double *Kb(int b) const { return K + b * NN(); }
    double *coefb(int b) const { return coef + b * NN(); }
    double *w2b(int b) const { return w2 + (i64)b * N; }
};
// every buffer a job touches is allocated here, on the calling thread, before the workers start
static int pc_job_buffers(Stage2 *p, PcJob &j, size_t hoff) {
    const i64 N = j.N, NN = N * N * j.nb;
    GET(w1, (j.tag + "_w1").c_str(), N * j.nb); GET(w2, (j.tag + "_w2").c_str(), N * j.nb);
    GET(coef, (j.tag + "_coef").c_str(), NN);
    GET(Ks, (j.tag + "_Ks").c_str(), NN); GET(Kp, (j.tag + "_Kp").c_str(), NN);
    GET(Tc, (j.tag + "_Tcaux").c_str(), 2 * NN);   // T then caux
    j.w1 = w1; j.w2 = w2; j.coef = coef; j.caux = Tc + NN; j.Ks = Ks; j.T = Tc; j.Kp = Kp;
    if (j.N >= 3) {                  // the factored metric (every job on the Cholesky path)
        GET(cw, (j.tag + "_cholw").c_str(), chol_workspace(j.N, 2 * j.nb));
        j.cholw = cw;
    }
    if (j.trd) {
        GET(bw, (j.tag + "_bandw").c_str(), band_workspace(j.N, j.nb));
        j.bandw = bw;
    }
    j.band = false;
    j.sbw = nullptr;
    if (!j.trd && j.N >= 1 && j.N <= SYEVB_NMAX) {
        GET(sb, (j.tag + "_sbw").c_str(), syevb_workspace(j.N) * j.nb);
        j.sbw = sb;
    }
    j.hw1 = p->hpin + hoff;
    if (j.N >= 3) {                  // scratch of eig_trd (every job) and of the K' reduction (trd jobs)
        GET(tm, (j.tag + "_trdm").c_str(), trd_workspace(j.N, j.nb));
        GET(te, (j.tag + "_trde").c_str(), (i64)j.N * j.nb);
        GET(tq, (j.tag + "_trdq").c_str(), trd_apply_workspace(j.N, j.nb, j.N));
        j.trdm = tm; j.trde = te; j.trdq = tq;
        if (j.trd) {
            GET(tw, (j.tag + "_trdw").c_str(), trd_workspace(j.N, j.nb));
            j.trdw = tw;
        }
    }
    for (int k = 0; k < 4; k++) p->hinfo[4 * j.slot + k] = 0;
    return 0;
}
// K' = Q T Q^T of a trd job
__global__ void k_identity_batched(double *Z, int N, i64 sZ) {
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (i64)N * N) Z[blockIdx.z * sZ + e] = (e % N) == (e / N) ? 1.0 : 0.0;
}
// eigen-decomposition of the nb matrices A + b*sA (order N, ld N, upper triangle read) as
// dsyevd('V','U') leaves it: eigenvectors in A, ascending eigenvalues at ev + b*sD, info words at
// dinfo_w[idx + b] (copied to hinfo).  A = Q T Q^T (nv2_trd.hip, one launch per column),
// T = Z W Z^T (rocsolver_dstedc on the tridiagonal matrix), U = Q Z with Q applied in compact WY
// form.  The job's T buffer holds Z.  On worker 1 the second problem's dstedc runs on the extra
// stream and host thread (the solver blocks its calling thread at times).
static bool stage2_timing();
// device time of the three steps of eig_trd for orders >= 500 (NV2_STAGE2_TIMING), printed at its end
struct EigTiming {
    hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
    bool on = false;
    EigTiming(int N) : on(stage2_timing() && N >= 500) {
        if (on) for (auto &x : e) if (hipEventCreate(&x) != hipSuccess) on = false;
    }
    void mark(int k, hipStream_t st) { if (on) (void)hipEventRecord(e[k], st); }
    void report(int N, int nbm) {
        if (!on) return;
        float t[3] = {0, 0, 0};
        (void)hipEventSynchronize(e[3]);
        for (int k = 0; k < 3; k++) (void)hipEventElapsedTime(&t[k], e[k], e[k + 1]);
        fprintf(stderr, "[eig_trd N=%d x%d] reduce %.2f  dstedc %.2f  Q %.2f ms\n", N, nbm, t[0], t[1], t[2]);
    }
    ~EigTiming() { for (auto &x : e) if (x) (void)hipEventDestroy(x); }
};
static int pc_trd_launch(Stage2 *p, PcJob &j) {
    EigTiming tm(j.nmax);
    tm.mark(0, p->ws[j.w]);
    NV2_HIP(hipEventRecord(p->evKr[0], p->ws[j.w]));
    TRYS(trd_reduce(p->ws[j.w], j.nmax, j.Kp, j.NN(), j.nb, j.trdw));
    NV2_HIP(hipEventRecord(p->evKr[1], p->ws[j.w]));
    p->kr_recorded = true;
    p->kr_N = j.nmax;
    p->kr_nb = j.nb;
    p->kr_form = 1;
    if (tm.on) {
        for (int k = 1; k < 4; k++) tm.mark(k, p->ws[j.w]);
        tm.report(j.nmax, j.nb);       // the reduction of K' only: the first figure
    }
    return 0;
}
static int eig_trd(Stage2 *p, PcJob &j, int N, double *A, i64 sA, double *ev, i64 sD, int idx) {
    hipStream_t st = p->ws[j.w];
    rocblas_handle h = p->wh[j.w];
    const int nbm = j.nb;
    EigTiming tm(N);
    tm.mark(0, st);
    TRYS(trd_reduce(st, N, A, sA, nbm, j.trdm));
    tm.mark(1, st);
    const TrdOut o = trd_outputs(N, nbm, j.trdm);
    for (int b = 0; b < nbm; b++)
        NV2_HIP(hipMemcpyAsync(ev + b * sD, o.d + (i64)b * N, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    NV2_HIP(hipMemcpyAsync(j.trde, o.e, sizeof(double) * N * nbm, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_identity_batched, dim3(nb((i64)N * N, 256), 1, nbm), dim3(256), 0, st, j.T, N, sA);
    NV2_LAUNCH_CHECK();
    const bool split = nbm > 1 && j.w == 1;
    int rc1 = 0;
    char err1[512] = "";
    std::thread t1;
    if (split) {
        NV2_HIP(hipEventRecord(p->xev[0], st));
        t1 = std::thread([&]() {
            try {
                if (hipStreamWaitEvent(p->xs, p->xev[0], 0) != hipSuccess ||
                    rocsolver_dstedc(p->xh, rocblas_evect_tridiagonal, N, ev + sD, j.trde + N, j.T + sA, N,
                                     p->dinfo_w + idx + 1) != rocblas_status_success ||
                    hipEventRecord(p->xev[1], p->xs) != hipSuccess) {
                    rc1 = -1;
                    snprintf(err1, sizeof(err1), "dstedc of the second problem (N=%d) failed", N);
                }
            } catch (...) {
                rc1 = -1;
                snprintf(err1, sizeof(err1), "dstedc of the second problem (N=%d): exception", N);
            }
        });
    }
    rocblas_status s0 = rocblas_status_success;
    for (int b = 0; b < (split ? 1 : nbm) && s0 == rocblas_status_success; b++)
        s0 = rocsolver_dstedc(h, rocblas_evect_tridiagonal, N, ev + b * sD, j.trde + (i64)b * N, j.T + b * sA, N,
                              p->dinfo_w + idx + b);
    if (t1.joinable()) t1.join();
    if (s0 != rocblas_status_success) { set_error("rocsolver_dstedc (N=%d): status %d", N, (int)s0); return -1; }
    if (rc1) { set_error("%s", err1); return rc1; }
    if (split) NV2_HIP(hipStreamWaitEvent(st, p->xev[1], 0));
    tm.mark(2, st);
    TRYS(trd_apply_q(h, st, N, A, sA, nbm, o.tau, false, N, j.T, sA, j.trdq));
    tm.mark(3, st);
    tm.report(N, nbm);
    for (int b = 0; b < nbm; b++)
        NV2_HIP(hipMemcpyAsync(A + b * sA, j.T + b * sA, sizeof(double) * N * N, hipMemcpyDeviceToDevice, st));
    NV2_HIP(hipMemcpyAsync(p->hinfo + idx, p->dinfo_w + idx, sizeof(int) * nbm, hipMemcpyDeviceToHost, st));
    return 0;
}
// The (+-1)' metric without its eigen-decomposition.  When no eigenvalue falls below thr, the
// thresholded inverse U_kept w^-1 U_kept^T of the reference is S^-1 itself, and every PC sum
// r^T (cauxT K caux + s)^-1 r, r = caux^T x, is unchanged when caux is replaced by any G with
// G G^T = S^-1 (G = caux O for an orthogonal O).  The Cholesky factor S = U^T U gives G = U^-1:
// K' = U^-T K U^-1 and r = U^-T x by triangular solves.  "No eigenvalue below thr" is decided by
// the Cholesky factorisation of S - thr*I succeeding; otherwise the eigen path runs.
// NV2_METRIC_EIGEN=1 forces the eigen path.
static bool chol_metric() {   // read at every solve, so that a caller can switch paths between runs
    const char *e = getenv("NV2_METRIC_EIGEN");
    return !(e && atoi(e) > 0);
}
// NV2_CHOL_EXACT=1: skip the trace test, so that every Cholesky-path metric takes the exact test
// (S - thr*I factored)
static bool chol_exact_only() {
    const char *e = getenv("NV2_CHOL_EXACT");
    return e && atoi(e) > 0;
}
// K' of the Cholesky path reduced to band form or to tridiagonal form.  The band form wins at the
// measured orders (per evaluation: CAS(14,14), N = 2744, 136 vs 165 ms; CAS(12,12) x 6 roots, N = 1728,
// 518 vs 556 ms; CAS(10,10), N = 1000, 22.6 vs 25.4 ms); below N = 512 the tridiagonal form's short
// chain stays.  NV2_K_TRIDIAG=1 / NV2_K_BAND=1 force one form.
static bool band_kprime(int N) {
    const char *t = getenv("NV2_K_TRIDIAG"), *b = getenv("NV2_K_BAND");
    if (t && atoi(t) > 0) return false;
    if (b && atoi(b) > 0) return true;
    return N >= 512;
}
__global__ void k_shift_diag(double *A, int N, i64 sA, int nbm, double d) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (i64)N * nbm) return;
    const i64 b = t / N, i = t % N;
    A[b * sA + i + (i64)N * i] += d;
}
// Partial sums of squares of the upper triangle of the nb matrices U (order N, ld N): workgroup
// (x, b) covers columns x, x + gridDim.x, ..; part[b * gridDim.x + x] receives its sum.
__global__ __launch_bounds__(256) void k_upper_sq_parts(const double *U, int N, i64 sU, double *part) {
    __shared__ double red[4];
    const int b = blockIdx.y;
    const double *Ub = U + b * sU;
    double acc = 0.0;
    for (int c = blockIdx.x; c < N; c += gridDim.x)
        for (int r = threadIdx.x; r <= c; r += 256) {
            const double v = Ub[r + (i64)N * c];
            acc += v * v;
        }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
// info[b] = 0 when trace(S_b^-1) = ||U_b^-1||_F^2 (the sum of the np parts) is below lim = 1/thr: then
// 1/lambda_min(S_b) <= trace(S_b^-1) < 1/thr, every eigenvalue is above thr.  1 otherwise (NaN included).
__global__ __launch_bounds__(256) void k_trace_test(const double *part, int np, double lim, int *info) {
    __shared__ double red[4];
    const int b = blockIdx.x;
    double acc = 0.0;
    for (int k = threadIdx.x; k < np; k += 256) acc += part[b * np + k];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) info[b] = (red[0] + red[1] + red[2] + red[3] < lim) ? 0 : 1;
}
#define TRACE_PARTS 512
// The factored metric: S_b = U_b^T U_b, U_b^-1 into T and the trace test of every eigenvalue
// above thr.  Info words: 4*slot + b for the eigenvalue test, + nb + b for the factorisation of S.
static int pc_metric_chol_launch(Stage2 *p, PcJob &j, double thr) {
    {
        j.chol = true;
        hipStream_t st = p->ws[j.w];
        rocblas_handle h = p->wh[j.w];
        const i64 NN = j.NN();
        NV2_HIP(hipMemcpyAsync(j.caux, j.S, sizeof(double) * NN * j.nb, hipMemcpyDeviceToDevice, st));
        EigTiming tm(j.N);
        tm.mark(0, st);
        // caux_b = U_b with S_b = U_b^T U_b: info words nb + b
        TRYS(chol_upper(h, st, j.N, j.caux, NN, j.nb, p->dinfo_w + 4 * j.slot + j.nb, j.cholw));
        // U^-1 into T while K is not there yet (K' then takes two GEMMs), and from it the test that every
        // eigenvalue of S is above thr: info words b (pc_kprime_launch factors S - thr*I when it fails)
        TRYS(chol_inverse_upper(h, st, j.N, j.caux, NN, j.nb, j.T, NN, j.cholw));
        hipLaunchKernelGGL(k_upper_sq_parts, dim3(TRACE_PARTS, j.nb), dim3(256), 0, st, j.T, j.N, NN, j.cholw);
        NV2_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_trace_test, dim3(j.nb), dim3(256), 0, st, j.cholw, TRACE_PARTS, chol_exact_only() ? 0.0 : 1.0 / thr,
                           p->dinfo_w + 4 * j.slot);
        NV2_LAUNCH_CHECK();
        for (int k = 1; k < 4; k++) tm.mark(k, st);
        tm.report(j.N, j.nb);            // the factorisations: the first figure
        NV2_HIP(hipMemcpyAsync(p->hinfo + 4 * j.slot, p->dinfo_w + 4 * j.slot, sizeof(int) * 4, hipMemcpyDeviceToHost,
                               st));
        return 0;
    }
}
static int pc_metric_launch(Stage2 *p, PcJob &j, double thr) {
    if (j.trd && j.N >= 3 && chol_metric()) return pc_metric_chol_launch(p, j, thr);
    if (j.N >= 3 && eigen_algo() == 0) TRYS(eig_trd(p, j, j.N, j.S, j.NN(), j.w1, j.N, 4 * j.slot));
    else TRYS(syevd_async(p, j.w, 4 * j.slot, true, j.N, j.S, j.NN(), j.w1, j.N, j.nb));
    if (j.N) NV2_HIP(hipMemcpyAsync(j.hw1, j.w1, sizeof(double) * j.N * j.nb, hipMemcpyDeviceToHost, p->ws[j.w]));
    return 0;
}
static int pc_kprime_form(Stage2 *p, PcJob &j, double thr);
// after the worker stream has finished: threshold the metric spectra, launch K' and its solves
static int pc_kprime_launch(Stage2 *p, PcJob &j, double thr) {
    const int N = j.N;
    if (j.chol) {                        // the factorisations of pc_metric_launch have finished
        bool ok = true, exact = false;
        for (int b = 0; b < j.nb; b++) {
            if (p->hinfo[4 * j.slot + j.nb + b] != 0) ok = false;       // S itself not positive definite
            if (p->hinfo[4 * j.slot + b] != 0) exact = true;            // the trace test left it open
            p->hinfo[4 * j.slot + b] = p->hinfo[4 * j.slot + j.nb + b] = 0;
        }
        hipStream_t st = p->ws[j.w];
        rocblas_handle h = p->wh[j.w];
        const i64 NN = j.NN();
        if (ok && exact) {
            // the exact test: S - thr*I positive definite, factored in the (still unused) K' buffer
            NV2_HIP(hipMemcpyAsync(j.Kp, j.S, sizeof(double) * NN * j.nb, hipMemcpyDeviceToDevice, st));
            hipLaunchKernelGGL(k_shift_diag, dim3(nb((i64)N * j.nb, 256)), dim3(256), 0, st, j.Kp, N, NN, j.nb, -thr);
            NV2_LAUNCH_CHECK();
            TRYS(chol_upper(h, st, N, j.Kp, NN, j.nb, p->dinfo_w + 4 * j.slot, j.cholw));
            NV2_HIP(hipMemcpyAsync(p->hinfo + 4 * j.slot, p->dinfo_w + 4 * j.slot, sizeof(int) * j.nb,
                                   hipMemcpyDeviceToHost, st));
            NV2_HIP(hipStreamSynchronize(st));
            for (int b = 0; b < j.nb; b++) {
                if (p->hinfo[4 * j.slot + b] != 0) ok = false;
                p->hinfo[4 * j.slot + b] = 0;
            }
        }
        if (stage2_timing() && N >= 500) fprintf(stderr, "[metric %s N=%d] Cholesky path: %s\n", j.tag.c_str(), N,
                                                 ok ? "every eigenvalue above thr" : "not positive definite above thr");
        if (ok) {
            j.nmax = N;
            for (int b = 0; b < j.nb; b++) { j.nnew[b] = N; j.minw[b] = thr; }   // thr: a lower bound
            EigTiming tk(N);             // K' formation: printed as the first figure
            tk.mark(0, st);
            std::swap(j.caux, j.T);      // caux = U^-1, so caux caux^T = S^-1; T is free scratch
            for (int b = 0; b < j.nb; b++)
                hipLaunchKernelGGL(k_sym_upper, dim3(nb(NN, 256)), dim3(256), 0, st, j.Kb(b), N, j.Ks + b * NN);
            NV2_LAUNCH_CHECK();
            const double one = 1.0, zero = 0.0;
            auto kprime = [&]() -> int {
                RB(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, N, N, N, &one, j.Ks,
                                                 N, NN, j.caux, N, NN, &zero, j.T, N, NN, j.nb));
                RB(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, N, N, N, &one,
                                                 j.caux, N, NN, j.T, N, NN, &zero, j.Kp, N, NN, j.nb));
                return 0;
            };
            TRYS(kprime());
            for (int k = 1; k < 4; k++) tk.mark(k, st);
            tk.report(N, j.nb);
            if (band_kprime(N)) {
                // flags in the K' info words (4*slot + 2 + b), read here and cleared: a panel whose
                // CholeskyQR2 failed sends both problems to the tridiagonal reduction of a fresh K'
                int *fl = p->dinfo_w + 4 * j.slot + 2;
                NV2_HIP(hipEventRecord(p->evKr[0], st));
                TRYS(band_reduce(h, st, N, j.Kp, NN, j.nb, j.bandw, fl));
                NV2_HIP(hipEventRecord(p->evKr[1], st));
                NV2_HIP(hipMemcpyAsync(p->hinfo + 4 * j.slot + 2, fl, sizeof(int) * j.nb, hipMemcpyDeviceToHost, st));
                NV2_HIP(hipStreamSynchronize(st));
                bool bok = true;
                for (int b = 0; b < j.nb; b++) {
                    if (p->hinfo[4 * j.slot + 2 + b] != 0) bok = false;
                    p->hinfo[4 * j.slot + 2 + b] = 0;
                }
                if (stage2_timing()) fprintf(stderr, "[K' %s N=%d] band reduction %s\n", j.tag.c_str(), N,
                                             bok ? "done" : "failed (CholeskyQR2): tridiagonal reduction");
                if (bok) {
                    // caux stays U^-1: the sums project y = caux^T x and apply Q^T to y (band_apply_qt)
                    j.band = true;
                    p->kr_recorded = true;
                    p->kr_N = N;
                    p->kr_nb = j.nb;
                    p->kr_form = 2;
                    return 0;
                }
                TRYS(kprime());
            }
            return pc_trd_launch(p, j);
        }
        j.chol = false;                  // some eigenvalue below thr: the eigen path after all
        TRYS(eig_trd(p, j, N, j.S, NN, j.w1, N, 4 * j.slot));
        NV2_HIP(hipMemcpyAsync(j.hw1, j.w1, sizeof(double) * N * j.nb, hipMemcpyDeviceToHost, st));
        NV2_HIP(hipStreamSynchronize(st));
    }
    TRYS(pc_kprime_form(p, j, thr));
    if (j.nmax == 0) return 0;
    if (j.trd) return pc_trd_launch(p, j);
    const int nmax = j.nmax;
    hipStream_t st = p->ws[j.w];
    rocblas_handle h = p->wh[j.w];
    const i64 NN = j.NN();
    const double one = 1.0, zero = 0.0;
    if (nmax >= 3 && eigen_algo() == 0) TRYS(eig_trd(p, j, nmax, j.Kp, NN, j.w2, N, 4 * j.slot + 2));
    else TRYS(syevd_async(p, j.w, 4 * j.slot + 2, true, nmax, j.Kp, NN, j.w2, N, j.nb));
    RB(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, N, nmax, nmax, &one, j.caux, N,
                                     NN, j.Kp, nmax, NN, &zero, j.coef, N, NN, j.nb));
    return 0;
}
// The thresholding of the metric spectra (host copies in hw1, info words in hinfo), caux =
// U_kept w^-1/2 and K' = caux^T K caux of a job on the eigen path, padded to a common order when
// the two problems keep different counts; nmax = 0 when nothing is kept.
static int pc_kprime_form(Stage2 *p, PcJob &j, double thr) {
    const int N = j.N;
    j.nmax = 0;
    for (int b = 0; b < j.nb; b++) {
        const int inf = p->hinfo[4 * j.slot + b];
        if (inf != 0) { set_error("syevd(N=%d, %s metric %d) failed, info=%d", N, j.tag.c_str(), b, inf); return -3; }
        int nnew = 0;
        for (int k = 0; k < N; k++) if (j.hw1[(i64)b * N + k] >= thr) nnew++;
        j.nnew[b] = nnew;
        j.minw[b] = N ? j.hw1[(i64)b * N] : 0.0;
        if (nnew > j.nmax) j.nmax = nnew;
    }
    const int nmax = j.nmax;
    if (stage2_timing() && N >= 500)
        fprintf(stderr, "[metric %s N=%d] kept %d / %d\n", j.tag.c_str(), N, j.nnew[0], j.nb > 1 ? j.nnew[1] : -1);
    if (nmax == 0) return 0;
    hipStream_t st = p->ws[j.w];
    rocblas_handle h = p->wh[j.w];
    const i64 NN = j.NN();
    NV2_HIP(hipMemsetAsync(j.caux, 0, sizeof(double) * NN * j.nb, st));
    for (int b = 0; b < j.nb; b++) {
        if (j.nnew[b])
            hipLaunchKernelGGL(k_caux, dim3(nb((i64)N * j.nnew[b], 256)), dim3(256), 0, st, j.Sb(b), j.w1 + (i64)b * N,
                               N, N - j.nnew[b], j.nnew[b], j.caux + b * NN);
        hipLaunchKernelGGL(k_sym_upper, dim3(nb(NN, 256)), dim3(256), 0, st, j.Kb(b), N, j.Ks + b * NN);
    }
    NV2_LAUNCH_CHECK();
    const double one = 1.0, zero = 0.0;
    RB(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, N, nmax, N, &one, j.Ks, N, NN,
                                     j.caux, N, NN, &zero, j.T, N, NN, j.nb));
    RB(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, nmax, nmax, N, &one,
                                     j.caux, N, NN, j.T, N, NN, &zero, j.Kp, nmax, NN, j.nb));
    if (j.nb > 1 && (j.nnew[0] < nmax || j.nnew[1] < nmax)) {
        hipLaunchKernelGGL(k_pad_gersh, dim3(1), dim3(256), 0, st, j.Kp, nmax, NN, j.nb, j.nnew[0], j.nnew[1], nmax);
        NV2_LAUNCH_CHECK();
    }
    return 0;
}
// one host thread per worker stream: the solver library blocks its calling thread at times,
// so separate threads are what lets the chains of different problems overlap on the device
// The jobs of worker w (the classes other than (+-1)') as two batched eigen-solves: every metric
// of every job in one syevb call, then every K' (after the host thresholding) in another, then
// coef = caux V per job.
static int pc_worker_batched(Stage2 *p, PcJob *J, int njob, int w, double thr, double *ms_done,
                             std::chrono::steady_clock::time_point t0) {
    hipStream_t st = p->ws[w];
    rocblas_handle h = p->wh[w];
    // metrics of order >= 3 factored first (S = U^T U, U^-1, trace test of every eigenvalue above thr)
    for (int k = 0; k < njob; k++)
        if (J[k].w == w && J[k].N >= 3 && chol_metric()) TRYS(pc_metric_chol_launch(p, J[k], thr));
    NV2_HIP(hipStreamSynchronize(st));
    // a job whose metric passes keeps caux = U^-1 (every eigenvalue kept); the others, and any metric
    // the trace test leaves open, take the eigen path: one batched solve of their metrics
    // the trace test is sufficient, not necessary: where it leaves a metric open that factors, the
    // exact test (S - thr*I positive definite) decides, as on the (+-1)' chain
    bool open = false;
    std::vector<char> pass(njob, 0);
    for (int k = 0; k < njob; k++) {
        PcJob &j = J[k];
        if (j.w != w || !j.chol) continue;
        bool ok = true, exact = false;
        for (int b = 0; b < j.nb; b++) {
            if (p->hinfo[4 * j.slot + j.nb + b] != 0) ok = false;
            if (p->hinfo[4 * j.slot + b] != 0) exact = true;
        }
        for (int q = 0; q < 4; q++) p->hinfo[4 * j.slot + q] = 0;   // all four words were copied
        pass[k] = ok && !exact;
        if (ok && exact) {
            const i64 NN = j.NN();
            NV2_HIP(hipMemcpyAsync(j.Kp, j.S, sizeof(double) * NN * j.nb, hipMemcpyDeviceToDevice, st));
            hipLaunchKernelGGL(k_shift_diag, dim3(nb((i64)j.N * j.nb, 256)), dim3(256), 0, st, j.Kp, j.N, NN, j.nb, -thr);
            NV2_LAUNCH_CHECK();
            TRYS(chol_upper(h, st, j.N, j.Kp, NN, j.nb, p->dinfo_w + 4 * j.slot, j.cholw));
            NV2_HIP(hipMemcpyAsync(p->hinfo + 4 * j.slot, p->dinfo_w + 4 * j.slot, sizeof(int) * j.nb,
                                   hipMemcpyDeviceToHost, st));
            pass[k] = 2;
            open = true;
        }
    }
    if (open) NV2_HIP(hipStreamSynchronize(st));
    std::vector<SyevbProb> pr;
    for (int k = 0; k < njob; k++) {
        PcJob &j = J[k];
        if (j.w != w) continue;
        if (j.chol) {
            bool ok = pass[k] == 1;
            if (pass[k] == 2) {
                ok = true;
                for (int b = 0; b < j.nb; b++) {
                    if (p->hinfo[4 * j.slot + b] != 0) ok = false;
                    p->hinfo[4 * j.slot + b] = 0;
                }
            }
            if (ok) {
                j.nmax = j.N;
                for (int b = 0; b < j.nb; b++) { j.nnew[b] = j.N; j.minw[b] = thr; }   // thr: a lower bound
                std::swap(j.caux, j.T);
                continue;
            }
            j.chol = false;
        }
        for (int b = 0; b < j.nb && j.N; b++)
            pr.push_back(SyevbProb{j.Sb(b), j.w1 + (i64)b * j.N, j.sbw + b * syevb_workspace(j.N), j.N});
    }
    TRYS(syevb(h, st, (int)pr.size(), pr.data(), p->plan(0)));
    for (int k = 0; k < njob; k++)
        if (J[k].w == w && J[k].N && !J[k].chol)
            NV2_HIP(hipMemcpyAsync(J[k].hw1, J[k].w1, sizeof(double) * J[k].N * J[k].nb, hipMemcpyDeviceToHost, st));
    NV2_HIP(hipStreamSynchronize(st));
    const double ms_metric = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    // K' of every job, then one batched solve of all of them
    pr.clear();
    const double one = 1.0, zero = 0.0;
    for (int k = 0; k < njob; k++) {
        PcJob &j = J[k];
        if (j.w != w) continue;
        if (j.chol) {                    // K' = U^-T K U^-1
            const int N = j.N;
            const i64 NN = j.NN();
            for (int b = 0; b < j.nb; b++)
                hipLaunchKernelGGL(k_sym_upper, dim3(nb(NN, 256)), dim3(256), 0, st, j.Kb(b), N, j.Ks + b * NN);
            NV2_LAUNCH_CHECK();
            RB(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, N, N, N, &one, j.Ks, N,
                                             NN, j.caux, N, NN, &zero, j.T, N, NN, j.nb));
            RB(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, N, N, N, &one,
                                             j.caux, N, NN, j.T, N, NN, &zero, j.Kp, N, NN, j.nb));
        } else {
            TRYS(pc_kprime_form(p, j, thr));
        }
        for (int b = 0; b < j.nb && j.nmax; b++)
            pr.push_back(SyevbProb{j.Kp + b * j.NN(), j.w2 + (i64)b * j.N, j.sbw + b * syevb_workspace(j.N), j.nmax});
    }
    TRYS(syevb(h, st, (int)pr.size(), pr.data(), p->plan(1)));
    for (int k = 0; k < njob; k++) {
        const PcJob &j = J[k];
        if (j.w != w || !j.nmax) continue;
        const i64 NN = j.NN();
        RB(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, j.N, j.nmax, j.nmax, &one,
                                         j.caux, j.N, NN, j.Kp, j.nmax, NN, &zero, j.coef, j.N, NN, j.nb));
    }
    NV2_HIP(hipStreamSynchronize(st));
    *ms_done = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stage2_timing()) {
        fprintf(stderr, "[worker %d batched] metrics at %.2f ms, K' spectra and coef at %.2f ms\n", w, ms_metric,
                *ms_done);
        for (int k = 0; k < njob; k++)
            if (J[k].w == w)
                fprintf(stderr, "[metric %s N=%d] %s\n", J[k].tag.c_str(), J[k].N,
                        J[k].chol ? "Cholesky path: every eigenvalue above thr" : "eigen path");
    }
    return 0;
}
static int pc_worker(Stage2 *p, PcJob *J, int njob, int w, double thr, double *ms_done,
                     std::chrono::steady_clock::time_point t0) {
    bool batched = syevb_on();
    for (int k = 0; k < njob; k++)
        if (J[k].w == w && (J[k].trd || !J[k].sbw)) batched = false;
    if (batched) return pc_worker_batched(p, J, njob, w, thr, ms_done, t0);
    for (int k = 0; k < njob; k++)
        if (J[k].w == w) TRYS(pc_metric_launch(p, J[k], thr));
    NV2_HIP(hipStreamSynchronize(p->ws[w]));
    for (int k = 0; k < njob; k++)
        if (J[k].w == w) TRYS(pc_kprime_launch(p, J[k], thr));
    NV2_HIP(hipStreamSynchronize(p->ws[w]));
    *ms_done = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

static int slice(Stage2 *p, const double *src, int s, i64 cnt, double *dst);

// The (+-1)' chain (n^3 metric, K', reduction) is the longest path of stage 2.  Its metric needs
// the 3-RDM only, so it can start right after the partial traces, before the Koopmans matrices
// exist; the worker thread then waits for K (filled by stage2_state once stage 1 is done) before
// forming K'.
struct Pm1pEarly {
    bool active = false;
    int state = -1;
    double thr = 0.0;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    bool k_ready = false, abort = false;
    bool k_filled = false;         // K filled and signalled ahead of stage2_state (stage2_pm1p_kready)
    bool kred_done = false;        // the K' reduction has finished on the device (or the chain has ended)
    bool qt_event = false;         // with kred_done: the held SC sums also wait for evQt on the device
    // Single-state runs on the band form: the PC sums go out on the chain's stream as soon as the
    // right-hand sides exist (x_ready, event evX), alongside the other classes on the main stream.
    bool sums_early = false, x_ready = false, sums_done = false;
    const double *X[2] = {nullptr, nullptr}, *eo[2] = {nullptr, nullptr};
    int next[2] = {0, 0};
    double *Ypc[2] = {nullptr, nullptr}, *Ub = nullptr, *yb = nullptr, *ptb = nullptr;
    int ncm = 0;                   // column stride of the two classes' right-hand sides (Ypc[0] = Ypc[1] + N3 ncm)
    double *qtw = nullptr;         // band_apply_qt scratch
    // .NATORB: density vectors T (N3 x n), their projections, and the density block of each class
    const double *Tr[2] = {nullptr, nullptr};
    double *Vr[2] = {nullptr, nullptr}, *rho[2] = {nullptr, nullptr};
    int nact = 0;
    int rc = 0;
    char err[512] = "";
    PcJob job;
    std::chrono::steady_clock::time_point t_start, t_metric, t_kready, t_end;   // for NV2_STAGE2_TIMING
};
static size_t pm1p_hoff(int n) {   // pinned-host offset of its metric eigenvalues: after the other jobs'
    const size_t nn = (size_t)n, m = nn * (nn + 1) / 2;
    return 2 * nn + 2 * nn * nn + 2 * m + 2 * nn * nn;
}
static void pm1p_stop(Stage2 *p) {   // abort a chain still waiting for K, and join it
    Pm1pEarly *E = p->early;
    if (!E || !E->active) return;
    {
        std::lock_guard<std::mutex> lk(E->mu);
        E->abort = true;
    }
    E->cv.notify_all();
    if (E->th.joinable()) E->th.join();
    E->active = false;
}
// .NATORB density block from the banded solutions u_x (Y = [u_x], N x next): -sum_i u_x[i] V[i, d] with
// V = (caux Q)^T T, as rho_ca[x, d] (nc x n) for (+1)' and rho_av[d, x] (n x nv) for (-1)'
// Releases the SC sums held for the (+-1)' K' reduction; with qt, they also wait for evQt on the device
static void early_kred_signal(Pm1pEarly *E, bool qt) {
    {
        std::lock_guard<std::mutex> lk(E->mu);
        E->qt_event = qt;
        E->kred_done = true;
    }
    E->cv.notify_all();
}
static int band_rho(rocblas_handle h, hipStream_t st, int plus, int N, int n, int next, const double *Y,
                    const double *V, double *rho) {
    if (plus) return gemm_on(h, st, true, false, next, n, N, -1.0, Y, N, V, N, 0.0, rho, next);
    return gemm_on(h, st, true, false, n, next, N, -1.0, V, N, Y, N, 0.0, rho, n);
}
// (-1)' and (+1)' PC sums on the band form, on the chain's stream: y = caux^T x, z = Q^T y, one banded
// solve per external orbital (both classes in one launch), block partials reduced into the classes
static int pm1p_band_sums(Stage2 *p, Pm1pEarly *E) {
    const PcJob &jb = E->job;
    hipStream_t st = p->ws[jb.w];
    rocblas_handle h = p->wh[jb.w];
    const int N = jb.nmax;
    const i64 NN = jb.NN();
    BandCls bc[2] = {};
    for (int plus = 0; plus < 2; plus++) {
        const int next = E->next[plus], bb = plus ? 0 : 1;
        if (!next || !E->X[plus]) continue;
        // y = caux^T x; z = Q^T y below, for both classes in the same GEMMs when both are there
        TRYS(gemm_on(h, st, true, false, N, next, N, 1.0, jb.caux + bb * NN, N, E->X[plus], N, 0.0, E->Ypc[plus], N));
        if (E->rho[plus]) {
            TRYS(gemm_on(h, st, true, false, N, E->nact, N, 1.0, jb.caux + bb * NN, N, E->Tr[plus], N, 0.0, E->Vr[plus], N));
            TRYS(band_apply_qt(st, N, jb.bandw, bb, 1, E->Vr[plus], N, 0, E->nact, E->qtw));
        }
        BandCls &c = bc[plus];
        c.Br = band_rows(N, jb.nb, bb, jb.bandw);
        c.Z = E->Ypc[plus];
        c.eo = E->eo[plus];
        c.plus = plus;
        c.next = next;
    }
    const int n0 = bc[0].next, n1 = bc[1].next;
    if (!n0 && !n1) {
        early_kred_signal(E, false);
        return 0;
    }
    if (n0 && n1) TRYS(band_apply_qt(st, N, jb.bandw, 0, 2, E->Ypc[1], N, (i64)N * E->ncm, E->ncm, E->qtw));
    else if (n0) TRYS(band_apply_qt(st, N, jb.bandw, 1, 1, E->Ypc[0], N, 0, n0, E->qtw));
    else TRYS(band_apply_qt(st, N, jb.bandw, 0, 1, E->Ypc[1], N, 0, n1, E->qtw));
    NV2_HIP(hipEventRecord(p->evQt, st));
    early_kred_signal(E, true);
    bc[0].U = E->Ub;
    bc[0].y = E->yb;
    bc[1].U = E->Ub + (i64)n0 * N * UR;
    bc[1].y = E->yb + (i64)n0 * N;
    hipLaunchKernelGGL(k_pm1p_bandr, dim3(n0 + n1), dim3(RW_T), 0, st, N, bc[0], bc[1], E->ptb);
    NV2_LAUNCH_CHECK();
    for (int c = 0; c < 2; c++) {
        if (!bc[c].next) continue;
        hipLaunchKernelGGL(k_reduce_parts, dim3(4), dim3(256), 0, st, E->ptb + 4 * (c ? n0 : 0), (i64)bc[c].next,
                           p->acc + 4 * (c ? 5 : 6));
        NV2_LAUNCH_CHECK();
        if (E->rho[c]) TRYS(band_rho(h, st, c, N, E->nact, bc[c].next, bc[c].y, E->Vr[c], E->rho[c]));
    }
    return 0;
}
static int pm1p_worker(Stage2 *p, Pm1pEarly *E) {
    PcJob &j = E->job;
    E->t_start = std::chrono::steady_clock::now();
    NV2_HIP(hipStreamWaitEvent(p->ws[j.w], p->evS, 0));
    TRYS(pc_metric_launch(p, j, E->thr));
    NV2_HIP(hipStreamSynchronize(p->ws[j.w]));
    E->t_metric = std::chrono::steady_clock::now();
    {
        std::unique_lock<std::mutex> lk(E->mu);
        E->cv.wait(lk, [E] { return E->k_ready || E->abort; });
        if (E->abort) return 0;
    }
    E->t_kready = std::chrono::steady_clock::now();
    NV2_HIP(hipStreamWaitEvent(p->ws[j.w], p->evK, 0));
    TRYS(pc_kprime_launch(p, j, E->thr));
    NV2_HIP(hipStreamSynchronize(p->ws[j.w]));
    E->t_end = std::chrono::steady_clock::now();
    // on the band form with the sums on this stream, the SC sums are released once Q^T is applied to
    // the right-hand sides (pm1p_band_sums): its workgroups need whole CUs, which the SC sums would hold
    if (!(E->sums_early && j.band)) early_kred_signal(E, false);
    if (E->sums_early && j.band) {
        {
            std::unique_lock<std::mutex> lk(E->mu);
            E->cv.wait(lk, [E] { return E->x_ready || E->abort; });
            if (E->abort) return 0;
        }
        NV2_HIP(hipStreamWaitEvent(p->ws[j.w], p->evX, 0));
        TRYS(pm1p_band_sums(p, E));
        E->sums_done = true;
        NV2_HIP(hipStreamSynchronize(p->ws[j.w]));
    }
    return 0;
}

void stage2_graph_launches(Stage2 *p, double *out) {
    for (int k = 0; k < 3; k++) out[k] = (double)syevb_plan_graph_launches(p->sbplan[k]);
}
void stage2_set_no4rdm(Stage2 *p, bool on) { p->no4 = on; }

void stage2_set_rho1(Stage2 *p, bool on) { p->rho1 = on; }

int stage2_rho1_get(Stage2 *p, double *can, double *mo, double *occ) {
    const char *nm[3] = {"rho1_R", "rho1_mo", "rho1_occ"};
    double *dst[3] = {can, mo, occ};
    const size_t cnt[3] = {(size_t)p->norb * p->norb, (size_t)p->norb * p->norb, (size_t)p->norb};
    for (int k = 0; k < 3; k++) {
        auto it = p->bufs.find(nm[k]);
        if (it == p->bufs.end() || !it->second.p) { set_error("no .NATORB density computed"); return -2; }
        if (dst[k]) NV2_HIP(hipMemcpyAsync(dst[k], it->second.p, sizeof(double) * cnt[k], hipMemcpyDeviceToHost, p->st));
    }
    NV2_HIP(hipStreamSynchronize(p->st));
    return 0;
}

int stage2_set_frozen(Stage2 *p, int nfrozen, const int *core_index) {
    std::vector<int> f;
    for (int k = 0; k < nfrozen; k++) {
        if (core_index[k] < 0 || core_index[k] >= p->nc) {
            set_error("frozen orbital %d is not a core orbital (ncore = %d)", core_index[k] + 1, p->nc);
            return -2;
        }
        f.push_back(core_index[k]);
    }
    p->frozen = f;
    return 0;
}

__global__ void k_freeze(double *eps, const int *idx, int nf) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nf) eps[idx[k]] = -1.0e30;
}

int stage2_pm1p_begin(Stage2 *p, int s, double thr) {
    if (p->no4 || !(p->in.amatIP && p->in.amatEA)) return 0;
    if (!p->early) p->early = new Pm1pEarly();
    Pm1pEarly *E = p->early;
    if (E->active) {
        if (E->state == s) return 0;
        pm1p_stop(p);
    }
    hipStream_t st = p->st;
    const int n = p->n;
    const i64 n2 = (i64)n * n, n4 = n2 * n2, n6 = n4 * n2, N3 = n2 * n;
    GET(da, "e_da", n2); GET(taa, "e_taa", n4); GET(ro3, "e_ro3", n6);
    TRYS(slice(p, p->in.dal, s, n2, da)); TRYS(slice(p, p->in.taa, s, n4, taa)); TRYS(slice(p, p->in.daaa, s, n6, ro3));
    Sl S{n, da, taa, ro3};
    PcJob &j = E->job;
    j = PcJob();
    j.tag = "pm1p"; j.N = (int)N3; j.nb = 2; j.slot = 4; j.w = 1; j.trd = true;
    GET(Sm, "pm1p_S", N3 * N3 * 2); GET(Km, "pm1p_K", N3 * N3 * 2);
    j.S = Sm;
    j.K = Km;
    TRYS(pc_job_buffers(p, j, pm1p_hoff(n)));
    GET(E1, "syevd_E1", N3 * 2);      // used only when the solver is forced to rocSOLVER (NV2_EIGEN)
    p->wE[1] = E1;
    for (int plus = 0; plus < 2; plus++) {   // metric, and the copy kept for the S-build
        const int bb = plus ? 0 : 1;
        GET(mx, plus ? "pm1_M_p" : "pm1_M_m", N3 * N3);
        hipLaunchKernelGGL(k_fill_pm1p_s, dim3(nb(N3 * N3, 256)), dim3(256), 0, st, S, plus, j.Sb(bb));
        NV2_LAUNCH_CHECK();
        NV2_HIP(hipMemcpyAsync(mx, j.Sb(bb), sizeof(double) * N3 * N3, hipMemcpyDeviceToDevice, st));
    }
    E->sums_early = p->M == 1 && band_kprime((int)N3) && chol_metric() && (p->nc + p->nv) > 0;
    if (E->sums_early) {              // the sums' buffers, allocated here on the calling thread
        const i64 nx = (i64)p->nc + p->nv;
        // both classes' right-hand sides in one buffer, problem b (the (+1)' class for b = 0) at b * N3 * ncm,
        // so that Q^T goes to both in the same batched GEMMs
        const i64 ncm = std::max<i64>(std::max(p->nc, p->nv), 1);
        GET(yy, "pm1_Ypcb", 2 * N3 * ncm);
        GET(ub, "pm1_bandU", nx * N3 * UR); GET(yb, "pm1_bandy", nx * N3); GET(pb, "pm1_bandparts", 4 * nx);
        E->Ypc[1] = yy; E->Ypc[0] = yy + N3 * ncm; E->ncm = (int)ncm;
        GET(qtw, "pm1_qtw", std::max(band_qt_workspace((int)N3, 2, (int)ncm), band_qt_workspace((int)N3, 1, n)));
        E->qtw = qtw;
        E->Ub = ub; E->yb = yb; E->ptb = pb;
    }
    E->x_ready = false;
    E->sums_done = false;
    NV2_HIP(hipEventRecord(p->evS, st));
    E->state = s;
    E->thr = thr;
    E->k_ready = false;
    E->k_filled = false;
    E->abort = false;
    E->rc = 0;
    E->err[0] = 0;
    E->kred_done = false;
    E->qt_event = false;
    E->active = true;
    E->th = std::thread([p, E]() {
        try {
            E->rc = pm1p_worker(p, E);
        } catch (const std::exception &x) {
            set_error("(+-1)' chain: %s", x.what());
            E->rc = -1;
        } catch (...) {
            set_error("(+-1)' chain: unknown exception");
            E->rc = -1;
        }
        if (E->rc) nv2_last_error(E->err, sizeof(E->err));
        {
            std::lock_guard<std::mutex> lk(E->mu);
            E->kred_done = true;       // also on every early exit: nothing may wait for it forever
        }
        E->cv.notify_all();
    });
    return 0;
}

// The chain's Koopmans matrix needs amat only: it is filled (and the chain released) as soon as bamat
// has finished, while the other Koopmans matrices of stage 1 are still being built.
int stage2_pm1p_kready(Stage2 *p, int s) {
    Pm1pEarly *E = p->early;
    if (!E || !E->active || E->state != s || E->k_filled) return 0;
    hipStream_t st = p->st;
    const int n = p->n;
    const i64 n2 = (i64)n * n, n6 = n2 * n2 * n2, N3 = n2 * n;
    for (int plus = 0; plus < 2; plus++) {
        const std::string t = plus ? "_p" : "_m";
        GET(am, ("k_amat" + t).c_str(), n6);
        TRYS(slice(p, plus ? p->in.amatEA : p->in.amatIP, s, n6, am));
        hipLaunchKernelGGL(k_fill_pm1p_k, dim3(nb(N3 * N3, 256)), dim3(256), 0, st, n, am, E->job.Kb(plus ? 0 : 1));
        NV2_LAUNCH_CHECK();
    }
    NV2_HIP(hipEventRecord(p->evK, st));
    {
        std::lock_guard<std::mutex> lk(E->mu);
        E->k_filled = true;
        E->k_ready = true;
    }
    E->cv.notify_all();
    return 0;
}

Stage2 *stage2_create(const S2Inputs &in, rocblas_handle h, hipStream_t st) {
    Stage2 *p = new Stage2();
    p->in = in;
    p->h = h;
    p->st = st;
    p->nc = in.nc; p->n = in.n; p->nv = in.nv; p->L = in.L; p->M = in.M;
    p->norb = in.nc + in.n + in.nv;
    if (hipMalloc(&p->dinfo, 2 * sizeof(int)) != hipSuccess) { delete p; return nullptr; }
    if (hipMalloc(&p->acc, sizeof(double) * 32) != hipSuccess) { delete p; return nullptr; }
    if (hipMalloc(&p->accq, sizeof(double) * 8 * 2 * QM) != hipSuccess) { delete p; return nullptr; }
    if (hipMalloc(&p->dinfo_w, sizeof(int) * 64) != hipSuccess) { delete p; return nullptr; }
    if (hipHostMalloc(&p->hinfo, sizeof(int) * 32, hipHostMallocDefault) != hipSuccess) { delete p; return nullptr; }
    {
        const i64 n = in.n;
        p->hpin_cap = (size_t)(2 * n * n * n + 4 * n * n + 2 * n * (n + 1) + 2 * n + 64);
        if (hipHostMalloc(&p->hpin, sizeof(double) * p->hpin_cap, hipHostMallocDefault) != hipSuccess) { delete p; return nullptr; }
    }
    for (int w = 0; w < NW; w++) {
        if (hipStreamCreateWithFlags(&p->ws[w], hipStreamNonBlocking) != hipSuccess) { delete p; return nullptr; }
        if (rocblas_create_handle(&p->wh[w]) != rocblas_status_success) { delete p; return nullptr; }
        rocblas_set_stream(p->wh[w], p->ws[w]);
        rocblas_set_pointer_mode(p->wh[w], rocblas_pointer_mode_host);
    }
    for (int e = 0; e < NW + 1; e++)
        if (hipEventCreateWithFlags(&p->wev[e], hipEventDisableTiming) != hipSuccess) { delete p; return nullptr; }
    if (hipEventCreateWithFlags(&p->evS, hipEventDisableTiming) != hipSuccess) { delete p; return nullptr; }
    if (hipEventCreateWithFlags(&p->evK, hipEventDisableTiming) != hipSuccess) { delete p; return nullptr; }
    if (hipEventCreateWithFlags(&p->evX, hipEventDisableTiming) != hipSuccess) { delete p; return nullptr; }
    if (hipEventCreateWithFlags(&p->evQt, hipEventDisableTiming) != hipSuccess) { delete p; return nullptr; }
    for (auto &e : p->evKr)
        if (hipEventCreate(&e) != hipSuccess) { delete p; return nullptr; }
    if (hipStreamCreateWithFlags(&p->xs, hipStreamNonBlocking) != hipSuccess) { delete p; return nullptr; }
    if (rocblas_create_handle(&p->xh) != rocblas_status_success) { delete p; return nullptr; }
    rocblas_set_stream(p->xh, p->xs);
    rocblas_set_pointer_mode(p->xh, rocblas_pointer_mode_host);
    for (int e = 0; e < 2; e++)
        if (hipEventCreateWithFlags(&p->xev[e], hipEventDisableTiming) != hipSuccess) { delete p; return nullptr; }
    // pair tables: core i<=j, core i<j, virt r<=s, virt r<s, active upper pairs
    std::vector<int2> pc, pcs, pv, pvs, up, pcd, pvd;
    for (int i = 0; i < in.nc; i++)
        for (int j = i; j < in.nc; j++) { pc.push_back({i, j}); if (j > i) pcs.push_back({i, j}); }
    for (int r = 0; r < in.nv; r++)
        for (int s = r; s < in.nv; s++) { pv.push_back({r, s}); if (s > r) pvs.push_back({r, s}); }
    for (int i = 0; i < in.nc; i++) pcd.push_back({i, i});
    for (int r = 0; r < in.nv; r++) pvd.push_back({r, r});
    for (int a = 0; a < in.n; a++)
        for (int b = a; b < in.n; b++) up.push_back({a, b});
    auto upl = [&](std::vector<int2> &v, int2 **d) {
        size_t b = sizeof(int2) * (v.size() ? v.size() : 1);
        if (hipMalloc(d, b) != hipSuccess) return false;
        if (v.size()) hipMemcpy(*d, v.data(), sizeof(int2) * v.size(), hipMemcpyHostToDevice);
        return true;
    };
    upl(pc, &p->pairs_c); upl(pcs, &p->pairs_cs); upl(pv, &p->pairs_v); upl(pvs, &p->pairs_vs); upl(up, &p->up);
    upl(pcd, &p->pairs_cd); upl(pvd, &p->pairs_vd);
    p->np_c = pc.size(); p->np_cs = pcs.size(); p->np_v = pv.size(); p->np_vs = pvs.size();
    return p;
}

void stage2_destroy(Stage2 *p) {
    if (!p) return;
    pm1p_stop(p);
    delete p->early;
    if (p->evS) (void)hipEventDestroy(p->evS);
    if (p->evK) (void)hipEventDestroy(p->evK);
    if (p->evX) (void)hipEventDestroy(p->evX);
    if (p->evQt) (void)hipEventDestroy(p->evQt);
    for (auto e : p->evKr) if (e) (void)hipEventDestroy(e);
    p->sync_own();      // this evaluation's streams only; other contexts keep running
    for (auto &pl : p->sbplan) syevb_plan_destroy(pl);
    for (auto &kv : p->bufs) if (kv.second.p) (void)hipFree(kv.second.p);
    (void)hipFree(p->dinfo); (void)hipFree(p->part); (void)hipFree(p->acc); (void)hipFree(p->accq);
    (void)hipFree(p->qpart); (void)hipFree(p->dinfo_w);
    if (p->hinfo) (void)hipHostFree(p->hinfo);
    if (p->hpin) (void)hipHostFree(p->hpin);
    for (int w = 0; w < NW; w++) {
        if (p->wh[w]) rocblas_destroy_handle(p->wh[w]);
        if (p->ws[w]) (void)hipStreamDestroy(p->ws[w]);
    }
    for (int e = 0; e < NW + 1; e++) if (p->wev[e]) (void)hipEventDestroy(p->wev[e]);
    if (p->xh) rocblas_destroy_handle(p->xh);
    if (p->xs) (void)hipStreamDestroy(p->xs);
    for (int e = 0; e < 2; e++) if (p->xev[e]) (void)hipEventDestroy(p->xev[e]);
    (void)hipFree(p->pairs_c); (void)hipFree(p->pairs_cs); (void)hipFree(p->pairs_v); (void)hipFree(p->pairs_vs);
    (void)hipFree(p->pairs_cd); (void)hipFree(p->pairs_vd);
    (void)hipFree(p->up);
    delete p;
}

static int slice(Stage2 *p, const double *src, int s, i64 cnt, double *dst) {
    hipLaunchKernelGGL(k_slice, dim3(nb(cnt, 256)), dim3(256), 0, p->st, src, p->M, s, cnt, dst);
    NV2_LAUNCH_CHECK();
    return 0;
}

// state-independent unpacking of the Cholesky blocks and h
static int prepare_blocks(Stage2 *p) {
    if (p->blocks_ready) return 0;
    const int nc = p->nc, n = p->n, nv = p->nv, L = p->L, norb = p->norb;
    const int c0 = 0, a0 = nc, v0 = nc + n;
    struct { const char *nm; int p0, np, q0, nq; } blk[] = {
        {"Bcc", c0, nc, c0, nc}, {"Bcv", c0, nc, v0, nv}, {"Bca", c0, nc, a0, n},
        {"Bac", a0, n, c0, nc},  {"Bav", a0, n, v0, nv},  {"Baa", a0, n, a0, n}};
    for (auto &b : blk) {
        i64 cnt = (i64)L * b.np * b.nq;
        GET(d, b.nm, cnt);
        if (cnt) {
            hipLaunchKernelGGL(k_unpack, dim3(nb(cnt, 256)), dim3(256), 0, p->st, p->in.B, L, b.p0, b.np, b.q0, b.nq, d);
            NV2_LAUNCH_CHECK();
        }
    }
    GET(hf, "hfull", (i64)norb * norb);
    hipLaunchKernelGGL(k_unpack_h, dim3(nb((i64)norb * norb, 256)), dim3(256), 0, p->st, p->in.oneint, norb, hf);
    NV2_LAUNCH_CHECK();
    p->blocks_ready = true;
    return 0;
}

static double *qd_fill(Stage2 *p, const Sl &S, int kind, int lt, i64 total, const std::string &name) {
    double *b = p->get(name, total);
    if (!b) { set_error("hipMalloc failed for %s", name.c_str()); return nullptr; }
    if (total > 0) {
        hipLaunchKernelGGL(k_qd_fill, dim3(nb(total, 256)), dim3(256), 0, p->st, S, kind, lt, p->up, b, total);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { set_error("k_qd_fill: %s", hipGetErrorString(e)); return nullptr; }
    }
    return b;
}
#define QFILL(var, kind, j, total, name) \
    double *var = qd_fill(p, Sj[j], (kind), (j) < s, (total), std::string(name) + std::to_string(j)); \
    if (!var) return -1;
#define QGET(var, name, j, cnt) GET(var, (std::string(name) + std::to_string(j)).c_str(), (cnt))

// NV2_STAGE2_TIMING=1 prints, per state, the host-side milestones of the solve layout to stderr
static bool stage2_timing() {
    static const bool on = getenv("NV2_STAGE2_TIMING") && atoi(getenv("NV2_STAGE2_TIMING")) > 0;
    return on;
}
// The SC sums other than (+-1)' wait until the (+-1)' K' band reduction has finished: they otherwise
// share the device with its chain of small launches and stretch it.  The (+-1)' SC sums, which also
// form the right-hand sides of the chain's PC sums, run before the wait.  NV2_HOLD_SC=0 starts every
// SC sum as soon as the integrals exist.
static bool hold_sc() {
    const char *e = getenv("NV2_HOLD_SC");
    return !(e && atoi(e) == 0);
}
static double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

int stage2_state(Stage2 *p, int s, double thr, int prec, double *out, S2Stats *stats, double *qd_out) {
    const auto t_begin = std::chrono::steady_clock::now();
    double ms_workers = 0.0, ms_wdone[NW] = {}, ms_scq = 0.0, ms_join = 0.0;
    (void)prec;
    const int nc = p->nc, n = p->n, nv = p->nv, L = p->L, norb = p->norb, M = p->M;
    const i64 n2 = (i64)n * n, n4 = n2 * n2, n6 = n4 * n2;
    hipStream_t st = p->st;
    RB(rocblas_set_stream(p->h, st));
    RB(rocblas_set_pointer_mode(p->h, rocblas_pointer_mode_host));
    TRYS(prepare_blocks(p));
    NV2_HIP(hipMemsetAsync(p->acc, 0, sizeof(double) * 32, st));
    S2Stats sd{};
    // ---- per-state slices
    GET(da, "da", n2); GET(taa, "taa", n4); GET(ro3, "ro3", n6);
    TRYS(slice(p, p->in.dal, s, n2, da)); TRYS(slice(p, p->in.taa, s, n4, taa)); TRYS(slice(p, p->in.daaa, s, n6, ro3));
    Sl S{n, da, taa, ro3};
    // ---- .NATORB: the first-order density blocks of (0)', (+1)' and (-1)' (zero where a class is absent)
    double *rho_cv = nullptr, *rho_ca = nullptr, *rho_av = nullptr;
    if (p->rho1) {
        rho_cv = p->get("rho1_cv", (i64)nc * nv);
        rho_ca = p->get("rho1_ca", (i64)nc * n);
        rho_av = p->get("rho1_av", (i64)n * nv);
        if (!rho_cv || !rho_ca || !rho_av) { set_error("hipMalloc failed for the .NATORB blocks"); return -1; }
        NV2_HIP(hipMemsetAsync(rho_cv, 0, sizeof(double) * ((i64)nc * nv > 0 ? (i64)nc * nv : 1), st));
        NV2_HIP(hipMemsetAsync(rho_ca, 0, sizeof(double) * ((i64)nc * n > 0 ? (i64)nc * n : 1), st));
        NV2_HIP(hipMemsetAsync(rho_av, 0, sizeof(double) * (i64)n * nv, st));
    }
    // ---- Koopmans matrices of this state
    GET(k_eako1, "k_eako1", n2); GET(k_ipko1, "k_ipko1", n2); GET(k_hrdm, "k_hrdm", n4); GET(k_eako2, "k_eako2", n4);
    GET(k_ipko2, "k_ipko2", n4); GET(k_ga, "k_ga", n4); GET(k_gd, "k_gd", n4); GET(k_gf, "k_gf", n2);
    TRYS(slice(p, p->in.eako1, s, n2, k_eako1)); TRYS(slice(p, p->in.ipko1, s, n2, k_ipko1));
    TRYS(slice(p, p->in.hrdm, s, n4, k_hrdm)); TRYS(slice(p, p->in.eako2, s, n4, k_eako2));
    TRYS(slice(p, p->in.ipko2, s, n4, k_ipko2)); TRYS(slice(p, p->in.gk0pa, s, n4, k_ga));
    TRYS(slice(p, p->in.gk0pd, s, n4, k_gd)); TRYS(slice(p, p->in.gk0pf, s, n2, k_gf));
    // ---- QD: transition RDMs of every pair (s, j), j != s (pair index of (max,min), qdnevpt.F:1912)
    QDArgs q{};
    std::vector<Sl> Sj(QM);
    const bool qd = qd_out && M > 1 && p->in.trdm1 && p->in.trdm2 && p->in.trdm3;
    if (qd) {
        if (M > QM) { set_error("QD-NEVPT2: at most %d states are supported (metat=%d)", QM, M); return -2; }
        q.M = M;
        q.self = s;
        NV2_HIP(hipMemsetAsync(p->accq, 0, sizeof(double) * 8 * 2 * QM, st));
        const int P = M * (M - 1) / 2;
        for (int j = 0; j < M; j++) {
            if (j == s) continue;
            const int hi = j > s ? j : s, lo = j > s ? s : j, pr = hi * (hi - 1) / 2 + lo;
            QGET(t1, "qd_d1_", j, n2); QGET(t2, "qd_d2_", j, n4); QGET(t3, "qd_d3_", j, n6);
            hipLaunchKernelGGL(k_slice, dim3(nb(n2, 256)), dim3(256), 0, st, p->in.trdm1, P, pr, n2, t1);
            hipLaunchKernelGGL(k_slice, dim3(nb(n4, 256)), dim3(256), 0, st, p->in.trdm2, P, pr, n4, t2);
            hipLaunchKernelGGL(k_slice, dim3(nb(n6, 256)), dim3(256), 0, st, p->in.trdm3, P, pr, n6, t3);
            NV2_LAUNCH_CHECK();
            Sj[j] = Sl{n, t1, t2, t3};
        }
    }

    // ================================================================ PC metric problems
    // The metric and Koopmans matrices of every PC class are filled on the main stream; their
    // eigen-solves (chains of small latency-bound solver kernels) then run concurrently on the
    // worker streams while the SC sums proceed on the main stream.
    // batched pairs: (+1)/(-1) n, (+2)/(-2) n^2, (+2)/(-2) diagonal pairs n(n+1)/2, (0)' alone 2n^2,
    // (+1)'/(-1)' n^3; problem b of a job: b = 0 for the "+" class, 1 for the "-" class
    enum { JPM1, JPM2, JPM2P, J0P, JPM1P, NJOB };
    PcJob J[NJOB];
    const int m = n * (n + 1) / 2;
    const bool has4 = !p->no4 && p->in.amatIP && p->in.amatEA;
    const i64 N3 = n2 * n;
    // the (+-1)' chain runs on worker 1 from its own start (stage2_pm1p_begin, possibly called
    // right after the partial traces); every return from here on stops it if it still waits
    if (has4) TRYS(stage2_pm1p_begin(p, s, thr));
    struct EarlyGuard {
        Stage2 *p;
        ~EarlyGuard() { pm1p_stop(p); }
    } eguard{p};
    {
        const char *tags[NJOB] = {"pm1", "pm2", "pm2p", "v0p", "pm1p"};
        const int Ns[NJOB] = {n, (int)n2, m, 2 * (int)n2, has4 ? (int)N3 : 0};
        for (int k = 0; k < NJOB; k++) {
            if (k == JPM1P && has4) {
                J[k] = p->early->job;          // its fields are final once worker 1 is joined
                continue;
            }
            J[k].tag = tags[k];
            J[k].N = Ns[k];
            J[k].nb = k == J0P ? 1 : 2;
            J[k].slot = k;
            J[k].w = k == JPM1P ? 1 : 0;
            J[k].trd = k == JPM1P;
            GET(Sm, (J[k].tag + "_S").c_str(), (i64)Ns[k] * Ns[k] * J[k].nb);
            GET(Km, (J[k].tag + "_K").c_str(), (i64)Ns[k] * Ns[k] * J[k].nb);
            J[k].S = Sm;
            J[k].K = Km;
        }
    }
    // (+1): S = 2 - da, K = 1-EAKO/2 ; (-1): S = da, K = 1-IPKO/2
    GET(ro_v1, "ro_v1", n2);
    hipLaunchKernelGGL(k_fill_diag_minus, dim3(nb(n2, 256)), dim3(256), 0, st, da, n, 2.0, J[JPM1].Sb(0));
    hipLaunchKernelGGL(k_fill_diag_minus, dim3(nb(n2, 256)), dim3(256), 0, st, da, n, 2.0, ro_v1);
    hipLaunchKernelGGL(k_scale, dim3(nb(n2, 256)), dim3(256), 0, st, k_eako1, n2, 0.5, J[JPM1].Kb(0));
    NV2_HIP(hipMemcpyAsync(J[JPM1].Sb(1), da, sizeof(double) * n2, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_scale, dim3(nb(n2, 256)), dim3(256), 0, st, k_ipko1, n2, 0.5, J[JPM1].Kb(1));
    NV2_LAUNCH_CHECK();
    // (+-2): fillcm2 / fillcm2p of (2-HRDM, 2-EAKO) and (2-RDM, 2-IPKO)
    double *Xm2[2], *Xm2p[2];
    for (int mode = 0; mode < 2; mode++) {
        const double *X = mode == 0 ? k_hrdm : taa;
        const double *K = mode == 0 ? k_eako2 : k_ipko2;
        PcJob &a = J[JPM2], &b = J[JPM2P];
        GET(xm, mode == 0 ? "pc_Xm_p2" : "pc_Xm_m2", n4);
        GET(xmp, mode == 0 ? "pc_Xmp_p2" : "pc_Xmp_m2", (i64)m * m);
        hipLaunchKernelGGL(k_fillcm2, dim3(nb(n4, 256)), dim3(256), 0, st, X, n, a.Sb(mode));
        hipLaunchKernelGGL(k_fillcm2, dim3(nb(n4, 256)), dim3(256), 0, st, K, n, a.Kb(mode));
        hipLaunchKernelGGL(k_fillcm2, dim3(nb(n4, 256)), dim3(256), 0, st, X, n, xm);
        hipLaunchKernelGGL(k_fillcm2p, dim3(nb((i64)m * m, 256)), dim3(256), 0, st, X, n, p->up, b.Sb(mode));
        hipLaunchKernelGGL(k_fillcm2p, dim3(nb((i64)m * m, 256)), dim3(256), 0, st, K, n, p->up, b.Kb(mode));
        hipLaunchKernelGGL(k_fillcm2p, dim3(nb((i64)m * m, 256)), dim3(256), 0, st, X, n, p->up, xmp);
        NV2_LAUNCH_CHECK();
        Xm2[mode] = xm;
        Xm2p[mode] = xmp;
    }
    // (0)': fillc0p
    hipLaunchKernelGGL(k_fillc0p, dim3(nb(4 * n4, 256)), dim3(256), 0, st, S, k_ga, k_gd, J[J0P].S, J[J0P].K);
    NV2_LAUNCH_CHECK();
    // (+-1)': n^3 metric (eee2 / eee2t) and amat, with a copy of the metric kept for the sums
    double *Mx[2] = {nullptr, nullptr}, *amk[2] = {nullptr, nullptr}, *bmk[2] = {nullptr, nullptr};
    double *cmk[2] = {nullptr, nullptr}, *dmk[2] = {nullptr, nullptr};
    if (has4) {
        for (int plus = 0; plus < 2; plus++) {
            PcJob &jb = J[JPM1P];
            const int bb = plus ? 0 : 1;
            const std::string t = plus ? "_p" : "_m";
            GET(am, ("k_amat" + t).c_str(), n6); GET(bm, ("k_bmat" + t).c_str(), n4);
            GET(cm, ("k_cmat" + t).c_str(), n4); GET(dm, ("k_dmat" + t).c_str(), n2);
            TRYS(slice(p, plus ? p->in.amatEA : p->in.amatIP, s, n6, am));
            TRYS(slice(p, plus ? p->in.bmatEA : p->in.bmatIP, s, n4, bm));
            TRYS(slice(p, plus ? p->in.cmatEA : p->in.cmatIP, s, n4, cm));
            TRYS(slice(p, plus ? p->in.dmatEA : p->in.dmatIP, s, n2, dm));
            GET(mx, ("pm1_M" + t).c_str(), N3 * N3);    // filled by stage2_pm1p_begin
            if (!p->early->k_filled) {
                hipLaunchKernelGGL(k_fill_pm1p_k, dim3(nb(N3 * N3, 256)), dim3(256), 0, st, n, am, jb.Kb(bb));
                NV2_LAUNCH_CHECK();
            }
            Mx[plus] = mx; amk[plus] = am; bmk[plus] = bm; cmk[plus] = cm; dmk[plus] = dm;
        }
        if (!p->early->k_filled) {
            NV2_HIP(hipEventRecord(p->evK, st));         // K is there: the chain may form K'
            {
                std::lock_guard<std::mutex> lk(p->early->mu);
                p->early->k_ready = true;
            }
            p->early->cv.notify_all();
        }
    }
    // start the eigen-solves on the worker threads/streams once the fills are done
    {
        size_t hoff = 0;
        int maxN[NW] = {};
        for (int k = 0; k < NJOB; k++) {
            if (k == JPM1P && has4) continue;          // allocated by stage2_pm1p_begin
            TRYS(pc_job_buffers(p, J[k], hoff));
            hoff += (size_t)J[k].N * J[k].nb;
            if (J[k].N * J[k].nb > maxN[J[k].w]) maxN[J[k].w] = J[k].N * J[k].nb;
        }
        for (int w = 0; w < NW; w++) {
            GET(E, ("syevd_E" + std::to_string(w)).c_str(), maxN[w] ? maxN[w] : 1);
            p->wE[w] = E;
        }
    }
    if (stage2_timing()) {           // fills done on the device
        NV2_HIP(hipStreamSynchronize(st));
        ms_workers = ms_since(t_begin);
    }
    NV2_HIP(hipEventRecord(p->wev[NW], st));
    for (int w = 0; w < NW; w++) NV2_HIP(hipStreamWaitEvent(p->ws[w], p->wev[NW], 0));
    int wrc[NW] = {};
    char werr[NW][512] = {};        // a worker's error message (the message buffer is per thread)
    std::vector<std::thread> workers;
    struct Joiner {     // every exit from here on waits for the worker threads first
        std::vector<std::thread> &v;
        ~Joiner() { for (auto &t : v) if (t.joinable()) t.join(); }
    } joiner{workers};
    for (int w = 0; w < (has4 ? 1 : NW); w++)          // with the (+-1)' chain, worker 1 is its thread
        workers.emplace_back([p, &J, w, thr, &wrc, &werr, &ms_wdone, t_begin]() {
            try {
                wrc[w] = pc_worker(p, J, NJOB, w, thr, &ms_wdone[w], t_begin);
            } catch (const std::exception &x) {
                set_error("stage-2 worker %d: %s", w, x.what());
                wrc[w] = -1;
            } catch (...) {
                set_error("stage-2 worker %d: unknown exception", w);
                wrc[w] = -1;
            }
            if (wrc[w]) nv2_last_error(werr[w], sizeof(werr[w]));
        });

    // ---- Fock operators and canonical orbitals (qdnevpt.F:369-698)
    GET(Jc, "Jc", L); GET(Ja, "Ja", L);
    hipLaunchKernelGGL(k_jvec, dim3(nb(L, 256)), dim3(256), 0, st, p->in.B, L, nc, n, da, Jc, Ja);
    NV2_LAUNCH_CHECK();
    double *hf = p->get("hfull", (i64)norb * norb);
    GET(eps, "eps", norb);
    NV2_HIP(hipMemsetAsync(eps, 0, sizeof(double) * norb, st));
    GET(Cc, "Ccore", (i64)nc * nc);
    GET(Cv, "Cvirt", (i64)nv * nv);
    GET(Jt, "Jtot", L);
    double *Bcc = p->get("Bcc", 0), *Bcv = p->get("Bcv", 0), *Bca = p->get("Bca", 0), *Bac = p->get("Bac", 0);
    double *Bav = p->get("Bav", 0), *Baa = p->get("Baa", 0);
    // Jt = 2 Jcore + Jact (Coulomb of both core and active densities)
    {
        const double one = 1.0, two = 2.0;
        NV2_HIP(hipMemcpyAsync(Jt, Ja, sizeof(double) * L, hipMemcpyDeviceToDevice, st));
        if (nc) RB(rocblas_daxpy(p->h, L, &two, Jc, 1, Jt, 1));
        (void)one;
    }
    if (nc) {
        NV2_HIP(hipMemcpy2DAsync(Cc, sizeof(double) * nc, hf, sizeof(double) * norb, sizeof(double) * nc, nc,
                                 hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_coul, dim3(nb((i64)nc * nc, 4)), dim3(256), 0, st, p->in.B, L, 0, nc, Jt, Cc);
        NV2_LAUNCH_CHECK();
        TRYS(gemm(p, true, false, nc, nc, L * nc, -1.0, Bcc, (i64)L * nc, Bcc, (i64)L * nc, 1.0, Cc, nc));
        GET(G, "fockG", (i64)L * n * (nc > nv ? nc : nv));
        TRYS(gemm_sb(p, L, n, n, Bac, L, (i64)L * n, da, n, 0, G, L, (i64)L * n, nc));
        TRYS(gemm(p, true, false, nc, nc, L * n, -0.5, G, (i64)L * n, Bac, (i64)L * n, 1.0, Cc, nc));
    }
    GET(wc, "wcore", nc);
    {
        NV2_HIP(hipMemcpy2DAsync(Cv, sizeof(double) * nv, hf + (i64)(nc + n) * norb + nc + n, sizeof(double) * norb,
                                 sizeof(double) * nv, nv, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_coul, dim3(nb((i64)nv * nv, 4)), dim3(256), 0, st, p->in.B, L, nc + n, nv, Jt, Cv);
        NV2_LAUNCH_CHECK();
        if (nc) TRYS(gemm(p, true, false, nv, nv, L * nc, -1.0, Bcv, (i64)L * nc, Bcv, (i64)L * nc, 1.0, Cv, nv));
        GET(G, "fockG", (i64)L * n * (nc > nv ? nc : nv));
        TRYS(gemm_sb(p, L, n, n, Bav, L, (i64)L * n, da, n, 0, G, L, (i64)L * n, nv));
        TRYS(gemm(p, true, false, nv, nv, L * n, -0.5, G, (i64)L * n, Bav, (i64)L * n, 1.0, Cv, nv));
    }
    // canonical core and virtual orbitals: dsyev('V','L') of both Fock blocks (qdnevpt.F:519, 671), as one
    // batched solve when both orders allow it
    GET(wv, "wvirt", nv);
    if (syevb_on() && nc <= SYEVB_NMAX && nv <= SYEVB_NMAX) {
        GET(fw, "fock_sbw", syevb_workspace(nc) + syevb_workspace(nv));
        SyevbProb fp[2] = {SyevbProb{Cc, wc, fw, nc, 1}, SyevbProb{Cv, wv, fw + syevb_workspace(nc), nv, 1}};
        TRYS(syevb(p->h, st, 2, fp, p->plan(2)));
    } else {
        if (nc) TRYS(syevd(p, false, nc, Cc, wc));
        TRYS(syevd(p, false, nv, Cv, wv));
    }
    if (nc) {
        NV2_HIP(hipMemcpyAsync(eps, wc, sizeof(double) * nc, hipMemcpyDeviceToDevice, st));
        if (!p->frozen.empty()) {   // .FROZEN: these canonical core orbitals are never holes
            const int nf = (int)p->frozen.size();
            int *fi = (int *)p->get("frozen_idx", (nf + 1) / 2);
            if (!fi) { set_error("hipMalloc failed for frozen_idx"); return -1; }
            NV2_HIP(hipMemcpyAsync(fi, p->frozen.data(), sizeof(int) * nf, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_freeze, dim3(1), dim3(256), 0, st, eps, fi, nf);
            NV2_LAUNCH_CHECK();
        }
    }
    NV2_HIP(hipMemcpyAsync(eps + nc + n, wv, sizeof(double) * nv, hipMemcpyDeviceToDevice, st));
    const double *ec = eps, *ev = eps + nc + n;
    // ---- Cholesky transform (qdnevpt.F:7599-7717)
    GET(T1, "T1", (i64)L * nc * (nv > nc ? nv : nc));
    GET(Tcc, "Tcc", (i64)L * nc * nc); GET(Tcv, "Tcv", (i64)L * nc * nv); GET(Tca, "Tca", (i64)L * nc * n);
    GET(Tav, "Tav", (i64)L * n * nv);
    if (nc) {
        TRYS(gemm(p, false, false, L * nc, nc, nc, 1.0, Bcc, (i64)L * nc, Cc, nc, 0.0, T1, (i64)L * nc));
        TRYS(gemm_sb(p, L, nc, nc, T1, L, (i64)L * nc, Cc, nc, 0, Tcc, L, (i64)L * nc, nc));
        TRYS(gemm(p, false, false, L * nc, nv, nv, 1.0, Bcv, (i64)L * nc, Cv, nv, 0.0, T1, (i64)L * nc));
        TRYS(gemm_sb(p, L, nc, nc, T1, L, (i64)L * nc, Cc, nc, 0, Tcv, L, (i64)L * nc, nv));
        TRYS(gemm_sb(p, L, nc, nc, Bca, L, (i64)L * nc, Cc, nc, 0, Tca, L, (i64)L * nc, n));
    }
    TRYS(gemm(p, false, false, L * n, nv, nv, 1.0, Bav, (i64)L * n, Cv, nv, 0.0, Tav, (i64)L * n));
    const double *Taa = Baa;
    // ---- integral blocks: I_xy[u,v] = sum_L T_x[L,u] T_y[L,v]
    const i64 Ncv = (i64)nc * nv, Nca = (i64)nc * n, Nav = (i64)n * nv, Naa = n2;
    GET(Icacv, "Icacv", Nca * Ncv); GET(Icvav, "Icvav", Ncv * Nav);
    GET(Icaca, "Icaca", Nca * Nca); GET(Iavav, "Iavav", Nav * Nav); GET(Icvaa, "Icvaa", Ncv * Naa);
    GET(Icaav, "Icaav", Nca * Nav); GET(Iaaav, "Iaaav", Naa * Nav); GET(Iaaca, "Iaaca", Naa * Nca);
    // class (0) needs (i r | j s) for i <= j only: per core orbital i, the blocks of j = i .. nc-1 in one
    // strided-batched GEMM, stored pair-major in pairs_c order
    const i64 nv2 = (i64)nv * nv;
    GET(V0b, "v0_pairblocks", (p->np_c ? p->np_c : 1) * nv2);
    if (nc && nv) {
        const double one = 1.0, zero = 0.0;
        i64 p0 = 0;
        for (int i = 0; i < nc; i++) {
            RB(rocblas_dgemm_strided_batched(p->h, rocblas_operation_transpose, rocblas_operation_none, nv, nv, L, &one,
                                             Tcv + (i64)L * i, L * nc, 0, Tcv + (i64)L * i, L * nc, L, &zero,
                                             V0b + p0 * nv2, nv, nv2, nc - i));
            p0 += nc - i;
        }
    }
    TRYS(gemm(p, true, false, Nca, Ncv, L, 1.0, Tca, L, Tcv, L, 0.0, Icacv, Nca));
    TRYS(gemm(p, true, false, Ncv, Nav, L, 1.0, Tcv, L, Tav, L, 0.0, Icvav, Ncv));
    TRYS(gemm(p, true, false, Nca, Nca, L, 1.0, Tca, L, Tca, L, 0.0, Icaca, Nca));
    TRYS(gemm(p, true, false, Nav, Nav, L, 1.0, Tav, L, Tav, L, 0.0, Iavav, Nav));
    TRYS(gemm(p, true, false, Ncv, Naa, L, 1.0, Tcv, L, Taa, L, 0.0, Icvaa, Ncv));
    TRYS(gemm(p, true, false, Nca, Nav, L, 1.0, Tca, L, Tav, L, 0.0, Icaav, Nca));
    TRYS(gemm(p, true, false, Naa, Nav, L, 1.0, Taa, L, Tav, L, 0.0, Iaaav, Naa));
    TRYS(gemm(p, true, false, Naa, Nca, L, 1.0, Taa, L, Tca, L, 0.0, Iaaca, Naa));
    // ---- one-electron integrals in the canonical bases and the Dyall-type heff vectors
    GET(Jc2, "Jc2", L);
    if (nc) {
        hipLaunchKernelGGL(k_jdiag, dim3(nb(L, 256)), dim3(256), 0, st, Tcc, L, nc, Jc2);
        NV2_LAUNCH_CHECK();
    } else NV2_HIP(hipMemsetAsync(Jc2, 0, sizeof(double) * L, st));
    GET(tmp, "h_tmp", (i64)norb * norb);
    // heff0p[k,r] = (C^T h_cv V)[k,r] + 2 (Jc2^T Tcv)[k,r] - sum_{L,j} Tcc[L,j,k] Tcv[L,j,r]
    GET(h0p, "heff0p", Ncv);
    GET(hm1p, "heffm1p", Nav);
    GET(hp1p, "heffp1p", Nca);
    if (nc) {
        TRYS(gemm(p, false, false, nc, nv, nv, 1.0, hf + (i64)(nc + n) * norb, norb, Cv, nv, 0.0, tmp, nc));
        TRYS(gemm(p, true, false, nc, nv, nc, 1.0, Cc, nc, tmp, nc, 0.0, h0p, nc));
        TRYS(gemm(p, true, false, 1, (int)Ncv, L, 2.0, Jc2, L, Tcv, L, 1.0, h0p, 1));
        TRYS(gemm(p, true, false, nc, nv, L * nc, -1.0, Tcc, (i64)L * nc, Tcv, (i64)L * nc, 1.0, h0p, nc));
    }
    // heffm1p[a,r] = (h_av V)[a,r] + 2 (Jc2^T Tav)[a,r] - sum_{L,j} Tca[L,j,a] Tcv[L,j,r] - sum_{L,b} Taa[L,b,a] Tav[L,b,r]
    TRYS(gemm(p, false, false, n, nv, nv, 1.0, hf + (i64)(nc + n) * norb + nc, norb, Cv, nv, 0.0, hm1p, n));
    TRYS(gemm(p, true, false, 1, (int)Nav, L, 2.0, Jc2, L, Tav, L, 1.0, hm1p, 1));
    if (nc) TRYS(gemm(p, true, false, n, nv, L * nc, -1.0, Tca, (i64)L * nc, Tcv, (i64)L * nc, 1.0, hm1p, n));
    TRYS(gemm(p, true, false, n, nv, L * n, -1.0, Taa, (i64)L * n, Tav, (i64)L * n, 1.0, hm1p, n));
    // heffp1p[i,a] = (C^T h_ca)[i,a] + 2 (Jc2^T Tca)[i,a] - sum_{L,j} Tcc[L,j,i] Tca[L,j,a]
    if (nc) {
        TRYS(gemm(p, true, false, nc, n, nc, 1.0, Cc, nc, hf + (i64)nc * norb, norb, 0.0, hp1p, nc));
        TRYS(gemm(p, true, false, 1, (int)Nca, L, 2.0, Jc2, L, Tca, L, 1.0, hp1p, 1));
        TRYS(gemm(p, true, false, nc, n, L * nc, -1.0, Tcc, (i64)L * nc, Tca, (i64)L * nc, 1.0, hp1p, nc));
    }
    // ================================================================ (-1)' and (+1)' SC (main stream)
    // first, and not held: the right-hand sides X of the (+-1)' PC sums are ready when the chain needs them
    double *Vpm[2] = {nullptr, nullptr}, *Hpm[2] = {nullptr, nullptr}, *Q2pm[2] = {nullptr, nullptr};
    double *Xpm[2] = {nullptr, nullptr};
    double *Trm[2] = {nullptr, nullptr}, *Vrm[2] = {nullptr, nullptr};   // .NATORB density vectors, projections
    for (int plus = 0; has4 && plus < 2; plus++) {
        const int cls = plus ? 5 : 6;
        const int next = plus ? nc : nv;
        const double *eo = plus ? ec : ev;
        const std::string t = plus ? "_p" : "_m";
        PcJob &jb = J[JPM1P];
        const int bb = plus ? 0 : 1;
        GET(Q2, ("pm1_Q2" + t).c_str(), N3 * n); GET(BC, ("pm1_BC" + t).c_str(), N3 * n);
        hipLaunchKernelGGL(k_fill_pm1p_q, dim3(nb(N3 * n, 256)), dim3(256), 0, st, S, plus, bmk[plus], cmk[plus], Q2, BC);
        NV2_LAUNCH_CHECK();
        Q2pm[plus] = Q2;
        if (!next) continue;
        if (p->rho1) {
            GET(Tr, ("rho1_T" + t).c_str(), N3 * n); GET(Vr, ("rho1_V" + t).c_str(), N3 * n);
            hipLaunchKernelGGL(k_fill_rho_t, dim3(nb(N3 * n, 256)), dim3(256), 0, st, S, plus, Tr);
            NV2_LAUNCH_CHECK();
            Trm[plus] = Tr;
            Vrm[plus] = Vr;
        }
        GET(V, ("pm1_V" + t).c_str(), N3 * next);
        hipLaunchKernelGGL(k_fill_pm1p_v, dim3(nb(N3 * next, 256)), dim3(256), 0, st, plus ? Iaaca : Iaaav, plus, n, nc,
                           next, V);
        NV2_LAUNCH_CHECK();
        double *H = hm1p;
        if (plus) {   // H[a,i] = heffp1p[i,a]
            GET(Hp, "pm1_H_p", (i64)n * next);
            hipLaunchKernelGGL(k_transpose, dim3(nb((i64)n * nc, 256)), dim3(256), 0, st, hp1p, nc, n, Hp);
            NV2_LAUNCH_CHECK();
            H = Hp;
        }
        Vpm[plus] = V;
        Hpm[plus] = H;
        GET(ro1, ("pm1_ro" + t).c_str(), n2);
        hipLaunchKernelGGL(k_fill_diag_minus, dim3(nb(n2, 256)), dim3(256), 0, st, da, n, plus ? 2.0 : 0.0, ro1);
        NV2_LAUNCH_CHECK();
        GET(W1, ("pm1_W1" + t).c_str(), (i64)n * next); GET(W2, ("pm1_W2" + t).c_str(), (i64)n * next);
        GET(Z1, ("pm1_Z1" + t).c_str(), N3 * next); GET(Z2, ("pm1_Z2" + t).c_str(), N3 * next);
        GET(Y1, ("pm1_Y1" + t).c_str(), N3 * next); GET(Y2, ("pm1_Y2" + t).c_str(), N3 * next);
        TRYS(gemm(p, false, false, n, next, n, 1.0, ro1, n, H, n, 0.0, W1, n));
        TRYS(gemm(p, false, false, n, next, n, 1.0, dmk[plus], n, H, n, 0.0, W2, n));
        TRYS(gemm(p, false, false, (int)N3, next, n, 1.0, Q2, N3, H, n, 0.0, Z1, N3));
        TRYS(gemm(p, false, false, (int)N3, next, n, 1.0, BC, N3, H, n, 0.0, Z2, N3));
        TRYS(gemm(p, false, false, (int)N3, next, (int)N3, 1.0, Mx[plus], N3, V, N3, 0.0, Y1, N3));   // M^T v
        TRYS(gemm(p, false, false, (int)N3, next, (int)N3, 1.0, jb.Kb(bb), N3, V, N3, 0.0, Y2, N3));  // Am^T v
        // PC right-hand sides X = M^T V + Q2 H (the S-build of qdnevpt.F:5277-5324 applied to the integrals)
        GET(Xpc, ("pm1_X" + t).c_str(), N3 * next);
        TRYS(gemm(p, true, false, (int)N3, next, (int)N3, 1.0, Mx[plus], N3, V, N3, 0.0, Xpc, N3));
        TRYS(gemm(p, false, false, (int)N3, next, n, 1.0, Q2, N3, H, n, 1.0, Xpc, N3));
        Xpm[plus] = Xpc;
        for (int j = 0; q.M && j < M; j++) {
            if (j == s) continue;
            QFILL(r1j, plus ? QK_P1P_R1 : QK_M1P_R1, j, n2, "qd_pm1r1_");
            QFILL(qj, plus ? QK_P1P_QSC : QK_M1P_QSC, j, N3 * n, "qd_pm1q_");
            double *mj = qd_fill(p, Sj[j], plus ? QK_P1P_MSC : QK_M1P_MSC, j < s, N3 * N3, "qd_pm1M");
            if (!mj) return -1;
            QGET(w1j, "qd_pm1w1_", j, (i64)n * next); QGET(z1j, "qd_pm1z1_", j, N3 * next);
            QGET(y1j, "qd_pm1y1_", j, N3 * next);
            TRYS(gemm(p, false, false, n, next, n, 1.0, r1j, n, H, n, 0.0, w1j, n));
            TRYS(gemm(p, false, false, (int)N3, next, n, 1.0, qj, N3, H, n, 0.0, z1j, N3));
            TRYS(gemm(p, false, false, (int)N3, next, (int)N3, 1.0, mj, N3, V, N3, 0.0, y1j, N3));
            q.a[j] = w1j; q.b[j] = z1j; q.c[j] = y1j;
        }
        double *pt = parts(p, next);
        TRYS(qd_prepare(p, q, next));
        hipLaunchKernelGGL(k_pm1p_sc, dim3(next), dim3(256), 0, st, plus, n, next, H, W1, W2, V, Z1, Z2, Y1, Y2, eo,
                           pt, q);
        NV2_LAUNCH_CHECK();
        TRYS(reduce_into(p, next, cls));
        TRYS(qd_reduce(p, q, next, cls));
    }
    if (has4 && p->early && p->early->sums_early) {   // release the chain's band-form sums
        Pm1pEarly *E = p->early;
        for (int plus = 0; plus < 2; plus++) {
            E->X[plus] = Xpm[plus];
            E->next[plus] = plus ? nc : nv;
            E->eo[plus] = plus ? ec : ev;
            E->Tr[plus] = Trm[plus];
            E->Vr[plus] = Vrm[plus];
            E->rho[plus] = Trm[plus] ? (plus ? rho_ca : rho_av) : nullptr;
        }
        E->nact = n;
        NV2_HIP(hipEventRecord(p->evX, st));
        {
            std::lock_guard<std::mutex> lk(E->mu);
            E->x_ready = true;
        }
        E->cv.notify_all();
    }

    // ================================================================ SC sums that need no PC solve (main stream)
    // the SC sums wait until the (+-1)' K' reduction has finished (hold_sc)
    if (hold_sc() && has4 && p->early && p->early->active) {
        Pm1pEarly *E = p->early;
        bool qt;
        {
            std::unique_lock<std::mutex> lk(E->mu);
            E->cv.wait(lk, [E] { return E->kred_done || E->abort; });
            qt = E->qt_event && !E->abort;
        }
        if (qt) NV2_HIP(hipStreamWaitEvent(st, p->evQt, 0));
    }
    // class (0): SC == PC
    if (nc && nv) {
        const int nt = (nv + V0T - 1) / V0T;
        dim3 g((unsigned)(nt * (nt + 1) / 2), (unsigned)p->np_c);
        double *pt = parts(p, (i64)g.x * g.y);
        if (!pt) { set_error("hipMalloc(parts) failed"); return -1; }
        hipLaunchKernelGGL(k_v0b, g, dim3(256), 0, st, V0b, p->pairs_c, nv, ec, ev, pt);
        NV2_LAUNCH_CHECK();
        TRYS(reduce_into(p, (i64)g.x * g.y, 0));
    }
    // (+2)/(-2) SC
    for (int mode = 0; mode < 2; mode++) {
        const int cls = mode == 0 ? 3 : 4;
        const double *X = mode == 0 ? k_hrdm : taa;
        const double *K = mode == 0 ? k_eako2 : k_ipko2;
        const double *I = mode == 0 ? Icaca : Iavav;
        const i64 ld = mode == 0 ? Nca : Nav;
        const int next = mode == 0 ? nc : nv;
        const double *eo = mode == 0 ? ec : ev;
        const int2 *pr = mode == 0 ? p->pairs_c : p->pairs_v;
        const i64 npr = mode == 0 ? p->np_c : p->np_v;
        for (int j = 0; q.M && j < M; j++) {
            if (j == s) continue;
            QFILL(xj, mode == 0 ? QK_P2_X : QK_M2_X, j, n4, "qd_pm2x_");
            QFILL(xsj, mode == 0 ? QK_P2_XS : QK_M2_XS, j, n4, "qd_pm2xs_");
            q.a[j] = xj; q.b[j] = xsj;
        }
        (void)pr;
        if (next && npr) {
            // diagonal pairs x == y by k_pm2_sc (symmetrised forms), pairs x < y by one GEMM + k_pm2_dot
            const int2 *prd = mode == 0 ? p->pairs_cd : p->pairs_vd, *prs = mode == 0 ? p->pairs_cs : p->pairs_vs;
            const i64 nd = next, ns = mode == 0 ? p->np_cs : p->np_vs;
            const int nqj = q.M ? M - 1 : 0, ncol = (int)n2 * (2 + nqj);
            const i64 nbs = (ns + 3) / 4;
            double *pt = parts(p, nd + nbs);
            if (!pt) { set_error("hipMalloc(parts) failed"); return -1; }
            TRYS(qd_prepare(p, q, nd + nbs));
            hipLaunchKernelGGL(k_pm2_sc, dim3((unsigned)nd), dim3(256), 0, st, I, ld, prd, mode, nc, n, X, K, eo, pt, q);
            NV2_LAUNCH_CHECK();
            if (ns) {
                GET(Tp, "pm2_T", ns * n2);
                GET(XK, "pm2_XK", n2 * ncol);
                GET(Up, "pm2_U", ns * ncol);
                NV2_HIP(hipMemcpyAsync(XK, X, sizeof(double) * n4, hipMemcpyDeviceToDevice, st));
                NV2_HIP(hipMemcpyAsync(XK + n4, K, sizeof(double) * n4, hipMemcpyDeviceToDevice, st));
                int col = 2;
                for (int j = 0; q.M && j < M; j++)
                    if (j != s) NV2_HIP(hipMemcpyAsync(XK + n4 * col++, q.a[j], sizeof(double) * n4, hipMemcpyDeviceToDevice, st));
                hipLaunchKernelGGL(k_pm2_tgather, dim3(nb(ns * n2, 256)), dim3(256), 0, st, I, ld, prs, ns, mode, nc, n, Tp);
                NV2_LAUNCH_CHECK();
                TRYS(gemm(p, true, false, ncol, (int)ns, (int)n2, 1.0, XK, n2, Tp, n2, 0.0, Up, ncol));
                QDArgs qs = q;
                if (q.M) qs.qpart = q.qpart + 2 * QM * nd;
                hipLaunchKernelGGL(k_pm2_dot, dim3((unsigned)nbs), dim3(256), 0, st, Tp, Up, (int)n2, ncol, prs, ns, mode,
                                   eo, pt + 4 * nd, qs);
                NV2_LAUNCH_CHECK();
            }
            TRYS(reduce_into(p, nd + nbs, cls));
            TRYS(qd_reduce(p, q, nd + nbs, cls));
        }
    }
    // (0)' SC
    const int nq0 = q.M ? M - 1 : 0;   // QD partners
    GET(XG, "v0p_xg", (4 + 2 * nq0) * n4);   // X1 | X2 | G1 | G2 | the partners' X1, X2: one GEMM operand
    double *X1 = XG, *X2 = XG + n4, *G1 = XG + 2 * n4, *G2 = XG + 3 * n4;
    hipLaunchKernelGGL(k_v0p_prep, dim3(nb(n4, 256)), dim3(256), 0, st, S, k_ga, k_gd, X1, X2, G1, G2);
    NV2_LAUNCH_CHECK();
    for (int j = 0; q.M && j < M; j++) {
        if (j == s) continue;
        QFILL(x1j, QK_0P_X1, j, n4, "qd_0px1_");
        QFILL(x2j, QK_0P_X2, j, n4, "qd_0px2_");
        QFILL(dsj, QK_0P_DS, j, n2, "qd_0pds_");
        q.a[j] = x1j; q.b[j] = x2j; q.c[j] = dsj;
    }
    if (nc && nv) {                  // GEMM over all (i, r) pairs
        const i64 npp = (i64)nc * nv, nbk = (npp + 3) / 4, ncw = (4 + 2 * nq0) * n2;
        int blk = 4;
        for (int j = 0; q.M && j < M; j++) {
            if (j == s) continue;
            NV2_HIP(hipMemcpyAsync(XG + n4 * blk++, q.a[j], sizeof(double) * n4, hipMemcpyDeviceToDevice, st));
            NV2_HIP(hipMemcpyAsync(XG + n4 * blk++, q.b[j], sizeof(double) * n4, hipMemcpyDeviceToDevice, st));
        }
        GET(DJK, "v0p_djk", npp * 2 * n2);
        GET(W, "v0p_W", npp * 2 * ncw);
        hipLaunchKernelGGL(k_v0p_gather, dim3(nb(npp * 2 * n2, 256)), dim3(256), 0, st, Icvaa, Ncv, Icaav, Nca, nc, n,
                           npp, 0, DJK);
        NV2_LAUNCH_CHECK();
        TRYS(gemm(p, true, false, (int)ncw, (int)(2 * npp), (int)n2, 1.0, XG, n2, DJK, n2, 0.0, W, ncw));
        double *pt = parts(p, nbk);
        if (!pt) { set_error("hipMalloc(parts) failed"); return -1; }
        TRYS(qd_prepare(p, q, nbk));
        hipLaunchKernelGGL(k_v0p_scdot, dim3((unsigned)nbk), dim3(256), 0, st, W, (int)ncw, DJK, n, npp, nc, da, k_gf,
                           h0p, ec, ev, pt, q);
        NV2_LAUNCH_CHECK();
        TRYS(reduce_into(p, nbk, 7));
        TRYS(qd_reduce(p, q, nbk, 7));
    }
    // ================================================================ join worker 0: every PC class but (+-1)'
    // (whose n^3 chain on worker 1 is the longest; its sums come last)
    if (stage2_timing()) {           // the SC sums are done on the device
        NV2_HIP(hipStreamSynchronize(st));
        ms_scq = ms_since(t_begin);
    }
    auto join_worker = [&](int w) -> int {
        if (w == 1 && has4) {
            Pm1pEarly *E = p->early;
            if (E->th.joinable()) E->th.join();
            E->active = false;
            if (E->rc) { set_error("(+-1)' PC solve: %s", E->err); return E->rc; }
            J[JPM1P] = E->job;
            ms_wdone[1] = ms_since(t_begin);
        } else {
            if (workers[w].joinable()) workers[w].join();
            if (wrc[w]) { set_error("PC solve worker %d: %s", w, werr[w]); return wrc[w]; }
        }
        NV2_HIP(hipEventRecord(p->wev[w], p->ws[w]));
        NV2_HIP(hipStreamWaitEvent(st, p->wev[w], 0));
        return 0;
    };
    TRYS(join_worker(0));

    // ================================================================ class (+1), SC and PC (fused kernel)
    {
        const PcJob &jb = J[JPM1];
        const int nnew = jb.nnew[0];
        double *coef = jb.coefb(0), *w = jb.w2b(0);
        GET(spc, "spc_v1", (i64)n * (nnew ? nnew : 1));
        TRYS(gemm(p, true, false, n, nnew, n, 1.0, ro_v1, n, coef, n, 0.0, spc, n));
        for (int j = 0; q.M && j < M; j++) {
            if (j == s) continue;
            QFILL(roj, QK_V1_RO, j, n2, "qd_v1ro_");
            QFILL(Rj, QK_V1_R, j, n2, "qd_v1R_");
            QGET(sj, "qd_v1s_", j, (i64)n * (nnew ? nnew : 1));
            TRYS(gemm(p, true, false, n, nnew, n, 1.0, Rj, n, coef, n, 0.0, sj, n));
            q.a[j] = roj; q.b[j] = sj;
        }
        if (nc && nv) {
            dim3 g(nb(nv, 64), (unsigned)p->np_c);
            double *pt = parts(p, (i64)g.x * g.y);
            TRYS(qd_prepare(p, q, (i64)g.x * g.y));
#define V1(NM) hipLaunchKernelGGL(k_v1<NM>, g, dim3(64), 0, st, Icacv, Nca, p->pairs_c, nc, n, nv, ec, ev, k_eako1, ro_v1, spc, w, nnew, pt, q)
            if (n <= 8) V1(8); else V1(16);
#undef V1
            NV2_LAUNCH_CHECK();
            TRYS(reduce_into(p, (i64)g.x * g.y, 1));
            TRYS(qd_reduce(p, q, (i64)g.x * g.y, 1));
        }
    }
    // ================================================================ class (-1), SC and PC (fused kernel)
    {
        const PcJob &jb = J[JPM1];
        const int nnew = jb.nnew[1];
        double *coef = jb.coefb(1), *w = jb.w2b(1);
        GET(spc, "spc_vm1", (i64)n * (nnew ? nnew : 1));
        TRYS(gemm(p, true, false, n, nnew, n, 1.0, da, n, coef, n, 0.0, spc, n));
        for (int j = 0; q.M && j < M; j++) {
            if (j == s) continue;
            QFILL(roj, QK_VM1_RO, j, n2, "qd_vm1ro_");
            QFILL(Rj, QK_VM1_R, j, n2, "qd_vm1R_");
            QGET(sj, "qd_vm1s_", j, (i64)n * (nnew ? nnew : 1));
            TRYS(gemm(p, true, false, n, nnew, n, 1.0, Rj, n, coef, n, 0.0, sj, n));
            q.a[j] = roj; q.b[j] = sj;
        }
        if (nc && nv) {
            const i64 tot = p->np_v * nc;
            double *pt = parts(p, nb(tot, 128));
            TRYS(qd_prepare(p, q, nb(tot, 128)));
#define VM1(NM) hipLaunchKernelGGL(k_vm1<NM>, dim3(nb(tot, 128)), dim3(128), 0, st, Icvav, Ncv, p->pairs_v, p->np_v, nc, n, ec, ev, da, k_ipko1, spc, w, nnew, pt, q)
            if (n <= 8) VM1(8); else VM1(16);
#undef VM1
            NV2_LAUNCH_CHECK();
            TRYS(reduce_into(p, nb(tot, 128), 2));
            TRYS(qd_reduce(p, q, nb(tot, 128), 2));
        }
    }
    // ================================================================ classes (+2) and (-2), PC
    for (int mode = 0; mode < 2; mode++) {
        const int cls = mode == 0 ? 3 : 4;
        const double *I = mode == 0 ? Icaca : Iavav;
        const i64 ld = mode == 0 ? Nca : Nav;
        const int next = mode == 0 ? nc : nv;
        const double *eo = mode == 0 ? ec : ev;
        const int2 *prs = mode == 0 ? p->pairs_cs : p->pairs_vs;
        const i64 nprs = mode == 0 ? p->np_cs : p->np_vs;
        {   // off-diagonal external pairs: n^2 metric
            const PcJob &jb = J[JPM2];
            const int nnew = jb.nnew[mode];
            double *coef = jb.coefb(mode), *w2 = jb.w2b(mode);
            GET(spc, "spc_pm2", n2 * (nnew ? nnew : 1));
            TRYS(gemm(p, true, false, (int)n2, nnew, (int)n2, 1.0, Xm2[mode], n2, coef, n2, 0.0, spc, n2));
            for (int j = 0; q.M && j < M; j++) {
                if (j == s) continue;
                QFILL(yj, mode == 0 ? QK_P2_Y : QK_M2_Y, j, n4, "qd_pm2y_");
                QGET(sj, "qd_pm2s_", j, n2 * (nnew ? nnew : 1));
                TRYS(gemm(p, true, false, (int)n2, nnew, (int)n2, 1.0, yj, n2, coef, n2, 0.0, sj, n2));
                q.a[j] = sj;
            }
            if (nprs && nnew) {
                const i64 tot = nprs * nnew;
                GET(Vt, "pm2_Vt", n2 * nprs);
                GET(Rp, "pm2_R", tot);
                hipLaunchKernelGGL(k_pm2_gather, dim3(nb(n2 * nprs, 256)), dim3(256), 0, st, I, ld, prs, nprs, mode, nc,
                                   n, Vt);
                NV2_LAUNCH_CHECK();
                TRYS(gemm(p, true, false, (int)nprs, nnew, (int)n2, 1.0, Vt, n2, spc, n2, 0.0, Rp, nprs));
                QDArgs qr = q;
                for (int j = 0; q.M && j < M; j++) {
                    if (j == s) continue;
                    QGET(rj, "qd_pm2r_", j, tot);
                    TRYS(gemm(p, true, false, (int)nprs, nnew, (int)n2, 1.0, Vt, n2, q.a[j], n2, 0.0, rj, nprs));
                    qr.a[j] = rj;
                }
                double *pt = parts(p, nb(tot, 256));
                TRYS(qd_prepare(p, qr, nb(tot, 256)));
                hipLaunchKernelGGL(k_pm2_denom, dim3(nb(tot, 256)), dim3(256), 0, st, Rp, prs, nprs, mode, w2, nnew, eo,
                                   pt, qr);
                NV2_LAUNCH_CHECK();
                TRYS(reduce_into(p, nb(tot, 256), cls));
                TRYS(qd_reduce(p, qr, nb(tot, 256), cls));
            }
        }
        {   // diagonal external pairs: n(n+1)/2 metric
            const PcJob &jb = J[JPM2P];
            const int nnew = jb.nnew[mode];
            double *coef = jb.coefb(mode), *w2 = jb.w2b(mode);
            GET(spc2, "spc_pm2p", (i64)m * (nnew ? nnew : 1));
            TRYS(gemm(p, true, false, m, nnew, m, 0.5, Xm2p[mode], m, coef, m, 0.0, spc2, m));
            for (int j = 0; q.M && j < M; j++) {
                if (j == s) continue;
                QFILL(yj, mode == 0 ? QK_P2P_Y : QK_M2P_Y, j, (i64)m * m, "qd_pm2py_");
                QGET(sj, "qd_pm2ps_", j, (i64)m * (nnew ? nnew : 1));
                TRYS(gemm(p, true, false, m, nnew, m, 1.0, yj, m, coef, m, 0.0, sj, m));
                q.a[j] = sj;
            }
            if (next && nnew) {
                const i64 tot = (i64)next * nnew;
                double *pt = parts(p, nb(tot, 128));
                TRYS(qd_prepare(p, q, nb(tot, 128)));
                hipLaunchKernelGGL(k_pm2p_pc, dim3(nb(tot, 128)), dim3(128), 0, st, I, ld, next, mode, nc, n, spc2,
                                   w2, nnew, eo, pt, q);
                NV2_LAUNCH_CHECK();
                TRYS(reduce_into(p, nb(tot, 128), cls));
                TRYS(qd_reduce(p, q, nb(tot, 128), cls));
            }
        }
    }
    // ================================================================ class (0)', PC
    {
        const PcJob &jb = J[J0P];
        const int nnew = jb.nnew[0], N = jb.N;
        double *coef = jb.coefb(0), *w = jb.w2b(0);
        if (nnew && nc && nv) {
            GET(V, "v0p_V", n4); GET(V2, "v0p_V2", n4);
            hipLaunchKernelGGL(k_v0p_vmats, dim3(nb(n4, 256)), dim3(256), 0, st, S, V, V2);
            NV2_LAUNCH_CHECK();
            GET(sm, "v0p_sm", n2 * nnew); GET(sp, "v0p_sp", n2 * nnew); GET(spp, "v0p_spp", n2 * nnew);
            TRYS(gemm(p, false, false, (int)n2, nnew, (int)n2, 1.0, V, n2, coef, N, 0.0, sm, n2));
            TRYS(gemm(p, false, false, (int)n2, nnew, (int)n2, 1.0, V, n2, coef + n2, N, 0.0, sp, n2));
            TRYS(gemm(p, false, false, (int)n2, nnew, (int)n2, 1.0, V2, n2, coef + n2, N, 0.0, spp, n2));
            GET(dc1, "v0p_dc1", nnew); GET(dc2, "v0p_dc2", nnew);
            hipLaunchKernelGGL(k_dacoef, dim3(nb(nnew, 128)), dim3(128), 0, st, da, n, coef, N, 0, nnew, dc1);
            hipLaunchKernelGGL(k_dacoef, dim3(nb(nnew, 128)), dim3(128), 0, st, da, n, coef, N, (int)n2, nnew, dc2);
            NV2_LAUNCH_CHECK();
            for (int j = 0; q.M && j < M; j++) {
                if (j == s) continue;
                QFILL(vj, QK_0P_V, j, n4, "qd_0pv_");
                QFILL(v2j, QK_0P_V2, j, n4, "qd_0pv2_");
                QFILL(roj, QK_0P_RO, j, n2, "qd_0pro_");
                QGET(smj, "qd_0psm_", j, n2 * nnew); QGET(spj, "qd_0psp_", j, n2 * nnew);
                QGET(sppj, "qd_0pspp_", j, n2 * nnew); QGET(r0j, "qd_0pr0_", j, nnew); QGET(r1j, "qd_0pr1_", j, nnew);
                TRYS(gemm(p, false, false, (int)n2, nnew, (int)n2, 1.0, vj, n2, coef, N, 0.0, smj, n2));
                TRYS(gemm(p, false, false, (int)n2, nnew, (int)n2, 1.0, vj, n2, coef + n2, N, 0.0, spj, n2));
                TRYS(gemm(p, false, false, (int)n2, nnew, (int)n2, 1.0, v2j, n2, coef + n2, N, 0.0, sppj, n2));
                TRYS(gemm(p, true, false, 1, nnew, (int)n2, 1.0, roj, n2, coef, N, 0.0, r0j, 1));
                TRYS(gemm(p, true, false, 1, nnew, (int)n2, 1.0, roj, n2, coef + n2, N, 0.0, r1j, 1));
                q.a[j] = smj; q.b[j] = spj; q.c[j] = sppj; q.d[j] = r0j; q.e[j] = r1j;
            }
            double *vx = nullptr;
            if (p->rho1) {
                vx = p->get("rho1_vx", nnew);
                if (!vx) { set_error("hipMalloc failed for rho1_vx"); return -1; }
                hipLaunchKernelGGL(k_rho_vaux0p, dim3(nb(nnew, 128)), dim3(128), 0, st, da, n, coef, N, nnew, vx);
                NV2_LAUNCH_CHECK();
            }
            {                        // CO = CM^T [dj | dk] over all (i, r) pairs, the QD partners' blocks after
                const int nq = q.M ? M - 1 : 0, nco = nnew * (1 + nq);
                const i64 npp = (i64)nc * nv, nbk = (npp + 3) / 4, ncm = 2 * n2 * nnew;
                GET(DJK, "v0p_djk_pc", npp * 2 * n2);
                GET(CM, "v0p_cm", ncm * (1 + nq));
                GET(CO, "v0p_co", npp * nco);
                hipLaunchKernelGGL(k_v0p_gather, dim3(nb(npp * 2 * n2, 256)), dim3(256), 0, st, Icvaa, Ncv, Icaav, Nca,
                                   nc, n, npp, 1, DJK);
                hipLaunchKernelGGL(k_v0p_cm, dim3(nb(ncm, 256)), dim3(256), 0, st, sm, sp, spp, (int)n2, nnew, CM);
                NV2_LAUNCH_CHECK();
                int blk = 1;
                for (int j = 0; q.M && j < M; j++) {
                    if (j == s) continue;
                    hipLaunchKernelGGL(k_v0p_cm, dim3(nb(ncm, 256)), dim3(256), 0, st, q.a[j], q.b[j], q.c[j], (int)n2,
                                       nnew, CM + ncm * blk++);
                    NV2_LAUNCH_CHECK();
                }
                TRYS(gemm(p, true, false, nco, (int)npp, (int)(2 * n2), 1.0, CM, 2 * n2, DJK, 2 * n2, 0.0, CO, nco));
                double *pt = parts(p, nbk);
                if (!pt) { set_error("hipMalloc(parts) failed"); return -1; }
                TRYS(qd_prepare(p, q, nbk));
                hipLaunchKernelGGL(k_v0p_pcden, dim3((unsigned)nbk), dim3(256), 0, st, CO, nnew, nco, npp, nc, dc1, dc2,
                                   w, h0p, ec, ev, pt, vx, p->rho1 ? rho_cv : nullptr, q);
                NV2_LAUNCH_CHECK();
                TRYS(reduce_into(p, nbk, 7));
                TRYS(qd_reduce(p, q, nbk, 7));
            }
        }
    }
    // ================================================================ join worker 1: the (+-1)' solves
    TRYS(join_worker(1));
    ms_join = ms_since(t_begin);
    if (p->kr_recorded) {            // device time of the K' reduction of this state's (+-1)' problems
        float t = 0.f;
        NV2_HIP(hipEventSynchronize(p->evKr[1]));
        NV2_HIP(hipEventElapsedTime(&t, p->evKr[0], p->evKr[1]));
        sd.kred_ms = t;
        sd.kred_N = p->kr_N;
        sd.kred_nb = p->kr_nb;
        sd.kred_form = p->kr_form;
        p->kr_recorded = false;
    }
    const int cls_of[NJOB][2] = {{1, 2}, {3, 4}, {3, 4}, {7, 7}, {5, 6}};
    for (int k = 0; k < NJOB; k++)
        for (int b = 0; b < J[k].nb; b++) {
            const int c = cls_of[k][b];
            if (k == JPM2P) { if (J[k].minw[b] < sd.min_metric_eig[c]) sd.min_metric_eig[c] = J[k].minw[b]; }
            else { sd.min_metric_eig[c] = J[k].minw[b]; sd.nnew[c] = J[k].nnew[b]; sd.chol[c] = J[k].chol ? 1 : 0; }
        }
    // ================================================================ classes (-1)' and (+1)', PC sums
    // sum_mu r_mu^2 / (w_mu + s_x) with r = C^T caux^T X_x and K' = C W C^T (qdnevpt.F:5351-5422,
    // 5130-5205) is y^T (K' + s_x)^-1 y with y = caux^T X_x; with K' = Q T Q^T it is
    // z^T (T + s_x)^-1 z, z = Q^T y: one tridiagonal solve per external orbital, no eigenvectors of K'.
    // On the band form the solves of both classes go out as one launch after the loop.
    BandCls bc[2] = {};
    int band_cls[2] = {-1, -1};
    const bool sums_done = has4 && p->early && p->early->sums_done;   // summed on the chain's stream
    for (int plus = 0; has4 && !sums_done && plus < 2; plus++) {
        const int cls = plus ? 5 : 6;
        const int next = plus ? nc : nv;
        const double *eo = plus ? ec : ev;
        const PcJob &jb = J[JPM1P];
        const int bb = plus ? 0 : 1;
        const int nnew = jb.nnew[bb], nmax = jb.nmax;
        if (!next || !nnew) continue;
        const i64 NN = jb.NN();
        const double *caux = jb.caux + bb * NN, *Kp = jb.Kp + bb * NN;
        const int ncq = 1 + (q.M ? M - 1 : 0);              // right-hand sides: this state, then the QD partners
        GET(Ypc, plus ? "pm1_Ypc1" : "pm1_Ypc0", (i64)nmax * next * ncq);
        // y = caux^T x (on the Cholesky path caux = U^-1 and nmax = N3)
        auto project = [&](const double *x, double *y) -> int {
            return gemm(p, true, false, nmax, next, (int)N3, 1.0, caux, N3, x, N3, 0.0, y, nmax);
        };
        TRYS(project(Xpm[plus], Ypc));
        int col = 1;
        for (int j = 0; q.M && j < M; j++) {
            if (j == s) continue;
            double *mj = qd_fill(p, Sj[j], plus ? QK_P1P_MPC : QK_M1P_MPC, j < s, N3 * N3, "qd_pm1M");
            if (!mj) return -1;
            QFILL(qj, plus ? QK_P1P_QPC : QK_M1P_QPC, j, N3 * n, "qd_pm1qp_");
            GET(xj, "qd_pm1X", N3 * next);
            TRYS(gemm(p, true, false, (int)N3, next, (int)N3, 1.0, mj, N3, Vpm[plus], N3, 0.0, xj, N3));
            TRYS(gemm(p, false, false, (int)N3, next, n, 1.0, qj, N3, Hpm[plus], n, 1.0, xj, N3));
            double *yj = Ypc + (i64)nmax * next * col++;
            TRYS(project(xj, yj));
            q.a[j] = yj;
        }
        if (jb.band) {                   // z = Q^T y in place; the banded solves follow the loop
            GET(qtw, "pm1_qtw", std::max(band_qt_workspace(nmax, 1, next * ncq), band_qt_workspace(nmax, 1, n)));
            TRYS(band_apply_qt(st, nmax, jb.bandw, bb, 1, Ypc, nmax, 0, next * ncq, qtw));
            if (Trm[plus]) {
                TRYS(gemm(p, true, false, nmax, n, (int)N3, 1.0, caux, N3, Trm[plus], N3, 0.0, Vrm[plus], nmax));
                TRYS(band_apply_qt(st, nmax, jb.bandw, bb, 1, Vrm[plus], nmax, 0, n, qtw));
            }
            BandCls &c = bc[plus];
            c.Br = band_rows(nmax, jb.nb, bb, jb.bandw);
            c.Z = Ypc;
            c.eo = eo;
            c.plus = plus;
            c.next = next;
            c.q = q;
            band_cls[plus] = cls;
            continue;
        }
        const TrdOut to = trd_outputs(nmax, jb.nb, jb.trdw);
        const double *dT = to.d + (i64)bb * nmax, *eT = to.e + (i64)bb * nmax, *tauT = to.tau + (i64)bb * nmax;
        GET(Wq, "pm1_applyq", trd_apply_workspace(nmax, 1, next * ncq));
        TRYS(trd_apply_q(p->h, st, nmax, Kp, NN, 1, tauT, true, next * ncq, Ypc, 0, Wq));   // z = Q^T y
        double *rho_b = nullptr;
        if (Trm[plus]) {                 // the density vectors take the path of the right-hand sides: Q^T caux^T T
            TRYS(gemm(p, true, false, nmax, n, (int)N3, 1.0, caux, N3, Trm[plus], N3, 0.0, Vrm[plus], nmax));
            GET(Wr, "pm1_applyq_rho", trd_apply_workspace(nmax, 1, n));
            TRYS(trd_apply_q(p->h, st, nmax, Kp, NN, 1, tauT, true, n, Vrm[plus], 0, Wr));
            rho_b = plus ? rho_ca : rho_av;
        }
        GET(rows, "pm1_rows", (i64)4 * nnew * next);
        const unsigned nbk = nb(next, 64);
        double *pt2 = parts(p, nbk);
        TRYS(qd_prepare(p, q, nbk));
        hipLaunchKernelGGL(k_pm1p_shift, dim3(nbk), dim3(64), 0, st, plus, nnew, next, dT, eT, Ypc, (i64)nmax, eo, rows,
                           pt2, q, Vrm[plus], n, rho_b, plus ? (i64)1 : (i64)n, plus ? (i64)nc : (i64)1);
        NV2_LAUNCH_CHECK();
        TRYS(reduce_into(p, nbk, cls));
        TRYS(qd_reduce(p, q, nbk, cls));
    }
    if (band_cls[0] >= 0 || band_cls[1] >= 0) {
        const int nmax = J[JPM1P].nmax;
        for (int c = 0; c < 2; c++) if (band_cls[c] < 0) bc[c].next = 0;
        const int n0 = bc[0].next, n1 = bc[1].next;
        GET(Ub, "pm1_bandU", (i64)(n0 + n1) * nmax * UR);
        GET(yb, "pm1_bandy", (i64)(n0 + n1) * nmax);
        bc[0].U = Ub;
        bc[0].y = yb;
        bc[1].U = Ub + (i64)n0 * nmax * UR;
        bc[1].y = yb + (i64)n0 * nmax;
        double *pt2 = parts(p, n0 + n1);
        if (!pt2) { set_error("hipMalloc(parts) failed"); return -1; }
        TRYS(qd_prepare(p, q, n0 + n1));
        bc[0].q.qpart = bc[1].q.qpart = q.qpart;
        bc[0].q.M = bc[1].q.M = q.M;
        bc[0].q.self = bc[1].q.self = q.self;
        hipLaunchKernelGGL(k_pm1p_bandr, dim3(n0 + n1), dim3(RW_T), 0, st, nmax, bc[0], bc[1], pt2);
        NV2_LAUNCH_CHECK();
        for (int c = 0; c < 2; c++) {
            if (!bc[c].next) continue;
            TRYS(reduce_into(p, bc[c].next, band_cls[c], c ? n0 : 0));
            TRYS(qd_reduce(p, q, bc[c].next, band_cls[c], c ? n0 : 0));
            if (Trm[c]) TRYS(band_rho(p->h, st, c, nmax, n, bc[c].next, bc[c].y, Vrm[c], c ? rho_ca : rho_av));
        }
    }
    if (stage2_timing()) {
        NV2_HIP(hipStreamSynchronize(st));
        fprintf(stderr, "[stage2 s=%d] fills %.2f  SC done %.2f  worker0 %.2f  worker1 %.2f  join %.2f  end %.2f ms\n", s,
                ms_workers, ms_scq, ms_wdone[0], ms_wdone[1], ms_join, ms_since(t_begin));
        if (has4) {
            auto rel = [&](std::chrono::steady_clock::time_point t) {
                return std::chrono::duration<double, std::milli>(t - t_begin).count();
            };
            fprintf(stderr, "[stage2 s=%d] (+-1)' chain: start %.2f  metric solved %.2f  K ready %.2f  K' reduced %.2f ms\n",
                    s, rel(p->early->t_start), rel(p->early->t_metric), rel(p->early->t_kready), rel(p->early->t_end));
        }
    }
    // ---- .NATORB: the density matrix in the canonical orbitals, back in the input orbitals, and its
    // eigenvalues (dsyev 'L', qdnevpt.F:1414)
    if (p->rho1) {
        const i64 no2 = (i64)norb * norb;
        GET(Rc, "rho1_R", no2); GET(Uf, "rho1_U", no2); GET(Rt, "rho1_tmp", no2); GET(Rmo, "rho1_mo", no2);
        GET(Rw, "rho1_w", no2); GET(occ, "rho1_occ", norb);
        hipLaunchKernelGGL(k_rho_assemble, dim3(nb(no2, 256)), dim3(256), 0, st, nc, n, nv, da, rho_cv,
                           has4 ? rho_ca : nullptr, has4 ? rho_av : nullptr, Rc, Uf, Cc, Cv);
        NV2_LAUNCH_CHECK();
        TRYS(gemm(p, false, false, norb, norb, norb, 1.0, Uf, norb, Rc, norb, 0.0, Rt, norb));
        TRYS(gemm(p, false, true, norb, norb, norb, 1.0, Rt, norb, Uf, norb, 0.0, Rmo, norb));
        NV2_HIP(hipMemcpyAsync(Rw, Rc, sizeof(double) * no2, hipMemcpyDeviceToDevice, st));
        TRYS(syevd(p, false, norb, Rw, occ));
    }
    // ---- results: (norm_sc, e_sc, norm_pc, e_pc) per class, SC==PC for class (0)
    double accs[32];
    NV2_HIP(hipMemcpyAsync(accs, p->acc, sizeof(accs), hipMemcpyDeviceToHost, st));
    NV2_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < NJOB; k++)
        for (int b = 0; b < J[k].nb; b++)
            if (J[k].nmax && p->hinfo[4 * J[k].slot + 2 + b] != 0) {
                set_error("syevd(N=%d, %s Koopmans %d) failed, info=%d", J[k].nmax, J[k].tag.c_str(), b,
                          p->hinfo[4 * J[k].slot + 2 + b]);
                return -3;
            }
    for (int k = 0; k < 8; k++) {
        out[4 * k + 0] = accs[4 * k + 1];
        out[4 * k + 1] = accs[4 * k + 0];
        out[4 * k + 2] = accs[4 * k + 3];
        out[4 * k + 3] = accs[4 * k + 2];
    }
    if (q.M) {   // qd_out[k*2M + j] = e2mp(j,s) of class k, qd_out[k*2M + M + j] = e2en(j,s)
        std::vector<double> hq(8 * 2 * QM);
        NV2_HIP(hipMemcpyAsync(hq.data(), p->accq, sizeof(double) * hq.size(), hipMemcpyDeviceToHost, st));
        NV2_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < 8; k++)
            for (int j = 0; j < M; j++) {
                qd_out[k * 2 * M + j] = j == s ? 0.0 : hq[k * 2 * QM + j];
                qd_out[k * 2 * M + M + j] = j == s ? 0.0 : hq[k * 2 * QM + QM + j];
            }
    } else if (qd_out) {
        for (int k = 0; k < 16 * M; k++) qd_out[k] = 0.0;
    }
    if (stats) *stats = sd;
    return 0;
}

}  // namespace nv2

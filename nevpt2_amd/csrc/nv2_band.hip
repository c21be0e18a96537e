// Orthogonal reduction of dense symmetric matrices to band form, half-bandwidth BAND_B, for the
// (+-1)' PC sums.  The reference diagonalises K' (qdnevpt.F:5351-5422, 5130-5205) and sums
// r_mu^2 / (w_mu + s); that sum is y^T (K' + s)^-1 y, which any orthogonal similarity keeps:
// with K' = Q B Q^T it is z^T (B + s)^-1 z, z = Q^T y.  A band matrix needs no bulge chasing to
// serve those shifted solves (a banded LU per shift does), and the dense-to-band step is made of
// GEMMs instead of one matrix-vector pass per column.
//
// Panel k (columns p0 = k*BAND_B ..+BAND_B, rows r0 = p0 + BAND_B .. N-1, m = N - r0 rows):
//   * CholeskyQR2 of the panel P (m x BAND_B): G = P^T P = R1^T R1, Q1 = P R1^-1, then the same on
//     Q1 (R2), P = Q2 R with R = R2 R1;
//   * Householder vectors reconstructed from Q2 (Ballard et al., "Reconstructing Householder vectors
//     from tall-skinny QR"): with S = -sign(diag(Q2_top)), the unpivoted LU Q2 - [S; 0] = Y U gives
//     the reflectors Y (unit lower trapezoidal) and T = -U S Y_top^-T, so that H = I - Y T Y^T has
//     H[:, :BAND_B] = Q2 S and H^T P = [S R; 0];
//   * the trailing block A22 <- H^T A22 H as W = A22 Y T, Z = W - 1/2 Y (T^T Y^T W),
//     A22 -= [Y Z] [Z Y]^T (rocBLAS GEMMs, the last one of rank 2 BAND_B; the BAND_B x BAND_B products
//     over the panel's rows P^T P, Q1^T Q1 and Y^T W are split over row blocks, k_atb_parts).
// The last panel (m <= BAND_B rows) is factored by Householder steps in LDS, with its orthogonal
// factor kept explicitly.  The panels of every matrix of the batch advance together.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "nv2_band.h"
#include "nv2_chol64.h"

namespace nv2 {

#define BB BAND_B
#define BWG 256
#define KC 128           // rows per workgroup of the A^T B products with BAND_B x BAND_B results
#define KS 32            // rows per LDS stage of those products
// keeps the compiler from hoisting the loads of later unrolled iterations (and their registers) upward
#define BAND_FENCE() __asm__ volatile("" ::: "memory")

struct BandLayout {      // offsets (doubles) inside one problem's part of the workspace
    int npan;            // panels with more than BAND_B rows (CholeskyQR2)
    i64 Y, T, Ql, Bl, Gp, R1, R1i, R2, R2i, Ui, V1, Q1, Q2, W, X, size;   // X = [Y | W2 | Y], m x 3 BAND_B
};
static BandLayout band_layout(int N) {
    BandLayout L;
    int np = 0;
    while ((i64)N - (i64)BB * (np + 1) > BB) np++;
    L.npan = np;
    const i64 n = N, sq = (i64)BB * BB, col = (i64)BB * n;
    i64 o = 0;
    L.Y = o; o += np * col;
    L.T = o; o += np * sq;
    L.Ql = o; o += sq;
    L.Bl = o; o += (2 * BB + 1) * n;
    L.Gp = o; o += sq * ((n + KC - 1) / KC);     // partial A^T B products, one per KC rows
    L.R1 = o; o += sq;
    L.R1i = o; o += sq;
    L.R2 = o; o += sq;
    L.R2i = o; o += sq;
    L.Ui = o; o += sq;
    L.V1 = o; o += sq;
    L.Q1 = o; o += col;
    L.Q2 = o; o += col;
    L.W = o; o += col;
    L.X = o; o += 3 * col;
    L.size = o;
    return L;
}
i64 band_workspace(int N, int nb) { return band_layout(N).size * nb; }
const double *band_rows(int N, int nb, int b, const double *work) {
    (void)nb;
    const BandLayout L = band_layout(N);
    return work + b * L.size + L.Bl;
}

__global__ void k_band_sym(double *A, int N, i64 sA) {   // lower <- upper
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (i64)N * N) return;
    const int i = (int)(e % N), j = (int)(e / N);
    double *Ab = A + blockIdx.z * sA;
    if (i > j) Ab[e] = Ab[j + (i64)N * i];
}
__global__ void k_band_zero_flags(int *f, int nb) {
    if (threadIdx.x < nb) f[threadIdx.x] = 0;
}

// R^T R = G (BAND_B x BAND_B) and R^-1, problem b = blockIdx.x.  G is the sum of the np partial
// products at Gp + b * sW (BAND_B^2 apart): the four wavefronts add them into LDS (consecutive threads
// on consecutive entries, every partial's loads in flight together), then factor it together
// (chol64_4w).  R and R^-1 are stored by rows (element (i, c) at i * BAND_B + c).  flag[b] is set by a
// pivot that is not positive and, when dev > 0, by an entry of G - I above dev: the second CholeskyQR
// pass then starts from a Q1 too far from orthogonal for Q2 to be orthogonal to rounding.
#define WAVE_LDS_SYNC() __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory")
#define BCH_T 256
__global__ __launch_bounds__(BCH_T) void k_bchol(const double *Gp, i64 sW, int np, double *R, double *Ri, int *flag,
                                                 double dev) {
    __shared__ double gs[BB * BB];
    __shared__ Chol64Lds L;
    const int b = blockIdx.x, t = threadIdx.x, w = t >> 6, c = t & 63;
    const double *Gb = Gp + b * sW;
    {
        double g[BB * BB / BCH_T];
#pragma unroll
        for (int u = 0; u < BB * BB / BCH_T; u++) g[u] = 0.0;
        for (int q = 0; q < np; q++)
#pragma unroll
            for (int u = 0; u < BB * BB / BCH_T; u++) g[u] += Gb[(i64)q * BB * BB + t + BCH_T * u];
#pragma unroll
        for (int u = 0; u < BB * BB / BCH_T; u++) gs[t + BCH_T * u] = g[u];
    }
    __syncthreads();
    double a[16], x[16];
    bool off = false;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int r = 16 * w + i;
        a[i] = gs[c + BB * r];
        if (dev > 0.0) off |= !(fabs(a[i] - (r == c ? 1.0 : 0.0)) <= dev);
    }
    const bool offall = __syncthreads_or(off) != 0;
    const bool bad = chol64_4w(a, x, L) != 0 || offall;
    double *Rb = R + b * sW, *Rib = Ri + b * sW;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int r = 16 * w + i;
        // R = R^-1 = I keeps the batch finite when the flag sends this problem elsewhere
        Rb[r * BB + c] = bad ? (r == c ? 1.0 : 0.0) : a[i];
        Rib[c * BB + r] = bad ? (r == c ? 1.0 : 0.0) : x[i];     // x[i] = R^-1[c][r]
    }
    if (bad && t == 0) flag[b] = 1;
}
// Partial products of A^T B for A, B with BAND_B columns (ld, m rows; problem b = blockIdx.y at A + b*sA,
// B + b*sB): workgroup c sums rows c*KC .. c*KC+KC-1 into Gp + b*sW + c*BAND_B^2.  rocBLAS gives a
// BAND_B x BAND_B result with thousands of rows to a single workgroup; here ceil(m / KC) of them share it.
// With Ac set, the rows of A it reads are also copied to Ac + b*sW and Ac + b*sW + 2*BAND_B*ld (the outer
// blocks of [Y | W2 | Y]).
__global__ __launch_bounds__(BWG) void k_atb_parts(const double *A, i64 sA, const double *B, i64 sB, int ld, int m,
                                                   double *Gp, i64 sW, double *Ac) {
    __shared__ double sa[KS][BB + 1];
    __shared__ double sb[KS][BB + 1];
    const int c = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
    const int k0 = c * KC, kn = m - k0 < KC ? m - k0 : KC;
    const double *Ab = A + b * sA + k0, *Bb = B + b * sB + k0;
    const int ti = t % 16, tj = t / 16;
    double acc[4][4] = {};
    for (int s0 = 0; s0 < kn; s0 += KS) {
        __syncthreads();                 // the previous stage is consumed
        for (int e = t; e < KS * BB; e += BWG) {
            const int k = e % KS, i = e / KS;
            const double av = s0 + k < kn ? Ab[s0 + k + (i64)ld * i] : 0.0;
            sa[k][i] = av;
            sb[k][i] = s0 + k < kn ? Bb[s0 + k + (i64)ld * i] : 0.0;
            if (Ac && s0 + k < kn) {
                double *Acb = Ac + b * sW + k0 + s0 + k + (i64)ld * i;
                Acb[0] = av;
                Acb[(i64)2 * BB * ld] = av;
            }
        }
        __syncthreads();
        for (int k = 0; k < KS; k++) {
            double av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) { av[u] = sa[k][ti + 16 * u]; bv[u] = sb[k][tj + 16 * u]; }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[u][v] += av[u] * bv[v];
        }
    }
    double *G = Gp + b * sW + (i64)c * BB * BB;
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) G[(ti + 16 * u) + BB * (tj + 16 * v)] = acc[u][v];
}
// V1 = T^T X with X the sum of the np partial products at Gp (problem b = blockIdx.x)
__global__ __launch_bounds__(BWG) void k_tx(const double *Gp, int np, const double *T, double *V1, i64 sW) {
    __shared__ double xs[BB][BB + 1];
    __shared__ double ts[BB][BB + 1];
    const int b = blockIdx.x, t = threadIdx.x;
    const double *Gb = Gp + b * sW, *Tb = T + b * sW;
    double g[BB * BB / BWG];
#pragma unroll
    for (int u = 0; u < BB * BB / BWG; u++) g[u] = 0.0;
    for (int q = 0; q < np; q++)
#pragma unroll
        for (int u = 0; u < BB * BB / BWG; u++) g[u] += Gb[(i64)q * BB * BB + t + BWG * u];
#pragma unroll
    for (int u = 0; u < BB * BB / BWG; u++) {
        const int e = t + BWG * u;
        xs[e % BB][e / BB] = g[u];
        ts[e % BB][e / BB] = Tb[e];
    }
    __syncthreads();
    const int ti = t % 16, tj = t / 16;
    double acc[4][4] = {};
    for (int k = 0; k < BB; k++) {
        double av[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) { av[u] = ts[k][ti + 16 * u]; bv[u] = xs[k][tj + 16 * u]; }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[u][v] += av[u] * bv[v];
    }
    double *Vb = V1 + b * sW;
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) Vb[(ti + 16 * u) + BB * (tj + 16 * v)] = acc[u][v];
}

// Householder form of panel k from its CholeskyQR2 factors, problem b = blockIdx.x, by the four
// wavefronts of a 256-thread workgroup (Ballard et al.): with S = -sign(diag(Q2_top)), the LU factors of
// Q2_top - S without pivoting give the reflectors Y_top = L (unit lower) and T = -U S L^-T.  Thread
// (w, c) holds rows 16 w .. 16 w + 15 of column c.  LU step k: the wavefront with row k publishes it,
// lane k of every wavefront publishes its rows of column k, and the rows below k are eliminated (one
// barrier per step, the published row and column double-buffered).  Then U^-1 (rows from the bottom)
// and L^-1 (rows from the top) in the same 64 steps, and the triangular products T and R~ = S R2 R1
// (written into the panel's top block of A; the rows below it are outside the band and never read
// again) from LDS.  R1, R2 and U^-1 are stored by rows.
__global__ __launch_bounds__(256) void k_brecon(double *A, int N, i64 sA, int p0, int r0, const double *Q2,
                                                const double *R1, const double *R2, double *Y, double *T,
                                                double *Ui, i64 sW) {
    __shared__ double lu[BB][BB + 1];    // U of Q2_top - S (zeros below the diagonal)
    __shared__ double ll[BB][BB + 1];    // L (strictly lower part, zeros elsewhere)
    __shared__ double xl[BB][BB + 1];    // L^-1; before the LU: R1
    __shared__ double r2[BB][BB + 1];    // R2
    __shared__ double rowb[2][BB], colb[2][BB], rub[2][BB], rlb[2][BB], sg[BB];
    const int b = blockIdx.x, t = threadIdx.x, c = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);     // uniform over the wavefront
    double *Ab = A + b * sA;
    const double *Qb = Q2 + b * sW, *R1b = R1 + b * sW, *R2b = R2 + b * sW;
    if (t < BB) sg[t] = Qb[t + (i64)N * t] >= 0.0 ? -1.0 : 1.0;
    for (int e = t; e < BB * BB; e += 256) {
        xl[e / BB][e % BB] = R1b[e];     // R1[k][c], stored by rows
        r2[e / BB][e % BB] = R2b[e];
    }
    __syncthreads();
    // R~[r][c] = S_r sum_{r <= k <= c} R2[r][k] R1[k][c] into the panel's top block (R1, R2 upper
    // triangular with zeros below the diagonal: the full sum)
    {
        double v[16];
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = 0.0;
#pragma unroll 4
        for (int k = 0; k < BB; k++) {
            const double bk = xl[k][c];
#pragma unroll
            for (int i = 0; i < 16; i++) v[i] += r2[16 * w + i][k] * bk;
        }
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int r = 16 * w + i;
            Ab[(r0 + r) + (i64)N * (p0 + c)] = r <= c ? sg[r] * v[i] : 0.0;
        }
    }
    double mm[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int r = 16 * w + i;
        mm[i] = Qb[r + (i64)N * c] - (r == c ? sg[c] : 0.0);
    }
    // LU without pivoting (|pivot| >= 1 by the choice of S)
#pragma unroll 1
    for (int kw = 0; kw < 4; kw++) {
#pragma unroll
        for (int ik = 0; ik < 16; ik++) {
            const int k = 16 * kw + ik, pb = ik & 1;
            if (w == kw) rowb[pb][c] = mm[ik];
            if (c == k)
#pragma unroll
                for (int i = 0; i < 16; i++) colb[pb][16 * w + i] = mm[i];
            __syncthreads();
            const double rp = 1.0 / rowb[pb][k], ukc = rowb[pb][c];
            if (w > kw) {
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    const double l = colb[pb][16 * w + i] * rp;
                    mm[i] = c > k ? mm[i] - l * ukc : (c == k ? l : mm[i]);
                }
            } else if (w == kw) {
#pragma unroll
                for (int i = ik + 1; i < 16; i++) {
                    const double l = colb[pb][16 * w + i] * rp;
                    mm[i] = c > k ? mm[i] - l * ukc : (c == k ? l : mm[i]);
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int r = 16 * w + i;
        lu[r][c] = r <= c ? mm[i] : 0.0;
        ll[r][c] = r > c ? mm[i] : 0.0;
    }
    __syncthreads();
    // U^-1 from the bottom row up and L^-1 from the top row down, one step of each per barrier
    double xu[16], xv[16];
#pragma unroll
    for (int i = 0; i < 16; i++) { xu[i] = 16 * w + i == c ? 1.0 : 0.0; xv[i] = xu[i]; }
#pragma unroll 1
    for (int sw = 0; sw < 4; sw++) {
#pragma unroll
        for (int is = 0; is < 16; is++) {
            const int kl = 16 * sw + is, ku = BB - 1 - kl, pb = is & 1;
            const int iu = 15 - is, wu = 3 - sw;          // owner of row ku
            if (w == wu) {
                const double x = c < ku ? 0.0 : xu[iu] / lu[ku][ku];
                xu[iu] = x;
                rub[pb][c] = x;
            }
            if (w == sw) rlb[pb][c] = xv[is];             // row kl of L^-1 is final (unit diagonal)
            __syncthreads();
            const double xuc = rub[pb][c], xlc = rlb[pb][c];
            if (w < wu) {                                 // rows below ku: every row of the block
#pragma unroll
                for (int i = 0; i < 16; i++) xu[i] -= lu[16 * w + i][ku] * xuc;
            } else if (w == wu) {
#pragma unroll
                for (int i = 0; i < iu; i++) xu[i] -= lu[16 * w + i][ku] * xuc;
            }
            if (w > sw) {                                 // rows above kl
#pragma unroll
                for (int i = 0; i < 16; i++) xv[i] -= ll[16 * w + i][kl] * xlc;
            } else if (w == sw) {
#pragma unroll
                for (int i = is + 1; i < 16; i++) xv[i] -= ll[16 * w + i][kl] * xlc;
            }
        }
    }
    __syncthreads();                     // xl (R1) is no longer read
#pragma unroll
    for (int i = 0; i < 16; i++) xl[16 * w + i][c] = xv[i];
    double *Yb = Y + b * sW, *Tb = T + b * sW, *Uib = Ui + b * sW;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int r = 16 * w + i;
        Uib[r * BB + c] = xu[i];
        Yb[r + (i64)N * c] = r > c ? mm[i] : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    // T[r][c] = -sum_{r <= k <= c} U[r][k] S_k L^-1[c][k] (U upper, L^-1 lower: the full sum)
    {
        double v[16];
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = 0.0;
#pragma unroll 4
        for (int k = 0; k < BB; k++) {
            const double zk = sg[k] * xl[c][k];
#pragma unroll
            for (int i = 0; i < 16; i++) v[i] -= lu[16 * w + i][k] * zk;
        }
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int r = 16 * w + i;
            Tb[r + BB * c] = r <= c ? v[i] : 0.0;
        }
    }
}

// The last panel (m = N - r0 <= BAND_B rows): Householder QR of P = A[r0:, p0:p0+BB] and
// A22 <- H^T A22 H in LDS; the product Qm = H_0 H_1 ... is kept (ld BAND_B) for the right-hand sides.
__global__ __launch_bounds__(BWG) void k_blast(double *A, int N, i64 sA, int p0, int r0, double *Ql, i64 sW) {
    __shared__ double P[BB][BB + 1], S[BB][BB + 1], Qm[BB][BB + 1];
    __shared__ double v[BB], w[BB], red[BWG / 64];
    const int b = blockIdx.x, m = N - r0;
    double *Ab = A + b * sA;
    for (int t = threadIdx.x; t < m * BB; t += BWG) P[t % m][t / m] = Ab[(r0 + t % m) + (i64)N * (p0 + t / m)];
    for (int t = threadIdx.x; t < m * m; t += BWG) {
        S[t % m][t / m] = Ab[(r0 + t % m) + (i64)N * (r0 + t / m)];
        Qm[t % m][t / m] = (t % m) == (t / m) ? 1.0 : 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    auto bsum = [&](double s) {
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        __syncthreads();
        if (lane == 0) red[wid] = s;
        __syncthreads();
        double r = 0.0;
        for (int q = 0; q < BWG / 64; q++) r += red[q];
        return r;
    };
    for (int k = 0; k < m - 1 && k < BB; k++) {
        // reflector of column k, rows k..m-1: v_k = 1, tau
        double xn = 0.0;
        for (int i = k + 1 + threadIdx.x; i < m; i += BWG) xn += P[i][k] * P[i][k];
        xn = bsum(xn);
        const double a1 = P[k][k];
        double beta, tau, scl;
        if (xn == 0.0) { beta = a1; tau = 0.0; scl = 0.0; }
        else {
            beta = -copysign(sqrt(a1 * a1 + xn), a1);
            tau = (beta - a1) / beta;
            scl = 1.0 / (a1 - beta);
        }
        for (int i = threadIdx.x; i < m; i += BWG) v[i] = i < k ? 0.0 : (i == k ? 1.0 : P[i][k] * scl);
        __syncthreads();
        if (threadIdx.x == 0) P[k][k] = beta;
        for (int i = k + 1 + threadIdx.x; i < m; i += BWG) P[i][k] = 0.0;
        // P[:, k+1:] -= tau v (v^T P[:, k+1:])
        for (int c = k + 1 + threadIdx.x; c < BB; c += BWG) {
            double s = 0.0;
            for (int i = k; i < m; i++) s += v[i] * P[i][c];
            s *= tau;
            for (int i = k; i < m; i++) P[i][c] -= s * v[i];
        }
        // S <- H S H: p = tau S v, K = tau/2 v^T p, w = p - K v, S -= v w^T + w v^T
        for (int i = threadIdx.x; i < m; i += BWG) {
            double s = 0.0;
            for (int j = k; j < m; j++) s += S[i][j] * v[j];
            w[i] = tau * s;
        }
        __syncthreads();
        double kv = 0.0;
        for (int i = threadIdx.x; i < m; i += BWG) kv += v[i] * w[i];
        kv = 0.5 * tau * bsum(kv);
        for (int i = threadIdx.x; i < m; i += BWG) w[i] -= kv * v[i];
        __syncthreads();
        for (int t = threadIdx.x; t < m * m; t += BWG) {
            const int i = t % m, j = t / m;
            S[i][j] -= v[i] * w[j] + w[i] * v[j];
        }
        // Qm <- Qm H: Qm[:, :] -= tau (Qm v) v^T
        __syncthreads();
        if (threadIdx.x < m) {
            double s = 0.0;
            for (int j = k; j < m; j++) s += Qm[threadIdx.x][j] * v[j];
            w[threadIdx.x] = tau * s;     // w is free again
        }
        __syncthreads();
        for (int t = threadIdx.x; t < m * m; t += BWG) {
            const int i = t % m, j = t / m;
            Qm[i][j] -= w[i] * v[j];
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < m * BB; t += BWG) Ab[(r0 + t % m) + (i64)N * (p0 + t / m)] = P[t % m][t / m];
    for (int t = threadIdx.x; t < m * m; t += BWG) {
        Ab[(r0 + t % m) + (i64)N * (r0 + t / m)] = S[t % m][t / m];
        Ql[b * sW + (t % m) + BB * (t / m)] = Qm[t % m][t / m];
    }
}

// rows of the band: Br[r * BAND_W + (c - r + BAND_B)] = B(r, c) for |c - r| <= BAND_B (zero outside
// the matrix); the lower triangle of A is the one kept up to date
__global__ void k_band_extract(const double *A, int N, i64 sA, double *Br, i64 sW) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (i64)N * BAND_W) return;
    const int b = blockIdx.z, r = (int)(t / BAND_W), c = r + (int)(t % BAND_W) - BB;
    double v = 0.0;
    if (c >= 0 && c < N) v = c <= r ? A[b * sA + r + (i64)N * c] : A[b * sA + c + (i64)N * r];
    Br[b * sW + t] = v;
}

// Q_b^T Z for right-hand sides Z (N x ncols per problem), panel by panel, two launches per panel.
// k_band_qt_step (ku, kp = ku + 1) covers rows base = BB (ku + 1) .. N-1 in blocks of QB rows and the
// columns in blocks of QB; workgroup (rb, cb, b) holds its Z block in LDS and
//   * (ku >= 0) updates it with panel ku: Z -= Y_ku W2, W2 = T_ku^T Y_ku^T Z from k_band_qt_reduce
//     (rows of the block relative to panel ku: QB rb ..);
//   * (kp < npan, rb >= 1) forms its partial Y_kp^T Z over its rows (relative to panel kp: QB (rb-1) ..).
// k_band_qt_reduce then adds the partials of each column block in row-block order (a quarter of the
// block's columns per workgroup) and forms W2 = T_kp^T (sum) for the next step.  No workgroup waits for
// another inside a launch.
#define QB 64
struct BandQtK {
    int N, npan, ku, kp, nrb, ncols, ncb, npmax;
    i64 sW, offY, offT, ldZ, sZ;
    const double *work;
    double *Z, *Gp, *W2;         // Gp: [b][cb][rb] QB x QB partials; W2: [b][cb] QB x QB
};
__global__ __launch_bounds__(256) void k_band_qt_step(BandQtK a) {
    __shared__ double zs[QB][QB + 1];
    __shared__ double ys[QB][BB + 1];
    __shared__ double ws[BB][QB + 1];
    const int rb = blockIdx.x, cb = blockIdx.y, b = blockIdx.z, t = threadIdx.x;
    const int N = a.N, base = BB * (a.ku + 1), rs = base + QB * rb;
    const int nr = N - rs < QB ? N - rs : QB, cs = QB * cb;
    const int ncl = a.ncols - cs < QB ? a.ncols - cs : QB;
    const double *w0 = a.work + b * a.sW;
    double *Zb = a.Z + b * a.sZ + a.ldZ * cs + rs;
    const int tr = t % 16, tc = t / 16;
    for (int e = t; e < QB * QB; e += 256) {
        const int r = e % QB, c = e / QB;
        zs[r][c] = r < nr && c < ncl ? Zb[r + a.ldZ * c] : 0.0;
    }
    if (a.ku >= 0) {
        const double *Yk = w0 + a.offY + (i64)a.ku * BB * N + QB * rb;
        const double *W2 = a.W2 + ((i64)b * a.ncb + cb) * QB * QB;
        for (int e = t; e < QB * BB; e += 256) {
            const int r = e % QB, i = e / QB;
            ys[r][i] = r < nr ? Yk[r + (i64)N * i] : 0.0;
        }
        for (int e = t; e < BB * QB; e += 256) ws[e % BB][e / BB] = W2[e];
        __syncthreads();
        double acc[4][4] = {};
        for (int i = 0; i < BB; i++) {
            double yv[4], wv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) { yv[u] = ys[tr + 16 * u][i]; wv[u] = ws[i][tc + 16 * u]; }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[u][v] += yv[u] * wv[v];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) zs[tr + 16 * u][tc + 16 * v] -= acc[u][v];
    }
    __syncthreads();
    if (a.ku >= 0)
        for (int e = t; e < QB * QB; e += 256) {
            const int r = e % QB, c = e / QB;
            if (r < nr && c < ncl) Zb[r + a.ldZ * c] = zs[r][c];
        }
    if (a.kp >= a.npan || rb == 0) return;
    {
        const double *Yk = w0 + a.offY + (i64)a.kp * BB * N + QB * (rb - 1);
        for (int e = t; e < QB * BB; e += 256) {
            const int r = e % QB, i = e / QB;
            ys[r][i] = r < nr ? Yk[r + (i64)N * i] : 0.0;
        }
    }
    __syncthreads();
    double acc[4][4] = {};
    for (int r = 0; r < QB; r++) {
        double yv[4], zv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) { yv[u] = ys[r][tr + 16 * u]; zv[u] = zs[r][tc + 16 * u]; }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[u][v] += yv[u] * zv[v];
    }
    double *G = a.Gp + (((i64)b * a.ncb + cb) * a.npmax + (rb - 1)) * QB * QB;
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) G[(tr + 16 * u) + BB * (tc + 16 * v)] = acc[u][v];
}
// W2 = T_kp^T (sum of the np partials) for columns 16 q .. 16 q + 15 of column block cb of problem b,
// workgroup (q, cb, b)
__global__ __launch_bounds__(256) void k_band_qt_reduce(BandQtK a, int np) {
    __shared__ double ts[BB][BB + 1];
    __shared__ double ws[BB][17];
    const int q = blockIdx.x, cb = blockIdx.y, b = blockIdx.z, t = threadIdx.x;
    const double *Gb = a.Gp + ((i64)b * a.ncb + cb) * a.npmax * QB * QB + (i64)BB * 16 * q;
    const double *Tk = a.work + b * a.sW + a.offT + (i64)a.kp * BB * BB;
    for (int e = t; e < BB * BB; e += 256) ts[e % BB][e / BB] = Tk[e];     // ts[j][i] = T[j][i]
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    int p = 0;
    for (; p + 4 <= np; p += 4) {
        double x[4][4];
#pragma unroll
        for (int pp = 0; pp < 4; pp++)
#pragma unroll
            for (int u = 0; u < 4; u++) x[pp][u] = Gb[(i64)(p + pp) * QB * QB + t + 256 * u];
#pragma unroll
        for (int pp = 0; pp < 4; pp++)
#pragma unroll
            for (int u = 0; u < 4; u++) g[u] += x[pp][u];
    }
    for (; p < np; p++)
#pragma unroll
        for (int u = 0; u < 4; u++) g[u] += Gb[(i64)p * QB * QB + t + 256 * u];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int e = t + 256 * u;                // entry (j, c) of the quarter: j = e % BB, c = e / BB
        ws[e % BB][e / BB] = g[u];
    }
    __syncthreads();
    const int i = t % BB, c0 = t / BB;             // outputs (i, c0 + 4 v)
    double o[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = 0; j < BB; j++) {
        const double tv = ts[j][i];
#pragma unroll
        for (int v = 0; v < 4; v++) o[v] += tv * ws[j][c0 + 4 * v];
    }
    double *W2 = a.W2 + ((i64)b * a.ncb + cb) * QB * QB + (i64)BB * 16 * q;
#pragma unroll
    for (int v = 0; v < 4; v++) W2[i + BB * (c0 + 4 * v)] = o[v];
}
// the last block: Z[r0:] <- Qm^T Z[r0:] (m <= BB rows), workgroup (column block, problem)
__global__ __launch_bounds__(256) void k_band_qt_last(BandQtK a, int r0, int m, i64 offQl) {
    __shared__ double zs[QB][QB + 1];
    __shared__ double qs[BB][BB + 1];
    const int cb = blockIdx.x, b = blockIdx.y, t = threadIdx.x, cs = QB * cb;
    const int ncl = a.ncols - cs < QB ? a.ncols - cs : QB;
    double *Zb = a.Z + b * a.sZ + a.ldZ * cs + r0;
    const double *Ql = a.work + b * a.sW + offQl;
    for (int e = t; e < QB * QB; e += 256) {
        const int r = e % QB, c = e / QB;
        zs[r][c] = r < m && c < ncl ? Zb[r + a.ldZ * c] : 0.0;
        qs[r][c] = r < m && c < m ? Ql[r + BB * c] : 0.0;        // qs[j][i] = Qm[j][i]
    }
    __syncthreads();
    const int tr = t % 16, tc = t / 16;
    double o[4][4] = {};
    for (int j = 0; j < BB; j++) {
        double qv[4], zv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) { qv[u] = qs[j][tr + 16 * u]; zv[u] = zs[j][tc + 16 * u]; }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) o[u][v] += qv[u] * zv[v];
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = tr + 16 * u, c = tc + 16 * v;
            if (i < m && c < ncl) Zb[i + a.ldZ * c] = o[u][v];
        }
}

#define RBB(call) do { if ((call) != rocblas_status_success) { set_error("rocBLAS failed: %s", #call); return -1; } } while (0)

static inline unsigned nbl(i64 tot, int bs) { return (unsigned)((tot + bs - 1) / bs); }

int band_reduce(rocblas_handle h, hipStream_t st, int N, double *A, i64 sA, int nb, double *work, int *dflag) {
    if (N < 1 || nb < 1 || nb > 32) { set_error("band_reduce: unsupported N=%d nb=%d", N, nb); return -2; }
    RBB(rocblas_set_stream(h, st));
    const BandLayout L = band_layout(N);
    const i64 sW = L.size;
    double *w0 = work;
    hipLaunchKernelGGL(k_band_zero_flags, dim3(1), dim3(64), 0, st, dflag, nb);
    hipLaunchKernelGGL(k_band_sym, dim3(nbl((i64)N * N, 256), 1, nb), dim3(256), 0, st, A, N, sA);
    NV2_LAUNCH_CHECK();
    const double one = 1.0, zero = 0.0, mone = -1.0, mhalf = -0.5;
    const rocblas_operation Nn = rocblas_operation_none, Tt = rocblas_operation_transpose;
    for (int k = 0; k < L.npan; k++) {
        const int p0 = BB * k, r0 = p0 + BB, m = N - r0;
        double *P = A + r0 + (i64)N * p0, *A22 = A + r0 + (i64)N * r0;
        double *Yk = w0 + L.Y + (i64)k * BB * N, *Tk = w0 + L.T + (i64)k * BB * BB;
        double *Gp = w0 + L.Gp, *R1 = w0 + L.R1, *R1i = w0 + L.R1i, *R2 = w0 + L.R2, *R2i = w0 + L.R2i;
        double *Ui = w0 + L.Ui, *V1 = w0 + L.V1, *Q1 = w0 + L.Q1, *Q2 = w0 + L.Q2;
        double *W = w0 + L.W, *X = w0 + L.X, *W2 = X + (i64)BB * N;
        const int np = (m + KC - 1) / KC;
        // CholeskyQR2 of P
        hipLaunchKernelGGL(k_atb_parts, dim3(np, nb), dim3(BWG), 0, st, P, sA, P, sA, N, m, Gp, sW, nullptr);
        hipLaunchKernelGGL(k_bchol, dim3(nb), dim3(BCH_T), 0, st, Gp, sW, np, R1, R1i, dflag, 0.0);
        RBB(rocblas_dgemm_strided_batched(h, Nn, Tt, m, BB, BB, &one, P, N, sA, R1i, BB, sW, &zero, Q1, N, sW, nb));
        hipLaunchKernelGGL(k_atb_parts, dim3(np, nb), dim3(BWG), 0, st, Q1, sW, Q1, sW, N, m, Gp, sW, nullptr);
        hipLaunchKernelGGL(k_bchol, dim3(nb), dim3(BCH_T), 0, st, Gp, sW, np, R2, R2i, dflag, 0.1);
        RBB(rocblas_dgemm_strided_batched(h, Nn, Tt, m, BB, BB, &one, Q1, N, sW, R2i, BB, sW, &zero, Q2, N, sW, nb));
        // reflectors: Y_top, T, U^-1, and S R into the panel
        hipLaunchKernelGGL(k_brecon, dim3(nb), dim3(256), 0, st, A, N, sA, p0, r0, Q2, R1, R2, Yk, Tk, Ui, sW);
        NV2_LAUNCH_CHECK();
        RBB(rocblas_dgemm_strided_batched(h, Nn, Tt, m - BB, BB, BB, &one, Q2 + BB, N, sW, Ui, BB, sW, &zero, Yk + BB,
                                          N, sW, nb));
        // A22 <- H^T A22 H
        RBB(rocblas_dgemm_strided_batched(h, Nn, Nn, m, BB, m, &one, A22, N, sA, Yk, N, sW, &zero, W, N, sW, nb));
        RBB(rocblas_dgemm_strided_batched(h, Nn, Nn, m, BB, BB, &one, W, N, sW, Tk, BB, sW, &zero, W2, N, sW, nb));
        // Y^T W2 partials; the same reads copy Y into the outer blocks of X = [Y | W2 | Y]
        hipLaunchKernelGGL(k_atb_parts, dim3(np, nb), dim3(BWG), 0, st, Yk, sW, W2, sW, N, m, Gp, sW, X);
        hipLaunchKernelGGL(k_tx, dim3(nb), dim3(BWG), 0, st, Gp, np, Tk, V1, sW);
        NV2_LAUNCH_CHECK();
        RBB(rocblas_dgemm_strided_batched(h, Nn, Nn, m, BB, BB, &mhalf, Yk, N, sW, V1, BB, sW, &one, W2, N, sW, nb));
        // A22 -= Y W2^T + W2 Y^T as one rank-2 BAND_B update: [Y | W2] [W2 | Y]^T, both operands inside
        // X = [Y | W2 | Y] (one pass over A22 instead of two)
        RBB(rocblas_dgemm_strided_batched(h, Nn, Tt, m, m, 2 * BB, &mone, X, N, sW, X + (i64)BB * N, N, sW, &one, A22, N,
                                          sA, nb));
    }
    {
        const int p0 = BB * L.npan, r0 = p0 + BB, m = N - r0;
        if (m >= 1) {
            hipLaunchKernelGGL(k_blast, dim3(nb), dim3(BWG), 0, st, A, N, sA, p0, r0, w0 + L.Ql, sW);
            NV2_LAUNCH_CHECK();
        }
    }
    hipLaunchKernelGGL(k_band_extract, dim3(nbl((i64)N * BAND_W, 256), 1, nb), dim3(256), 0, st, A, N, sA,
                       w0 + L.Bl, sW);
    NV2_LAUNCH_CHECK();
    return 0;
}

static inline int qt_ncb(int ncols) { return (ncols + QB - 1) / QB; }
static inline int qt_npmax(int N) { return (N + QB - 1) / QB - 1; }
i64 band_qt_workspace(int N, int nb, int ncols) {
    const i64 ncb = qt_ncb(ncols);
    return (i64)nb * ncb * qt_npmax(N) * QB * QB + (i64)nb * ncb * QB * QB;
}
// Z_b <- Q_b^T Z_b in place for the nb problems b0 .. b0+nb-1 (Z_b = Z + (b - b0) sZ: N x ncols, ld ldZ).
// Q_b = H_0 H_1 ... H_{np-1} diag(I, Qm), so Q_b^T applies H_0^T first: per panel k, Z[r0:] -= Y_k T_k^T
// Y_k^T Z[r0:] (k_band_qt_step and k_band_qt_reduce), then Z[r0:] <- Qm^T Z[r0:] (k_band_qt_last).
int band_apply_qt(hipStream_t st, int N, const double *work, int b0, int nb, double *Z, i64 ldZ, i64 sZ, int ncols,
                  double *scratch) {
    if (nb < 1 || nb > 65535 || ncols < 0) { set_error("band_apply_qt: nb=%d ncols=%d", nb, ncols); return -2; }
    if (ncols == 0) return 0;
    const BandLayout L = band_layout(N);
    BandQtK a{};
    a.N = N;
    a.npan = L.npan;
    a.ncols = ncols;
    a.ncb = qt_ncb(ncols);
    a.npmax = qt_npmax(N);
    a.sW = L.size;
    a.offY = L.Y;
    a.offT = L.T;
    a.ldZ = ldZ;
    a.sZ = sZ;
    a.work = work + b0 * L.size;
    a.Z = Z;
    a.Gp = scratch;
    a.W2 = scratch + (i64)nb * a.ncb * a.npmax * QB * QB;
    for (int ku = -1; ku < L.npan; ku++) {
        a.ku = ku;
        a.kp = ku + 1;
        a.nrb = (N - BB * (ku + 1) + QB - 1) / QB;
        hipLaunchKernelGGL(k_band_qt_step, dim3(a.nrb, a.ncb, nb), dim3(256), 0, st, a);
        NV2_LAUNCH_CHECK();
        if (a.kp < L.npan) {
            hipLaunchKernelGGL(k_band_qt_reduce, dim3(QB / 16, a.ncb, nb), dim3(256), 0, st, a, a.nrb - 1);
            NV2_LAUNCH_CHECK();
        }
    }
    const int r0 = BB * (L.npan + 1), m = N - r0;
    if (m >= 1) {
        hipLaunchKernelGGL(k_band_qt_last, dim3(a.ncb, nb), dim3(256), 0, st, a, r0, m, L.Ql);
        NV2_LAUNCH_CHECK();
    }
    return 0;
}

}  // namespace nv2

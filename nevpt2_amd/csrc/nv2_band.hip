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
//     A22 -= Y Z^T + Z Y^T (rocBLAS GEMMs; the BAND_B x BAND_B products over the panel's rows
//     P^T P, Q1^T Q1 and Y^T W are split over row blocks, k_atb_parts).
// The last panel (m <= BAND_B rows) is factored by Householder steps in LDS, with its orthogonal
// factor kept explicitly.  The panels of every matrix of the batch advance together.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "nv2_band.h"

namespace nv2 {

#define BB BAND_B
#define BWG 256
#define KC 128           // rows per workgroup of the A^T B products with BAND_B x BAND_B results
#define KS 32            // rows per LDS stage of those products
// keeps the compiler from hoisting the loads of later unrolled iterations (and their registers) upward
#define BAND_FENCE() __asm__ volatile("" ::: "memory")

struct BandLayout {      // offsets (doubles) inside one problem's part of the workspace
    int npan;            // panels with more than BAND_B rows (CholeskyQR2)
    i64 Y, T, Ql, Bl, Gp, R1, R1i, R2, R2i, Ui, V1, Q1, Q2, W, W2, size;
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
    L.W2 = o; o += col;
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
// products at Gp + b * sW (BAND_B^2 apart): the four wavefronts of the workgroup add them into LDS
// (consecutive threads on consecutive entries, every partial's loads in flight together), then one
// wavefront factors: lane c keeps column c in registers and row j travels through LDS at step j (one
// wavefront, in-order LDS: no workgroup barrier).  R and R^-1 are stored by rows (element (i, c) at
// i * BAND_B + c).  flag[b] is set by a pivot that is not positive and, when dev > 0, by an entry of
// G - I above dev: the second CholeskyQR pass then starts from a Q1 too far from orthogonal for Q2 to
// be orthogonal to rounding.
#define WAVE_LDS_SYNC() __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory")
#define BCH_T 256
__global__ __launch_bounds__(BCH_T) void k_bchol(const double *Gp, i64 sW, int np, double *R, double *Ri, int *flag,
                                                 double dev) {
    __shared__ double gs[BB * BB];
    __shared__ double rb[BB];
    const int b = blockIdx.x, t = threadIdx.x, c = t & 63;
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
    if (t >= 64) return;
    double a[BB];
#pragma unroll
    for (int r = 0; r < BB; r++) a[r] = gs[c + BB * r];   // G[r][c] (symmetric: row c)
    bool off = false;
    if (dev > 0.0)
#pragma unroll
        for (int r = 0; r < BB; r++) off |= !(fabs(a[r] - (r == c ? 1.0 : 0.0)) <= dev);
    bool bad = __ballot(off) != 0;
#pragma unroll
    for (int j = 0; j < BB; j++) {          // right-looking: a[r] = G[r][c] for j <= r <= c
        rb[c] = a[j];
        WAVE_LDS_SYNC();
        const double d = rb[j];
        bad |= !(d > 0.0);                  // uniform
        const double u = sqrt(d > 0.0 ? d : 1.0), inv = 1.0 / u;
        const double rjc = c == j ? u : a[j] * inv;   // R[j][c]
        a[j] = rjc;
#pragma unroll
        for (int r = j + 1; r < BB; r++) a[r] -= (rb[r] * inv) * rjc;
        WAVE_LDS_SYNC();
    }
    double *Rb = R + b * sW, *Rib = Ri + b * sW;
#pragma unroll
    for (int i = 0; i < BB; i++) Rb[i * BB + c] = i <= c ? a[i] : 0.0;
#pragma unroll
    for (int i = BB - 1; i >= 0; i--) {     // X = R^-1 in place, rows from the bottom: a[k] = X[k][c], k > i
        rb[c] = a[i];
        WAVE_LDS_SYNC();
        double s = c == i ? 1.0 : 0.0;
#pragma unroll
        for (int k = i + 1; k < BB; k++) s -= rb[k] * a[k];
        a[i] = s / rb[i];
        WAVE_LDS_SYNC();
    }
#pragma unroll
    for (int i = 0; i < BB; i++) Rib[i * BB + c] = a[i];
    if (bad) {                              // R = R^-1 = I keeps the batch finite; the flag sends this problem elsewhere
#pragma unroll
        for (int i = 0; i < BB; i++) { Rb[i * BB + c] = i == c ? 1.0 : 0.0; Rib[i * BB + c] = i == c ? 1.0 : 0.0; }
        if (c == 0) flag[b] = 1;
    }
}
// Partial products of A^T B for A, B with BAND_B columns (ld, m rows; problem b = blockIdx.y at A + b*sA,
// B + b*sB): workgroup c sums rows c*KC .. c*KC+KC-1 into Gp + b*sW + c*BAND_B^2.  rocBLAS gives a
// BAND_B x BAND_B result with thousands of rows to a single workgroup; here ceil(m / KC) of them share it.
__global__ __launch_bounds__(BWG) void k_atb_parts(const double *A, i64 sA, const double *B, i64 sB, int ld, int m,
                                                   double *Gp, i64 sW) {
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
            sa[k][i] = s0 + k < kn ? Ab[s0 + k + (i64)ld * i] : 0.0;
            sb[k][i] = s0 + k < kn ? Bb[s0 + k + (i64)ld * i] : 0.0;
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

// Householder form of panel k from its CholeskyQR2 factors, problem b = blockIdx.x, by two wavefronts
// (lane c of a wavefront keeps column c of the matrix at hand in registers; rows and columns that
// other lanes need travel through LDS).  Wavefront 0 factors Q2_top - S = L U without pivoting and
// inverts U (stored by rows); wavefront 1 meanwhile forms R~ = S R2 R1 into the panel's top block of A
// (the rows below it are outside the band and never read again), then, once L and U are there, Y_top = L and T = -U S L^-T.  R1, R2 come stored by
// rows.  The phases of one wavefront order their LDS traffic by program order (one wavefront, in-order
// LDS); the two meet at one workgroup barrier.
__global__ __launch_bounds__(128) void k_brecon(double *A, int N, i64 sA, int p0, int r0, const double *Q2,
                                                const double *R1, const double *R2, double *Y, double *T,
                                                double *Ui, i64 sW) {
    __shared__ double mt[BB][BB + 1];    // the LU factors
    __shared__ double r2[BB][BB + 1];    // R2
    __shared__ double sgs[BB], rb[BB], cb[BB];
    const int b = blockIdx.x, c = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double *Ab = A + b * sA;
    const double *Qb = Q2 + b * sW, *R1b = R1 + b * sW, *R2b = R2 + b * sW;
    const double sgc = Qb[c + (i64)N * c] >= 0.0 ? -1.0 : 1.0;
    double m[BB];
    if (wv == 0) {
        sgs[c] = sgc;
        // LU without pivoting of Q2_top - S (|pivot| >= 1 by the choice of S); m = column c
#pragma unroll
        for (int r = 0; r < BB; r++) m[r] = Qb[r + (i64)N * c] - (r == c ? sgc : 0.0);
#pragma unroll
        for (int k = 0; k < BB; k++) {
            rb[c] = m[k];                // row k: U[k][c] for c >= k
            if (c == k) {
#pragma unroll
                for (int i = k + 1; i < BB; i++) cb[i] = m[i];   // column k below the pivot
            }
            WAVE_LDS_SYNC();
            const double rp = 1.0 / rb[k], ukc = m[k];
#pragma unroll
            for (int i = k + 1; i < BB; i++) {
                const double l = cb[i] * rp;
                m[i] = c > k ? m[i] - l * ukc : (c == k ? l : m[i]);
            }
            WAVE_LDS_SYNC();
        }
#pragma unroll
        for (int r = 0; r < BB; r++) mt[r][c] = m[r];
    } else {
#pragma unroll
        for (int r = 0; r < BB; r++) { r2[r][c] = R2b[r * BB + c]; m[r] = R1b[r * BB + c]; }   // m = column c of R1
        WAVE_LDS_SYNC();
        // R~[r][c] = S_r sum_{r <= k <= c} R2[r][k] R1[k][c] (R1[k][c] = 0 for k > c)
        const double *Qd = Qb;
#pragma unroll
        for (int r = 0; r < BB; r++) {
            double v = 0.0;
#pragma unroll
            for (int k = r; k < BB; k++) v += r2[r][k] * m[k];
            const double sr = Qd[r + (i64)N * r] >= 0.0 ? -1.0 : 1.0;
            Ab[(r0 + r) + (i64)N * (p0 + c)] = r <= c ? sr * v : 0.0;
            BAND_FENCE();
        }
        // rows r0 + BB .. N-1 of the panel's columns lie outside the band: nothing reads them again
    }
    __syncthreads();                     // L, U and S in LDS
    double *Yb = Y + b * sW, *Tb = T + b * sW, *Uib = Ui + b * sW;
    if (wv == 0) {
        // U^-1 by rows from the bottom, in m: m[k] = X[k][c] for k > i (zero for k > c)
#pragma unroll
        for (int i = BB - 1; i >= 0; i--) {
            double v = c == i ? 1.0 : 0.0;
#pragma unroll
            for (int k = i + 1; k < BB; k++) v -= mt[i][k] * m[k];
            m[i] = v / mt[i][i];
            BAND_FENCE();
        }
#pragma unroll
        for (int i = 0; i < BB; i++) Uib[i * BB + c] = m[i];
    } else {
#pragma unroll
        for (int r = 0; r < BB; r++) Yb[r + (i64)N * c] = r > c ? mt[r][c] : (r == c ? 1.0 : 0.0);
        // row c of L^-1: L^T w = e_c, from the bottom (w[k] = 0 for k > c)
#pragma unroll
        for (int k = BB - 1; k >= 0; k--) {
            double v = k == c ? 1.0 : 0.0;
#pragma unroll
            for (int t = k + 1; t < BB; t++) v -= mt[t][k] * m[t];
            m[k] = v;
            BAND_FENCE();
        }
        // T[r][c] = -sum_{r <= k <= c} U[r][k] S_k L^-1[c][k]
#pragma unroll
        for (int r = 0; r < BB; r++) {
            double v = 0.0;
#pragma unroll
            for (int k = r; k < BB; k++) v -= mt[r][k] * sgs[k] * m[k];
            Tb[r + BB * c] = r <= c ? v : 0.0;
            BAND_FENCE();
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

// z = Q_b^T y for column sets of right-hand sides (one launch for every set; QT_C columns per
// workgroup, y read and written in place in global memory).  Q_b = H_0 H_1 ... H_{np-1} diag(I, Qm)
// from band_reduce, so Q_b^T y applies H_0^T first: for panel k (rows r0 = BB (k+1) .. N-1, m of them)
// w = Y_k^T y[r0:], w' = T_k^T w, y[r0:] -= Y_k w'; then y[r0:] <- Qm^T y[r0:] for the last block.
// Thread (i = t % 64, c = t / 64) owns w[i][c]; the m x BB reflector block streams through LDS in chunks
// of QT_R rows, once for the product and once for the update.  Every workgroup of a set walks the
// same panels in the same order, so a chunk is shared through L2.
#define QT_C 4
#define QT_R 64
__global__ __launch_bounds__(256) void k_band_qt(BandQtArgs a) {
    __shared__ double yk[QT_R][BB + 1];
    __shared__ double tk[BB][BB + 1];
    __shared__ double ys[QT_R][QT_C];
    __shared__ double wv[BB][QT_C], wv2[BB][QT_C];
    int si = 0;
    while (si + 1 < a.nset && (int)blockIdx.x >= a.s[si + 1].wg0) si++;
    const BandQtSet S = a.s[si];
    const int N = a.N, t = threadIdx.x, i = t & 63, c = t >> 6;
    const int c0 = ((int)blockIdx.x - S.wg0) * QT_C;
    const int nc = S.ncols - c0 < QT_C ? S.ncols - c0 : QT_C;
    double *Y = S.Y + S.ld * c0;
    for (int k = 0; k < a.npan; k++) {
        const int r0 = BB * (k + 1), m = N - r0;
        const double *Yk = S.w + a.offY + (i64)k * BB * N, *Tk = S.w + a.offT + (i64)k * BB * BB;
        double acc = 0.0;
        for (int rc = 0; rc < m; rc += QT_R) {
            __syncthreads();
#pragma unroll
            for (int u = 0; u < QT_R * BB / 256; u++) {
                const int e = t + 256 * u, rr = e % QT_R, ii = e / QT_R;
                yk[rr][ii] = rc + rr < m ? Yk[(rc + rr) + (i64)N * ii] : 0.0;
            }
            {
                const int rr = t % QT_R, cc = t / QT_R;
                ys[rr][cc] = rc + rr < m && cc < nc ? Y[(r0 + rc + rr) + S.ld * cc] : 0.0;
            }
            __syncthreads();
#pragma unroll 16
            for (int rr = 0; rr < QT_R; rr++) acc += yk[rr][i] * ys[rr][c];
        }
#pragma unroll
        for (int u = 0; u < BB * BB / 256; u++) {
            const int e = t + 256 * u, jj = e % BB, ii = e / BB;
            tk[jj][ii] = Tk[jj + BB * ii];
        }
        wv[i][c] = acc;
        __syncthreads();
        double w2 = 0.0;
#pragma unroll 16
        for (int jj = 0; jj < BB; jj++) w2 += tk[jj][i] * wv[jj][c];
        wv2[i][c] = w2;
        for (int rc = 0; rc < m; rc += QT_R) {
            __syncthreads();
#pragma unroll
            for (int u = 0; u < QT_R * BB / 256; u++) {
                const int e = t + 256 * u, rr = e % QT_R, ii = e / QT_R;
                yk[rr][ii] = rc + rr < m ? Yk[(rc + rr) + (i64)N * ii] : 0.0;
            }
            __syncthreads();
            const int rr = t % QT_R, cc = t / QT_R;
            if (rc + rr < m && cc < nc) {
                double v = 0.0;
#pragma unroll 16
                for (int ii = 0; ii < BB; ii++) v += yk[rr][ii] * wv2[ii][cc];
                Y[(r0 + rc + rr) + S.ld * cc] -= v;
            }
        }
        __syncthreads();
    }
    const int r0 = BB * (a.npan + 1), m = N - r0;
    if (m >= 1) {
        const double *Ql = S.w + a.offQl;
#pragma unroll
        for (int u = 0; u < BB * BB / 256; u++) {
            const int e = t + 256 * u, jj = e % BB, ii = e / BB;
            tk[jj][ii] = jj < m && ii < m ? Ql[jj + BB * ii] : 0.0;
        }
        {
            const int rr = t % QT_R, cc = t / QT_R;
            ys[rr][cc] = rr < m && cc < nc ? Y[(r0 + rr) + S.ld * cc] : 0.0;
        }
        __syncthreads();
        if (i < m && c < nc) {
            double v = 0.0;
            for (int jj = 0; jj < m; jj++) v += tk[jj][i] * ys[jj][c];
            Y[(r0 + i) + S.ld * c] = v;
        }
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
        double *W = w0 + L.W, *W2 = w0 + L.W2;
        const int np = (m + KC - 1) / KC;
        // CholeskyQR2 of P
        hipLaunchKernelGGL(k_atb_parts, dim3(np, nb), dim3(BWG), 0, st, P, sA, P, sA, N, m, Gp, sW);
        hipLaunchKernelGGL(k_bchol, dim3(nb), dim3(BCH_T), 0, st, Gp, sW, np, R1, R1i, dflag, 0.0);
        RBB(rocblas_dgemm_strided_batched(h, Nn, Tt, m, BB, BB, &one, P, N, sA, R1i, BB, sW, &zero, Q1, N, sW, nb));
        hipLaunchKernelGGL(k_atb_parts, dim3(np, nb), dim3(BWG), 0, st, Q1, sW, Q1, sW, N, m, Gp, sW);
        hipLaunchKernelGGL(k_bchol, dim3(nb), dim3(BCH_T), 0, st, Gp, sW, np, R2, R2i, dflag, 0.1);
        RBB(rocblas_dgemm_strided_batched(h, Nn, Tt, m, BB, BB, &one, Q1, N, sW, R2i, BB, sW, &zero, Q2, N, sW, nb));
        // reflectors: Y_top, T, U^-1, and S R into the panel
        hipLaunchKernelGGL(k_brecon, dim3(nb), dim3(128), 0, st, A, N, sA, p0, r0, Q2, R1, R2, Yk, Tk, Ui, sW);
        NV2_LAUNCH_CHECK();
        RBB(rocblas_dgemm_strided_batched(h, Nn, Tt, m - BB, BB, BB, &one, Q2 + BB, N, sW, Ui, BB, sW, &zero, Yk + BB,
                                          N, sW, nb));
        // A22 <- H^T A22 H
        RBB(rocblas_dgemm_strided_batched(h, Nn, Nn, m, BB, m, &one, A22, N, sA, Yk, N, sW, &zero, W, N, sW, nb));
        RBB(rocblas_dgemm_strided_batched(h, Nn, Nn, m, BB, BB, &one, W, N, sW, Tk, BB, sW, &zero, W2, N, sW, nb));
        hipLaunchKernelGGL(k_atb_parts, dim3(np, nb), dim3(BWG), 0, st, Yk, sW, W2, sW, N, m, Gp, sW);
        hipLaunchKernelGGL(k_tx, dim3(nb), dim3(BWG), 0, st, Gp, np, Tk, V1, sW);
        NV2_LAUNCH_CHECK();
        RBB(rocblas_dgemm_strided_batched(h, Nn, Nn, m, BB, BB, &mhalf, Yk, N, sW, V1, BB, sW, &one, W2, N, sW, nb));
        RBB(rocblas_dgemm_strided_batched(h, Nn, Tt, m, m, BB, &mone, Yk, N, sW, W2, N, sW, &one, A22, N, sA, nb));
        RBB(rocblas_dgemm_strided_batched(h, Nn, Tt, m, m, BB, &mone, W2, N, sW, Yk, N, sW, &one, A22, N, sA, nb));
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

// z = Q_b^T y in place for up to BAND_QT_SETS column sets (set s: problem b_s's Q, ncols_s columns of
// order N at Y_s, leading dimension ld_s), one launch
int band_apply_qt(hipStream_t st, int N, const double *work, int nset, const int *b, double *const *Y,
                  const i64 *ld, const int *ncols) {
    if (nset < 0 || nset > BAND_QT_SETS) { set_error("band_apply_qt: %d column sets", nset); return -2; }
    const BandLayout L = band_layout(N);
    BandQtArgs a{};
    a.N = N;
    a.npan = L.npan;
    a.offY = L.Y;
    a.offT = L.T;
    a.offQl = L.Ql;
    int wg = 0;
    for (int q = 0; q < nset; q++) {
        if (ncols[q] <= 0) continue;
        BandQtSet &S = a.s[a.nset++];
        S.w = work + b[q] * L.size;
        S.Y = Y[q];
        S.ld = ld[q];
        S.ncols = ncols[q];
        S.wg0 = wg;
        wg += (ncols[q] + QT_C - 1) / QT_C;
    }
    if (!wg) return 0;
    hipLaunchKernelGGL(k_band_qt, dim3(wg), dim3(256), 0, st, a);
    NV2_LAUNCH_CHECK();
    return 0;
}

}  // namespace nv2

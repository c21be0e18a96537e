// Householder tridiagonalisation of a batch of dense symmetric matrices and the two ways
// the stage-2 PC solves use it (the dsyev('V','U') calls of qdnevpt.F:5723-5797 and
// 6303-6410, N = n^3):
//   * eigenvectors U = Q Z of the metric, with Z from rocsolver_dstedc on T;
//   * Q^T Y for the Koopmans matrix K', whose eigenvectors are never formed: the PC sums
//     sum_mu r_mu^2 / (lambda_mu + eps) are y^T (K' + eps)^-1 y, a tridiagonal solve.
//
// A = Q T Q^T in the LAPACK dsytrd 'L' storage: T on the diagonal and first subdiagonal,
// the reflector v_k = e_{k+1} + A(k+2:N, k) below it, H_k = I - tau_k v_k v_k^T and
// Q = H_0 H_1 ... H_{N-2}.  The input is symmetrised from its upper triangle first (the
// triangle dsyev('V','U') reads).
//
// Two launches per column k, both batched over the matrices: k_trd_cols (one wavefront per
// trailing column) applies the rank-2 update of step k-1 and forms (A v_k)_c as a column
// dot product (A is symmetric, so there are no partial sums to reduce); k_trd_house (one
// workgroup per matrix) forms w_k and generates reflector k+1.  Fusing the two through a
// last-workgroup-done counter needs an agent-scope release per workgroup, which writes back
// the freshly updated columns from every L2 and measured twice as slow.  The 2(N-2)+3
// launches are captured in a hipGraph once per shape and buffer set, then replayed.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "nv2_common.h"
#include "nv2_trd.h"

namespace nv2 {

#define TRYT(x) do { int _r = (x); if (_r) return _r; } while (0)

#define TRD_WG 256      // threads of a k_trd_cols workgroup (4 wavefronts)
#define TRD_HOUSE 512   // threads of the k_trd_house workgroup
#define TRD_U 4         // independent loads in flight per lane

__global__ void k_trd_sym_upper(double *A, int N, i64 sA) {
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (i64)N * N) return;
    const int i = (int)(e % N), j = (int)(e / N);
    double *Ab = A + blockIdx.z * sA;
    if (i > j) Ab[e] = Ab[j + (i64)N * i];
}

__device__ inline double trd_wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

__device__ inline double trd_block_sum(double x, double *red) {
    x = trd_wave_sum(x);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();                   // red[] may still be read by a previous call
    if (lane == 0) red[wid] = x;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
    return s;
}

// Householder step of matrix b by one workgroup.  k = -1: reflector of column 0 from the
// input.  k >= 0: w_k = p - (tau_k/2)(p^T v_k) v_k from p = tau_k A v_k, then reflector k+1
// of column k+1 brought up to date with step k (stored values carry steps <= k-1).
__device__ void trd_house(double *Ab, int N, int k, const double *vk, double *wk, double *vn, const double *pb,
                          double *db, double *eb, double *taub, double *red, double *sc) {
    const int c1 = k + 1, nt = blockDim.x;
    double alpha = 0.0;
    if (k >= 0) {
        double s = 0.0;
        for (int r = c1 + threadIdx.x; r < N; r += nt) s += pb[r] * vk[r];
        s = trd_block_sum(s, red);
        alpha = -0.5 * taub[k] * s;
        for (int r = c1 + threadIdx.x; r < N; r += nt) wk[r] = pb[r] + alpha * vk[r];
    }
    const double vc1 = k >= 0 ? vk[c1] : 0.0, wc1 = k >= 0 ? pb[c1] + alpha * vc1 : 0.0;
    double xn = 0.0;                    // |x|^2, x = rows c1+2.. of the column
    for (int r = c1 + 2 + threadIdx.x; r < N; r += nt) {
        double a = Ab[r + (i64)N * c1];
        if (k >= 0) a -= vk[r] * wc1 + (pb[r] + alpha * vk[r]) * vc1;
        xn += a * a;
    }
    xn = trd_block_sum(xn, red);
    if (threadIdx.x == 0) {
        double a0 = Ab[c1 + (i64)N * c1], a1 = Ab[(c1 + 1) + (i64)N * c1];
        if (k >= 0) {
            const double v1 = vk[c1 + 1], w1 = pb[c1 + 1] + alpha * v1;
            a0 -= 2.0 * vc1 * wc1;
            a1 -= v1 * wc1 + w1 * vc1;
        }
        double beta, t, scale;
        if (xn == 0.0) { beta = a1; t = 0.0; scale = 0.0; }
        else {
            beta = -copysign(sqrt(a1 * a1 + xn), a1);
            t = (beta - a1) / beta;
            scale = 1.0 / (a1 - beta);
        }
        db[c1] = a0;
        eb[c1] = beta;
        taub[c1] = t;
        *sc = scale;
        Ab[c1 + (i64)N * c1] = a0;
        Ab[(c1 + 1) + (i64)N * c1] = beta;
        if (c1 + 2 == N) {              // last reflector: T(N-1,N-1) as well
            double an = Ab[(N - 1) + (i64)N * (N - 1)];
            if (k >= 0) an -= 2.0 * vk[N - 1] * (pb[N - 1] + alpha * vk[N - 1]);
            db[N - 1] = an;
        }
    }
    __syncthreads();
    const double scale = *sc;
    for (int r = threadIdx.x; r < N; r += nt) {   // each element read and written by one thread
        double v = 0.0;
        if (r == c1 + 1) v = 1.0;
        else if (r >= c1 + 2) {
            double a = Ab[r + (i64)N * c1];
            if (k >= 0) a -= vk[r] * wc1 + (pb[r] + alpha * vk[r]) * vc1;
            v = a * scale;
            Ab[r + (i64)N * c1] = v;
        }
        vn[r] = v;
    }
}

// step k >= 0, phase 1, matrix b = blockIdx.z: every wavefront takes q trailing columns c
// (rows and columns o = k+1 .. N-1), applies the pending rank-2 update of step k-1 to the
// stored column and forms p_c = tau_k (A v_k)_c as the column's dot product with v_k (A is
// symmetric, so there are no partial sums to combine).
__global__ __launch_bounds__(TRD_WG) void k_trd_cols(double *A, int N, i64 sA, int k, int q, const double *vbuf,
                                                    const double *wbuf, i64 sV, const double *tau, double *p) {
    extern __shared__ double sh[];
    const int b = blockIdx.z, o = k + 1, m = N - o;
    double *Ab = A + b * sA;
    const double *vk = vbuf + b * 2 * sV + (i64)(k & 1) * sV;          // v_k
    const double *vp = vbuf + b * 2 * sV + (i64)((k + 1) & 1) * sV;    // v_{k-1}
    const double *wp = wbuf + b * 2 * sV + (i64)((k + 1) & 1) * sV;    // w_{k-1}
    const bool upd = k >= 1;
    double *sv = sh, *svp = sh + m, *swp = sh + 2 * m;
    for (int t = threadIdx.x; t < m; t += TRD_WG) {
        sv[t] = vk[o + t];
        svp[t] = upd ? vp[o + t] : 0.0;
        swp[t] = upd ? wp[o + t] : 0.0;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double tk = tau[(i64)b * N + k];
    for (int j = 0; j < q; j++) {
        const int cl = ((int)blockIdx.x * (TRD_WG / 64) + wave) * q + j;
        if (cl >= m) break;                                            // uniform per wavefront
        double *col = Ab + (i64)N * (o + cl) + o;
        const double vcl = svp[cl], wcl = swp[cl];
        double acc = 0.0;
        for (int r0 = lane; r0 < m; r0 += 64 * TRD_U) {
            double a[TRD_U];
#pragma unroll
            for (int u = 0; u < TRD_U; u++) {
                const int r = r0 + 64 * u;
                a[u] = r < m ? col[r] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < TRD_U; u++) {
                const int r = r0 + 64 * u;
                if (r < m) {
                    if (upd) {
                        a[u] -= svp[r] * wcl + swp[r] * vcl;
                        col[r] = a[u];
                    }
                    acc += a[u] * sv[r];
                }
            }
        }
        acc = trd_wave_sum(acc);
        if (lane == 0) p[b * sV + o + cl] = tk * acc;
    }
}

// step k, phase 2 (k = -1: the first reflector), one workgroup per matrix
__global__ __launch_bounds__(TRD_HOUSE) void k_trd_house(double *A, int N, i64 sA, int k, double *vbuf, double *wbuf,
                                                        i64 sV, const double *p, double *d, double *e,
                                                        double *tau) {
    __shared__ double red[TRD_HOUSE / 64];
    __shared__ double sc;
    const int b = blockIdx.x;
    trd_house(A + b * sA, N, k, vbuf + b * 2 * sV + (i64)(k & 1) * sV, wbuf + b * 2 * sV + (i64)(k & 1) * sV,
              vbuf + b * 2 * sV + (i64)((k + 1) & 1) * sV, p + b * sV, d + (i64)b * N, e + (i64)b * N,
              tau + (i64)b * N, red, &sc);
}

__global__ void k_trd_zero(double *x, i64 cnt) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < cnt) x[t] = 0.0;
}

__global__ void k_trd_small(double *A, int N, i64 sA, double *d, double *e, double *tau);

static inline unsigned nblk(i64 tot, int bs) { return (unsigned)((tot + bs - 1) / bs); }

// work: vbuf (2N per matrix), wbuf (2N), p (N), d, e, tau (N each)
i64 trd_workspace(int N, int nb) { return (i64)8 * N * nb; }
TrdOut trd_outputs(int N, int nb, double *work) {
    TrdOut o;
    o.d = work + (i64)5 * N * nb;
    o.e = o.d + (i64)N * nb;
    o.tau = o.e + (i64)N * nb;
    return o;
}

static int trd_launches(hipStream_t st, int N, double *A, i64 sA, int nb, double *work) {
    const i64 sV = N;
    double *vbuf = work, *wbuf = vbuf + (i64)nb * 2 * sV, *p = wbuf + (i64)nb * 2 * sV;
    const TrdOut o = trd_outputs(N, nb, work);
    hipLaunchKernelGGL(k_trd_zero, dim3(nblk((i64)5 * N * nb, 256)), dim3(256), 0, st, work, (i64)5 * N * nb);
    hipLaunchKernelGGL(k_trd_sym_upper, dim3(nblk((i64)N * N, 256), 1, nb), dim3(256), 0, st, A, N, sA);
    hipLaunchKernelGGL(k_trd_house, dim3(nb), dim3(TRD_HOUSE), 0, st, A, N, sA, -1, vbuf, wbuf, sV, p, o.d, o.e,
                       o.tau);
    for (int k = 0; k <= N - 3; k++) {
        const int m = N - k - 1;
        int q = (int)(((i64)m * nb + 1535) / 1536);          // about 1536 wavefronts over the batch
        if (q < 1) q = 1;
        const unsigned gx = nblk(m, (TRD_WG / 64) * q);
        hipLaunchKernelGGL(k_trd_cols, dim3(gx, 1, nb), dim3(TRD_WG), sizeof(double) * 3 * m, st, A, N, sA, k, q,
                           vbuf, wbuf, sV, o.tau, p);
        hipLaunchKernelGGL(k_trd_house, dim3(nb), dim3(TRD_HOUSE), 0, st, A, N, sA, k, vbuf, wbuf, sV, p, o.d, o.e,
                           o.tau);
    }
    return 0;
}

struct TrdPlan {
    int N = 0, nb = 0;
    double *A = nullptr, *work = nullptr;
    i64 sA = 0;
    hipStream_t st = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
};
void trd_plan_destroy(TrdPlan *pl) {
    if (!pl) return;
    if (pl->ge) (void)hipGraphExecDestroy(pl->ge);
    if (pl->g) (void)hipGraphDestroy(pl->g);
    delete pl;
}

int trd_reduce(hipStream_t st, int N, double *A, i64 sA, int nb, double *work, TrdPlan **plan) {
    if (N < 1 || nb < 1 || nb > 65535) { set_error("trd_reduce: unsupported N=%d nb=%d", N, nb); return -2; }
    if (N <= 2) {
        const TrdOut o = trd_outputs(N, nb, work);
        hipLaunchKernelGGL(k_trd_small, dim3(nb), dim3(64), 0, st, A, N, sA, o.d, o.e, o.tau);
        NV2_LAUNCH_CHECK();
        return 0;
    }
    const int lds_max = 160 * 1024;   // the LDS of one CU
    if (sizeof(double) * 3 * (size_t)N > (size_t)lds_max) { set_error("trd_reduce: N=%d exceeds the LDS staging", N); return -2; }
    static const hipError_t attr =    // once per process (thread-safe static initialisation)
        hipFuncSetAttribute((const void *)k_trd_cols, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    NV2_HIP(attr);
    if (!plan) {
        TRYT(trd_launches(st, N, A, sA, nb, work));
        NV2_LAUNCH_CHECK();
        return 0;
    }
    TrdPlan *pl = *plan;
    if (!pl || pl->N != N || pl->nb != nb || pl->A != A || pl->sA != sA || pl->work != work || pl->st != st) {
        // the N-1 launches are captured once per shape and buffer set, then replayed
        trd_plan_destroy(pl);
        *plan = pl = new TrdPlan();
        pl->N = N; pl->nb = nb; pl->A = A; pl->sA = sA; pl->work = work; pl->st = st;
        NV2_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = trd_launches(st, N, A, sA, nb, work);
        const hipError_t ec = hipStreamEndCapture(st, &pl->g);
        if (rc) return rc;
        NV2_HIP(ec);
        NV2_HIP(hipGraphInstantiate(&pl->ge, pl->g, nullptr, nullptr, 0));
    }
    NV2_HIP(hipGraphLaunch(pl->ge, st));
    return 0;
}

// ---------------------------------------------------------------- applying Q
// Q = H_0 ... H_{N-2} = I - V T V^T with V the unit lower trapezoidal matrix of the reflectors
// and T upper triangular, T^-1 = diag(1/tau) + striu(V^T V) (the compact WY form built in one
// piece).  A reflector with tau = 0 is the identity: its column of V is zeroed and its
// diagonal entry set to 1, so it drops out.
__global__ void k_trd_v(const double *A, int N, i64 sA, const double *tau, double *V, i64 sVb) {
    const int K = N - 1;
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (i64)N * K) return;
    const int b = blockIdx.z, r = (int)(t % N), k = (int)(t / N);
    const double *Ab = A + b * sA;
    double v = r == k + 1 ? 1.0 : (r > k + 1 ? Ab[r + (i64)N * k] : 0.0);
    if (tau[(i64)b * N + k] == 0.0) v = 0.0;
    V[b * sVb + t] = v;
}
__global__ void k_trd_tinv_diag(double *S, int K, i64 sS, const double *tau, int N) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const double t = tau[(i64)blockIdx.z * N + k];
    S[blockIdx.z * sS + k + (i64)K * k] = t != 0.0 ? 1.0 / t : 1.0;
}
// order <= 2: T is the matrix itself (upper triangle), no reflectors
__global__ void k_trd_small(double *A, int N, i64 sA, double *d, double *e, double *tau) {
    const int b = blockIdx.x;
    const double *Ab = A + b * sA;
    if (threadIdx.x) return;
    d[(i64)b * N] = Ab[0];
    e[(i64)b * N] = N > 1 ? Ab[N] : 0.0;
    tau[(i64)b * N] = 0.0;
    if (N > 1) { d[(i64)b * N + 1] = Ab[N + 1]; tau[(i64)b * N + 1] = 0.0; }
}

#define RBT(call) do { if ((call) != rocblas_status_success) { set_error("rocBLAS failed: %s", #call); return -1; } } while (0)

i64 trd_apply_workspace(int N, int nb, int ncol) {
    const i64 K = N > 1 ? N - 1 : 1;
    return (i64)nb * ((i64)N * K + K * K + K * (ncol > 0 ? ncol : 1));
}

// X_b (N x ncol, ld N, stride sX) <- Q_b X_b (trans = false) or Q_b^T X_b (trans = true):
// W = V^T X, then W <- T W = S^-1 W (or T^T W = S^-T W) by a triangular solve, X -= V W.
int trd_apply_q(rocblas_handle h, hipStream_t st, int N, const double *A, i64 sA, int nb, const double *tau,
                bool trans, int ncol, double *X, i64 sX, double *scratch) {
    if (ncol <= 0 || N <= 2) return 0;                 // N <= 2: no reflectors (Q = I)
    const int K = N - 1;
    const i64 sVb = (i64)N * K, sS = (i64)K * K, sW = (i64)K * ncol;
    double *V = scratch, *S = V + nb * sVb, *W = S + nb * sS;
    const double one = 1.0, zero = 0.0, mone = -1.0;
    hipLaunchKernelGGL(k_trd_v, dim3(nblk((i64)N * K, 256), 1, nb), dim3(256), 0, st, A, N, sA, tau, V, sVb);
    NV2_LAUNCH_CHECK();
    RBT(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, K, K, N, &one, V, N, sVb,
                                      V, N, sVb, &zero, S, K, sS, nb));
    hipLaunchKernelGGL(k_trd_tinv_diag, dim3(nblk(K, 256), 1, nb), dim3(256), 0, st, S, K, sS, tau, N);
    NV2_LAUNCH_CHECK();
    RBT(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, K, ncol, N, &one, V, N,
                                      sVb, X, N, sX, &zero, W, K, sW, nb));
    RBT(rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_upper,
                                      trans ? rocblas_operation_transpose : rocblas_operation_none,
                                      rocblas_diagonal_non_unit, K, ncol, &one, S, K, sS, W, K, sW, nb));
    RBT(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, N, ncol, K, &mone, V, N, sVb,
                                      W, K, sW, &one, X, N, sX, nb));
    return 0;
}

}  // namespace nv2

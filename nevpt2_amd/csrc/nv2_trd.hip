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
// One launch per column k (k_trd_col), batched over the matrices.  Every workgroup repeats the
// O(N) Householder step of column k itself instead of waiting for one workgroup to publish it
// (a last-workgroup-done handshake needs an agent-scope release per workgroup, which writes back
// the freshly updated columns from every L2 and measured twice as slow); then one wavefront per
// trailing column applies the rank-2 update of step k-1 and forms (A v_k)_c as a column dot
// product (A is symmetric, so there are no partial sums to reduce).  The launches are not
// captured in a hipGraph: the callers run on worker threads beside others that synchronise the
// device, which invalidates a capture in progress (and replay saved only ~4 % per column).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdlib>

#include "nv2_common.h"
#include "nv2_trd.h"

namespace nv2 {

#define TRYT(x) do { int _r = (x); if (_r) return _r; } while (0)


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

#define TRD_NMAX 4096    // largest order (three vectors of it staged in LDS)

// Launch k of the reduction, matrix b = blockIdx.z.  Every workgroup first repeats the
// Householder step of column k (cheap: O(N - k) loads and two block sums), so no workgroup
// ever waits for another:
//   k >= 1: w_{k-1} = p_{k-1} - (tau_{k-1}/2)(p_{k-1}^T v_{k-1}) v_{k-1}, with
//           p_{k-1} = tau_{k-1} A v_{k-1} from launch k-1; column k brought up to date with
//           step k-1 (its stored values carry steps <= k-2);
//   then    reflector v_k, tau_k from that column (d_k, e_k of T).
// Then every wavefront takes q trailing columns c >= k+1 (rows >= k+1), applies step k-1 to
// the stored column and forms p_k[c] = tau_k (A v_k)_c as the column's dot product with v_k.
// Workgroup 0 publishes v_k, d_k, e_k, tau_k and writes column k-1 of A in the dsytrd 'L'
// storage (column k is still being read by the other workgroups of this launch).
template <int TRD_WG, int TRD_PRE>
__global__ __launch_bounds__(TRD_WG) void k_trd_col(double *A, int N, i64 sA, int k, int q, double *vbuf,
                                                   double *pbuf, i64 sV, double *d, double *e, double *tau) {
    extern __shared__ double sh[];
    __shared__ double red[TRD_WG / 64];
    __shared__ double sc[4];
    const int b = blockIdx.z, m0 = N - k;                         // rows k .. N-1
    double *Ab = A + b * sA;
    double *db = d + (i64)b * N, *eb = e + (i64)b * N, *taub = tau + (i64)b * N;
    const double *vp = vbuf + b * 2 * sV + (i64)((k + 1) & 1) * sV;   // v_{k-1}
    double *vk_out = vbuf + b * 2 * sV + (i64)(k & 1) * sV;           // v_k
    const double *pp = pbuf + b * 2 * sV + (i64)((k + 1) & 1) * sV;   // p_{k-1}
    double *pk = pbuf + b * 2 * sV + (i64)(k & 1) * sV;               // p_k
    double *svp = sh, *swp = sh + m0, *sv = sh + 2 * m0;              // v_{k-1}, w_{k-1}, v_k at row k + t
    const bool upd = k >= 1;
    const bool pub = blockIdx.x == 0;
    const int m = m0 - 1;                                          // trailing columns / rows k+1 .. N-1
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // the first rows of this wavefront's first trailing column, loaded before the Householder step
    // so that both memory latencies overlap
    const int cl0 = ((int)blockIdx.x * (TRD_WG / 64) + wave) * q;
    double pre[TRD_PRE];
#pragma unroll
    for (int u = 0; u < TRD_PRE; u++) {
        const int r = lane + 64 * u;
        pre[u] = cl0 < m && r < m ? Ab[(i64)N * (k + 1 + cl0) + (k + 1) + r] : 0.0;
    }
    // ---- Householder step of column k, rows k..N-1 staged in LDS at index t = row - k:
    // svp = v_{k-1}, swp = p_{k-1} then w_{k-1}, sv = column k then v_k
    double s = 0.0;
    for (int t = threadIdx.x; t < m0; t += TRD_WG) {
        const double c = Ab[(k + t) + (i64)N * k];
        const double v = upd ? vp[k + t] : 0.0, pv = upd ? pp[k + t] : 0.0;
        svp[t] = v;
        swp[t] = pv;
        sv[t] = c;
        s += pv * v;
        if (t == 0) { sc[0] = v; sc[1] = pv; }
    }
    double alpha = 0.0;
    if (upd) alpha = -0.5 * taub[k - 1] * trd_block_sum(s, red);
    else __syncthreads();
    const double vkk = sc[0], wkk = sc[1] + alpha * sc[0];         // v_{k-1}, w_{k-1} at row k
    double xn = 0.0;
    for (int t = threadIdx.x; t < m0; t += TRD_WG) {                // same t -> thread map as above
        const double w = swp[t] + alpha * svp[t];
        swp[t] = w;
        double c = sv[t];
        if (upd) c -= svp[t] * wkk + w * vkk;
        sv[t] = c;
        if (t >= 2) xn += c * c;
    }
    xn = trd_block_sum(xn, red);                                   // its barriers publish sv[0], sv[1]
    const double a0 = sv[0], a1 = m0 > 1 ? sv[1] : 0.0;
    double beta, tk, scale;
    if (xn == 0.0) { beta = a1; tk = 0.0; scale = 0.0; }
    else {
        beta = -copysign(sqrt(a1 * a1 + xn), a1);
        tk = (beta - a1) / beta;
        scale = 1.0 / (a1 - beta);
    }
    __syncthreads();                                               // every thread has read sv[0], sv[1]
    for (int t = threadIdx.x; t < m0; t += TRD_WG) {
        const double v = t == 1 ? 1.0 : (t >= 2 ? sv[t] * scale : 0.0);
        sv[t] = v;
        if (pub) vk_out[k + t] = v;
    }
    if (pub) {
        for (int r = threadIdx.x; r < k; r += TRD_WG) vk_out[r] = 0.0;
        if (threadIdx.x == 0) {
            db[k] = a0;
            eb[k] = beta;
            taub[k] = tk;
            if (upd) {                                             // column k-1 in dsytrd 'L' storage
                Ab[(k - 1) + (i64)N * (k - 1)] = db[k - 1];
                Ab[k + (i64)N * (k - 1)] = eb[k - 1];
            }
        }
        if (upd)
            for (int t = 1 + threadIdx.x; t < m0; t += TRD_WG) Ab[(k + t) + (i64)N * (k - 1)] = vp[k + t];
    }
    __syncthreads();
    // ---- trailing columns c = k+1 .. N-1 (local index cl = c - k - 1), rows k+1 .. N-1,
    // 64*TRD_PRE rows per pass with all loads of a pass in flight together
    for (int jq = 0; jq < q; jq++) {
        const int cl = cl0 + jq;
        if (cl >= m) break;                                            // uniform per wavefront
        double *col = Ab + (i64)N * (k + 1 + cl) + (k + 1);
        const double vcl = svp[1 + cl], wcl = swp[1 + cl];
        double acc = 0.0;
        for (int r0 = 0; r0 < m; r0 += 64 * TRD_PRE) {
            double a[TRD_PRE];
            if (jq == 0 && r0 == 0) {
#pragma unroll
                for (int u = 0; u < TRD_PRE; u++) a[u] = pre[u];
            } else {
#pragma unroll
                for (int u = 0; u < TRD_PRE; u++) {
                    const int r = r0 + lane + 64 * u;
                    a[u] = r < m ? col[r] : 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < TRD_PRE; u++) {
                const int r = r0 + lane + 64 * u;
                if (r < m) {
                    if (upd) {
                        a[u] -= svp[1 + r] * wcl + swp[1 + r] * vcl;
                        col[r] = a[u];
                    }
                    acc += a[u] * sv[1 + r];
                }
            }
        }
        acc = trd_wave_sum(acc);
        if (lane == 0) pk[k + 1 + cl] = tk * acc;
    }
}

// after the last launch (k = N-2): T(N-1,N-1) and column N-2 of A
__global__ void k_trd_last(double *A, int N, i64 sA, double *d, const double *e) {
    const int b = blockIdx.x;
    double *Ab = A + b * sA;
    if (threadIdx.x) return;
    d[(i64)b * N + N - 1] = Ab[(N - 1) + (i64)N * (N - 1)];
    Ab[(N - 2) + (i64)N * (N - 2)] = d[(i64)b * N + N - 2];
    Ab[(N - 1) + (i64)N * (N - 2)] = e[(i64)b * N + N - 2];
}

__global__ void k_trd_zero(double *x, i64 cnt) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < cnt) x[t] = 0.0;
}

__global__ void k_trd_small(double *A, int N, i64 sA, double *d, double *e, double *tau);

static inline unsigned nblk(i64 tot, int bs) { return (unsigned)((tot + bs - 1) / bs); }

// work: vbuf and pbuf (2N per matrix each, double-buffered by launch parity), then d, e, tau
i64 trd_workspace(int N, int nb) { return (i64)7 * N * nb; }
TrdOut trd_outputs(int N, int nb, double *work) {
    TrdOut o;
    o.d = work + (i64)4 * N * nb;
    o.e = o.d + (i64)N * nb;
    o.tau = o.e + (i64)N * nb;
    return o;
}

typedef void (*TrdColFn)(double *, int, i64, int, int, double *, double *, i64, double *, double *, double *);
struct TrdCfg {
    TrdColFn fn;
    int wg, waves;
};
// launch shape of k_trd_col: workgroup size, rows per lane in flight, wavefronts per launch
// (NV2_TRD_CFG / NV2_TRD_WAVES select among the compiled variants for measurement)
static const TrdCfg &trd_cfg() {
    static const TrdCfg cfg = [] {
        // 16 wavefronts per workgroup and 16 rows per lane in flight: measured 12% faster than 4 and 8
        // at order 2744 (about 4096 wavefronts per launch; 8192 gained nothing more)
        const TrdColFn fns[2] = {k_trd_col<1024, 16>, k_trd_col<256, 8>};
        const int wgs[2] = {1024, 256};
        const char *e = getenv("NV2_TRD_CFG");
        int i = e ? atoi(e) : 0;
        if (i < 0 || i > 1) i = 0;
        const char *w = getenv("NV2_TRD_WAVES");
        int waves = w ? atoi(w) : 4096;
        if (waves < 64) waves = 64;
        for (auto f : fns) (void)hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
        return TrdCfg{fns[i], wgs[i], waves};
    }();
    return cfg;
}

static int trd_launches(hipStream_t st, int N, double *A, i64 sA, int nb, double *work) {
    const i64 sV = N;
    double *vbuf = work, *pbuf = vbuf + (i64)nb * 2 * sV;
    const TrdOut o = trd_outputs(N, nb, work);
    hipLaunchKernelGGL(k_trd_zero, dim3(nblk((i64)4 * N * nb, 256)), dim3(256), 0, st, work, (i64)4 * N * nb);
    hipLaunchKernelGGL(k_trd_sym_upper, dim3(nblk((i64)N * N, 256), 1, nb), dim3(256), 0, st, A, N, sA);
    const TrdCfg &c = trd_cfg();
    for (int k = 0; k <= N - 2; k++) {
        const int m = N - k - 1;                              // trailing columns
        int q = (int)(((i64)m * nb + c.waves - 1) / c.waves);  // about c.waves wavefronts over the batch
        if (q < 1) q = 1;
        const unsigned gx = nblk(m, (c.wg / 64) * q);
        hipLaunchKernelGGL(c.fn, dim3(gx, 1, nb), dim3(c.wg), sizeof(double) * 3 * (N - k), st, A, N, sA, k, q, vbuf,
                           pbuf, sV, o.d, o.e, o.tau);
    }
    hipLaunchKernelGGL(k_trd_last, dim3(nb), dim3(64), 0, st, A, N, sA, o.d, o.e);
    return 0;
}

int trd_reduce(hipStream_t st, int N, double *A, i64 sA, int nb, double *work) {
    if (N < 1 || nb < 1 || nb > 65535) { set_error("trd_reduce: unsupported N=%d nb=%d", N, nb); return -2; }
    if (N <= 2) {
        const TrdOut o = trd_outputs(N, nb, work);
        hipLaunchKernelGGL(k_trd_small, dim3(nb), dim3(64), 0, st, A, N, sA, o.d, o.e, o.tau);
        NV2_LAUNCH_CHECK();
        return 0;
    }
    if (N > TRD_NMAX) { set_error("trd_reduce: order %d above %d", N, TRD_NMAX); return -2; }
    TRYT(trd_launches(st, N, A, sA, nb, work));
    NV2_LAUNCH_CHECK();
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

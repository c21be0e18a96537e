// Blocked Cholesky factorisation S = U^T U (upper triangle, column-major) of a batch of
// symmetric matrices, for the factored PC metric of the stage-2 solves.  The reference has no
// Cholesky step (it diagonalises every metric, qdnevpt.F:5723-5797); the factorisation stands in
// for that eigen-decomposition when no eigenvalue falls below the threshold.
//
// Right-looking by block rows of CHOL_NB, four launches per block row, all batched over the
// matrices:
//   k_chol_diag   factors the diagonal block in LDS (one workgroup per matrix) and writes
//                 U11 and its inverse;
//   GEMM          P = U11^-T A12 into a panel buffer (the triangular solve as a product with the
//                 explicit inverse of a 64 x 64 block: rocBLAS trsm costs several helper kernels
//                 per call);
//   GEMM          A22 -= P^T P over the whole trailing block (its strict lower triangle is never
//                 read again; rocBLAS syrk measured 2-3x slower here than the MFMA GEMM kernels);
//   k_chol_put    copies P into the block row of U.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "nv2_chol.h"
#include "nv2_chol64.h"

namespace nv2 {

#define CHOL_NB 64
#define CHOL_WG 256

// Factor the kb x kb diagonal block at (k0, k0) of matrix b in LDS and invert it.  info[b] =
// k0 + j + 1 for the first pivot that is not positive (NaN included); a matrix whose info is
// already set is left alone (its panel is then scratch).  Uinv + b * CHOL_NB^2 receives U11^-1
// (upper triangular, ld CHOL_NB).
__global__ __launch_bounds__(CHOL_WG) void k_chol_diag(double *A, int N, i64 sA, int k0, int kb, int *info,
                                                       double *Uinv) {
    __shared__ double a[CHOL_NB][CHOL_NB + 1];      // a[r][c]: U11 in the upper triangle
    __shared__ double x[CHOL_NB][CHOL_NB + 1];      // x[r][c]: U11^-1
    __shared__ int bad;
    const int b = blockIdx.x;
    if (info[b] != 0) return;                       // uniform over the workgroup
    double *Ab = A + b * sA + k0 + (i64)N * k0;
    for (int t = threadIdx.x; t < kb * kb; t += CHOL_WG) {
        const int r = t % kb, c = t / kb;
        a[r][c] = Ab[r + (i64)N * c];
    }
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    // One barrier per step: step j updates the trailing block with the unscaled row j
    // (a[r][c] -= a[j][r] a[j][c] / d) and scales row j - 1 into U, which no thread reads or writes
    // in step j; the last row is scaled after the loop.
    double dprev = 0.0;
    bool brk = false;
    for (int j = 0; j < kb; j++) {
        const double d = a[j][j];
        if (!(d > 0.0)) {                           // every thread sees the same value
            if (threadIdx.x == 0) bad = k0 + j + 1;
            brk = true;
            break;
        }
        if (j > 0) {                                // row j - 1 of U
            const double u = sqrt(dprev), inv = 1.0 / u;
            for (int c = j - 1 + threadIdx.x; c < kb; c += CHOL_WG) a[j - 1][c] = c == j - 1 ? u : a[j - 1][c] * inv;
        }
        const double rd = 1.0 / d;
        const int m = kb - j - 1;                   // trailing (r, c), j < r <= c
        for (int t = threadIdx.x; t < m * m; t += CHOL_WG) {
            const int r = j + 1 + t % m, c = j + 1 + t / m;
            if (r <= c) a[r][c] -= a[j][r] * a[j][c] * rd;
        }
        dprev = d;
        __syncthreads();
    }
    if (!brk && threadIdx.x == 0) a[kb - 1][kb - 1] = sqrt(dprev);
    __syncthreads();
    if (bad) {
        if (threadIdx.x == 0) info[b] = bad;
        return;
    }
    // U11^-1
    for (int t = threadIdx.x; t < kb * kb; t += CHOL_WG) x[t % kb][t / kb] = 0.0;
    __syncthreads();
    upper_inverse_lds<CHOL_NB>(a, x, kb);
    __syncthreads();
    double *Ub = Uinv + (i64)b * CHOL_NB * CHOL_NB;
    for (int t = threadIdx.x; t < kb * kb; t += CHOL_WG) {
        const int r = t % kb, c = t / kb;
        if (r <= c) Ab[r + (i64)N * c] = a[r][c];
        Ub[r + CHOL_NB * c] = x[r][c];
    }
}

// The same by the four wavefronts of a 256-thread workgroup with the block in registers (chol64_4w;
// kb < CHOL_NB: padded with the identity, so the factor of the pad is the identity).  The 256-thread
// LDS form above spent most of its time in 64 barrier-separated steps over LDS; this one is the form
// the (+-1)' chain and the smaller classes' factorisations use.
__global__ __launch_bounds__(256) void k_chol_diag1(double *A, int N, i64 sA, int k0, int kb, int *info,
                                                    double *Uinv) {
    __shared__ Chol64Lds L;
    const int b = blockIdx.x, w = threadIdx.x >> 6, c = threadIdx.x & 63;
    if (info[b] != 0) return;                       // uniform over the workgroup
    double *Ab = A + b * sA + k0 + (i64)N * k0;
    double a[16], x[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int r = 16 * w + i;
        double v = r == c ? 1.0 : 0.0;
        if (r < kb && c < kb) v = r <= c ? Ab[r + (i64)N * c] : Ab[c + (i64)N * r];
        a[i] = v;
    }
    const int bad = chol64_4w(a, x, L);
    if (bad) {
        if (threadIdx.x == 0) info[b] = k0 + bad;
        return;
    }
    double *Ub = Uinv + (i64)b * CHOL_NB * CHOL_NB;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int r = 16 * w + i;
        if (r <= c && c < kb && r < kb) Ab[r + (i64)N * c] = a[i];   // U in the upper triangle of the block
        if (c < kb && r < kb) Ub[c + CHOL_NB * r] = x[i];       // x[i] = U^-1[c][r]
    }
}

// block row of U: A[k0 + r, k0 + kb + c] = P[r + CHOL_NB * c], r < kb, c < m
__global__ void k_chol_put(double *A, int N, i64 sA, int k0, int kb, int m, const double *P, i64 sP) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (i64)kb * m) return;
    const int r = (int)(t % kb), c = (int)(t / kb), b = blockIdx.z;
    A[b * sA + (k0 + r) + (i64)N * (k0 + kb + c)] = P[b * sP + r + (i64)CHOL_NB * c];
}

// U_jj^-1 for every diagonal block jj = blockIdx.x of matrix b = blockIdx.z, into the same block of
// Ui (upper triangle and diagonal; Ui was cleared before)
__global__ __launch_bounds__(CHOL_WG) void k_trinv_diag(const double *U, int N, i64 sU, double *Ui, i64 sUi) {
    __shared__ double a[CHOL_NB][CHOL_NB + 1];
    __shared__ double x[CHOL_NB][CHOL_NB + 1];
    const int b = blockIdx.z, k0 = blockIdx.x * CHOL_NB, kb = N - k0 < CHOL_NB ? N - k0 : CHOL_NB;
    const double *Ub = U + b * sU + k0 + (i64)N * k0;
    double *Xb = Ui + b * sUi + k0 + (i64)N * k0;
    for (int t = threadIdx.x; t < kb * kb; t += CHOL_WG) {
        const int r = t % kb, c = t / kb;
        a[r][c] = r <= c ? Ub[r + (i64)N * c] : 0.0;
        x[r][c] = 0.0;
    }
    __syncthreads();
    upper_inverse_lds<CHOL_NB>(a, x, kb);
    __syncthreads();
    for (int t = threadIdx.x; t < kb * kb; t += CHOL_WG) {
        const int r = t % kb, c = t / kb;
        if (r <= c) Xb[r + (i64)N * c] = x[r][c];
    }
}

__global__ void k_chol_clear(double *x, i64 cnt) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < cnt) x[t] = 0.0;
}

__global__ void k_chol_zero_info(int *info, int nb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nb) info[t] = 0;
}

#define RBH(call) do { if ((call) != rocblas_status_success) { set_error("rocBLAS failed: %s", #call); return -1; } } while (0)

i64 chol_workspace(int N, int nb) { return (i64)nb * CHOL_NB * ((i64)CHOL_NB + N); }

// Ui = U^-1 by block columns: the diagonal blocks are inverted in LDS, then for each block column j
// (columns k0 .. k0+kb) Ui[0:k0, j] = -(Ui[0:k0, 0:k0] U[0:k0, j]) Ui[j, j], two GEMMs on the
// already inverted leading block.
int chol_inverse_upper(rocblas_handle h, hipStream_t st, int N, const double *U, i64 sU, int nb, double *Ui, i64 sUi,
                       double *work) {
    if (N < 1 || nb < 1 || nb > 65535) { set_error("chol_inverse_upper: unsupported N=%d nb=%d", N, nb); return -2; }
    RBH(rocblas_set_stream(h, st));
    for (int b = 0; b < nb; b++) {
        hipLaunchKernelGGL(k_chol_clear, dim3((unsigned)(((i64)N * N + 255) / 256)), dim3(256), 0, st, Ui + b * sUi,
                           (i64)N * N);
        NV2_LAUNCH_CHECK();
    }
    const int nblk = (N + CHOL_NB - 1) / CHOL_NB;
    hipLaunchKernelGGL(k_trinv_diag, dim3(nblk, 1, nb), dim3(CHOL_WG), 0, st, U, N, sU, Ui, sUi);
    NV2_LAUNCH_CHECK();
    double *P = work;                            // k0 x kb panels, ld N, one per matrix
    const i64 sP = (i64)CHOL_NB * N;
    const double one = 1.0, mone = -1.0, zero = 0.0;
    for (int k0 = CHOL_NB; k0 < N; k0 += CHOL_NB) {
        const int kb = N - k0 < CHOL_NB ? N - k0 : CHOL_NB;
        RBH(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, k0, kb, k0, &one, Ui, N,
                                          sUi, U + (i64)N * k0, N, sU, &zero, P, N, sP, nb));
        RBH(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, k0, kb, kb, &mone, P, N,
                                          sP, Ui + k0 + (i64)N * k0, N, sUi, &zero, Ui + (i64)N * k0, N, sUi, nb));
    }
    return 0;
}

int chol_upper(rocblas_handle h, hipStream_t st, int N, double *A, i64 sA, int nb, int *info, double *work) {
    if (N < 1 || nb < 1 || nb > 65535) { set_error("chol_upper: unsupported N=%d nb=%d", N, nb); return -2; }
    RBH(rocblas_set_stream(h, st));
    double *Uinv = work, *P = work + (i64)nb * CHOL_NB * CHOL_NB;
    const i64 sU = (i64)CHOL_NB * CHOL_NB, sP = (i64)CHOL_NB * N;
    hipLaunchKernelGGL(k_chol_zero_info, dim3((nb + 255) / 256), dim3(256), 0, st, info, nb);
    NV2_LAUNCH_CHECK();
    const double one = 1.0, mone = -1.0, zero = 0.0;
    for (int k0 = 0; k0 < N; k0 += CHOL_NB) {
        const int kb = N - k0 < CHOL_NB ? N - k0 : CHOL_NB;
        hipLaunchKernelGGL(k_chol_diag1, dim3(nb), dim3(256), 0, st, A, N, sA, k0, kb, info, Uinv);
        NV2_LAUNCH_CHECK();
        const int m = N - k0 - kb;
        if (m <= 0) break;
        double *A12 = A + k0 + (i64)N * (k0 + kb), *A22 = A + (k0 + kb) + (i64)N * (k0 + kb);
        RBH(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, kb, m, kb, &one, Uinv,
                                          CHOL_NB, sU, A12, N, sA, &zero, P, CHOL_NB, sP, nb));
        RBH(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, m, m, kb, &mone, P,
                                          CHOL_NB, sP, P, CHOL_NB, sP, &one, A22, N, sA, nb));
        hipLaunchKernelGGL(k_chol_put, dim3((unsigned)(((i64)kb * m + 255) / 256), 1, nb), dim3(256), 0, st, A, N, sA,
                           k0, kb, m, P, sP);
        NV2_LAUNCH_CHECK();
    }
    return 0;
}

}  // namespace nv2

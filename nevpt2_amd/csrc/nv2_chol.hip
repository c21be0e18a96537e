// Blocked Cholesky factorisation S = U^T U (upper triangle, column-major) of a batch of
// symmetric matrices, for the factored PC metric of the stage-2 solves.  The reference has no
// Cholesky step (it diagonalises every metric, qdnevpt.F:5723-5797); the factorisation stands in
// for that eigen-decomposition when no eigenvalue falls below the threshold.
//
// Right-looking by block rows of CHOL_NB: the diagonal block is factored in LDS by one workgroup
// per matrix (k_chol_diag), the block row to its right becomes U11^-T A12 (rocBLAS trsm), and the
// trailing matrix takes the update A22 -= U12^T U12 as a full GEMM (its strict lower triangle is
// never read again; rocBLAS syrk measured 2-3x slower here than the MFMA GEMM kernels).  About
// 3 N / CHOL_NB launches instead of the several launches per column of an unblocked panel.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "nv2_chol.h"

namespace nv2 {

#define CHOL_NB 64
#define CHOL_WG 256

// Factor the kb x kb diagonal block at (k0, k0) of matrix b in LDS.  info[b] = k0 + j + 1 for
// the first pivot that is not positive (NaN included); a matrix whose info is already set is
// left alone.
__global__ __launch_bounds__(CHOL_WG) void k_chol_diag(double *A, int N, i64 sA, int k0, int kb, int *info) {
    __shared__ double a[CHOL_NB][CHOL_NB + 1];      // a[r][c], upper triangle used
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
    for (int j = 0; j < kb; j++) {
        const double d = a[j][j];
        if (!(d > 0.0)) {                           // every thread sees the same value
            if (threadIdx.x == 0) bad = k0 + j + 1;
            break;
        }
        const double u = sqrt(d), inv = 1.0 / u;
        __syncthreads();                            // everyone has read a[j][j]
        if (threadIdx.x == 0) a[j][j] = u;
        for (int c = j + 1 + threadIdx.x; c < kb; c += CHOL_WG) a[j][c] *= inv;   // row j of U
        __syncthreads();
        const int m = kb - j - 1;                   // trailing (r, c), j < r <= c
        for (int t = threadIdx.x; t < m * m; t += CHOL_WG) {
            const int r = j + 1 + t % m, c = j + 1 + t / m;
            if (r <= c) a[r][c] -= a[j][r] * a[j][c];
        }
        __syncthreads();
    }
    __syncthreads();
    if (bad) {
        if (threadIdx.x == 0) info[b] = bad;
        return;
    }
    for (int t = threadIdx.x; t < kb * kb; t += CHOL_WG) {
        const int r = t % kb, c = t / kb;
        if (r <= c) Ab[r + (i64)N * c] = a[r][c];
    }
}

__global__ void k_chol_zero_info(int *info, int nb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nb) info[t] = 0;
}

#define RBH(call) do { if ((call) != rocblas_status_success) { set_error("rocBLAS failed: %s", #call); return -1; } } while (0)

int chol_upper(rocblas_handle h, hipStream_t st, int N, double *A, i64 sA, int nb, int *info) {
    if (N < 1 || nb < 1 || nb > 65535) { set_error("chol_upper: unsupported N=%d nb=%d", N, nb); return -2; }
    RBH(rocblas_set_stream(h, st));
    hipLaunchKernelGGL(k_chol_zero_info, dim3((nb + 255) / 256), dim3(256), 0, st, info, nb);
    NV2_LAUNCH_CHECK();
    const double one = 1.0, mone = -1.0;
    for (int k0 = 0; k0 < N; k0 += CHOL_NB) {
        const int kb = N - k0 < CHOL_NB ? N - k0 : CHOL_NB;
        hipLaunchKernelGGL(k_chol_diag, dim3(nb), dim3(CHOL_WG), 0, st, A, N, sA, k0, kb, info);
        NV2_LAUNCH_CHECK();
        const int m = N - k0 - kb;
        if (m <= 0) break;
        double *U11 = A + k0 + (i64)N * k0, *A12 = A + k0 + (i64)N * (k0 + kb), *A22 = A + (k0 + kb) + (i64)N * (k0 + kb);
        RBH(rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                                          rocblas_diagonal_non_unit, kb, m, &one, U11, N, sA, A12, N, sA, nb));
        RBH(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, m, m, kb, &mone, A12,
                                          N, sA, A12, N, sA, &one, A22, N, sA, nb));
    }
    return 0;
}

}  // namespace nv2

// Batched Householder tridiagonalisation (nv2_trd.hip)
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "nv2_common.h"

namespace nv2 {

struct TrdOut {
    double *d, *e, *tau;   // N per matrix each, matrix b at + b*N
};
// scratch of trd_reduce, in doubles; d, e, tau live inside it (trd_outputs)
i64 trd_workspace(int N, int nb);
TrdOut trd_outputs(int N, int nb, double *work);
// A_b = Q_b T_b Q_b^T for the nb matrices A + b*sA (ld N), from their upper triangles;
// reflectors left below the subdiagonal as dsytrd('L') does
int trd_reduce(hipStream_t st, int N, double *A, i64 sA, int nb, double *work);
// X_b <- Q_b X_b or Q_b^T X_b for X_b = X + b*sX (N x ncol, ld N); h must run on st
int trd_apply_q(rocblas_handle h, hipStream_t st, int N, const double *A, i64 sA, int nb, const double *tau,
                bool trans, int ncol, double *X, i64 sX, double *scratch);
i64 trd_apply_workspace(int N, int nb, int ncol);

// (T + s I) u = z for the symmetric tridiagonal T (diagonal d[0..n), off-diagonal e[0..n-1)), by
// Gaussian elimination with partial pivoting (the row interchanges of LAPACK dgtsv).  z is read at
// z[i*zs]; the eliminated rows go to rows[(4*i + c)*rs] (c: diagonal, first and second
// superdiagonal, right-hand side).  f(i, u_i) is called for i = n-1 .. 0 by the back substitution.
template <typename F>
__device__ inline void trd_shift_solve(int n, const double *d, const double *e, double s, const double *z, i64 zs,
                                       double *rows, i64 rs, F f) {
    double di = d[0] + s, dui = n > 1 ? e[0] : 0.0, bi = z[0];
    for (int i = 0; i < n - 1; i++) {
        const double dl = e[i], dn = d[i + 1] + s, dun = i < n - 2 ? e[i + 1] : 0.0, bn = z[(i64)(i + 1) * zs];
        double *ri = rows + (i64)4 * i * rs;
        if (fabs(di) >= fabs(dl)) {          // no interchange: row i+1 -= (dl/di) row i
            const double fact = di != 0.0 ? dl / di : 0.0;
            ri[0] = di; ri[rs] = dui; ri[2 * rs] = 0.0; ri[3 * rs] = bi;
            di = dn - fact * dui; dui = dun; bi = bn - fact * bi;
        } else {                             // rows i and i+1 interchanged
            const double fact = di / dl;
            ri[0] = dl; ri[rs] = dn; ri[2 * rs] = dun; ri[3 * rs] = bn;
            di = dui - fact * dn; dui = -fact * dun; bi = bi - fact * bn;
        }
    }
    double u1 = bi / di, u2 = 0.0;
    f(n - 1, u1);
    for (int i = n - 2; i >= 0; i--) {
        const double *ri = rows + (i64)4 * i * rs;
        const double u = (ri[3 * rs] - ri[rs] * u1 - ri[2 * rs] * u2) / ri[0];
        f(i, u);
        u2 = u1;
        u1 = u;
    }
}

}  // namespace nv2

// Batched blocked Cholesky factorisation (nv2_chol.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "nv2_common.h"

namespace nv2 {

// S_b = U_b^T U_b for the nb matrices A + b*sA (order N, ld N, upper triangle read and
// overwritten by U; the strict lower triangle is overwritten with scratch values).  info[b] = 0 on
// success, else the 1-based column of the first pivot that is not positive.  work holds
// chol_workspace(N, nb) doubles.  Asynchronous on st.
i64 chol_workspace(int N, int nb);
int chol_upper(rocblas_handle h, hipStream_t st, int N, double *A, i64 sA, int nb, int *info, double *work);
// Ui_b = U_b^-1 for the nb upper triangular factors U + b*sU (order N, ld N; their strict lower
// triangle is not read) into Ui + b*sUi (ld N, strict lower triangle zero).  work holds
// chol_workspace(N, nb) doubles.  Asynchronous on st.
int chol_inverse_upper(rocblas_handle h, hipStream_t st, int N, const double *U, i64 sU, int nb, double *Ui, i64 sUi,
                       double *work);

}  // namespace nv2

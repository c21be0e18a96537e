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

}  // namespace nv2

// Orthogonal reduction of dense symmetric matrices to band form (nv2_band.hip), for the (+-1)' PC
// sums: y^T (K' + s)^-1 y = z^T (B + s)^-1 z with K' = Q B Q^T, z = Q^T y, B banded.
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "nv2_common.h"

namespace nv2 {

#define BAND_B 64             // half-bandwidth of the reduced matrices
#define BAND_W (2 * BAND_B + 1)

// scratch of band_reduce for nb problems of order N, in doubles
i64 band_workspace(int N, int nb);
// Reduces the nb symmetric matrices A + b*sA (order N, ld N, upper triangle read; overwritten) to
// B_b = Q_b^T A_b Q_b of half-bandwidth BAND_B and keeps Q_b (as panels of compact WY reflectors)
// and the band of B_b in work.  dflag[b] becomes nonzero when a panel's CholeskyQR2 loses
// positive definiteness (the caller then reduces that matrix some other way).  Asynchronous.
int band_reduce(rocblas_handle h, hipStream_t st, int N, double *A, i64 sA, int nb, double *work, int *dflag);
// Z_b <- Q_b^T Z_b in place for problems b0 .. b0+nb-1 of band_reduce's workspace (Z_b = Z + (b - b0) sZ,
// N x ncols, leading dimension ldZ; read only workspace); scratch of band_qt_workspace doubles.
// Asynchronous.
i64 band_qt_workspace(int N, int nb, int ncols);
int band_apply_qt(hipStream_t st, int N, const double *work, int b0, int nb, double *Z, i64 ldZ, i64 sZ, int ncols,
                  double *scratch);
// the rows of B_b: entry (r, c), |c - r| <= BAND_B, at [r * BAND_W + (c - r + BAND_B)] (zero outside 0..N-1)
const double *band_rows(int N, int nb, int b, const double *work);

}  // namespace nv2

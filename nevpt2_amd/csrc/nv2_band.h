// Orthogonal reduction of dense symmetric matrices to band form (nv2_band.hip), for the (+-1)' PC
// sums: y^T (K' + s)^-1 y = z^T (B + s)^-1 z with K' = Q B Q^T, z = Q^T y, B banded.
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "nv2_common.h"

namespace nv2 {

#define BAND_B 64             // half-bandwidth of the reduced matrices
#define BAND_W (2 * BAND_B + 1)
#define BAND_QT_SETS 4        // column sets of one band_apply_qt launch

// kernel arguments of band_apply_qt: one set of right-hand sides, and the launch
struct BandQtSet {
    const double *w;
    double *Y;
    i64 ld;
    int ncols, wg0;
};
struct BandQtArgs {
    int N, npan, nset;
    i64 offY, offT, offQl;
    BandQtSet s[BAND_QT_SETS];
};

// scratch of band_reduce for nb problems of order N, in doubles
i64 band_workspace(int N, int nb);
// Reduces the nb symmetric matrices A + b*sA (order N, ld N, upper triangle read; overwritten) to
// B_b = Q_b^T A_b Q_b of half-bandwidth BAND_B and keeps Q_b (as panels of compact WY reflectors)
// and the band of B_b in work.  dflag[b] becomes nonzero when a panel's CholeskyQR2 loses
// positive definiteness (the caller then reduces that matrix some other way).  Asynchronous.
int band_reduce(rocblas_handle h, hipStream_t st, int N, double *A, i64 sA, int nb, double *work, int *dflag);
// z = Q_b^T y in place, for up to BAND_QT_SETS sets of right-hand sides in one launch: set q holds
// ncols[q] columns of order N at Y[q] (leading dimension ld[q]) and uses problem b[q]'s Q
// (band_reduce's workspace, read only).  Asynchronous.
int band_apply_qt(hipStream_t st, int N, const double *work, int nset, const int *b, double *const *Y,
                  const i64 *ld, const int *ncols);
// the rows of B_b: entry (r, c), |c - r| <= BAND_B, at [r * BAND_W + (c - r + BAND_B)] (zero outside 0..N-1)
const double *band_rows(int N, int nb, int b, const double *work);

}  // namespace nv2

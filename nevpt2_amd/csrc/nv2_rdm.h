// Spin-free RDMs from CI vectors on the device (nv2_rdm.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "nv2_common.h"

namespace nv2 {

// Number of occupation strings of k electrons in n orbitals.
i64 ci_nstrings(int n, int k);

// Packed 4-RDM of each of `nroots` states (ro4[nw, n, n, n, n, nroots]) and, for nroots >= 2 and
// trdm3 != nullptr, the transition 3-/2-/1-RDMs of the state pairs (i > j) at index i(i-1)/2 + j,
// state-pair index fastest; dC holds the CI vectors C[root][Ia][Ib] on the device.  Synchronous.
int ci_fill_rdms(int n, int na, int nb, int nroots, const double *dC, rocblas_handle h, hipStream_t st, double *ro4,
                 double *trdm3, double *trdm2, double *trdm1);

}  // namespace nv2

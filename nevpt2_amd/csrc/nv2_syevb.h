// Batched symmetric eigensolver for many small matrices of different orders (nv2_syevb.hip)
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "nv2_common.h"

namespace nv2 {

#define SYEVB_NMAX 1024     // largest order handled (one workgroup per matrix)

// One problem: A (order N, ld N, device) holds the matrix in its upper triangle (lower = 1: its
// lower triangle) on entry and the eigenvectors (columns) on exit; w receives the eigenvalues in ascending order; work is this
// problem's own device scratch of syevb_workspace(N) doubles.  The result is what
// dsyev('V','U') returns, up to the sign of each eigenvector and the choice of basis inside a
// degenerate eigenspace.
struct SyevbProb {
    double *A, *w, *work;
    int N;
    int lower = 0;
};
i64 syevb_workspace(int N);
// Launch shapes and captured launches of one call site: with a plan, the first call with a list of
// problems times the candidate shapes (once per list of orders in the process) and solves; the
// second call with the same buffers captures its launches in a hipGraph, and later ones replay it.
struct SyevbPlan;
SyevbPlan *syevb_plan_create();
void syevb_plan_destroy(SyevbPlan *plan);
i64 syevb_plan_graph_launches(const SyevbPlan *plan);   // hipGraph launches so far (0 for no plan)
// Every problem of P (any number, any mix of orders 1..SYEVB_NMAX) on stream st: five kernel
// launches per 16 problems.  h is not used (no library call is made).  Without a plan the launches
// go to st directly (shapes from the process's tuning table as well).
int syevb(rocblas_handle h, hipStream_t st, int np, const SyevbProb *P, SyevbPlan *plan = nullptr);

}  // namespace nv2

// Cholesky factor and its inverse of one 64 x 64 symmetric positive definite matrix by the four
// wavefronts of a 256-thread workgroup.  Thread (w = wavefront, c = lane) holds rows 16 w .. 16 w + 15
// of column c in registers.  Step j of the factorisation (right-looking, G = R^T R with R upper):
// the wavefront that holds row j publishes it through LDS, every thread forms R[j][c] = G[j][c] /
// sqrt(G[j][j]) and updates the rows it holds below j.  The inverse X = R^-1 is formed from the
// bottom row up: row k of X is B[k][:] / R[k][k], after which every row r < k drops R[r][k] X[k][:]
// (B starts as the identity).  One workgroup barrier per step, the published row double-buffered;
// the steps of one wavefront's rows are unrolled, the four row blocks are a loop, so the code stays
// small (a fully unrolled one-wavefront form is instruction-fetch bound).
#pragma once
#include "nv2_common.h"

namespace nv2 {

struct Chol64Lds {
    double rb[2][64];       // the published row of the current step
    double rl[64][65];      // R, for the inverse's multipliers R[r][k]
};

// a: in G[16 w + i][c], out R[16 w + i][c] (zero below the diagonal); x: out X[16 w + i][c].
// Returns the 1-based first step whose pivot is not positive (0 when the matrix factors); then a and x
// hold finite values without meaning.  Uniform over the workgroup.  Every thread of the 256 calls it.
__device__ inline int chol64_4w(double (&a)[16], double (&x)[16], Chol64Lds &L) {
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), c = threadIdx.x & 63;   // w uniform
    int bad = 0;
#pragma unroll 1
    for (int jw = 0; jw < 4; jw++) {
#pragma unroll
        for (int ij = 0; ij < 16; ij++) {
            const int j = 16 * jw + ij;
            double *rb = L.rb[ij & 1];
            if (w == jw) rb[c] = a[ij];
            __syncthreads();
            const double d = rb[j];
            if (!(d > 0.0) && bad == 0) bad = j + 1;
            const double u = sqrt(d > 0.0 ? d : 1.0), inv = 1.0 / u;
            const double rjc = c < j ? 0.0 : (c == j ? u : rb[c] * inv);   // R[j][c]
            if (w > jw) {
#pragma unroll
                for (int i = 0; i < 16; i++) a[i] -= (rb[16 * w + i] * inv) * rjc;
            } else if (w == jw) {
#pragma unroll
                for (int i = ij + 1; i < 16; i++) a[i] -= (rb[16 * w + i] * inv) * rjc;
                a[ij] = rjc;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int r = 16 * w + i;
        if (r > c) a[i] = 0.0;
        L.rl[r][c] = a[i];
        x[i] = r == c ? 1.0 : 0.0;
    }
    __syncthreads();
#pragma unroll 1
    for (int kw = 3; kw >= 0; kw--) {
#pragma unroll
        for (int ik = 15; ik >= 0; ik--) {
            const int k = 16 * kw + ik;
            double *rb = L.rb[ik & 1];
            const double rkk = L.rl[k][k];
            if (w == kw) {
                const double xk = c < k ? 0.0 : x[ik] / rkk;
                x[ik] = xk;
                rb[c] = xk;
            }
            __syncthreads();
            const double xkc = rb[c];
            if (w < kw) {
#pragma unroll
                for (int i = 0; i < 16; i++) x[i] -= L.rl[16 * w + i][k] * xkc;
            } else if (w == kw) {
#pragma unroll
                for (int i = 0; i < ik; i++) x[i] -= L.rl[16 * w + i][k] * xkc;
            }
        }
    }
    return bad;
}

}  // namespace nv2

// Cholesky factor and its inverse of one 64 x 64 symmetric positive definite matrix by the four
// wavefronts of a 256-thread workgroup.  Thread (w = wavefront, c = lane) holds rows 16 w .. 16 w + 15
// of column c in registers.  Step j (right-looking, G = R^T R with R upper): the wavefront that holds
// row j publishes it through LDS, every thread forms R[j][c] = G[j][c] / sqrt(G[j][j]) and updates the
// rows it holds below j.  The same row operations applied to the identity give, in the same 64 steps,
// X = R^-T (lower triangular; row j of X is scaled by 1/sqrt(G[j][j]) when it is final), so R^-1 = X^T
// needs no second sweep.  One workgroup barrier per step, the published rows double-buffered; the steps
// of one wavefront's rows are unrolled, the four row blocks are a loop, so the code stays small (a fully
// unrolled one-wavefront form is instruction-fetch bound).
#pragma once
#include "nv2_common.h"

namespace nv2 {

struct Chol64Lds {
    double rb[2][64];       // row j of the partly reduced matrix
    double rx[2][64];       // row j of X
};

// a: in G[16 w + i][c], out R[16 w + i][c] (zero below the diagonal); x: out X[16 w + i][c] = R^-1[c][16 w + i]
// (zero above the diagonal).  Returns the 1-based first step whose pivot is not positive (0 when the matrix
// factors); then a and x hold finite values without meaning.  Uniform over the workgroup.  Every thread of
// the 256 calls it.
__device__ inline int chol64_4w(double (&a)[16], double (&x)[16], Chol64Lds &L) {
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), c = threadIdx.x & 63;   // w uniform
    int bad = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = 16 * w + i == c ? 1.0 : 0.0;
#pragma unroll 1
    for (int jw = 0; jw < 4; jw++) {
#pragma unroll
        for (int ij = 0; ij < 16; ij++) {
            const int j = 16 * jw + ij;
            double *rb = L.rb[ij & 1], *rx = L.rx[ij & 1];
            if (w == jw) { rb[c] = a[ij]; rx[c] = x[ij]; }
            __syncthreads();
            const double d = rb[j];
            if (!(d > 0.0) && bad == 0) bad = j + 1;
            const double u = sqrt(d > 0.0 ? d : 1.0), inv = 1.0 / u;
            const double rjc = c < j ? 0.0 : (c == j ? u : rb[c] * inv);   // R[j][c]
            const double xjc = rx[c] * inv;                                  // X[j][c]
            if (w > jw) {
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    const double l = rb[16 * w + i] * inv;
                    a[i] -= l * rjc;
                    x[i] -= l * xjc;
                }
            } else if (w == jw) {
#pragma unroll
                for (int i = ij + 1; i < 16; i++) {
                    const double l = rb[16 * w + i] * inv;
                    a[i] -= l * rjc;
                    x[i] -= l * xjc;
                }
                a[ij] = rjc;
                x[ij] = xjc;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 16; i++)
        if (16 * w + i > c) a[i] = 0.0;
    return bad;
}

}  // namespace nv2

// Lower-order (3-RDM and below) part of bamat (koopro4QD.F90:4009-4094) as GEMMs.
//
// With E3(a..f) = eee2(a..f), E3t = eee2t, E2 = ee2 materialised per state and
// K1(d,a) = sum_i w(d,i,i,a), K2(c,i) = sum_d w(c,d,d,i) (one array: both are sum_m w(x,m,m,y)), the reference's loop
// body regroups (P=ap, Q=bp, R=cp) into
//   amat = G1 + sum_d E3(R,P,Q,b,d,c)(F(d,a) - K1(d,a)/2) - sum_i E3(R,P,Q,b,a,i)(F(c,i) + K2(c,i)/2)
//        + sum_i E3(R,P,Q,i,a,c)(K2(b,i)/2 - F(b,i)) + C1 - C2 - C3
//   atmat = sum_d E3t(R,P,Q,b,d,c)F(d,a) - sum_i E3t(R,P,Q,b,a,i)F(c,i) + sum_i E3t(R,P,Q,i,a,c)(F(b,i) + K2(b,i)/2)
//        + 2 G5 - X2[Q<->b] + 2 d(b,Q)(Y2 - sum_d E2(R,P,d,c)K1(d,a)/2 - sum_i E2(R,P,a,i)K2(c,i)/2 + Z3)
//        - C1[Q<->b] + C2[Q<->b] - C4
// with the n^8 pieces as GEMMs:
//   G1[(R,P,Q,b),(c,a)] = sum_{d,i} E3(R,P,Q,b,d,i) w(c,i,d,a)
//   G5[(R,P),(b,Q),(a,c)] = sum_{d,j} E3(R,P,d,j,a,c) w(d,j,b,Q)
//   Y2[(R,P),(c,a)] = sum_{d,i} E2(R,P,d,i) w(c,i,d,a)
//   X2 = G1 - sum_d E3(R,P,Q,b,d,c)K1(d,a)/2 - sum_i E3(R,P,Q,b,a,i)K2(c,i)/2
//   Z3(P,R,a,c) = sum_{d,j,i} [E3(R,P,d,j,i,c) w(d,j,i,a) - E3(R,P,d,j,a,i) w(d,j,c,i)]
// and the n^7 pieces inside one combine kernel.  C1..C4 are the 4-RDM contractions.
#include <rocblas/rocblas.h>

#include "nv2_common.h"
#include "nv2_stage1.h"

namespace nv2 {

#define AT(p, q, r, t) atwo[(p) + n * ((q) + n * ((r) + n * (t)))]
#define FF(p, q) F[(p) + n * (q)]

// E3f/E3tf (n^6) and E2f (n^4), column-major over their own index order
__global__ void k_bamat_tensors(Rdm R, int s, double *E3f, double *E3tf, double *E2f) {
    const int n = R.n;
    const i64 n6 = (i64)n * n * n * n * n * n;
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n6) return;
    int x[6];
    i64 t = e;
    for (int k = 0; k < 6; k++) { x[k] = t % n; t /= n; }
    E3f[e] = eee2(R, s, x[0], x[1], x[2], x[3], x[4], x[5]);
    E3tf[e] = eee2t(R, s, x[0], x[1], x[2], x[3], x[4], x[5]);
    if (e < (i64)n * n * n * n) E2f[e] = ee2(R, s, x[0], x[1], x[2], x[3]);
}
// Bx[(d + n i), (c + n a)] = w(c,i,d,a); K(x,y) = sum_m w(x,m,m,y) serves as both K1 and K2
__global__ void k_bamat_wmats(int n, const double *atwo, double *Bx, double *K) {
    const i64 n4 = (i64)n * n * n * n;
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n4) {
        int d = e % n, i = (e / n) % n, c = (e / n / n) % n, a = e / n / n / n;
        Bx[e] = AT(c, i, d, a);
    }
    if (e < (i64)n * n) {
        int x = e % n, y = e / n;
        double k = 0.0;
        for (int m = 0; m < n; m++) k += AT(x, m, m, y);
        K[e] = k;
    }
}
__global__ void k_bamat_z3(int n, const double *E3f, const double *atwo, double *Z3) {
    const i64 n4 = (i64)n * n * n * n;
    i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n4) return;
    const int P = e % n, R = (e / n) % n, a = (e / n / n) % n, c = e / n / n / n;
    double z = 0.0;
    for (int d = 0; d < n; d++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const i64 base = R + (i64)n * (P + (i64)n * (d + (i64)n * j));
                const i64 n4l = n4;
                z += E3f[base + n4l * (i + (i64)n * c)] * AT(d, j, i, a) - E3f[base + n4l * (a + (i64)n * i)] * AT(d, j, c, i);
            }
    Z3[e] = z;     // Z3 at P + n(R + n(a + n c))
}

__global__ void k_bamat_combine(int n, int M, int s, const double *E3f, const double *E3tf, const double *E2f,
                                const double *F, const double *K, const double *G1,
                                const double *G5, const double *Y2, const double *Z3, const double *C1,
                                const double *C2, const double *C3, const double *C4, double *amat, double *atmat) {
    const i64 N = n, n2 = N * N, n4 = n2 * n2, n6 = n4 * n2;
    // threads run R fastest (the leading index of every E3/G operand below, so their loads are
    // contiguous), then P, Q, a, b, c; a block's outputs (P-fastest rows of n) are completed by its
    // own lanes, so the strided stores still fill whole lines in L2
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n6) return;
    const int R = t % n, P = (t / n) % n, Q = (t / n2) % n, a = (t / (n2 * N)) % n, b = (t / n4) % n, c = t / (n4 * N);
    const i64 e = P + N * (Q + N * (R + N * (a + N * (b + N * (i64)c))));
    // E3 index helper: (x0..x5) -> x0 + n(x1 + n(x2 + ...))
#define I6(x0, x1, x2, x3, x4, x5) ((x0) + N * ((x1) + N * ((x2) + N * ((x3) + N * ((x4) + N * (i64)(x5))))))
    double am = G1[(R + N * (P + N * (Q + N * b))) + n4 * (c + N * a)];
    double at = 0.0;
    double xs = G1[(R + N * (P + N * (b + N * Q))) + n4 * (c + N * a)];     // X2 at (P, b, R, a, Q, c)
    for (int m = 0; m < n; m++) {
        const double k1ma = K[m + N * a], k2cm = K[c + N * m], k2bm = K[b + N * m];
        const double fma = FF(m, a), fcm = FF(c, m), fbm = FF(b, m);
        am += E3f[I6(R, P, Q, b, m, c)] * (fma - 0.5 * k1ma) - E3f[I6(R, P, Q, b, a, m)] * (fcm + 0.5 * k2cm)
              + E3f[I6(R, P, Q, m, a, c)] * (0.5 * k2bm - fbm);
        at += E3tf[I6(R, P, Q, b, m, c)] * fma - E3tf[I6(R, P, Q, b, a, m)] * fcm
              + E3tf[I6(R, P, Q, m, a, c)] * (fbm + 0.5 * k2bm);
        xs += -0.5 * E3f[I6(R, P, b, Q, m, c)] * k1ma - 0.5 * E3f[I6(R, P, b, Q, a, m)] * k2cm;
    }
    at += 2.0 * G5[(R + N * P) + n2 * (b + N * Q) + n4 * (a + N * c)] - xs;
    if (b == Q) {
        double y = Y2[(R + N * P) + n2 * (c + N * a)] + Z3[P + N * (R + N * (a + N * c))];
        for (int m = 0; m < n; m++)
            y += -0.5 * E2f[R + N * (P + N * (m + N * c))] * K[m + N * a] - 0.5 * E2f[R + N * (P + N * (a + N * m))] * K[c + N * m];
        at += 2.0 * y;
    }
#undef I6
    const i64 o = s + (i64)M * e;
    const i64 sw = s + (i64)M * (P + N * (b + N * (R + N * (a + N * (Q + N * c)))));
    amat[o] = am + C1[o] - C2[o] - C3[o];
    atmat[o] = at - C1[sw] + C2[sw] - C4[o];
}

static inline unsigned nbk(i64 tot, int bs) { return (unsigned)((tot + bs - 1) / bs); }

int launch_bamat_lo(rocblas_handle h, const Rdm &R, int s, const double *atwo, const double *Fhalf, double *work,
                    const double *C1, const double *C2, const double *C3, const double *C4, double *amat,
                    double *atmat, hipStream_t st) {
    const int n = R.n;
    const i64 N = n, n2 = N * N, n4 = n2 * n2, n6 = n4 * n2;
    // workspace layout (bamat_lo_workspace doubles)
    double *E3f = work, *E3tf = E3f + n6, *E2f = E3tf + n6, *Bx = E2f + n4, *K = Bx + n4;
    double *G1 = K + n2, *G5 = G1 + n6, *Y2 = G5 + n6, *Z3 = Y2 + n4;
    hipLaunchKernelGGL(k_bamat_tensors, dim3(nbk(n6, 256)), dim3(256), 0, st, R, s, E3f, E3tf, E2f);
    NV2_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bamat_wmats, dim3(nbk(n4, 256)), dim3(256), 0, st, n, atwo, Bx, K);
    NV2_LAUNCH_CHECK();
    if (rocblas_set_stream(h, st) != rocblas_status_success) { set_error("rocblas_set_stream"); return -1; }
    const double one = 1.0, zero = 0.0;
    auto ok = [](rocblas_status x) { return x == rocblas_status_success; };
    // G1 = E3f[(R,P,Q,b),(d,i)] x Bx
    if (!ok(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (int)n4, (int)n2, (int)n2, &one, E3f,
                          (int)n4, Bx, (int)n2, &zero, G1, (int)n4))) { set_error("bamat_lo G1 gemm"); return -1; }
    // G5 per (a,c): E3f[(R,P),(d,j)] x w[(d,j),(b,Q)]
    if (!ok(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, (int)n2, (int)n2,
                                          (int)n2, &one, E3f, (int)n2, n4, atwo, (int)n2, 0, &zero, G5, (int)n2, n4,
                                          (int)n2))) { set_error("bamat_lo G5 gemm"); return -1; }
    // Y2 = E2f[(R,P),(d,i)] x Bx
    if (!ok(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (int)n2, (int)n2, (int)n2, &one, E2f,
                          (int)n2, Bx, (int)n2, &zero, Y2, (int)n2))) { set_error("bamat_lo Y2 gemm"); return -1; }
    hipLaunchKernelGGL(k_bamat_z3, dim3(nbk(n4, 128)), dim3(128), 0, st, n, E3f, atwo, Z3);
    NV2_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bamat_combine, dim3(nbk(n6, 256)), dim3(256), 0, st, n, R.M, s, E3f, E3tf, E2f, Fhalf, K,
                       G1, G5, Y2, Z3, C1, C2, C3, C4, amat, atmat);
    NV2_LAUNCH_CHECK();
    return 0;
}

i64 bamat_lo_workspace(int n) {
    const i64 N = n, n2 = N * N, n4 = n2 * n2, n6 = n4 * n2;
    return 4 * n6 + 4 * n4 + n2;   // E3f E3tf G1 G5 | E2f Bx Y2 Z3 | K
}

}  // namespace nv2

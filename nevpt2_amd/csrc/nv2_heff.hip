// QD-NEVPT2 effective Hamiltonian post-processing (qdnevpt.F:1046-1204 and
// hermitian_heff, qdnevpt.F:8089-8184), for M <= 16 states, in one work-item.
//
//   A = Heff2 + diag(E0)                      (e2mp(i,i) += e(i), :1050-1052)
//   right eigenpairs of A, ascending          (dgeev + reordering, :1068-1092)
//   S = V^T V, C' = V S^{-1/2}, Heff~ = C' diag(w) C'^T, eigenvalues of Heff~ (dsyev 'U')
//
// Eigenvalues of the non-symmetric A come from shifted QR iterations on its
// Hessenberg form (Wilkinson shifts, deflation), eigenvectors from inverse
// iteration on A; complex eigenvalue pairs are reported as an error, as the
// hermitisation assumes a real spectrum.  Roots whose eigenvalues coincide
// (degenerate states are the common QD case) start from different vectors and
// are re-orthogonalised against the vectors already found in their cluster at
// every iteration, so a degenerate eigenspace yields independent vectors; a
// numerically singular S = V^T V (a defective Heff) is reported as an error.  Symmetric sub-problems use cyclic
// Jacobi.  Right eigenvectors are normalised to unit length as dgeev returns
// them; the hermitised matrix does not depend on their signs.
#include <hip/hip_runtime.h>

#include "nv2_common.h"
#include "../../include/nevpt2_amd.h"

namespace nv2 {

#define HM 16
#define A2(X, i, j) X[(i) + HM * (j)]

__device__ static void jacobi_sym(int m, double *A, double *U, double *w) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) A2(U, i, j) = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                tot += A2(A, i, j) * A2(A, i, j);
                if (i != j) off += A2(A, i, j) * A2(A, i, j);
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < m - 1; p++)
            for (int q = p + 1; q < m; q++) {
                const double apq = A2(A, p, q);
                if (apq == 0.0) continue;
                const double th = (A2(A, q, q) - A2(A, p, p)) / (2.0 * apq);
                const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; k++) {          // A <- A J
                    const double akp = A2(A, k, p), akq = A2(A, k, q);
                    A2(A, k, p) = c * akp - s * akq;
                    A2(A, k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < m; k++) {          // A <- J^T A
                    const double apk = A2(A, p, k), aqk = A2(A, q, k);
                    A2(A, p, k) = c * apk - s * aqk;
                    A2(A, q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < m; k++) {          // U <- U J
                    const double ukp = A2(U, k, p), ukq = A2(U, k, q);
                    A2(U, k, p) = c * ukp - s * ukq;
                    A2(U, k, q) = s * ukp + c * ukq;
                }
            }
    }
    for (int i = 0; i < m; i++) w[i] = A2(A, i, i);
}

// real eigenvalues of a general matrix (copied into H); returns 0, or 1 for a complex pair / no convergence
__device__ static int qr_eigenvalues(int m, double *H, double *w) {
    // Householder reduction to upper Hessenberg form
    for (int k = 0; k < m - 2; k++) {
        double v[HM], alpha = 0.0;
        for (int i = k + 1; i < m; i++) alpha += A2(H, i, k) * A2(H, i, k);
        alpha = sqrt(alpha);
        if (alpha == 0.0) continue;
        if (A2(H, k + 1, k) > 0.0) alpha = -alpha;
        for (int i = 0; i < m; i++) v[i] = 0.0;
        v[k + 1] = A2(H, k + 1, k) - alpha;
        for (int i = k + 2; i < m; i++) v[i] = A2(H, i, k);
        double vn = 0.0;
        for (int i = k + 1; i < m; i++) vn += v[i] * v[i];
        if (vn == 0.0) continue;
        for (int j = 0; j < m; j++) {          // H <- (I - 2vv^T/vn) H
            double d = 0.0;
            for (int i = k + 1; i < m; i++) d += v[i] * A2(H, i, j);
            d *= 2.0 / vn;
            for (int i = k + 1; i < m; i++) A2(H, i, j) -= d * v[i];
        }
        for (int i = 0; i < m; i++) {          // H <- H (I - 2vv^T/vn)
            double d = 0.0;
            for (int j = k + 1; j < m; j++) d += A2(H, i, j) * v[j];
            d *= 2.0 / vn;
            for (int j = k + 1; j < m; j++) A2(H, i, j) -= d * v[j];
        }
    }
    int hi = m - 1, iter = 0;
    while (hi >= 0) {
        if (hi == 0) { w[0] = A2(H, 0, 0); hi--; continue; }
        // deflation test on the last subdiagonal element
        const double sub = fabs(A2(H, hi, hi - 1));
        const double dsum = fabs(A2(H, hi, hi)) + fabs(A2(H, hi - 1, hi - 1));
        if (sub <= 1e-15 * (dsum > 0.0 ? dsum : 1.0)) { w[hi] = A2(H, hi, hi); hi--; iter = 0; continue; }
        // find the start of the active unreduced block
        int lo = hi - 1;
        while (lo > 0) {
            const double s2 = fabs(A2(H, lo, lo - 1));
            const double d2 = fabs(A2(H, lo, lo)) + fabs(A2(H, lo - 1, lo - 1));
            if (s2 <= 1e-15 * (d2 > 0.0 ? d2 : 1.0)) break;
            lo--;
        }
        if (lo == hi - 1) {       // isolated 2x2 block: solve it directly
            const double a = A2(H, hi - 1, hi - 1), b = A2(H, hi - 1, hi), c = A2(H, hi, hi - 1), d = A2(H, hi, hi);
            const double tr = a + d, det = a * d - b * c, disc = 0.25 * tr * tr - det;
            if (disc < 0.0) return 1;
            const double r = sqrt(disc);
            w[hi - 1] = 0.5 * tr - r;
            w[hi] = 0.5 * tr + r;
            hi -= 2;
            iter = 0;
            continue;
        }
        if (++iter > 200) return 1;
        // Wilkinson shift from the trailing 2x2 block (real part when complex)
        const double a = A2(H, hi - 1, hi - 1), b = A2(H, hi - 1, hi), c = A2(H, hi, hi - 1), d = A2(H, hi, hi);
        const double tr = a + d, det = a * d - b * c, disc = 0.25 * tr * tr - det;
        double mu = d;
        if (disc >= 0.0) {
            const double r = sqrt(disc), l1 = 0.5 * tr - r, l2 = 0.5 * tr + r;
            mu = fabs(l1 - d) < fabs(l2 - d) ? l1 : l2;
        } else {
            mu = 0.5 * tr;
        }
        if (iter % 11 == 10) mu += sub;    // exceptional shift
        // one shifted QR step on the block [lo, hi] with Givens rotations
        double cs[HM], sn[HM];
        for (int i = lo; i <= hi; i++) A2(H, i, i) -= mu;
        for (int k = lo; k < hi; k++) {
            const double x = A2(H, k, k), y = A2(H, k + 1, k);
            const double rr = sqrt(x * x + y * y);
            const double cc = rr == 0.0 ? 1.0 : x / rr, ss = rr == 0.0 ? 0.0 : y / rr;
            cs[k] = cc; sn[k] = ss;
            for (int j = k; j < m; j++) {
                const double h1 = A2(H, k, j), h2 = A2(H, k + 1, j);
                A2(H, k, j) = cc * h1 + ss * h2;
                A2(H, k + 1, j) = -ss * h1 + cc * h2;
            }
        }
        for (int k = lo; k < hi; k++) {
            const double cc = cs[k], ss = sn[k];
            const int top = (k + 2 < hi ? k + 2 : hi);
            for (int i = 0; i <= top; i++) {
                const double h1 = A2(H, i, k), h2 = A2(H, i, k + 1);
                A2(H, i, k) = cc * h1 + ss * h2;
                A2(H, i, k + 1) = -ss * h1 + cc * h2;
            }
        }
        for (int i = lo; i <= hi; i++) A2(H, i, i) += mu;
    }
    return 0;
}

// eigenvector of A for eigenvalue lam by inverse iteration (LU with partial pivoting), started from a
// vector that depends on the root index k and kept orthogonal to the ncl vectors V(:, cl[0..ncl-1]) of
// the same eigenvalue cluster
__device__ static void inverse_iteration(int m, const double *A, double lam, int k, const double *V, const int *cl,
                                         int ncl, double *x) {
    double L[HM * HM];
    int piv[HM];
    double scale = 0.0;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) scale = fmax(scale, fabs(A2(A, i, j)));
    const double shift = lam + 1e-13 * (scale > 0.0 ? scale : 1.0);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) A2(L, i, j) = A2(A, i, j) - (i == j ? shift : 0.0);
    for (int c = 0; c < m; c++) {
        int p = c;
        for (int i = c + 1; i < m; i++)
            if (fabs(A2(L, i, c)) > fabs(A2(L, p, c))) p = i;
        piv[c] = p;
        if (p != c)
            for (int j = 0; j < m; j++) { double t = A2(L, c, j); A2(L, c, j) = A2(L, p, j); A2(L, p, j) = t; }
        if (A2(L, c, c) == 0.0) A2(L, c, c) = 1e-300;
        for (int i = c + 1; i < m; i++) {
            const double f = A2(L, i, c) / A2(L, c, c);
            A2(L, i, c) = f;
            for (int j = c + 1; j < m; j++) A2(L, i, j) -= f * A2(L, c, j);
        }
    }
    for (int i = 0; i < m; i++) x[i] = 1.0 + 0.5 * sin(1.7 * (i + 1) * (k + 1)) + (i == k % m ? 1.0 : 0.0);
    for (int it = 0; it < 6; it++) {
        for (int c = 0; c < ncl; c++) {      // stay orthogonal to the cluster's vectors found so far
            const double *v = &A2(V, 0, cl[c]);
            double d = 0.0;
            for (int i = 0; i < m; i++) d += v[i] * x[i];
            for (int i = 0; i < m; i++) x[i] -= d * v[i];
        }
        for (int c = 0; c < m; c++) { const int p = piv[c]; if (p != c) { double t = x[c]; x[c] = x[p]; x[p] = t; } }
        for (int i = 0; i < m; i++)
            for (int c = 0; c < i; c++) x[i] -= A2(L, i, c) * x[c];
        for (int i = m - 1; i >= 0; i--) {
            for (int c = i + 1; c < m; c++) x[i] -= A2(L, i, c) * x[c];
            x[i] /= A2(L, i, i);
        }
        for (int c = 0; c < ncl; c++) {
            const double *v = &A2(V, 0, cl[c]);
            double d = 0.0;
            for (int i = 0; i < m; i++) d += v[i] * x[i];
            for (int i = 0; i < m; i++) x[i] -= d * v[i];
        }
        double nrm = 0.0;
        for (int i = 0; i < m; i++) nrm += x[i] * x[i];
        nrm = sqrt(nrm);
        if (!(nrm > 0.0)) nrm = 1.0;
        for (int i = 0; i < m; i++) x[i] /= nrm;
    }
}

// status[0]: 0 ok, 1 complex or unconverged eigenvalues, 2 linearly dependent eigenvectors (defective Heff)
__global__ void k_qd_heff(int m, const double *h2, const double *e0, double *heff, double *energies, int *status) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double A[HM * HM], H[HM * HM], V[HM * HM], S[HM * HM], U[HM * HM], w[HM], sw[HM];
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) A2(A, i, j) = h2[i + m * j] + (i == j ? e0[i] : 0.0);
    for (int i = 0; i < HM * HM; i++) H[i] = A[i];
    status[0] = qr_eigenvalues(m, H, w);
    if (status[0]) return;
    // ascending order (the reference's exchange sort, qdnevpt.F:1070-1092)
    for (int i = 0; i < m; i++)
        for (int j = i + 1; j < m; j++)
            if (w[j] < w[i]) { double t = w[i]; w[i] = w[j]; w[j] = t; }
    double scale = 0.0;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) scale = fmax(scale, fabs(A2(A, i, j)));
    for (int k = 0; k < m; k++) {
        int cl[HM], ncl = 0;
        for (int j = 0; j < k; j++)
            if (fabs(w[j] - w[k]) <= 1e-9 * (scale > 0.0 ? scale : 1.0)) cl[ncl++] = j;
        inverse_iteration(m, A, w[k], k, V, cl, ncl, &A2(V, 0, k));
    }
    // S = V^T V and S^{-1/2}
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) {
            double d = 0.0;
            for (int k = 0; k < m; k++) d += A2(V, k, i) * A2(V, k, j);
            A2(S, i, j) = d;
        }
    jacobi_sym(m, S, U, sw);
    double swmax = 0.0, swmin = 1e300;
    for (int i = 0; i < m; i++) { swmax = fmax(swmax, sw[i]); swmin = fmin(swmin, sw[i]); }
    if (!(swmin > 1e-10 * swmax)) { status[0] = 2; return; }
    double C[HM * HM];
    for (int i = 0; i < m; i++)          // C' = V U diag(1/sqrt(sw)) U^T
        for (int j = 0; j < m; j++) {
            double d = 0.0;
            for (int k = 0; k < m; k++) {
                double t = 0.0;
                for (int l = 0; l < m; l++) t += A2(V, i, l) * A2(U, l, k);
                d += t / sqrt(sw[k]) * A2(U, j, k);
            }
            A2(C, i, j) = d;
        }
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) {
            double d = 0.0;
            for (int k = 0; k < m; k++) d += A2(C, i, k) * w[k] * A2(C, j, k);
            A2(H, i, j) = d;
            heff[i + m * j] = d;
        }
    jacobi_sym(m, H, U, sw);
    for (int i = 0; i < m; i++)
        for (int j = i + 1; j < m; j++)
            if (sw[j] < sw[i]) { double t = sw[i]; sw[i] = sw[j]; sw[j] = t; }
    for (int i = 0; i < m; i++) energies[i] = sw[i];
}

}  // namespace nv2

using namespace nv2;

extern "C" int nv2_qd_heff(int64_t metat, const double *heff2, const double *e0, double *heff, double *energies) {
    if (metat < 1 || metat > HM) { set_error("nv2_qd_heff: metat=%lld out of range 1..%d", (long long)metat, HM); return -2; }
    const int m = (int)metat;
    double *d = nullptr;
    int *ds = nullptr;
    const size_t mm = (size_t)m * m;
    if (hipMalloc(&d, sizeof(double) * (2 * mm + 2 * m)) != hipSuccess) { set_error("hipMalloc failed"); return -1; }
    if (hipMalloc(&ds, sizeof(int)) != hipSuccess) { (void)hipFree(d); set_error("hipMalloc failed"); return -1; }
    int rc = 0, status = 0;
    if (hipMemcpy(d, heff2, sizeof(double) * mm, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + mm, e0, sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("nv2_qd_heff: upload failed"); rc = -1;
    }
    if (!rc) {
        hipLaunchKernelGGL(k_qd_heff, dim3(1), dim3(64), 0, 0, m, d, d + mm, d + mm + m, d + 2 * mm + m, ds);
        if (hipDeviceSynchronize() != hipSuccess) { set_error("nv2_qd_heff: kernel failed"); rc = -1; }
    }
    if (!rc && (hipMemcpy(&status, ds, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(heff, d + mm + m, sizeof(double) * mm, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(energies, d + 2 * mm + m, sizeof(double) * m, hipMemcpyDeviceToHost) != hipSuccess)) {
        set_error("nv2_qd_heff: download failed"); rc = -1;
    }
    if (!rc && status == 1) { set_error("nv2_qd_heff: Heff has complex or unconverged eigenvalues"); rc = -3; }
    if (!rc && status == 2) {
        set_error("nv2_qd_heff: the eigenvectors of Heff are linearly dependent (defective matrix)");
        rc = -3;
    }
    (void)hipFree(d);
    (void)hipFree(ds);
    return rc;
}

// Stage 1 of the NEVPT2 hot path on gfx950: Koopmans-matrix builds.
//
// O(n^7) builds (koopE, koop2E, koopman0pE, bbmat, bcmat, bdmat) run one
// wavefront per output element, the lanes sharing the reference's inner loop
// nest; RDMs stay resident in HBM/MALL.  Partial traces and the 2-hole RDM run
// one thread per element.  The O(n^9)
// 4-RDM part of bamat is split off: the full <E E E E> tensor E4 is
// materialised once per state (k_build_e4) and contracted with the active ERIs
// on the matrix cores (nv2_contract.hip); everything lower-order in bamat runs
// as GEMMs plus a combining kernel (nv2_bamat_lo.hip).
#include "nv2_common.h"
#include "nv2_stage1.h"

namespace nv2 {

static __device__ double ro3t(const Rdm &R, int s, int a, int b, int c, int ap, int bp, int cp) {
    // three-hole density (eexx_functions.F90:164-207)
    double v = -D3(s, ap, bp, cp, a, b, c);
    if (ap == a) v += 2.0 * D2(s, bp, cp, b, c);
    if (bp == b) v += 2.0 * D2(s, ap, cp, a, c);
    if (cp == c) v += 2.0 * D2(s, ap, bp, a, b);
    if (bp == c) v -= D2(s, ap, cp, a, b);
    if (ap == c) v -= D2(s, cp, bp, a, b);
    if (cp == b) v -= D2(s, bp, ap, c, a);
    if (ap == b) v -= D2(s, bp, cp, a, c);
    if (cp == a) v -= D2(s, ap, bp, c, b);
    if (bp == a) v -= D2(s, ap, cp, b, c);
    if (cp == c && bp == b) v -= 4.0 * D1(s, ap, a);
    if (cp == b && bp == c) v += 2.0 * D1(s, ap, a);
    if (cp == c && ap == a) v -= 4.0 * D1(s, bp, b);
    if (cp == a && ap == c) v += 2.0 * D1(s, bp, b);
    if (ap == a && bp == b) v -= 4.0 * D1(s, cp, c);
    if (ap == b && bp == a) v += 2.0 * D1(s, cp, c);
    if (cp == c && ap == b) v += 2.0 * D1(s, bp, a);
    if (cp == b && ap == c) v -= D1(s, bp, a);
    if (cp == c && bp == a) v += 2.0 * D1(s, ap, b);
    if (cp == a && bp == c) v -= D1(s, ap, b);
    if (bp == b && ap == c) v += 2.0 * D1(s, cp, a);
    if (bp == c && ap == b) v -= D1(s, cp, a);
    if (bp == b && cp == a) v += 2.0 * D1(s, ap, c);
    if (bp == a && cp == b) v -= D1(s, ap, c);
    if (ap == a && cp == b) v += 2.0 * D1(s, bp, c);
    if (ap == b && cp == a) v -= D1(s, bp, c);
    if (ap == a && bp == c) v += 2.0 * D1(s, cp, b);
    if (ap == c && bp == a) v -= D1(s, cp, b);
    if (ap == a && bp == b && cp == c) v += 8.0;
    if (ap == b && bp == a && cp == c) v -= 4.0;
    if (ap == a && bp == c && cp == b) v -= 4.0;
    if (ap == c && bp == b && cp == a) v -= 4.0;
    if (ap == b && bp == c && cp == a) v += 2.0;
    if (ap == c && bp == a && cp == b) v += 2.0;
    return v;
}

#define AT(p, q, r, t) atwo[(p) + n * ((q) + n * ((r) + n * (t)))]
#define FF(p, q) F[(p) + n * (q)]

// decode a flat Fortran index with the state fastest: out = s + M*(i0 + n*(i1 + ...))
__device__ __forceinline__ void decode(i64 idx, int M, int n, int k, int *s, int *ix) {
    *s = (int)(idx % M);
    idx /= M;
    for (int t = 0; t < k; t++) { ix[t] = (int)(idx % n); idx /= n; }
}

// ---------------------------------------------------------------- partial traces
__global__ void k_bro3(Rdm R, int nele, double *daaa) {
    const int n = R.n, M = R.M;
    i64 tot = (i64)M * n * n * n * n * n * n;
    i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    int s, x[6];
    decode(idx, M, n, 6, &s, x);
    double v = 0.0;
    for (int ll = 0; ll < n; ll++) v += g4(R, s, x[0], x[1], x[2], ll, x[3], x[4], x[5], ll);
    daaa[idx] = v / (double)(nele - 3);
}
__global__ void k_bro2(Rdm R, int nele, double *taa) {
    const int n = R.n, M = R.M;
    i64 tot = (i64)M * n * n * n * n;
    i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    int s, x[4];
    decode(idx, M, n, 4, &s, x);
    double v = 0.0;
    for (int ll = 0; ll < n; ll++) v += D3(s, x[0], x[1], ll, x[2], x[3], ll);
    taa[idx] = v / (double)(nele - 2);
}
__global__ void k_bro1(Rdm R, int nele, double *dal) {
    const int n = R.n, M = R.M;
    i64 tot = (i64)M * n * n;
    i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    int s, x[2];
    decode(idx, M, n, 2, &s, x);
    double v = 0.0;
    for (int ll = 0; ll < n; ll++) v += D2(s, x[0], ll, x[1], ll);
    dal[idx] = v / (double)(nele - 1);
}

// ---------------------------------------------------------------- Dyall h, active ERIs (koopro4QD.F90:3900-3937)
__global__ void k_atwo(int n, int ncore, int L, const double *B, double *atwo) {
    i64 tot = (i64)n * n * n * n;
    i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    int i = idx % n, j = (idx / n) % n, k = (idx / n / n) % n, l = idx / n / n / n;
    const double *u = B + (i64)L * tri(i + ncore, j + ncore), *v = B + (i64)L * tri(k + ncore, l + ncore);
    double x = 0.0;
    for (int m = 0; m < L; m++) x += u[m] * v[m];
    atwo[idx] = x;
}
// one wavefront per active pair a<=b; lanes split the Cholesky index
__global__ void k_aj(int n, int ncore, int L, const double *B, double *aj) {
    const int pr = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (pr >= n * n) return;
    const int a = pr / n, b = pr % n;
    if (b < a) return;                          // uniform per wavefront
    const i64 ac = a + ncore, bc = b + ncore, jab = tri(ac, bc);
    double acc = 0.0;
    for (int j = 0; j < ncore; j++) {
        const double *u = B + (i64)L * tri(ac, bc), *v = B + (i64)L * tri(j, j);
        const double *u2 = B + (i64)L * tri(ac, j), *v2 = B + (i64)L * tri(j, bc);
        double x = 0, y = 0;
        for (int l = lane; l < L; l += 64) { x += u[l] * v[l]; y += u2[l] * v2[l]; }
        x = wave_sum(x);
        y = wave_sum(y);
        acc = acc + 2.0 * x - y;
    }
    if (lane == 0) aj[jab] = aj[jab] + acc;
}

// active f(ab) = aj(ab) - alpha * sum_c atwo(a,c,c,b) (e.g. koopro4QD.F90:3993-4005)
__global__ void k_make_f(int n, int ncore, const double *aj, const double *atwo, double alpha, double *F) {
    int a = blockIdx.x, b = threadIdx.x;
    if (b < a || b >= n) return;
    double v = aj[tri(a + ncore, b + ncore)];
    for (int c = 0; c < n; c++) v = v - AT(a, c, c, b) * alpha;
    FF(a, b) = v;
    FF(b, a) = v;
}
// active block of aj as an n x n matrix
__global__ void k_aj_active(int n, int ncore, const double *aj, double *A) {
    int a = blockIdx.x, b = threadIdx.x;
    if (b >= n) return;
    A[a + n * b] = aj[tri(a + ncore, b + ncore)];
}

// The O(n^7) builds below run one wavefront per output element; the 64 lanes
// split the flattened n^3 inner loop of the reference and reduce with shuffles.
#define WAVE_SETUP(ntot)                                                          \
    const i64 wid = ((i64)blockIdx.x * blockDim.x + threadIdx.x) >> 6;          \
    const int lane = threadIdx.x & 63;                                            \
    if (wid >= (ntot)) return;

// ---------------------------------------------------------------- koopE (koop_matrices.F90:165-250); F = aj
__global__ void k_koopE(Rdm R, const double *atwo, const double *F, double *ipa, double *eaa) {
    const int n = R.n, M = R.M;
    WAVE_SETUP((i64)M * n * n);
    int s, x[2];
    decode(wid, M, n, 2, &s, x);
    const int ap = x[0], a = x[1];
    double ip = 0.0, dumm = 0.0;
    const int n3 = n * n * n;
    for (int t = lane; t < n3; t += 64) {
        const int c = t / (n * n), d = (t / n) % n, ie = t % n;
        if (d == 0 && ie == 0) ip = ip - FF(a, c) * D1(s, ap, c);
        ip = ip - AT(c, ie, a, d) * D2(s, ap, c, d, ie);
        if (c == 0) dumm = dumm + (2.0 * AT(d, ie, ap, a) - AT(ap, ie, d, a)) * D1(s, d, ie);   // (d,ie) = t
    }
    ip = 2.0 * wave_sum(ip);
    dumm = 2.0 * FF(ap, a) + wave_sum(dumm);
    if (lane == 0) {
        ipa[wid] = ip;
        eaa[wid] = ip + 2.0 * dumm;
    }
}

// ---------------------------------------------------------------- 2-hole RDM (koop_matrices.F90:78-88)
__global__ void k_hrdm(Rdm R, double *out) {
    const int n = R.n, M = R.M;
    i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (i64)M * n * n * n * n) return;
    int s, x[4];
    decode(idx, M, n, 4, &s, x);
    out[idx] = ro2t(R, s, x[0], x[1], x[2], x[3]);
}

// ---------------------------------------------------------------- koop2E (koop_matrices.F90:253-386)
__global__ void k_koop2E(Rdm R, const double *atwo, const double *F, double *koopaa, double *koopeaa) {
    const int n = R.n, M = R.M;
    WAVE_SETUP((i64)M * n * n * n * n);
    int s, x[4];
    decode(wid, M, n, 4, &s, x);
    const int ap = x[0], bp = x[1], a = x[2], b = x[3];
    double ip = 0.0, ea = 0.0;
    const int n3 = n * n * n;
    for (int t = lane; t < n3; t += 64) {
        const int d = t / (n * n), ie = (t / n) % n, ief = t % n;
        if (ie == 0 && ief == 0) ip = ip - FF(a, d) * D2(s, ap, bp, d, b) - FF(b, d) * D2(s, ap, bp, a, d);
        double dint = AT(d, ie, a, ief), dint2 = AT(d, ie, b, ief);
        double dum = D3(s, ap, bp, d, ief, b, ie);
        if (d == ief) dum = dum + 0.5 * D2(s, ap, bp, ie, b);
        if (d == b) dum = dum + 0.5 * D2(s, ap, bp, ief, ie);
        ip = ip - dint * dum;
        dum = D3(s, ap, bp, d, a, ief, ie);
        if (d == a) dum = dum + 0.5 * D2(s, ap, bp, ie, ief);
        if (d == ief) dum = dum + 0.5 * D2(s, ap, bp, a, ie);
        ip = ip - dint2 * dum;
        // EA part: loop indices (c, d2, ie2) = (d, ie, ief)
        const int c = d, dd = ie, e2 = ief;
        if (dd == 0 && e2 == 0) ea = ea + FF(a, c) * ro2t(R, s, ap, bp, c, b) + FF(b, c) * ro2t(R, s, ap, bp, a, c);
        dint = AT(c, e2, dd, a);
        dum = -ro3t(R, s, ap, bp, e2, dd, b, c);
        if (c == e2) dum = dum + 2.0 * ro2t(R, s, ap, bp, dd, b);
        if (dd == e2) dum = dum - 0.5 * ro2t(R, s, ap, bp, c, b);
        if (b == e2) dum = dum - 0.5 * ro2t(R, s, ap, bp, dd, c);
        ea = ea + dint * dum;
        dint = AT(c, e2, dd, b);
        dum = -ro3t(R, s, ap, bp, e2, a, dd, c);
        if (c == e2) dum = dum + 2.0 * ro2t(R, s, ap, bp, a, dd);
        if (a == e2) dum = dum - 0.5 * ro2t(R, s, ap, bp, c, dd);
        if (dd == e2) dum = dum - 0.5 * ro2t(R, s, ap, bp, a, c);
        ea = ea + dint * dum;
    }
    ip = wave_sum(ip);
    ea = wave_sum(ea);
    if (lane == 0) {
        koopaa[wid] = ip;
        koopeaa[wid] = ea;
    }
}

// ---------------------------------------------------------------- koopman0pE (koop_matrices.F90:389-588)
// wave index (s,bp,ap,a,b) matches the output layout koopaa(s,bp,ap,a,b)
__global__ void k_koop0p(Rdm R, const double *atwo, const double *F, double *koopaa, double *koopbb) {
    const int n = R.n, M = R.M;
    WAVE_SETUP((i64)M * n * n * n * n);
    int s, x[4];
    decode(wid, M, n, 4, &s, x);
    const int bp = x[0], ap = x[1], a = x[2], b = x[3];
    double ka = 0.0, kb = 0.0;
    const int n3 = n * n * n;
    for (int t = lane; t < n3; t += 64) {
        const int c = t / (n * n), d = (t / n) % n, ie = t % n;
        if (d == 0 && ie == 0) {
            ka = ka + 2.0 * (FF(c, a) * ee2(R, s, bp, ap, c, b) - FF(b, c) * ee2(R, s, bp, ap, a, c));
            double dum1 = -ee2(R, s, a, ap, bp, c);
            if (a == ap) dum1 = dum1 + 2.0 * D1(s, bp, c);
            if (bp == ap) dum1 = dum1 + D1(s, a, c);
            double dum2 = -ee2(R, s, c, ap, bp, b);
            if (ap == c) dum2 = dum2 + 2.0 * D1(s, bp, b);
            if (bp == ap) dum2 = dum2 + D1(s, c, b);
            kb = kb - FF(b, c) * dum1 + FF(c, a) * dum2;
        }
        {
            double dum1 = eee2(R, s, bp, ap, c, ie, d, b) + eee2(R, s, bp, ap, d, b, c, ie);
            double dum2 = eee2(R, s, bp, ap, a, ie, c, d) + eee2(R, s, bp, ap, c, d, a, ie);
            ka = ka + AT(c, ie, d, a) * dum1 - AT(b, ie, c, d) * dum2;
        }
        double e1 = -eee2(R, s, c, ie, a, ap, bp, d) - eee2(R, s, a, ap, bp, d, c, ie);
        if (ap == a) e1 = e1 + 2.0 * (ee2(R, s, c, ie, bp, d) + ee2(R, s, bp, d, c, ie));
        if (bp == ap) e1 = e1 + ee2(R, s, c, ie, a, d) + ee2(R, s, a, d, c, ie);
        if (ap == c) {
            e1 = e1 - ee2(R, s, a, ie, bp, d);
            if (ie == a) e1 = e1 + 2.0 * D1(s, bp, d);
        }
        if (bp == ie) {
            e1 = e1 + ee2(R, s, a, ap, c, d);
            if (ap == a) e1 = e1 - 2.0 * D1(s, c, d);
        }
        double e2 = -eee2(R, s, c, ie, d, ap, bp, b) - eee2(R, s, d, ap, bp, b, c, ie);
        if (ap == d) e2 = e2 + 2.0 * (ee2(R, s, c, ie, bp, b) + ee2(R, s, bp, b, c, ie));
        if (bp == ap) e2 = e2 + ee2(R, s, c, ie, d, b) + ee2(R, s, d, b, c, ie);
        if (ap == c) {
            e2 = e2 - ee2(R, s, d, ie, bp, b);
            if (d == ie) e2 = e2 + 2.0 * D1(s, bp, b);
        }
        if (bp == ie) {
            e2 = e2 + ee2(R, s, d, ap, c, b);
            if (d == ap) e2 = e2 - 2.0 * D1(s, c, b);
        }
        kb = kb - 0.5 * AT(c, ie, b, d) * e1 + 0.5 * AT(c, ie, d, a) * e2;
    }
    ka = wave_sum(ka);
    kb = wave_sum(kb);
    if (lane == 0) {
        koopaa[wid] = ka;
        koopbb[wid] = kb;
    }
}
// koop2E's two n^4 matrices as GEMMs (the same sums as k_koop2E): with atwo3[(d,ie,ief), a] =
// atwo(d,ie,a,ief) (k_k0p_at3),
//   C1[(ap,bp,b),a] = sum_(d,ie,ief) Y1 atwo(d,ie,a,ief),  Y1 = the kernel's first IP dum
//   C2[(ap,bp,a),b] = sum_(d,ie,ief) Y2 atwo(d,ie,b,ief),  Y2 = its second IP dum
//   C3[(ap,bp,b),a] = sum_(c,e2,dd) Z1 atwo(c,e2,dd,a),    Z1, Z2 = its EA dums
//   C4[(ap,bp,a),b] = sum_(c,e2,dd) Z2 atwo(c,e2,dd,b)
//   koopaa = fock_ip - C1 - C2,  koopeaa = fock_ea + C3 + C4
__global__ void k_k2e_x(Rdm R, int s, int which, double *X) {
    const int n = R.n;
    const i64 n3 = (i64)n * n * n, e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n3 * n3) return;
    const int r = (int)(e % n3), cl = (int)(e / n3);
    const int ap = r % n, bp = (r / n) % n, r2 = r / (n * n);     // rows (ap,bp,b) or (ap,bp,a)
    const int q0 = cl % n, q1 = (cl / n) % n, q2 = cl / (n * n);
    double v;
    if (which == 0) {                 // Y1, columns (d,ie,ief)
        const int b = r2, d = q0, ie = q1, ief = q2;
        v = D3(s, ap, bp, d, ief, b, ie);
        if (d == ief) v = v + 0.5 * D2(s, ap, bp, ie, b);
        if (d == b) v = v + 0.5 * D2(s, ap, bp, ief, ie);
    } else if (which == 1) {          // Y2, columns (d,ie,ief)
        const int a = r2, d = q0, ie = q1, ief = q2;
        v = D3(s, ap, bp, d, a, ief, ie);
        if (d == a) v = v + 0.5 * D2(s, ap, bp, ie, ief);
        if (d == ief) v = v + 0.5 * D2(s, ap, bp, a, ie);
    } else if (which == 2) {          // Z1, columns (c,e2,dd)
        const int b = r2, c = q0, e2 = q1, dd = q2;
        v = -ro3t(R, s, ap, bp, e2, dd, b, c);
        if (c == e2) v = v + 2.0 * ro2t(R, s, ap, bp, dd, b);
        if (dd == e2) v = v - 0.5 * ro2t(R, s, ap, bp, c, b);
        if (b == e2) v = v - 0.5 * ro2t(R, s, ap, bp, dd, c);
    } else {                          // Z2, columns (c,e2,dd)
        const int a = r2, c = q0, e2 = q1, dd = q2;
        v = -ro3t(R, s, ap, bp, e2, a, dd, c);
        if (c == e2) v = v + 2.0 * ro2t(R, s, ap, bp, a, dd);
        if (a == e2) v = v - 0.5 * ro2t(R, s, ap, bp, c, dd);
        if (dd == e2) v = v - 0.5 * ro2t(R, s, ap, bp, a, c);
    }
    X[e] = v;
}
__global__ void k_k2e_combine(Rdm R, int s, const double *F, const double *C, double *koopaa, double *koopeaa) {
    const int n = R.n, M = R.M;
    const i64 n3 = (i64)n * n * n, n4 = n3 * n, e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n4) return;
    const int ap = (int)(e % n), bp = (int)((e / n) % n), a = (int)((e / (n * n)) % n), b = (int)(e / n3);
    const double *C1 = C, *C2 = C + n4, *C3 = C + 2 * n4, *C4 = C + 3 * n4;
    double ip = 0.0, ea = 0.0;
    for (int d = 0; d < n; d++) {     // the kernel's ie = ief = 0 (dd = e2 = 0) terms
        ip = ip - FF(a, d) * D2(s, ap, bp, d, b) - FF(b, d) * D2(s, ap, bp, a, d);
        ea = ea + FF(a, d) * ro2t(R, s, ap, bp, d, b) + FF(b, d) * ro2t(R, s, ap, bp, a, d);
    }
    const i64 rb = ap + n * (bp + (i64)n * b), ra = ap + n * (bp + (i64)n * a);
    ip = ip - C1[rb + n3 * a] - C2[ra + n3 * b];
    ea = ea + C3[rb + n3 * a] + C4[ra + n3 * b];
    const i64 o = s + (i64)M * e;
    koopaa[o] = ip;
    koopeaa[o] = ea;
}
// koopman0pE's two n^4 matrices as GEMMs (the same sums as k_koop0p, qdnevpt.F koopman0pE): each of
// its four n^7 terms is sum_{three indices} atwo(..) * X(..) with X built from the 2- and 3-RDM, so X
// is materialised for one state as an n^3 x n^3 matrix and multiplied with atwo (n^3 x n); the O(n^5)
// Fock terms and the final index order go to k_k0p_combine.
//   C1[(bp,ap,b),a] = sum_(c,ie,d) X1 atwo(c,ie,d,a),  X1 = eee2(bp,ap,c,ie,d,b) + eee2(bp,ap,d,b,c,ie)
//   C2[(bp,ap,a),b] = sum_(ie,c,d) X2 atwo(b,ie,c,d),  X2 = eee2(bp,ap,a,ie,c,d) + eee2(bp,ap,c,d,a,ie)
//   C3[(a,ap,bp),b] = sum_(c,ie,d) E1 atwo(c,ie,b,d),  E1 = the kernel's e1
//   C4[(b,ap,bp),a] = sum_(c,ie,d) E2 atwo(c,ie,d,a),  E2 = the kernel's e2
//   koopaa = fock_a + C1 - C2,  koopbb = fock_b - C3/2 + C4/2
__global__ void k_k0p_x(Rdm R, int s, int which, double *X) {
    const int n = R.n;
    const i64 n3 = (i64)n * n * n, e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n3 * n3) return;
    const int r = (int)(e % n3), cl = (int)(e / n3);
    const int r0 = r % n, r1 = (r / n) % n, r2 = r / (n * n);
    const int q0 = cl % n, q1 = (cl / n) % n, q2 = cl / (n * n);
    double v;
    if (which == 0) {                 // rows (bp,ap,b), columns (c,ie,d)
        const int bp = r0, ap = r1, b = r2, c = q0, ie = q1, d = q2;
        v = eee2(R, s, bp, ap, c, ie, d, b) + eee2(R, s, bp, ap, d, b, c, ie);
    } else if (which == 1) {          // rows (bp,ap,a), columns (ie,c,d)
        const int bp = r0, ap = r1, a = r2, ie = q0, c = q1, d = q2;
        v = eee2(R, s, bp, ap, a, ie, c, d) + eee2(R, s, bp, ap, c, d, a, ie);
    } else if (which == 2) {          // rows (a,ap,bp), columns (c,ie,d): e1
        const int a = r0, ap = r1, bp = r2, c = q0, ie = q1, d = q2;
        v = -eee2(R, s, c, ie, a, ap, bp, d) - eee2(R, s, a, ap, bp, d, c, ie);
        if (ap == a) v = v + 2.0 * (ee2(R, s, c, ie, bp, d) + ee2(R, s, bp, d, c, ie));
        if (bp == ap) v = v + ee2(R, s, c, ie, a, d) + ee2(R, s, a, d, c, ie);
        if (ap == c) {
            v = v - ee2(R, s, a, ie, bp, d);
            if (ie == a) v = v + 2.0 * D1(s, bp, d);
        }
        if (bp == ie) {
            v = v + ee2(R, s, a, ap, c, d);
            if (ap == a) v = v - 2.0 * D1(s, c, d);
        }
    } else {                          // rows (b,ap,bp), columns (c,ie,d): e2
        const int b = r0, ap = r1, bp = r2, c = q0, ie = q1, d = q2;
        v = -eee2(R, s, c, ie, d, ap, bp, b) - eee2(R, s, d, ap, bp, b, c, ie);
        if (ap == d) v = v + 2.0 * (ee2(R, s, c, ie, bp, b) + ee2(R, s, bp, b, c, ie));
        if (bp == ap) v = v + ee2(R, s, c, ie, d, b) + ee2(R, s, d, b, c, ie);
        if (ap == c) {
            v = v - ee2(R, s, d, ie, bp, b);
            if (d == ie) v = v + 2.0 * D1(s, bp, b);
        }
        if (bp == ie) {
            v = v + ee2(R, s, d, ap, c, b);
            if (d == ap) v = v - 2.0 * D1(s, c, b);
        }
    }
    X[e] = v;
}
// AT3[(c,ie,d), b] = atwo(c,ie,b,d)
__global__ void k_k0p_at3(int n, const double *atwo, double *AT3) {
    const i64 n2 = (i64)n * n, e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n2 * n2) return;
    const int c = (int)(e % n), ie = (int)((e / n) % n), d = (int)((e / n2) % n), b = (int)(e / (n2 * n));
    AT3[e] = atwo[c + n * (ie + n * (b + (i64)n * d))];
}
__global__ void k_k0p_combine(Rdm R, int s, const double *F, const double *C, double *koopaa, double *koopbb) {
    const int n = R.n, M = R.M;
    const i64 n3 = (i64)n * n * n, n4 = n3 * n, e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n4) return;
    const int bp = (int)(e % n), ap = (int)((e / n) % n), a = (int)((e / (n * n)) % n), b = (int)(e / n3);
    const double *C1 = C, *C2 = C + n4, *C3 = C + 2 * n4, *C4 = C + 3 * n4;
    double ka = 0.0, kb = 0.0;
    for (int c = 0; c < n; c++) {     // the terms of the kernel's d = ie = 0 branch
        ka = ka + 2.0 * (FF(c, a) * ee2(R, s, bp, ap, c, b) - FF(b, c) * ee2(R, s, bp, ap, a, c));
        double dum1 = -ee2(R, s, a, ap, bp, c);
        if (a == ap) dum1 = dum1 + 2.0 * D1(s, bp, c);
        if (bp == ap) dum1 = dum1 + D1(s, a, c);
        double dum2 = -ee2(R, s, c, ap, bp, b);
        if (ap == c) dum2 = dum2 + 2.0 * D1(s, bp, b);
        if (bp == ap) dum2 = dum2 + D1(s, c, b);
        kb = kb - FF(b, c) * dum1 + FF(c, a) * dum2;
    }
    ka = ka + C1[(bp + n * (ap + (i64)n * b)) + n3 * a] - C2[(bp + n * (ap + (i64)n * a)) + n3 * b];
    kb = kb - 0.5 * C3[(a + n * (ap + (i64)n * bp)) + n3 * b] + 0.5 * C4[(b + n * (ap + (i64)n * bp)) + n3 * a];
    const i64 o = s + (i64)M * e;
    koopaa[o] = ka;
    koopbb[o] = kb;
}
i64 koop0p_workspace(int n) {   // one n^6 term matrix, atwo3, six n^4 products (bcmat's C1..C3t)
    const i64 n3 = (i64)n * n * n;
    return n3 * n3 + 7 * n3 * n;
}
static int launch_koop0p_gemm(const Rdm &R, const double *atwo, const double *F, rocblas_handle h, double *work,
                              double *koopaa, double *koopbb, hipStream_t st) {
    const int n = R.n;
    const i64 n3 = (i64)n * n * n, n4 = n3 * n, n6 = n3 * n3;
    double *X = work, *AT3 = work + n6, *C = AT3 + n4;
    if (rocblas_set_stream(h, st) != rocblas_status_success) { set_error("rocblas_set_stream failed"); return -1; }
    hipLaunchKernelGGL(k_k0p_at3, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, n, atwo, AT3);
    NV2_LAUNCH_CHECK();
    const double one = 1.0, zero = 0.0;
    const rocblas_operation Nn = rocblas_operation_none, Tt = rocblas_operation_transpose;
    const int m3 = (int)n3;
    for (int s = 0; s < R.M; s++) {
        for (int which = 0; which < 4; which++) {
            hipLaunchKernelGGL(k_k0p_x, dim3((unsigned)((n6 + 255) / 256)), dim3(256), 0, st, R, s, which, X);
            NV2_LAUNCH_CHECK();
            rocblas_status rs;
            double *Cw = C + which * n4;
            if (which == 1)           // B[(ie,c,d), b] = atwo[b + n (ie,c,d)]: atwo as n x n^3, transposed
                rs = rocblas_dgemm(h, Nn, Tt, m3, n, m3, &one, X, m3, atwo, n, &zero, Cw, m3);
            else
                rs = rocblas_dgemm(h, Nn, Nn, m3, n, m3, &one, X, m3, which == 2 ? AT3 : atwo, m3, &zero, Cw, m3);
            if (rs != rocblas_status_success) { set_error("koop0p GEMM failed (%d)", (int)rs); return -1; }
        }
        hipLaunchKernelGGL(k_k0p_combine, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, R, s, F, C, koopaa,
                           koopbb);
        NV2_LAUNCH_CHECK();
    }
    return 0;
}
static int launch_koop2e_gemm(const Rdm &R, const double *atwo, const double *F, rocblas_handle h, double *work,
                              double *koopaa, double *koopeaa, hipStream_t st) {
    const int n = R.n;
    const i64 n3 = (i64)n * n * n, n4 = n3 * n, n6 = n3 * n3;
    double *X = work, *AT3 = work + n6, *C = AT3 + n4;
    if (rocblas_set_stream(h, st) != rocblas_status_success) { set_error("rocblas_set_stream failed"); return -1; }
    hipLaunchKernelGGL(k_k0p_at3, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, n, atwo, AT3);
    NV2_LAUNCH_CHECK();
    const double one = 1.0, zero = 0.0;
    const rocblas_operation Nn = rocblas_operation_none;
    const int m3 = (int)n3;
    for (int s = 0; s < R.M; s++) {
        for (int which = 0; which < 4; which++) {
            hipLaunchKernelGGL(k_k2e_x, dim3((unsigned)((n6 + 255) / 256)), dim3(256), 0, st, R, s, which, X);
            NV2_LAUNCH_CHECK();
            if (rocblas_dgemm(h, Nn, Nn, m3, n, m3, &one, X, m3, which < 2 ? AT3 : atwo, m3, &zero, C + which * n4, m3) !=
                rocblas_status_success) { set_error("koop2E GEMM failed"); return -1; }
        }
        hipLaunchKernelGGL(k_k2e_combine, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, R, s, F, C, koopaa,
                           koopeaa);
        NV2_LAUNCH_CHECK();
    }
    return 0;
}
__global__ void k_koop0pf(Rdm R, const double *atwo, const double *F, double *koopab) {
    const int n = R.n, M = R.M;
    WAVE_SETUP((i64)M * n * n);
    int s, x[2];
    decode(wid, M, n, 2, &s, x);
    const int a = x[0], b = x[1];
    double v = 0.0;
    const int n3 = n * n * n;
    for (int t = lane; t < n3; t += 64) {
        const int c = t / (n * n), d = (t / n) % n, ie = t % n;
        if (d == 0 && ie == 0) v = v + 2.0 * (FF(c, a) * D1(s, c, b) - FF(b, c) * D1(s, a, c));
        double dum1 = ee2(R, s, c, ie, d, b) + ee2(R, s, d, b, c, ie);
        double dum2 = ee2(R, s, a, ie, c, d) + ee2(R, s, c, d, a, ie);
        v = v + AT(c, ie, d, a) * dum1 - AT(b, ie, c, d) * dum2;
    }
    v = wave_sum(v);
    if (lane == 0) koopab[wid] = v;
}

// ---------------------------------------------------------------- bbmat (koopro4QD.F90:4231-4299); F alpha=1
__global__ void k_bbmat(Rdm R, const double *atwo, const double *F, const double *AJ, double *bmat, double *btmat) {
    const int n = R.n, M = R.M;
    WAVE_SETUP((i64)M * n * n * n * n);
    int s, x[4];
    decode(wid, M, n, 4, &s, x);
    const int ap = x[0], bp = x[1], cp = x[2], a = x[3];
    double bv = 0.0, bt = 0.0;
    const int n3 = n * n * n;
    for (int t = lane; t < n3; t += 64) {
        const int c = t / (n * n), ie = (t / n) % n, ief = t % n;
        if (ie == 0 && ief == 0) {
            bv = bv - FF(c, a) * ee2(R, s, cp, ap, bp, c);
            bt = bt + AJ[c + n * a] * ee2tr(R, s, cp, ap, bp, c);
        }
        double dint = AT(a, ie, c, ief);
        bv = bv - eee2(R, s, cp, ap, bp, ie, c, ief) * dint;
        bt = bt + eee2t(R, s, cp, ap, bp, ie, c, ief) * dint;
    }
    bv = wave_sum(bv);
    bt = wave_sum(bt);
    if (lane == 0) {
        bmat[wid] = bv;
        btmat[wid] = bt;
    }
}

// ---------------------------------------------------------------- bcmat (koopro4QD.F90:4301-4397); F alpha=1/2
__global__ void k_bcmat(Rdm R, const double *atwo, const double *F, double *cmat, double *ctmat) {
    const int n = R.n, M = R.M;
    WAVE_SETUP((i64)M * n * n * n * n);
    int s, x[4];
    decode(wid, M, n, 4, &s, x);
    const int ap = x[0], a = x[1], b = x[2], c = x[3];
    double cv = 0.0, ct = 0.0;
    const int n3 = n * n * n;
    for (int t = lane; t < n3; t += 64) {
        const int d = t / (n * n), ie = (t / n) % n, ief = t % n;
        if (ie == 0 && ief == 0) {
            double fda = FF(d, a), fcd = FF(c, d), fbd = FF(b, d);
            cv = cv + fda * ee2(R, s, ap, b, d, c) - fcd * ee2(R, s, ap, b, a, d) - fbd * ee2(R, s, ap, d, a, c);
            ct = ct + fda * ee2t(R, s, ap, b, d, c) - fcd * ee2t(R, s, ap, b, a, d) + fbd * ee2t(R, s, ap, d, a, c);
        }
        if (ief == 0) {
            double dint1 = AT(c, ie, d, a), dint2 = AT(d, ie, ie, a), dint3 = AT(c, d, d, ie), dint4 = AT(b, d, d, ie);
            double dum1 = ee2(R, s, ap, b, d, ie), dum2 = ee2(R, s, ap, b, d, c);
            double dum3 = ee2(R, s, ap, b, a, ie), dum4 = ee2(R, s, ap, ie, a, c);
            double dum1t = ee2t(R, s, ap, b, d, ie), dum2t = ee2t(R, s, ap, b, d, c);
            double dum3t = ee2t(R, s, ap, b, a, ie), dum4t = ee2t(R, s, ap, ie, a, c);
            cv = cv + dum1 * dint1 - 0.5 * (dum2 * dint2 + dum3 * dint3 - dum4 * dint4);
            ct = ct + dum1t * dint1 - 0.5 * (dum2t * dint2 + dum3t * dint3 - dum4t * dint4);
        }
        double di1 = AT(d, ief, ie, a), di2 = AT(d, ief, c, ie), di3 = AT(d, ief, b, ie);
        double du1 = eee2(R, s, ap, b, d, ief, ie, c), du1t = eee2tl(R, s, ap, b, d, ief, ie, c);
        double du2 = eee2(R, s, ap, b, d, ief, a, ie), du2t = eee2tl(R, s, ap, b, d, ief, a, ie);
        double du3 = eee2(R, s, ap, ie, d, ief, a, c), du3t = eee2tl(R, s, ap, ie, d, ief, a, c);
        cv = cv + du1 * di1 - du2 * di2 - du3 * di3;
        ct = ct + du1t * di1 - du2t * di2 + du3t * di3;
    }
    cv = wave_sum(cv);
    ct = wave_sum(ct);
    if (lane == 0) {
        cmat[wid] = cv;
        ctmat[wid] = ct;
    }
}

// bbmat and bcmat as GEMMs (the same sums as k_bbmat, k_bcmat): the n^7 terms are products of n^3 x
// n^3 term matrices with atwo, atwo^T or atwo3; the O(n^6) and O(n^5) terms and the index order go to
// the combine kernels.
//   bbmat:  Cb[(ap,bp,cp),a] = sum_(ie,c,ief) eee2(cp,ap,bp,ie,c,ief) atwo(a,ie,c,ief)  (and eee2t: Cbt)
//   bcmat:  C1[(ap,b,c),a] = sum_(d,ief,ie) eee2(ap,b,d,ief,ie,c) atwo(d,ief,ie,a)
//           C2[(ap,b,a),c] = sum_(d,ief,ie) eee2(ap,b,d,ief,a,ie) atwo(d,ief,c,ie)
//           C3[(ap,a,c),b] = sum_(d,ief,ie) eee2(ap,ie,d,ief,a,c) atwo(d,ief,b,ie)   (and eee2tl: C1t..C3t)
__global__ void k_bbc_x(Rdm R, int s, int which, double *X) {
    const int n = R.n;
    const i64 n3 = (i64)n * n * n, e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n3 * n3) return;
    const int r = (int)(e % n3), cl = (int)(e / n3);
    const int r0 = r % n, r1 = (r / n) % n, r2 = r / (n * n);
    const int q0 = cl % n, q1 = (cl / n) % n, q2 = cl / (n * n);
    double v;
    switch (which) {
        case 0: v = eee2(R, s, r2, r0, r1, q0, q1, q2); break;     // rows (ap,bp,cp), columns (ie,c,ief)
        case 1: v = eee2t(R, s, r2, r0, r1, q0, q1, q2); break;
        case 2: v = eee2(R, s, r0, r1, q0, q1, q2, r2); break;     // rows (ap,b,c), columns (d,ief,ie)
        case 3: v = eee2tl(R, s, r0, r1, q0, q1, q2, r2); break;
        case 4: v = eee2(R, s, r0, r1, q0, q1, r2, q2); break;     // rows (ap,b,a), columns (d,ief,ie)
        case 5: v = eee2tl(R, s, r0, r1, q0, q1, r2, q2); break;
        case 6: v = eee2(R, s, r0, q2, q0, q1, r1, r2); break;     // rows (ap,a,c), columns (d,ief,ie)
        default: v = eee2tl(R, s, r0, q2, q0, q1, r1, r2); break;
    }
    X[e] = v;
}
__global__ void k_bb_combine(Rdm R, int s, const double *F, const double *AJ, const double *C, double *bmat,
                             double *btmat) {
    const int n = R.n, M = R.M;
    const i64 n3 = (i64)n * n * n, n4 = n3 * n, e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n4) return;
    const int ap = (int)(e % n), bp = (int)((e / n) % n), cp = (int)((e / (n * n)) % n), a = (int)(e / n3);
    double bv = 0.0, bt = 0.0;
    for (int c = 0; c < n; c++) {
        bv = bv - FF(c, a) * ee2(R, s, cp, ap, bp, c);
        bt = bt + AJ[c + n * a] * ee2tr(R, s, cp, ap, bp, c);
    }
    const i64 o = s + (i64)M * e;
    bmat[o] = bv - C[e];
    btmat[o] = bt + C[n4 + e];
}
__global__ void k_bc_combine(Rdm R, int s, const double *F, const double *atwo, const double *C, double *cmat,
                             double *ctmat) {
    const int n = R.n, M = R.M;
    const i64 n3 = (i64)n * n * n, n4 = n3 * n, e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n4) return;
    const int ap = (int)(e % n), a = (int)((e / n) % n), b = (int)((e / (n * n)) % n), c = (int)(e / n3);
    double cv = 0.0, ct = 0.0;
    for (int d = 0; d < n; d++) {
        const double fda = FF(d, a), fcd = FF(c, d), fbd = FF(b, d);
        cv = cv + fda * ee2(R, s, ap, b, d, c) - fcd * ee2(R, s, ap, b, a, d) - fbd * ee2(R, s, ap, d, a, c);
        ct = ct + fda * ee2t(R, s, ap, b, d, c) - fcd * ee2t(R, s, ap, b, a, d) + fbd * ee2t(R, s, ap, d, a, c);
        for (int ie = 0; ie < n; ie++) {   // the kernel's ief = 0 terms
            const double dint1 = AT(c, ie, d, a), dint2 = AT(d, ie, ie, a), dint3 = AT(c, d, d, ie), dint4 = AT(b, d, d, ie);
            const double dum1 = ee2(R, s, ap, b, d, ie), dum2 = ee2(R, s, ap, b, d, c);
            const double dum3 = ee2(R, s, ap, b, a, ie), dum4 = ee2(R, s, ap, ie, a, c);
            const double dum1t = ee2t(R, s, ap, b, d, ie), dum2t = ee2t(R, s, ap, b, d, c);
            const double dum3t = ee2t(R, s, ap, b, a, ie), dum4t = ee2t(R, s, ap, ie, a, c);
            cv = cv + dum1 * dint1 - 0.5 * (dum2 * dint2 + dum3 * dint3 - dum4 * dint4);
            ct = ct + dum1t * dint1 - 0.5 * (dum2t * dint2 + dum3t * dint3 - dum4t * dint4);
        }
    }
    const double *C1 = C, *C1t = C + n4, *C2 = C + 2 * n4, *C2t = C + 3 * n4, *C3 = C + 4 * n4, *C3t = C + 5 * n4;
    const i64 i1 = (ap + n * (b + (i64)n * c)) + n3 * a, i2 = (ap + n * (b + (i64)n * a)) + n3 * c;
    const i64 i3 = (ap + n * (a + (i64)n * c)) + n3 * b;
    const i64 o = s + (i64)M * e;
    cmat[o] = cv + C1[i1] - C2[i2] - C3[i3];
    ctmat[o] = ct + C1t[i1] - C2t[i2] + C3t[i3];
}
static int launch_bbc_gemm(const Rdm &R, const double *atwo, const double *Fone, const double *Fhalf, const double *AJ,
                           rocblas_handle h, double *work, Stage1Out &o, hipStream_t st) {
    const int n = R.n;
    const i64 n3 = (i64)n * n * n, n4 = n3 * n, n6 = n3 * n3;
    double *X = work, *AT3 = work + n6, *C = AT3 + n4;
    if (rocblas_set_stream(h, st) != rocblas_status_success) { set_error("rocblas_set_stream failed"); return -1; }
    hipLaunchKernelGGL(k_k0p_at3, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, n, atwo, AT3);
    NV2_LAUNCH_CHECK();
    const double one = 1.0, zero = 0.0;
    const rocblas_operation Nn = rocblas_operation_none, Tt = rocblas_operation_transpose;
    const int m3 = (int)n3;
    const unsigned g6 = (unsigned)((n6 + 255) / 256), g4 = (unsigned)((n4 + 255) / 256);
    for (int s = 0; s < R.M; s++) {
        for (int which = 0; which < 8; which++) {
            hipLaunchKernelGGL(k_bbc_x, dim3(g6), dim3(256), 0, st, R, s, which, X);
            NV2_LAUNCH_CHECK();
            rocblas_status rs;
            if (which < 2)            // B[(ie,c,ief), a] = atwo[a + n (ie,c,ief)]
                rs = rocblas_dgemm(h, Nn, Tt, m3, n, m3, &one, X, m3, atwo, n, &zero, C + which * n4, m3);
            else                      // bcmat: C1 with atwo, C2 and C3 with atwo3
                rs = rocblas_dgemm(h, Nn, Nn, m3, n, m3, &one, X, m3, which < 4 ? atwo : AT3, m3, &zero,
                                   C + (which - 2) * n4, m3);   // bbmat's two products are consumed by then
            if (rs != rocblas_status_success) { set_error("bbmat/bcmat GEMM failed (%d)", (int)rs); return -1; }
            if (which == 1) {
                hipLaunchKernelGGL(k_bb_combine, dim3(g4), dim3(256), 0, st, R, s, Fone, AJ, C, o.bmat, o.btmat);
                NV2_LAUNCH_CHECK();
            }
        }
        hipLaunchKernelGGL(k_bc_combine, dim3(g4), dim3(256), 0, st, R, s, Fhalf, atwo, C, o.cmat, o.ctmat);
        NV2_LAUNCH_CHECK();
    }
    return 0;
}

// ---------------------------------------------------------------- bdmat (koopro4QD.F90:4399-4459); F alpha=1
__global__ void k_bdmat(Rdm R, const double *atwo, const double *F, const double *AJ, double *dmat, double *dtmat) {
    const int n = R.n, M = R.M;
    WAVE_SETUP((i64)M * n * n);
    int s, x[2];
    decode(wid, M, n, 2, &s, x);
    const int ap = x[0], a = x[1];
    double dv = 0.0, dt = 0.0;
    const int n3 = n * n * n;
    for (int t = lane; t < n3; t += 64) {
        const int c = t / (n * n), ie = (t / n) % n, ief = t % n;
        if (ie == 0 && ief == 0) {
            dv = dv - FF(c, a) * D1(s, ap, c);
            double dum = -D1(s, ap, c);
            if (ap == c) dum = dum + 2.0;
            dt = dt + AJ[c + n * a] * dum;
        }
        double dint = AT(a, ie, c, ief);
        dv = dv - ee2(R, s, ap, ie, c, ief) * dint;
        dt = dt + ee2t(R, s, ap, ie, c, ief) * dint;
    }
    dv = wave_sum(dv);
    dt = wave_sum(dt);
    if (lane == 0) {
        dmat[wid] = dv;
        dtmat[wid] = dt;
    }
}

// ---------------------------------------------------------------- bamat: E4 = <E E E E> for one state
// RDM reads with the index arithmetic in 32 bits for a compile-time n (n^6 < 2^31 for n <= 16); the
// state offset stays 64-bit.
template <int NN> __device__ __forceinline__ double d2n(const Rdm &R, int s, int p, int q, int r, int t) {
    return R.d2[s + (i64)R.M * (unsigned)(p + NN * (q + NN * (r + NN * t)))];
}
template <int NN> __device__ __forceinline__ double d1n(const Rdm &R, int s, int p, int q) {
    return R.d1[s + (i64)R.M * (unsigned)(p + NN * q)];
}
// lindice of a descending tuple (indices_norm.F90:41-44), 0-based, in 32 bits
__device__ __forceinline__ unsigned lind32(int a, int b, int c, int d) {
    const unsigned A = a, B = b, C = c;
    return A * (A + 1) * (A + 2) * (A + 3) / 24u + B * (B + 1) * (B + 2) / 6u + C * (C + 1) / 2u + (unsigned)d;
}
// g4 (nv2_common.h) with each side sorted once: ord8 sorts the side whose lindice is larger (the creation
// side on a tie) and reads ro4 at that word with the other side co-permuted (ord_utils.F90:71-129).  Both
// sides are sorted by the same selection sort, and lindice order on descending tuples is their
// lexicographic order, so the entry read is the one g4 reads.
template <int NN>
__device__ __forceinline__ double g4n(const Rdm &R, int s, int a, int c, int e, int g, int b, int d, int f, int h) {
    int p[4] = {a, c, e, g}, q[4] = {b, d, f, h}, P[4] = {b, d, f, h}, Q[4] = {a, c, e, g};
    sel_sort4(p, q);
    sel_sort4(P, Q);
    const bool cw = p[0] != P[0] ? p[0] > P[0] : p[1] != P[1] ? p[1] > P[1] : p[2] != P[2] ? p[2] > P[2] : p[3] >= P[3];
    int u[4], v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        u[k] = cw ? p[k] : P[k];
        v[k] = cw ? q[k] : Q[k];
    }
    const unsigned off = lind32(u[0], u[1], u[2], u[3]) +
                         (unsigned)R.nw * (unsigned)(v[0] + NN * (v[1] + NN * (v[2] + NN * v[3])));
    return R.ro4[off + (i64)s * R.nw * (NN * NN * NN * NN)];
}
// D3T[q1 + n(p1 + n(p2 + n(p3 + n(q2 + n q3))))] = D3(s; p1,p2,p3, q1,q2,q3): the 3-RDM of state s with
// q1 fastest, so a read whose first pair is (g, h) runs contiguously along an E4 row (x7, x8).
template <int NN>
__global__ __launch_bounds__(256) void k_e4_d3t(Rdm R, int s, double *D3T) {
    constexpr int N6 = NN * NN * NN * NN * NN * NN;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= N6) return;
    const int q1 = e % NN, r = e / NN, p1 = r % NN, r2 = r / NN;
    const int p2 = r2 % NN, p3 = (r2 / NN) % NN, q23 = r2 / (NN * NN);
    D3T[e] = R.d3[s + (i64)R.M * (unsigned)(p1 + NN * (p2 + NN * (p3 + NN * (q1 + NN * q23))))];
}
template <int NN> __device__ __forceinline__ double d3t(const double *D3T, int p1, int p2, int p3, int q1, int q2, int q3) {
    return D3T[q1 + NN * (p1 + NN * (p2 + NN * (p3 + NN * (q2 + NN * q3))))];
}
__device__ __forceinline__ unsigned tet3(unsigned x) { return x * (x + 1) * (x + 2) / 6u; }
__device__ __forceinline__ unsigned tri2(unsigned x) { return x * (x + 1) / 2u; }

// E4 is stored C-ordered over (x1..x8), x8 fastest, i.e. over the pairs P1 = (x1,x2), .., P4 = (x7,x8)
// (P = creation * n + annihilation): offset ((P1 N + P2) N + P3) N + P4, N = n^2.  E4 is Gamma4 plus
// delta terms in D3/D2/D1 (eexx_functions.F90 eeee).  Gamma4 is symmetric under permutations of the
// four pairs, so a workgroup takes E4B_MS multisets P0 >= P1 >= P2 (ranked tet3(P0) + tri2(P1) + P2),
// gathers Gamma4(P0, P1, P2, P4) once for all N values of P4 into LDS, and writes the rows
// (X0, X1, X2, :) of every distinct arrangement of the three pairs: up to 6 contiguous rows per
// gather of a row, each with its own delta terms.  The D3 terms are read at the pair-permuted position
// whose first pair holds h (the 3-RDM is symmetric under pair permutations) from the q1-fastest copy D3T.
// Gamma4 is read at one representative per pair-permutation orbit; the RDMs' symmetry makes it the
// value g4 reads at every position, up to the rounding of the stored entries.  Workgroups are dealt to
// the 8 XCDs round-robin, so the rank is remapped to keep neighbouring multisets (which share Gamma4
// lines) on one XCD.  Measured at n = 14 (FP64): 7.3 ms, against 22.4 ms for one gather per element.
#define E4B_MS 4
template <typename T, int NN>
__global__ __launch_bounds__(256) void k_build_e4(Rdm R, int s, const double *__restrict__ D3T, T *__restrict__ E4,
                                                  unsigned G, unsigned NMS, unsigned B0) {
    constexpr int NP = NN * NN;
    __shared__ double L[E4B_MS][NP];
    __shared__ int X[E4B_MS][6][4];   // pairs at positions 0..2 of arrangement k, and whether it is written
    const unsigned G8 = G / 8, i0 = B0 + blockIdx.x;
    const unsigned blk = i0 < 8 * G8 ? (i0 % 8) * G8 + i0 / 8 : i0;
    if (threadIdx.x < E4B_MS) {
        const int m = threadIdx.x;
        unsigned w = blk * E4B_MS + m;
        const bool ok = w < NMS;
        int P0 = 0, P1 = 0, P2 = 0;
        if (ok) {
            int x = (int)cbrt(6.0 * w);
            while (x > 0 && tet3(x) > w) x--;
            while (tet3(x + 1) <= w) x++;
            w -= tet3(x);
            int y = (int)sqrt(2.0 * w);
            while (y > 0 && tri2(y) > w) y--;
            while (tri2(y + 1) <= w) y++;
            P0 = x; P1 = y; P2 = (int)(w - tri2(y));
        }
        const int Pc[3] = {P0, P1, P2};
        for (int k = 0; k < 6; k++) {   // slot s0 at position 0, the other two in order (k even) or swapped
            const int s0 = k >> 1, o1 = s0 == 0 ? 1 : 0, o2 = s0 == 2 ? 1 : 2;
            const int s1 = (k & 1) ? o2 : o1, s2 = (k & 1) ? o1 : o2;
            int pos[3];
            pos[s0] = 0; pos[s1] = 1; pos[s2] = 2;
            // equal pairs give equal rows: keep the arrangement that leaves them in order
            const bool keep = ok && !((P0 == P1 && pos[0] > pos[1]) || (P1 == P2 && pos[1] > pos[2]));
            X[m][k][0] = Pc[s0]; X[m][k][1] = Pc[s1]; X[m][k][2] = Pc[s2]; X[m][k][3] = keep;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < E4B_MS * NP; e += 256) {
        const int m = e / NP, p4 = e - m * NP;
        if (!X[m][0][3]) continue;
        const int P0 = X[m][0][0], P1 = X[m][0][1], P2 = X[m][0][2];
        L[m][p4] = g4n<NN>(R, s, P0 / NN, P1 / NN, P2 / NN, p4 / NN, P0 % NN, P1 % NN, P2 % NN, p4 % NN);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < E4B_MS * 6 * NP; e += 256) {
        const int mk = e / NP, p4 = e - mk * NP, m = mk / 6, k = mk - 6 * m;
        if (!X[m][k][3]) continue;
        const int X0 = X[m][k][0], X1 = X[m][k][1], X2 = X[m][k][2];
        const int a = X0 / NN, b = X0 % NN, c = X1 / NN, d = X1 % NN, ee = X2 / NN, f = X2 % NN, g = p4 / NN, h = p4 % NN;
        double v = L[m][p4];
        if (b == c) v += d3t<NN>(D3T, g, a, ee, h, d, f);
        if (b == ee) v += d3t<NN>(D3T, g, a, c, h, f, d);
        if (b == g) v += d3t<NN>(D3T, a, c, ee, h, d, f);
        if (d == ee) v += d3t<NN>(D3T, g, a, c, h, b, f);
        if (d == g) v += d3t<NN>(D3T, c, a, ee, h, b, f);
        if (f == g) v += d3t<NN>(D3T, ee, a, c, h, b, d);
        if (b == c) {
            if (d == ee) v += d2n<NN>(R, s, a, g, f, h);
            if (d == g) v += d2n<NN>(R, s, a, ee, h, f);
            if (f == g) v += d2n<NN>(R, s, a, ee, d, h);
        }
        if (b == ee && d == g) v += d2n<NN>(R, s, a, c, f, h);
        if (b == ee && f == g) v += d2n<NN>(R, s, a, c, h, d);
        if (b == g && d == ee) v += d2n<NN>(R, s, a, c, h, f);
        if (d == ee && f == g) v += d2n<NN>(R, s, a, c, b, h);
        if (b == c && d == ee && f == g) v += d1n<NN>(R, s, a, h);
        E4[(((i64)X0 * NP + X1) * NP + X2) * NP + p4] = (T)v;
    }
}

static inline unsigned nblk(i64 tot, int bs) { return (unsigned)((tot + bs - 1) / bs); }

int launch_bro(const Rdm &R, int nele, double *daaa, double *taa, double *dal, hipStream_t st) {
    const i64 n = R.n, M = R.M;
    hipLaunchKernelGGL(k_bro3, dim3(nblk(M * n * n * n * n * n * n, 256)), dim3(256), 0, st, R, nele, daaa);
    NV2_LAUNCH_CHECK();
    Rdm R3 = R; R3.d3 = daaa;
    hipLaunchKernelGGL(k_bro2, dim3(nblk(M * n * n * n * n, 256)), dim3(256), 0, st, R3, nele, taa);
    NV2_LAUNCH_CHECK();
    Rdm R2 = R3; R2.d2 = taa;
    hipLaunchKernelGGL(k_bro1, dim3(nblk(M * n * n, 256)), dim3(256), 0, st, R2, nele, dal);
    NV2_LAUNCH_CHECK();
    return 0;
}

int launch_bro21(const Rdm &R, int nele, double *taa, double *dal, hipStream_t st) {
    const i64 n = R.n, M = R.M;
    hipLaunchKernelGGL(k_bro2, dim3(nblk(M * n * n * n * n, 256)), dim3(256), 0, st, R, nele, taa);
    NV2_LAUNCH_CHECK();
    Rdm R2 = R; R2.d2 = taa;
    hipLaunchKernelGGL(k_bro1, dim3(nblk(M * n * n, 256)), dim3(256), 0, st, R2, nele, dal);
    NV2_LAUNCH_CHECK();
    return 0;
}

int launch_aj_atwo(int n, int ncore, int norb, int L, const double *B, const double *oneint, double *aj,
                   double *atwo, hipStream_t st) {
    NV2_HIP(hipMemcpyAsync(aj, oneint, sizeof(double) * (i64)norb * (norb + 1) / 2, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_aj, dim3(nblk((i64)n * n * 64, 256)), dim3(256), 0, st, n, ncore, L, B, aj);
    NV2_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_atwo, dim3(nblk((i64)n * n * n * n, 256)), dim3(256), 0, st, n, ncore, L, B, atwo);
    NV2_LAUNCH_CHECK();
    return 0;
}

int launch_make_f(int n, int ncore, const double *aj, const double *atwo, double alpha, double *F, hipStream_t st) {
    hipLaunchKernelGGL(k_make_f, dim3(n), dim3(64), 0, st, n, ncore, aj, atwo, alpha, F);
    NV2_LAUNCH_CHECK();
    return 0;
}

int launch_koop_prep(int n, int ncore, const double *aj, const double *atwo, Stage1Scratch &w, hipStream_t st) {
    hipLaunchKernelGGL(k_aj_active, dim3(n), dim3(64), 0, st, n, ncore, aj, w.AJ);
    NV2_LAUNCH_CHECK();
    if (launch_make_f(n, ncore, aj, atwo, 0.5, w.Fhalf, st)) return -1;
    if (launch_make_f(n, ncore, aj, atwo, 1.0, w.Fone, st)) return -1;
    return 0;
}

int launch_koop_rest(const Rdm &R, const double *atwo, Stage1Scratch &w, Stage1Out &o, hipStream_t st) {
    const i64 n = R.n, M = R.M;
    const i64 t2 = M * n * n, t4 = M * n * n * n * n;
    // koopE with f = aj
    hipLaunchKernelGGL(k_koopE, dim3(nblk(t2 * 64, 256)), dim3(256), 0, st, R, atwo, w.AJ, o.ipko1, o.eako1);
    NV2_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_hrdm, dim3(nblk(t4, 256)), dim3(256), 0, st, R, o.hrdm);
    NV2_LAUNCH_CHECK();
    if (w.h && w.k0p) {
        if (launch_koop2e_gemm(R, atwo, w.Fhalf, w.h, w.k0p, o.ipko2, o.eako2, st)) return -1;
    } else {
        hipLaunchKernelGGL(k_koop2E, dim3(nblk(t4 * 64, 256)), dim3(256), 0, st, R, atwo, w.Fhalf, o.ipko2, o.eako2);
        NV2_LAUNCH_CHECK();
    }
    if (w.h && w.k0p) {
        if (launch_koop0p_gemm(R, atwo, w.Fhalf, w.h, w.k0p, o.gk0pa, o.gk0pd, st)) return -1;
    } else {
        hipLaunchKernelGGL(k_koop0p, dim3(nblk(t4 * 64, 256)), dim3(256), 0, st, R, atwo, w.Fhalf, o.gk0pa, o.gk0pd);
        NV2_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_koop0pf, dim3(nblk(t2 * 64, 256)), dim3(256), 0, st, R, atwo, w.Fhalf, o.gk0pf);
    NV2_LAUNCH_CHECK();
    if (o.bmat && w.h && w.k0p) {
        if (launch_bbc_gemm(R, atwo, w.Fone, w.Fhalf, w.AJ, w.h, w.k0p, o, st)) return -1;
        hipLaunchKernelGGL(k_bdmat, dim3(nblk(t2 * 64, 256)), dim3(256), 0, st, R, atwo, w.Fone, w.AJ, o.dmat, o.dtmat);
        NV2_LAUNCH_CHECK();
    } else if (o.bmat) {
        hipLaunchKernelGGL(k_bbmat, dim3(nblk(t4 * 64, 256)), dim3(256), 0, st, R, atwo, w.Fone, w.AJ, o.bmat, o.btmat);
        NV2_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_bcmat, dim3(nblk(t4 * 64, 256)), dim3(256), 0, st, R, atwo, w.Fhalf, o.cmat, o.ctmat);
        NV2_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_bdmat, dim3(nblk(t2 * 64, 256)), dim3(256), 0, st, R, atwo, w.Fone, w.AJ, o.dmat, o.dtmat);
        NV2_LAUNCH_CHECK();
    }
    return 0;
}

int launch_koop_small(const Rdm &R, int ncore, const double *aj, const double *atwo, Stage1Scratch &w,
                      Stage1Out &o, hipStream_t st) {
    if (launch_koop_prep(R.n, ncore, aj, atwo, w, st)) return -1;
    return launch_koop_rest(R, atwo, w, o, st);
}

// Launches the big stage-1 kernels are split into (NV2_E4_CHUNKS for the E4 build, NV2_PASS_CHUNKS for
// each contraction pass; default 1): between two launches the dispatcher serves the other streams'
// pending kernels
static int env_chunks(const char *name, int dflt) {
    const char *e = getenv(name);
    const int c = e ? atoi(e) : dflt;
    return c < 1 ? 1 : (c > 64 ? 64 : c);
}
int stage1_chunks() { return env_chunks("NV2_PASS_CHUNKS", 1); }
// The E4 build goes out as 32 launches (about 10 000 workgroups each at n = 14): between two launches
// the (+-1)' chain's pending kernels, which a single launch of 318 000 workgroups held back for its
// whole duration, get the CUs (CAS(14,14): 64.2 against 64.9 ms per evaluation).
static int e4_chunks() { return env_chunks("NV2_E4_CHUNKS", 32); }
template <typename T>
static int build_e4_n(const Rdm &R, int s, T *E4, double *d3t, hipStream_t st) {
    const int nch = e4_chunks();
    switch (R.n) {
#define CASE(k)                                                                                                 \
    case k: {                                                                                                   \
        constexpr unsigned np = k * k, nms = np * (np + 1) * (np + 2) / 6, g = (nms + E4B_MS - 1) / E4B_MS;    \
        const unsigned chunk = ((g + nch - 1) / nch + 7) / 8 * 8;   /* whole XCD rounds per launch */       \
        hipLaunchKernelGGL((k_e4_d3t<k>), dim3((k * k * k * k * k * k + 255) / 256), dim3(256), 0, st, R, s, d3t);    \
        NV2_LAUNCH_CHECK();                                                                                     \
        for (unsigned b0 = 0; b0 < g; b0 += chunk) {                                                            \
            const unsigned gc = g - b0 < chunk ? g - b0 : chunk;                                               \
            hipLaunchKernelGGL((k_build_e4<T, k>), dim3(gc), dim3(256), 0, st, R, s, d3t, E4, g, nms, b0);     \
        }                                                                                                       \
        break;                                                                                                  \
    }
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12)
        CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
        default: set_error("E4 build supports 1 <= nact <= 16 (got %d)", R.n); return -2;
    }
    NV2_LAUNCH_CHECK();
    return 0;
}

int launch_build_e4(const Rdm &R, int s, void *E4, int prec, double *d3t, hipStream_t st) {
    if (!d3t) { set_error("E4 build: no n^6 scratch for the transposed 3-RDM"); return -1; }
    if (prec == 0) return build_e4_n<double>(R, s, (double *)E4, d3t, st);
    return build_e4_n<float>(R, s, (float *)E4, d3t, st);
}

}  // namespace nv2

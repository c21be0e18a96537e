/*
 * Oracle: CPU restatement of NEVPT2 stage 1 (Koopmans-matrix builds).
 *
 * TEST INFRASTRUCTURE ONLY -- loaded by tests/ and by the cpu_baseline leg of
 * bench.py as the checker / timed CPU baseline; never part of the product.
 *
 * Each routine follows the reference loop nest line by line (same loop order,
 * same accumulation order per output element), with OpenMP over loop indices
 * whose outputs are disjoint instead of the reference's ATOMIC updates.
 * Arrays use the reference layouts: Fortran column-major, 0-based here, state
 * index first (d1[s,p,q], d2[s,p,q,r,t], d3[s,...6]); ro4[word,i5,i6,i7,i8,s];
 * one-electron quantities packed over the norb triangle (tri(p,q)).
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

typedef int64_t i64;

typedef struct {
    i64 n, M, nw;
    const double *d3, *d2, *d1, *ro4;
} rdm_t;

static inline i64 ix2(const rdm_t *R, i64 s, i64 p, i64 q) { return s + R->M * (p + R->n * q); }
static inline i64 ix4(const rdm_t *R, i64 s, i64 p, i64 q, i64 r, i64 t) {
    i64 n = R->n; return s + R->M * (p + n * (q + n * (r + n * t)));
}
static inline i64 ix6(const rdm_t *R, i64 s, i64 p, i64 q, i64 r, i64 t, i64 u, i64 v) {
    i64 n = R->n; return s + R->M * (p + n * (q + n * (r + n * (t + n * (u + n * v)))));
}
#define D1(R, s, p, q) ((R)->d1[ix2(R, s, p, q)])
#define D2(R, s, p, q, r, t) ((R)->d2[ix4(R, s, p, q, r, t)])
#define D3(R, s, p, q, r, t, u, v) ((R)->d3[ix6(R, s, p, q, r, t, u, v)])

static inline i64 tri(i64 p, i64 q) { i64 hi = p > q ? p : q, lo = p > q ? q : p; return hi * (hi + 1) / 2 + lo; }

/* ---------------- packed 4-RDM gather (indices_norm.F90:41-106, ord_utils.F90:71-129) */
static inline i64 lind0(i64 a, i64 b, i64 c, i64 d) {
    a++; b++; c++; d++;
    return (a - 1) * a * (a + 1) * (a + 2) / 24 + (b - 1) * b * (b + 1) / 6 + (c - 1) * c / 2 + d - 1;
}
static inline i64 norm4(const i64 *t) {
    i64 v[4] = {t[0], t[1], t[2], t[3]};
    for (int i = 0; i < 4; i++) {
        int jm = i; i64 im = v[i];
        for (int j = i + 1; j < 4; j++) if (v[j] > im) { im = v[j]; jm = j; }
        if (jm != i) { i64 x = v[i]; v[i] = v[jm]; v[jm] = x; }
    }
    return lind0(v[0], v[1], v[2], v[3]);
}
static inline double g4(const rdm_t *R, i64 s, i64 a, i64 c, i64 e, i64 g, i64 b, i64 d, i64 f, i64 h) {
    i64 c4[4] = {a, c, e, g}, a4[4] = {b, d, f, h};
    i64 nv1[4], nv2[4];
    if (norm4(c4) >= norm4(a4)) { memcpy(nv1, c4, sizeof nv1); memcpy(nv2, a4, sizeof nv2); }
    else { memcpy(nv1, a4, sizeof nv1); memcpy(nv2, c4, sizeof nv2); }
    for (int i = 0; i < 4; i++) {
        i64 im = nv1[i]; int jm = i;
        for (int j = i + 1; j < 4; j++) if (nv1[j] > im) { im = nv1[j]; jm = j; }
        if (jm != i) {
            i64 x = nv1[i]; nv1[i] = nv1[jm]; nv1[jm] = x;
            x = nv2[i]; nv2[i] = nv2[jm]; nv2[jm] = x;
        }
    }
    i64 n = R->n, w = lind0(nv1[0], nv1[1], nv1[2], nv1[3]);
    return R->ro4[w + R->nw * (nv2[0] + n * (nv2[1] + n * (nv2[2] + n * (nv2[3] + n * s))))];
}

/* ---------------- accessors (eexx_functions.F90) */
static inline double ee2(const rdm_t *R, i64 s, i64 a, i64 b, i64 c, i64 d) {   /* :125-135 */
    double v = D2(R, s, a, c, b, d);
    if (b == c) v += D1(R, s, a, d);
    return v;
}
static inline double ee2t(const rdm_t *R, i64 s, i64 a, i64 b, i64 c, i64 d) {  /* :138-148 */
    double v = -ee2(R, s, b, a, c, d);
    if (a == b) v += 2.0 * D1(R, s, c, d);
    return v;
}
static inline double ee2tr(const rdm_t *R, i64 s, i64 a, i64 b, i64 c, i64 d) { /* :151-161 */
    double v = -ee2(R, s, a, b, d, c);
    if (c == d) v += 2.0 * D1(R, s, a, b);
    return v;
}
static inline double eee2(const rdm_t *R, i64 s, i64 a, i64 b, i64 c, i64 d, i64 e, i64 f) { /* :79-94 */
    double v = D3(R, s, a, c, e, b, d, f);
    if (e == d) v += D2(R, s, a, c, b, f);
    if (e == b) v += D2(R, s, a, c, f, d);
    if (b == c) v += D2(R, s, a, e, d, f);
    if (b == c && e == d) v += D1(R, s, a, f);
    return v;
}
static inline double eee2t(const rdm_t *R, i64 s, i64 a, i64 b, i64 c, i64 d, i64 e, i64 f) { /* :97-108 */
    double v = -eee2(R, s, a, b, d, c, e, f);
    if (c == d) v += 2.0 * ee2(R, s, a, b, e, f);
    return v;
}
static inline double eee2tl(const rdm_t *R, i64 s, i64 a, i64 b, i64 c, i64 d, i64 e, i64 f) { /* :111-122 */
    double v = -eee2(R, s, b, a, c, d, e, f);
    if (a == b) v += 2.0 * ee2(R, s, c, d, e, f);
    return v;
}
static inline double eeee(const rdm_t *R, i64 s, i64 a, i64 b, i64 c, i64 d, i64 e, i64 f, i64 g, i64 h) { /* :25-61 */
    double v = g4(R, s, a, c, e, g, b, d, f, h);
    if (b == c) v += D3(R, s, a, e, g, d, f, h);
    if (b == e) v += D3(R, s, a, c, g, f, d, h);
    if (b == g) v += D3(R, s, a, c, e, h, d, f);
    if (d == e) v += D3(R, s, a, g, c, b, h, f);
    if (d == g) v += D3(R, s, a, c, e, b, h, f);
    if (f == g) v += D3(R, s, a, e, c, b, h, d);
    if (b == c) {
        if (d == e) v += D2(R, s, a, g, f, h);
        if (d == g) v += D2(R, s, a, e, h, f);
        if (f == g) v += D2(R, s, a, e, d, h);
    }
    if (b == e && d == g) v += D2(R, s, a, c, f, h);
    if (b == e && f == g) v += D2(R, s, a, c, h, d);
    if (b == g && d == e) v += D2(R, s, a, c, h, f);
    if (d == e && f == g) v += D2(R, s, a, c, b, h);
    if (b == c && d == e && f == g) v += D1(R, s, a, h);
    return v;
}
static inline double eeeet(const rdm_t *R, i64 s, i64 a, i64 b, i64 c, i64 d, i64 e, i64 f, i64 g, i64 h) { /* :64-76 */
    double v = -eeee(R, s, a, b, d, c, e, f, g, h);
    if (c == d) v += 2.0 * eee2(R, s, a, b, e, f, g, h);
    return v;
}
static inline double ro2t(const rdm_t *R, i64 s, i64 ap, i64 bp, i64 a, i64 b) {  /* :210-225 */
    double v = D2(R, s, a, b, ap, bp);
    if (ap == a) v -= 2.0 * D1(R, s, b, bp);
    if (bp == b) v -= 2.0 * D1(R, s, a, ap);
    if (ap == b) v += D1(R, s, a, bp);
    if (bp == a) v += D1(R, s, b, ap);
    if (ap == a && bp == b) v += 4.0;
    if (ap == b && bp == a) v -= 2.0;
    return v;
}
static double ro3t(const rdm_t *R, i64 s, i64 a, i64 b, i64 c, i64 ap, i64 bp, i64 cp) { /* :164-207 */
    double v = -D3(R, s, ap, bp, cp, a, b, c);
    if (ap == a) v += 2.0 * D2(R, s, bp, cp, b, c);
    if (bp == b) v += 2.0 * D2(R, s, ap, cp, a, c);
    if (cp == c) v += 2.0 * D2(R, s, ap, bp, a, b);
    if (bp == c) v -= D2(R, s, ap, cp, a, b);
    if (ap == c) v -= D2(R, s, cp, bp, a, b);
    if (cp == b) v -= D2(R, s, bp, ap, c, a);
    if (ap == b) v -= D2(R, s, bp, cp, a, c);
    if (cp == a) v -= D2(R, s, ap, bp, c, b);
    if (bp == a) v -= D2(R, s, ap, cp, b, c);
    if (cp == c && bp == b) v -= 4.0 * D1(R, s, ap, a);
    if (cp == b && bp == c) v += 2.0 * D1(R, s, ap, a);
    if (cp == c && ap == a) v -= 4.0 * D1(R, s, bp, b);
    if (cp == a && ap == c) v += 2.0 * D1(R, s, bp, b);
    if (ap == a && bp == b) v -= 4.0 * D1(R, s, cp, c);
    if (ap == b && bp == a) v += 2.0 * D1(R, s, cp, c);
    if (cp == c && ap == b) v += 2.0 * D1(R, s, bp, a);
    if (cp == b && ap == c) v -= D1(R, s, bp, a);
    if (cp == c && bp == a) v += 2.0 * D1(R, s, ap, b);
    if (cp == a && bp == c) v -= D1(R, s, ap, b);
    if (bp == b && ap == c) v += 2.0 * D1(R, s, cp, a);
    if (bp == c && ap == b) v -= D1(R, s, cp, a);
    if (bp == b && cp == a) v += 2.0 * D1(R, s, ap, c);
    if (bp == a && cp == b) v -= D1(R, s, ap, c);
    if (ap == a && cp == b) v += 2.0 * D1(R, s, bp, c);
    if (ap == b && cp == a) v -= D1(R, s, bp, c);
    if (ap == a && bp == c) v += 2.0 * D1(R, s, cp, b);
    if (ap == c && bp == a) v -= D1(R, s, cp, b);
    if (ap == a && bp == b && cp == c) v += 8.0;
    if (ap == b && bp == a && cp == c) v -= 4.0;
    if (ap == a && bp == c && cp == b) v -= 4.0;
    if (ap == c && bp == b && cp == a) v -= 4.0;
    if (ap == b && bp == c && cp == a) v += 2.0;
    if (ap == c && bp == a && cp == b) v += 2.0;
    return v;
}

/* exported scalar accessors (for pinning against the reference's own modules) */
#define MKR(n_, M_, d3_, d2_, d1_, ro4_) rdm_t R = {n_, M_, (n_) * ((n_) + 1) * ((n_) + 2) * ((n_) + 3) / 24, d3_, d2_, d1_, ro4_}
double orc_eeee(i64 n, i64 M, const double *ro4, const double *d3, const double *d2, const double *d1, i64 s, const i64 *ix) {
    MKR(n, M, d3, d2, d1, ro4);
    return eeee(&R, s, ix[0], ix[1], ix[2], ix[3], ix[4], ix[5], ix[6], ix[7]);
}
double orc_eeeet(i64 n, i64 M, const double *ro4, const double *d3, const double *d2, const double *d1, i64 s, const i64 *ix) {
    MKR(n, M, d3, d2, d1, ro4);
    return eeeet(&R, s, ix[0], ix[1], ix[2], ix[3], ix[4], ix[5], ix[6], ix[7]);
}
double orc_eee2(i64 n, i64 M, const double *d3, const double *d2, const double *d1, i64 s, const i64 *ix) {
    MKR(n, M, d3, d2, d1, NULL);
    return eee2(&R, s, ix[0], ix[1], ix[2], ix[3], ix[4], ix[5]);
}
double orc_eee2t(i64 n, i64 M, const double *d3, const double *d2, const double *d1, i64 s, const i64 *ix) {
    MKR(n, M, d3, d2, d1, NULL);
    return eee2t(&R, s, ix[0], ix[1], ix[2], ix[3], ix[4], ix[5]);
}
double orc_eee2tl(i64 n, i64 M, const double *d3, const double *d2, const double *d1, i64 s, const i64 *ix) {
    MKR(n, M, d3, d2, d1, NULL);
    return eee2tl(&R, s, ix[0], ix[1], ix[2], ix[3], ix[4], ix[5]);
}
double orc_ro3t(i64 n, i64 M, const double *d3, const double *d2, const double *d1, i64 s, const i64 *ix) {
    MKR(n, M, d3, d2, d1, NULL);
    return ro3t(&R, s, ix[0], ix[1], ix[2], ix[3], ix[4], ix[5]);
}

/* ---------------- partial traces (koopro4QD.F90:4881-5049) */
void orc_bro3(i64 n, i64 M, i64 nele, const double *ro4, double *daaa) {
    MKR(n, M, NULL, NULL, NULL, ro4);
    i64 n6 = n * n * n * n * n * n;
    memset(daaa, 0, sizeof(double) * M * n6);
#pragma omp parallel for collapse(3)
    for (i64 i = 0; i < n; i++)
        for (i64 j = 0; j < n; j++)
            for (i64 k = 0; k < n; k++)
                for (i64 ii = 0; ii < n; ii++)
                    for (i64 jj = 0; jj < n; jj++)
                        for (i64 kk = 0; kk < n; kk++) {
                            i64 o = M * (i + n * (j + n * (k + n * (ii + n * (jj + n * kk)))));
                            for (i64 ll = 0; ll < n; ll++)
                                for (i64 s = 0; s < M; s++) daaa[o + s] += g4(&R, s, i, j, k, ll, ii, jj, kk, ll);
                            for (i64 s = 0; s < M; s++) daaa[o + s] /= (double)(nele - 3);
                        }
}
void orc_bro2(i64 n, i64 M, i64 nele, const double *daaa, double *taa) {
    MKR(n, M, daaa, NULL, NULL, NULL);
    memset(taa, 0, sizeof(double) * M * n * n * n * n);
    for (i64 i = 0; i < n; i++)
        for (i64 j = 0; j < n; j++)
            for (i64 ii = 0; ii < n; ii++)
                for (i64 jj = 0; jj < n; jj++) {
                    i64 o = M * (i + n * (j + n * (ii + n * jj)));
                    for (i64 ll = 0; ll < n; ll++)
                        for (i64 s = 0; s < M; s++) taa[o + s] += D3(&R, s, i, j, ll, ii, jj, ll);
                    for (i64 s = 0; s < M; s++) taa[o + s] /= (double)(nele - 2);
                }
}
void orc_bro1(i64 n, i64 M, i64 nele, const double *taa, double *dal) {
    MKR(n, M, NULL, taa, NULL, NULL);
    memset(dal, 0, sizeof(double) * M * n * n);
    for (i64 i = 0; i < n; i++)
        for (i64 j = 0; j < n; j++) {
            i64 o = M * (i + n * j);
            for (i64 ll = 0; ll < n; ll++)
                for (i64 s = 0; s < M; s++) dal[o + s] += D2(&R, s, i, ll, j, ll);
            for (i64 s = 0; s < M; s++) dal[o + s] /= (double)(nele - 1);
        }
}

/* ---------------- Dyall h and active ERIs (koopro4QD.F90:3900-3971) */
void orc_get_aj_atwo(i64 ncore, i64 nact, i64 norb, i64 L, const double *B, const double *oneint,
                     double *aj, double *atwo) {
    i64 np = norb * (norb + 1) / 2;
    memcpy(aj, oneint, sizeof(double) * np);
#define AI(p, q, r, t) ai_chol(B, L, p, q, r, t)
    for (i64 a = 0; a < nact; a++)
        for (i64 b = a; b < nact; b++) {
            i64 ac = a + ncore, bc = b + ncore, jab = tri(ac, bc);
            for (i64 j = 0; j < ncore; j++) {
                double x = 0, y = 0;
                const double *u = B + L * tri(ac, bc), *v = B + L * tri(j, j);
                for (i64 l = 0; l < L; l++) x += u[l] * v[l];
                u = B + L * tri(ac, j); v = B + L * tri(j, bc);
                for (i64 l = 0; l < L; l++) y += u[l] * v[l];
                aj[jab] = aj[jab] + 2.0 * x - y;
            }
        }
    for (i64 i = 0; i < nact; i++)
        for (i64 j = 0; j < nact; j++)
            for (i64 k = 0; k < nact; k++)
                for (i64 l = 0; l < nact; l++) {
                    const double *u = B + L * tri(i + ncore, j + ncore), *v = B + L * tri(k + ncore, l + ncore);
                    double x = 0;
                    for (i64 m = 0; m < L; m++) x += u[m] * v[m];
                    atwo[i + nact * (j + nact * (k + nact * l))] = x;
                }
#undef AI
}

#define AT(p, q, r, t) atwo[(p) + n * ((q) + n * ((r) + n * (t)))]
/* active Dyall f: f(ab) = aj(ab) - alpha * sum_c atwo(a,c,c,b), active a<=b, as a full n x n matrix */
static void make_f(i64 n, i64 ncore, const double *aj, const double *atwo, double alpha, double *F) {
    for (i64 a = 0; a < n; a++)
        for (i64 b = a; b < n; b++) {
            double v = aj[tri(a + ncore, b + ncore)];
            for (i64 c = 0; c < n; c++) v = v - AT(a, c, c, b) * alpha;
            F[a + n * b] = F[b + n * a] = v;
        }
}
#define FF(p, q) F[(p) + n * (q)]

/* ---------------- koopE (koop_matrices.F90:165-250) */
void orc_koopE(i64 n, i64 ncore, i64 M, const double *atwo, const double *taa, const double *dal,
               const double *aj, double *ipa, double *eaa) {
    MKR(n, M, NULL, taa, dal, NULL);
    double *F = malloc(sizeof(double) * n * n);
    make_f(n, ncore, aj, atwo, 0.0, F);   /* f = aj (no two-electron correction) */
    for (i64 a = 0; a < n; a++)
        for (i64 ap = 0; ap < n; ap++) {
            i64 o = M * (ap + n * a);
            for (i64 s = 0; s < M; s++) ipa[o + s] = 0.0;
            for (i64 c = 0; c < n; c++) {
                double dum = FF(a, c);
                for (i64 s = 0; s < M; s++) ipa[o + s] = ipa[o + s] - dum * D1(&R, s, ap, c);
                for (i64 d = 0; d < n; d++)
                    for (i64 ie = 0; ie < n; ie++)
                        for (i64 s = 0; s < M; s++) ipa[o + s] = ipa[o + s] - AT(c, ie, a, d) * D2(&R, s, ap, c, d, ie);
            }
            for (i64 s = 0; s < M; s++) ipa[o + s] = 2.0 * ipa[o + s];
        }
    double *dumm = malloc(sizeof(double) * M);
    for (i64 ap = 0; ap < n; ap++)
        for (i64 a = 0; a < n; a++) {
            for (i64 s = 0; s < M; s++) dumm[s] = 2.0 * FF(ap, a);
            for (i64 d = 0; d < n; d++)
                for (i64 ie = 0; ie < n; ie++) {
                    double dum2 = 2.0 * AT(d, ie, ap, a) - AT(ap, ie, d, a);
                    for (i64 s = 0; s < M; s++) dumm[s] = dumm[s] + dum2 * D1(&R, s, d, ie);
                }
            i64 o = M * (ap + n * a);
            for (i64 s = 0; s < M; s++) eaa[o + s] = ipa[o + s] + 2.0 * dumm[s];
        }
    free(dumm); free(F);
}

/* ---------------- 2-hole RDM (koop_matrices.F90:78-88) */
void orc_hrdm(i64 n, i64 M, const double *taa, const double *dal, double *out) {
    MKR(n, M, NULL, taa, dal, NULL);
    for (i64 i = 0; i < n; i++)
        for (i64 j = 0; j < n; j++)
            for (i64 k = 0; k < n; k++)
                for (i64 l = 0; l < n; l++)
                    for (i64 s = 0; s < M; s++)
                        out[s + M * (i + n * (j + n * (k + n * l)))] = ro2t(&R, s, i, j, k, l);
}

/* ---------------- koop2E (koop_matrices.F90:253-386), C1 symmetry */
void orc_koop2E(i64 n, i64 ncore, i64 M, const double *atwo, const double *daaa, const double *taa,
                const double *dal, const double *aj, double *koopaa, double *koopeaa) {
    MKR(n, M, daaa, taa, dal, NULL);
    double *F = malloc(sizeof(double) * n * n);
    make_f(n, ncore, aj, atwo, 0.5, F);
#pragma omp parallel for collapse(2)
    for (i64 a = 0; a < n; a++)
        for (i64 b = 0; b < n; b++)
            for (i64 ap = 0; ap < n; ap++)
                for (i64 bp = 0; bp < n; bp++) {
                    i64 o = M * (ap + n * (bp + n * (a + n * b)));
                    for (i64 s = 0; s < M; s++) koopaa[o + s] = 0.0;
                    for (i64 d = 0; d < n; d++) {
                        for (i64 s = 0; s < M; s++)
                            koopaa[o + s] = koopaa[o + s] - FF(a, d) * D2(&R, s, ap, bp, d, b) - FF(b, d) * D2(&R, s, ap, bp, a, d);
                        for (i64 ie = 0; ie < n; ie++)
                            for (i64 ief = 0; ief < n; ief++) {
                                double dint = AT(d, ie, a, ief), dint2 = AT(d, ie, b, ief);
                                for (i64 s = 0; s < M; s++) {
                                    double dum = D3(&R, s, ap, bp, d, ief, b, ie);
                                    if (d == ief) dum = dum + 0.5 * D2(&R, s, ap, bp, ie, b);
                                    if (d == b) dum = dum + 0.5 * D2(&R, s, ap, bp, ief, ie);
                                    koopaa[o + s] = koopaa[o + s] - dint * dum;
                                    dum = D3(&R, s, ap, bp, d, a, ief, ie);
                                    if (d == a) dum = dum + 0.5 * D2(&R, s, ap, bp, ie, ief);
                                    if (d == ief) dum = dum + 0.5 * D2(&R, s, ap, bp, a, ie);
                                    koopaa[o + s] = koopaa[o + s] - dint2 * dum;
                                }
                            }
                    }
                }
#pragma omp parallel for collapse(2)
    for (i64 a = 0; a < n; a++)
        for (i64 b = 0; b < n; b++)
            for (i64 ap = 0; ap < n; ap++)
                for (i64 bp = 0; bp < n; bp++) {
                    i64 o = M * (ap + n * (bp + n * (a + n * b)));
                    for (i64 s = 0; s < M; s++) koopeaa[o + s] = 0.0;
                    for (i64 c = 0; c < n; c++) {
                        for (i64 s = 0; s < M; s++) {
                            double dum1 = ro2t(&R, s, ap, bp, c, b), dum2 = ro2t(&R, s, ap, bp, a, c);
                            koopeaa[o + s] = koopeaa[o + s] + FF(a, c) * dum1 + FF(b, c) * dum2;
                        }
                        for (i64 d = 0; d < n; d++)
                            for (i64 ie = 0; ie < n; ie++) {
                                double dint = AT(c, ie, d, a);
                                for (i64 s = 0; s < M; s++) {
                                    double dum = -ro3t(&R, s, ap, bp, ie, d, b, c);
                                    if (c == ie) dum = dum + 2.0 * ro2t(&R, s, ap, bp, d, b);
                                    if (d == ie) dum = dum - 0.5 * ro2t(&R, s, ap, bp, c, b);
                                    if (b == ie) dum = dum - 0.5 * ro2t(&R, s, ap, bp, d, c);
                                    koopeaa[o + s] = koopeaa[o + s] + dint * dum;
                                }
                                dint = AT(c, ie, d, b);
                                for (i64 s = 0; s < M; s++) {
                                    double dum = -ro3t(&R, s, ap, bp, ie, a, d, c);
                                    if (c == ie) dum = dum + 2.0 * ro2t(&R, s, ap, bp, a, d);
                                    if (a == ie) dum = dum - 0.5 * ro2t(&R, s, ap, bp, c, d);
                                    if (d == ie) dum = dum - 0.5 * ro2t(&R, s, ap, bp, a, c);
                                    koopeaa[o + s] = koopeaa[o + s] + dint * dum;
                                }
                            }
                    }
                }
    free(F);
}

/* ---------------- koopman0pE (koop_matrices.F90:389-588) */
void orc_koopman0pE(i64 n, i64 ncore, i64 M, const double *atwo, const double *daaa, const double *taa,
                    const double *dal, const double *aj, double *koopaa, double *koopbb, double *koopab) {
    MKR(n, M, daaa, taa, dal, NULL);
    double *F = malloc(sizeof(double) * n * n);
    make_f(n, ncore, aj, atwo, 0.5, F);
#pragma omp parallel for collapse(2)
    for (i64 a = 0; a < n; a++)
        for (i64 b = 0; b < n; b++) {
            for (i64 ap = 0; ap < n; ap++)
                for (i64 bp = 0; bp < n; bp++) {
                    i64 o = M * (bp + n * (ap + n * (a + n * b)));
                    for (i64 s = 0; s < M; s++) koopaa[o + s] = 0.0;
                    for (i64 c = 0; c < n; c++) {
                        for (i64 s = 0; s < M; s++)
                            koopaa[o + s] = koopaa[o + s] + 2.0 * (FF(c, a) * ee2(&R, s, bp, ap, c, b) - FF(b, c) * ee2(&R, s, bp, ap, a, c));
                        for (i64 d = 0; d < n; d++)
                            for (i64 ie = 0; ie < n; ie++)
                                for (i64 s = 0; s < M; s++) {
                                    double dum1 = eee2(&R, s, bp, ap, c, ie, d, b) + eee2(&R, s, bp, ap, d, b, c, ie);
                                    double dum2 = eee2(&R, s, bp, ap, a, ie, c, d) + eee2(&R, s, bp, ap, c, d, a, ie);
                                    koopaa[o + s] = koopaa[o + s] + AT(c, ie, d, a) * dum1 - AT(b, ie, c, d) * dum2;
                                }
                    }
                    for (i64 s = 0; s < M; s++) koopbb[o + s] = 0.0;
                    for (i64 c = 0; c < n; c++) {
                        for (i64 s = 0; s < M; s++) {
                            double dum1 = -ee2(&R, s, a, ap, bp, c);
                            if (a == ap) dum1 = dum1 + 2.0 * D1(&R, s, bp, c);
                            if (bp == ap) dum1 = dum1 + D1(&R, s, a, c);
                            double dum2 = -ee2(&R, s, c, ap, bp, b);
                            if (ap == c) dum2 = dum2 + 2.0 * D1(&R, s, bp, b);
                            if (bp == ap) dum2 = dum2 + D1(&R, s, c, b);
                            koopbb[o + s] = koopbb[o + s] - FF(b, c) * dum1 + FF(c, a) * dum2;
                        }
                        for (i64 d = 0; d < n; d++)
                            for (i64 ie = 0; ie < n; ie++)
                                for (i64 s = 0; s < M; s++) {
                                    double dum1 = -eee2(&R, s, c, ie, a, ap, bp, d) - eee2(&R, s, a, ap, bp, d, c, ie);
                                    if (ap == a) dum1 = dum1 + 2.0 * (ee2(&R, s, c, ie, bp, d) + ee2(&R, s, bp, d, c, ie));
                                    if (bp == ap) dum1 = dum1 + ee2(&R, s, c, ie, a, d) + ee2(&R, s, a, d, c, ie);
                                    if (ap == c) {
                                        dum1 = dum1 - ee2(&R, s, a, ie, bp, d);
                                        if (ie == a) dum1 = dum1 + 2.0 * D1(&R, s, bp, d);
                                    }
                                    if (bp == ie) {
                                        dum1 = dum1 + ee2(&R, s, a, ap, c, d);
                                        if (ap == a) dum1 = dum1 - 2.0 * D1(&R, s, c, d);
                                    }
                                    double dum2 = -eee2(&R, s, c, ie, d, ap, bp, b) - eee2(&R, s, d, ap, bp, b, c, ie);
                                    if (ap == d) dum2 = dum2 + 2.0 * (ee2(&R, s, c, ie, bp, b) + ee2(&R, s, bp, b, c, ie));
                                    if (bp == ap) dum2 = dum2 + ee2(&R, s, c, ie, d, b) + ee2(&R, s, d, b, c, ie);
                                    if (ap == c) {
                                        dum2 = dum2 - ee2(&R, s, d, ie, bp, b);
                                        if (d == ie) dum2 = dum2 + 2.0 * D1(&R, s, bp, b);
                                    }
                                    if (bp == ie) {
                                        dum2 = dum2 + ee2(&R, s, d, ap, c, b);
                                        if (d == ap) dum2 = dum2 - 2.0 * D1(&R, s, c, b);
                                    }
                                    koopbb[o + s] = koopbb[o + s] - 0.5 * AT(c, ie, b, d) * dum1 + 0.5 * AT(c, ie, d, a) * dum2;
                                }
                    }
                }
            i64 o2 = M * (a + n * b);
            for (i64 s = 0; s < M; s++) koopab[o2 + s] = 0.0;
            for (i64 c = 0; c < n; c++) {
                for (i64 s = 0; s < M; s++) {
                    double dum1 = D1(&R, s, c, b), dum2 = D1(&R, s, a, c);
                    koopab[o2 + s] = koopab[o2 + s] + 2.0 * (FF(c, a) * dum1 - FF(b, c) * dum2);
                }
                for (i64 d = 0; d < n; d++)
                    for (i64 ie = 0; ie < n; ie++)
                        for (i64 s = 0; s < M; s++) {
                            double dum1 = ee2(&R, s, c, ie, d, b) + ee2(&R, s, d, b, c, ie);
                            double dum2 = ee2(&R, s, a, ie, c, d) + ee2(&R, s, c, d, a, ie);
                            koopab[o2 + s] = koopab[o2 + s] + AT(c, ie, d, a) * dum1 - AT(b, ie, c, d) * dum2;
                        }
            }
        }
    free(F);
}

/* ---------------- bamat, C1 branch (koopro4QD.F90:3974-4104) */
#define A6(arr, s, p, q, r, t, u, v) arr[(s) + M * ((p) + n * ((q) + n * ((r) + n * ((t) + n * ((u) + n * (v))))))]
void orc_bamat(i64 n, i64 ncore, i64 M, const double *atwo, const double *ro4, const double *daaa,
               const double *taa, const double *dal, const double *aj, double *amat, double *atmat, i64 pair_limit) {
    MKR(n, M, daaa, taa, dal, ro4);
    double *F = malloc(sizeof(double) * n * n);
    make_f(n, ncore, aj, atwo, 0.5, F);
    i64 n6 = n * n * n * n * n * n;
    memset(amat, 0, sizeof(double) * M * n6);
    memset(atmat, 0, sizeof(double) * M * n6);
    /* a and c occupy the same slots in every target (amat(.,ap,bp,cp,a,b,c), atmat(.,ap,b,cp,a,bp,c)),
       so iterations with different (a,c) never touch the same element */
    /* pair_limit >= 0 restricts the work to the first pair_limit (a,c) pairs (timing samples only) */
    const i64 npairs = (pair_limit >= 0 && pair_limit < n * n) ? pair_limit : n * n;
#pragma omp parallel for schedule(dynamic)
    for (i64 ac = 0; ac < npairs; ac++) {
        const i64 a = ac / n, c = ac % n;
            for (i64 b = 0; b < n; b++)
                for (i64 ap = 0; ap < n; ap++)
                    for (i64 bp = 0; bp < n; bp++)
                        for (i64 cp = 0; cp < n; cp++)
                            for (i64 d = 0; d < n; d++) {
                                for (i64 s = 0; s < M; s++) {
                                    A6(amat, s, ap, bp, cp, a, b, c) = A6(amat, s, ap, bp, cp, a, b, c)
                                        + FF(d, a) * eee2(&R, s, cp, ap, bp, b, d, c)
                                        - FF(c, d) * eee2(&R, s, cp, ap, bp, b, a, d)
                                        - FF(b, d) * eee2(&R, s, cp, ap, bp, d, a, c);
                                    A6(atmat, s, ap, bp, cp, a, b, c) = A6(atmat, s, ap, bp, cp, a, b, c)
                                        + FF(d, a) * eee2t(&R, s, cp, ap, bp, b, d, c)
                                        - FF(c, d) * eee2t(&R, s, cp, ap, bp, b, a, d)
                                        + FF(b, d) * eee2t(&R, s, cp, ap, bp, d, a, c);
                                }
                                for (i64 ie = 0; ie < n; ie++) {
                                    double dint1 = AT(c, ie, d, a), dint2 = AT(d, ie, ie, a);
                                    double dint3 = AT(c, d, d, ie), dint4 = AT(b, d, d, ie);
                                    for (i64 s = 0; s < M; s++) {
                                        double dum1 = eee2(&R, s, cp, ap, bp, b, d, ie);
                                        double dum2 = eee2(&R, s, cp, ap, bp, b, d, c);
                                        double dum3 = eee2(&R, s, cp, ap, bp, b, a, ie);
                                        double dum4 = eee2(&R, s, cp, ap, bp, ie, a, c);
                                        double dum1t = -dum1, dum2t = -dum2, dum3t = -dum3;
                                        if (bp == b) {
                                            dum1t = dum1t + 2.0 * ee2(&R, s, cp, ap, d, ie);
                                            dum2t = dum2t + 2.0 * ee2(&R, s, cp, ap, d, c);
                                            dum3t = dum3t + 2.0 * ee2(&R, s, cp, ap, a, ie);
                                        }
                                        A6(atmat, s, ap, b, cp, a, bp, c) = A6(atmat, s, ap, b, cp, a, bp, c)
                                            + dum1t * dint1 - 0.5 * (dum2t * dint2 + dum3t * dint3);
                                        double dum4t = eee2t(&R, s, cp, ap, bp, ie, a, c);
                                        A6(amat, s, ap, bp, cp, a, b, c) = A6(amat, s, ap, bp, cp, a, b, c)
                                            + dum1 * dint1 - 0.5 * (dum2 * dint2 + dum3 * dint3 - dum4 * dint4);
                                        A6(atmat, s, ap, bp, cp, a, b, c) = A6(atmat, s, ap, bp, cp, a, b, c) + 0.5 * dum4t * dint4;
                                    }
                                    for (i64 ief = 0; ief < n; ief++)
                                        for (i64 s = 0; s < M; s++) {
                                            double di1 = AT(d, ief, ie, a);
                                            double du1 = eeee(&R, s, cp, ap, bp, b, d, ief, ie, c);
                                            double du1t = -du1;
                                            if (bp == b) du1t = du1t + 2.0 * eee2(&R, s, cp, ap, d, ief, ie, c);
                                            A6(amat, s, ap, bp, cp, a, b, c) = A6(amat, s, ap, bp, cp, a, b, c) + du1 * di1;
                                            A6(atmat, s, ap, b, cp, a, bp, c) = A6(atmat, s, ap, b, cp, a, bp, c) + du1t * di1;
                                            double du2 = eeee(&R, s, cp, ap, bp, b, d, ief, a, ie);
                                            double du2t = -du2;
                                            if (bp == b) du2t = du2t + 2.0 * eee2(&R, s, cp, ap, d, ief, a, ie);
                                            double di2 = AT(d, ief, c, ie);
                                            A6(amat, s, ap, bp, cp, a, b, c) = A6(amat, s, ap, bp, cp, a, b, c) - du2 * di2;
                                            A6(atmat, s, ap, b, cp, a, bp, c) = A6(atmat, s, ap, b, cp, a, bp, c) - du2t * di2;
                                            double du3 = eeee(&R, s, cp, ap, bp, ie, d, ief, a, c);
                                            double du3t = eeeet(&R, s, cp, ap, bp, ie, d, ief, a, c);
                                            double di3 = AT(d, ief, b, ie);
                                            A6(amat, s, ap, bp, cp, a, b, c) = A6(amat, s, ap, bp, cp, a, b, c) - du3 * di3;
                                            A6(atmat, s, ap, bp, cp, a, b, c) = A6(atmat, s, ap, bp, cp, a, b, c) + du3t * di3;
                                        }
                                }
                            }
    }
    free(F);
}

/* ---------------- bbmat (koopro4QD.F90:4231-4299) */
#define A4(arr, s, p, q, r, t) arr[(s) + M * ((p) + n * ((q) + n * ((r) + n * (t))))]
void orc_bbmat(i64 n, i64 ncore, i64 M, const double *atwo, const double *daaa, const double *taa,
               const double *dal, const double *aj, double *bmat, double *btmat) {
    MKR(n, M, daaa, taa, dal, NULL);
    double *F = malloc(sizeof(double) * n * n);
    make_f(n, ncore, aj, atwo, 1.0, F);
#pragma omp parallel for collapse(2)
    for (i64 ap = 0; ap < n; ap++)
        for (i64 bp = 0; bp < n; bp++)
            for (i64 cp = 0; cp < n; cp++)
                for (i64 a = 0; a < n; a++) {
                    for (i64 s = 0; s < M; s++) { A4(bmat, s, ap, bp, cp, a) = 0.0; A4(btmat, s, ap, bp, cp, a) = 0.0; }
                    for (i64 c = 0; c < n; c++) {
                        double fac = FF(c, a), ajac = aj[tri(c + ncore, a + ncore)];
                        for (i64 s = 0; s < M; s++) {
                            A4(bmat, s, ap, bp, cp, a) = A4(bmat, s, ap, bp, cp, a) - fac * ee2(&R, s, cp, ap, bp, c);
                            A4(btmat, s, ap, bp, cp, a) = A4(btmat, s, ap, bp, cp, a) + ajac * ee2tr(&R, s, cp, ap, bp, c);
                        }
                        for (i64 ie = 0; ie < n; ie++)
                            for (i64 ief = 0; ief < n; ief++) {
                                double dint = AT(a, ie, c, ief);
                                for (i64 s = 0; s < M; s++) {
                                    double dum = eee2(&R, s, cp, ap, bp, ie, c, ief);
                                    double dumt = eee2t(&R, s, cp, ap, bp, ie, c, ief);
                                    A4(bmat, s, ap, bp, cp, a) = A4(bmat, s, ap, bp, cp, a) - dum * dint;
                                    A4(btmat, s, ap, bp, cp, a) = A4(btmat, s, ap, bp, cp, a) + dumt * dint;
                                }
                            }
                    }
                }
    free(F);
}

/* ---------------- bcmat (koopro4QD.F90:4301-4397) */
void orc_bcmat(i64 n, i64 ncore, i64 M, const double *atwo, const double *daaa, const double *taa,
               const double *dal, const double *aj, double *cmat, double *ctmat) {
    MKR(n, M, daaa, taa, dal, NULL);
    double *F = malloc(sizeof(double) * n * n);
    make_f(n, ncore, aj, atwo, 0.5, F);
    memset(cmat, 0, sizeof(double) * M * n * n * n * n);
    memset(ctmat, 0, sizeof(double) * M * n * n * n * n);
#pragma omp parallel for collapse(2)
    for (i64 ap = 0; ap < n; ap++)
        for (i64 a = 0; a < n; a++)
            for (i64 b = 0; b < n; b++)
                for (i64 c = 0; c < n; c++)
                    for (i64 d = 0; d < n; d++) {
                        double fda = FF(d, a), fcd = FF(c, d), fbd = FF(b, d);
                        for (i64 s = 0; s < M; s++) {
                            A4(cmat, s, ap, a, b, c) = A4(cmat, s, ap, a, b, c) + fda * ee2(&R, s, ap, b, d, c)
                                - fcd * ee2(&R, s, ap, b, a, d) - fbd * ee2(&R, s, ap, d, a, c);
                            A4(ctmat, s, ap, a, b, c) = A4(ctmat, s, ap, a, b, c) + fda * ee2t(&R, s, ap, b, d, c)
                                - fcd * ee2t(&R, s, ap, b, a, d) + fbd * ee2t(&R, s, ap, d, a, c);
                        }
                        for (i64 ie = 0; ie < n; ie++) {
                            double dint1 = AT(c, ie, d, a), dint2 = AT(d, ie, ie, a), dint3 = AT(c, d, d, ie), dint4 = AT(b, d, d, ie);
                            for (i64 s = 0; s < M; s++) {
                                double dum1 = ee2(&R, s, ap, b, d, ie), dum2 = ee2(&R, s, ap, b, d, c);
                                double dum3 = ee2(&R, s, ap, b, a, ie), dum4 = ee2(&R, s, ap, ie, a, c);
                                double dum1t = ee2t(&R, s, ap, b, d, ie), dum2t = ee2t(&R, s, ap, b, d, c);
                                double dum3t = ee2t(&R, s, ap, b, a, ie), dum4t = ee2t(&R, s, ap, ie, a, c);
                                A4(cmat, s, ap, a, b, c) = A4(cmat, s, ap, a, b, c) + dum1 * dint1 - 0.5 * (dum2 * dint2 + dum3 * dint3 - dum4 * dint4);
                                A4(ctmat, s, ap, a, b, c) = A4(ctmat, s, ap, a, b, c) + dum1t * dint1 - 0.5 * (dum2t * dint2 + dum3t * dint3 - dum4t * dint4);
                            }
                            for (i64 ief = 0; ief < n; ief++) {
                                double di1 = AT(d, ief, ie, a), di2 = AT(d, ief, c, ie), di3 = AT(d, ief, b, ie);
                                for (i64 s = 0; s < M; s++) {
                                    double du1 = eee2(&R, s, ap, b, d, ief, ie, c), du1t = eee2tl(&R, s, ap, b, d, ief, ie, c);
                                    double du2 = eee2(&R, s, ap, b, d, ief, a, ie), du2t = eee2tl(&R, s, ap, b, d, ief, a, ie);
                                    double du3 = eee2(&R, s, ap, ie, d, ief, a, c), du3t = eee2tl(&R, s, ap, ie, d, ief, a, c);
                                    A4(cmat, s, ap, a, b, c) = A4(cmat, s, ap, a, b, c) + du1 * di1 - du2 * di2 - du3 * di3;
                                    A4(ctmat, s, ap, a, b, c) = A4(ctmat, s, ap, a, b, c) + du1t * di1 - du2t * di2 + du3t * di3;
                                }
                            }
                        }
                    }
    free(F);
}

/* ---------------- bdmat (koopro4QD.F90:4399-4459) */
void orc_bdmat(i64 n, i64 ncore, i64 M, const double *atwo, const double *taa, const double *dal,
               const double *aj, double *dmat, double *dtmat) {
    MKR(n, M, NULL, taa, dal, NULL);
    double *F = malloc(sizeof(double) * n * n);
    make_f(n, ncore, aj, atwo, 1.0, F);
    for (i64 ap = 0; ap < n; ap++)
        for (i64 a = 0; a < n; a++) {
            i64 o = M * (ap + n * a);
            for (i64 s = 0; s < M; s++) { dmat[o + s] = 0.0; dtmat[o + s] = 0.0; }
            for (i64 c = 0; c < n; c++) {
                double fac = FF(c, a), ajac = aj[tri(c + ncore, a + ncore)];
                for (i64 s = 0; s < M; s++) {
                    dmat[o + s] = dmat[o + s] - fac * D1(&R, s, ap, c);
                    double dum = -D1(&R, s, ap, c);
                    if (ap == c) dum = dum + 2.0;
                    dtmat[o + s] = dtmat[o + s] + ajac * dum;
                }
                for (i64 ie = 0; ie < n; ie++)
                    for (i64 ief = 0; ief < n; ief++) {
                        double dint = AT(a, ie, c, ief);
                        for (i64 s = 0; s < M; s++) {
                            double dum = ee2(&R, s, ap, ie, c, ief), dumt = ee2t(&R, s, ap, ie, c, ief);
                            dmat[o + s] = dmat[o + s] - dum * dint;
                            dtmat[o + s] = dtmat[o + s] + dumt * dint;
                        }
                    }
            }
        }
    free(F);
}

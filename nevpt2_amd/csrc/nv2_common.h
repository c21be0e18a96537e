// Shared device helpers for the NEVPT2 hot path on gfx950.
//
// RDM views use the reference's array conventions so that accessor code reads
// like the closed forms in eexx_functions.F90: column-major, 0-based, state
// index fastest (d1[s,p,q], d2[s,p,q,r,t], d3[s,p1,p2,p3,q1,q2,q3]); the packed
// 4-RDM is ro4[word,i5,i6,i7,i8,s] addressed through ord8 + lindice
// (ord_utils.F90:71-129, indices_norm.F90:41-44).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

typedef int64_t i64;

namespace nv2 {

void set_error(const char *fmt, ...);

#define NV2_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess) {                                                            \
            ::nv2::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(_e)); \
            return -1;                                                                     \
        }                                                                                  \
    } while (0)

#define NV2_LAUNCH_CHECK() NV2_HIP(hipGetLastError())

struct Rdm {
    int n, M;
    i64 nw;
    const double *d3, *d2, *d1, *ro4;
};

__host__ __device__ inline i64 tri(i64 p, i64 q) {
    i64 hi = p > q ? p : q, lo = p > q ? q : p;
    return hi * (hi + 1) / 2 + lo;
}

__device__ __forceinline__ i64 ix2(const Rdm &R, int s, int p, int q) { return s + (i64)R.M * (p + R.n * q); }
__device__ __forceinline__ i64 ix4(const Rdm &R, int s, int p, int q, int r, int t) {
    const i64 n = R.n;
    return s + (i64)R.M * (p + n * (q + n * (r + n * t)));
}
__device__ __forceinline__ i64 ix6(const Rdm &R, int s, int p, int q, int r, int t, int u, int v) {
    const i64 n = R.n;
    return s + (i64)R.M * (p + n * (q + n * (r + n * (t + n * (u + n * v)))));
}
#define D1(s, p, q) (R.d1[ix2(R, s, p, q)])
#define D2(s, p, q, r, t) (R.d2[ix4(R, s, p, q, r, t)])
#define D3(s, p, q, r, t, u, v) (R.d3[ix6(R, s, p, q, r, t, u, v)])

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// lindice of a 0-based descending tuple, 0-based result (indices_norm.F90:41-44)
__host__ __device__ __forceinline__ i64 lind0(i64 a, i64 b, i64 c, i64 d) {
    a++; b++; c++; d++;
    return (a - 1) * a * (a + 1) * (a + 2) / 24 + (b - 1) * b * (b + 1) / 6 + (c - 1) * c / 2 + d - 1;
}

// ord8's descending selection sort: first maximum wins ties (ord_utils.F90:71-129)
__device__ __forceinline__ void sel_sort4(int *v, int *w) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int im = v[i], jm = i;
#pragma unroll
        for (int j = i + 1; j < 4; j++)
            if (v[j] > im) { im = v[j]; jm = j; }
        if (jm != i) {
            int x = v[i]; v[i] = v[jm]; v[jm] = x;
            x = w[i]; w[i] = w[jm]; w[jm] = x;
        }
    }
}

__device__ __forceinline__ i64 norm4(int a, int b, int c, int d) {
    int v[4] = {a, b, c, d}, w[4] = {0, 0, 0, 0};
    sel_sort4(v, w);
    return lind0(v[0], v[1], v[2], v[3]);
}

// ro4 gather for creation (a,c,e,g), annihilation (b,d,f,h)
__device__ __forceinline__ double g4(const Rdm &R, int s, int a, int c, int e, int g, int b, int d, int f, int h) {
    int v1[4], v2[4];
    if (norm4(a, c, e, g) >= norm4(b, d, f, h)) {
        v1[0] = a; v1[1] = c; v1[2] = e; v1[3] = g; v2[0] = b; v2[1] = d; v2[2] = f; v2[3] = h;
    } else {
        v1[0] = b; v1[1] = d; v1[2] = f; v1[3] = h; v2[0] = a; v2[1] = c; v2[2] = e; v2[3] = g;
    }
    sel_sort4(v1, v2);
    const i64 n = R.n;
    i64 w = lind0(v1[0], v1[1], v1[2], v1[3]);
    return R.ro4[w + R.nw * (v2[0] + n * (v2[1] + n * (v2[2] + n * (v2[3] + n * s))))];
}

// ---------------- accessors (eexx_functions.F90)
__device__ __forceinline__ double ee2(const Rdm &R, int s, int a, int b, int c, int d) {
    double v = D2(s, a, c, b, d);
    if (b == c) v += D1(s, a, d);
    return v;
}
__device__ __forceinline__ double ee2t(const Rdm &R, int s, int a, int b, int c, int d) {
    double v = -ee2(R, s, b, a, c, d);
    if (a == b) v += 2.0 * D1(s, c, d);
    return v;
}
__device__ __forceinline__ double ee2tr(const Rdm &R, int s, int a, int b, int c, int d) {
    double v = -ee2(R, s, a, b, d, c);
    if (c == d) v += 2.0 * D1(s, a, b);
    return v;
}
__device__ __forceinline__ double eee2(const Rdm &R, int s, int a, int b, int c, int d, int e, int f) {
    double v = D3(s, a, c, e, b, d, f);
    if (e == d) v += D2(s, a, c, b, f);
    if (e == b) v += D2(s, a, c, f, d);
    if (b == c) v += D2(s, a, e, d, f);
    if (b == c && e == d) v += D1(s, a, f);
    return v;
}
__device__ __forceinline__ double eee2t(const Rdm &R, int s, int a, int b, int c, int d, int e, int f) {
    double v = -eee2(R, s, a, b, d, c, e, f);
    if (c == d) v += 2.0 * ee2(R, s, a, b, e, f);
    return v;
}
__device__ __forceinline__ double eee2tl(const Rdm &R, int s, int a, int b, int c, int d, int e, int f) {
    double v = -eee2(R, s, b, a, c, d, e, f);
    if (a == b) v += 2.0 * ee2(R, s, c, d, e, f);
    return v;
}
__device__ __forceinline__ double eeee(const Rdm &R, int s, int a, int b, int c, int d, int e, int f, int g, int h) {
    double v = g4(R, s, a, c, e, g, b, d, f, h);
    if (b == c) v += D3(s, a, e, g, d, f, h);
    if (b == e) v += D3(s, a, c, g, f, d, h);
    if (b == g) v += D3(s, a, c, e, h, d, f);
    if (d == e) v += D3(s, a, g, c, b, h, f);
    if (d == g) v += D3(s, a, c, e, b, h, f);
    if (f == g) v += D3(s, a, e, c, b, h, d);
    if (b == c) {
        if (d == e) v += D2(s, a, g, f, h);
        if (d == g) v += D2(s, a, e, h, f);
        if (f == g) v += D2(s, a, e, d, h);
    }
    if (b == e && d == g) v += D2(s, a, c, f, h);
    if (b == e && f == g) v += D2(s, a, c, h, d);
    if (b == g && d == e) v += D2(s, a, c, h, f);
    if (d == e && f == g) v += D2(s, a, c, b, h);
    if (b == c && d == e && f == g) v += D1(s, a, h);
    return v;
}
__device__ __forceinline__ double ro2t(const Rdm &R, int s, int ap, int bp, int a, int b) {
    double v = D2(s, a, b, ap, bp);
    if (ap == a) v -= 2.0 * D1(s, b, bp);
    if (bp == b) v -= 2.0 * D1(s, a, ap);
    if (ap == b) v += D1(s, a, bp);
    if (bp == a) v += D1(s, b, ap);
    if (ap == a && bp == b) v += 4.0;
    if (ap == b && bp == a) v -= 2.0;
    return v;
}
// x = a^-1 for the upper triangular kb x kb block a (kb <= NBK, row stride NBK + 1) in LDS, by a workgroup of
// at least 4 * kb threads (x cleared and a barrier passed by the caller),
// by columns, all at once: x[i][c] = (delta_ic - sum_{i<k<=c} a[i][k] x[k][c]) / a[i][i] for i = c down
// to 0.  Four lanes per column (threads 4c .. 4c+3, always active together) split each sum by k mod 4 and
// meet by two shuffles; a column lives in one wavefront, whose in-order LDS (the wait after each store)
// orders x[i][c] before its reads, so the columns need no barrier.  The caller's barrier follows.
template <int NBK>
__device__ __forceinline__ void upper_inverse_lds(const double (*a)[NBK + 1], double (*x)[NBK + 1], int kb) {
    const int c = threadIdx.x >> 2, sub = threadIdx.x & 3;
    if (c >= kb) return;
    for (int i = c; i >= 0; i--) {
        double sm = 0.0;
        for (int k = i + 1 + sub; k <= c; k += 4) sm += a[i][k] * x[k][c];
        sm += __shfl_xor(sm, 1, 64);
        sm += __shfl_xor(sm, 2, 64);
        const double xi = ((c == i ? 1.0 : 0.0) - sm) / a[i][i];
        if (sub == 0) x[i][c] = xi;
        __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    }
}

}  // namespace nv2

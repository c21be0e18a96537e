// The O(n^9) part of bamat: four contractions of E4 = <E E E E> with the active ERIs
// (koopro4QD.F90:4069-4094), on the matrix cores.
//
// With E4 stored x8 fastest (E4[x1..x8], offset sum x_p n^(8-p)) and cp = x1, ap = x2:
//   C1(ap,bp,cp,a,b,c) = sum_{x5 x6 x7} E4[cp,ap,bp,b,x5,x6,x7,c] B1[(x5,x6,x7), a]
//   C2(ap,bp,cp,a,b,c) = sum_{x5 x6 x8} E4[cp,ap,bp,b,x5,x6,a,x8] B2[(x5,x6,x8), c]
//   C3(ap,bp,cp,a,b,c) = sum_{x4 x5 x6} E4[cp,ap,bp,x4,x5,x6,a,c] B3[(x4,x5,x6), b]
//   C4(ap,bp,cp,a,b,c) = sum_{x3 x5 x6} E4[cp,ap,x3,bp,x5,x6,a,c] B3[(x3,x5,x6), b]
// (B1[(d,j,i),a] = w(d,j,i,a), B2[(d,j,i),c] = w(d,j,c,i), B3[(i,d,j),b] = w(d,j,b,i)).
//
// Every index runs over n <= 16 active orbitals, so each contraction is a set of 16x16 MFMA
// tiles.  Pass A (one workgroup per (cp,ap,bp), four wavefronts) streams the contiguous
// sub-slab E4[cp,ap,bp,:,:,:,:,:] (n^5 values) once through LDS in chunks (x4,x5) of n^3
// values and feeds C1, C2 and C3 from each chunk (the K ranges of C1 and C2 split over two
// wavefronts each, the C3 tiles over all four); pass B (one
// workgroup per (cp,ap,bp'), four wavefronts) streams E4[cp,ap,:,bp',...] for C4, whose sum
// runs over x3.  E4 is read twice instead of four times.  FP64 uses v_mfma_f64_16x16x4f64;
// the FP32 form stores E4 in float and uses v_mfma_f32_16x16x4f32 with FP32 accumulation; the
// split-BF16x3 form stores E4 in float too and splits every operand exactly into three bf16 terms
// (x = h + m + l, by truncation: 8 + 8 + 8 significant bits), forming the six products of order
// up to 2^-16 -- hh, hm, mh, mm, hl, lh -- with two v_mfma_f32_16x16x16_bf16 per K-step of four and
// FP32 accumulation (the three dropped products are of order 2^-24, the FP32 rounding level).
// Outputs land directly in amat's layout (state fastest).
#include "nv2_common.h"
#include "nv2_stage1.h"

namespace nv2 {

typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef short mf_s4 __attribute__((ext_vector_type(4)));

// A operand of lane l is A[l % 16][l / 16], B operand B[l / 16][l % 16] (one K step of 4);
// accumulator register r of lane l holds D[row(l, r)][l % 16]
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef mf_d4 V;
    __device__ static V mma(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    typedef mf_f4 V;
    __device__ static V mma(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// split-BF16x3 on the bf16 matrix cores.  v_mfma_f32_16x16x16_bf16 takes four consecutive K values per
// lane (lane l: A[l % 16][4 (l / 16) + t], B[4 (l / 16) + t][l % 16], t = 0..3) and keeps the FP32
// accumulator layout of v_mfma_f32_16x16x4f32, so one K index of the FP32 form (lane group l / 16)
// spreads its split products over the four slots t of its lane group.
struct MfmaSplit3 {
    typedef mf_f4 V;
    __device__ static void split(float x, short &h, short &m, short &l) {
        const unsigned uh = __float_as_uint(x) & 0xffff0000u;
        const float r1 = x - __uint_as_float(uh);              // exact: at most 16 significant bits
        const unsigned um = __float_as_uint(r1) & 0xffff0000u;
        const float r2 = r1 - __uint_as_float(um);             // exact: at most 8 significant bits
        const unsigned ul = __float_as_uint(r2) & 0xffff0000u;   // == r2
        h = (short)(uh >> 16); m = (short)(um >> 16); l = (short)(ul >> 16);
    }
    __device__ static V mma(float a, float b, V c) {
        short ah, am, al, bh, bm, bl;
        split(a, ah, am, al);
        split(b, bh, bm, bl);
        const mf_s4 A0 = {ah, ah, am, am}, B0 = {bh, bm, bh, bm};      // hh + hm + mh + mm
        const mf_s4 A1 = {ah, al, 0, 0}, B1 = {bl, bh, 0, 0};          // hl + lh
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A0, B0, c, 0, 0, 0);
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A1, B1, c, 0, 0, 0);
    }
    __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

struct ContractPlan {
    int n, M;
    double *Bk[3];
};

// B matrices from atwo (Fortran layout): B1[(d,j,i),a]=w(d,j,i,a) B2[(d,j,i),c]=w(d,j,c,i) B3[(i,d,j),b]=w(d,j,b,i),
// row-major [k][col]
__global__ void k_make_bk(int n, const double *atwo, double *B1, double *B2, double *B3) {
    const i64 tot = (i64)n * n * n * n;
    i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    int col = idx % n;
    i64 k = idx / n;
    int x2 = k % n, x1 = (k / n) % n, x0 = k / n / n;        // k = (x0*n + x1)*n + x2
#define W(p, q, r, t) atwo[(p) + n * ((q) + n * ((r) + (i64)n * (t)))]
    B1[idx] = W(x0, x1, x2, col);          // (d,j,i) = (x0,x1,x2), col a
    B2[idx] = W(x0, x1, col, x2);          // (d,j,i), col c
    B3[idx] = W(x1, x2, col, x0);          // (i,d,j) = (x0,x1,x2), col b
#undef W
}

template <typename V> __device__ inline V vzero() { V z = {0, 0, 0, 0}; return z; }

// pass A: workgroup g = (cp*n + ap)*n + bp.  The chunks (x4,x5) are visited x5-outer, so the B1 and
// B2 slices of one x5 are staged in LDS once and serve the n chunks (x4,x5); C1 and C2 therefore
// keep one accumulator tile per x4 in registers.  Wavefronts 0 and 1 split the K range of C1, 2 and
// 3 that of C2 (their halves are added through LDS at the end), and all four share the C3 row tiles,
// so every wavefront issues about the same number of MFMAs per chunk.
// amdgpu_waves_per_eu(2): two workgroups per CU (the compiler otherwise takes 284 VGPR + AGPR per lane at
// n = 14 in FP64, one wavefront per SIMD, and every barrier and load of a chunk goes unhidden; with the
// hint it fits 217 without spilling, and two 70 KB LDS footprints fit a CU)
template <typename T, int NN, typename MM = Mfma<T>>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_e4_pass_a(const T *__restrict__ E4, const double *__restrict__ B1,
                                                   const double *__restrict__ B2, const double *__restrict__ B3,
                                                   int M, int s, double *C1, double *C2, double *C3) {
    constexpr int N2 = NN * NN, N3 = N2 * NN;
    constexpr int KA = (N2 + 3) / 4;          // K steps of C1 / C2 per chunk (K = n^2)
    constexpr int KH = (KA + 1) / 2;          // K steps of one half
    constexpr int KC = (NN + 3) / 4;          // K steps of C3 per chunk (K = n)
    constexpr int T3 = (N2 + 15) / 16;        // row tiles of C3 (rows (a,c))
    constexpr int TW = (T3 + 3) / 4;          // C3 tiles per wavefront
    typedef typename MM::V V;
    __shared__ T ech[N3], b1s[N3], b2s[N3];
    __shared__ T b3s[N2];
    __shared__ T part[2][256];                // the upper K halves of C1 and C2, one x4 tile at a time
    const int g = blockIdx.x;
    const int bp = g % NN, ap = (g / NN) % NN, cp = g / N2;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const bool c2w = wave >= 2;               // this wavefront's half belongs to C2 (else C1)
    const int k0 = (wave & 1) * KH, k1 = (wave & 1) ? KA : KH;
    const T *slab = E4 + (i64)g * N2 * N3;   // E4[cp,ap,bp,...]
    const i64 obase = ap + (i64)NN * bp + (i64)N2 * cp;
    V accA[NN], acc3[TW];
#pragma unroll
    for (int u = 0; u < NN; u++) accA[u] = vzero<V>();
#pragma unroll
    for (int u = 0; u < TW; u++) acc3[u] = vzero<V>();
    // the next E4 chunk is loaded into registers while the current one is multiplied
    constexpr int PF = (N3 + 255) / 256;
    T pf[PF];
#pragma unroll
    for (int u = 0; u < PF; u++) {
        const int e = threadIdx.x + 256 * u;
        pf[u] = e < N3 ? slab[e] : (T)0;          // chunk (x4 = 0, x5 = 0)
    }
    // and the B3 slice of the next chunk (N2 <= 256 values: one per thread)
    static_assert(N2 <= 256, "one B3 value per thread");
    T pb3 = threadIdx.x < N2 ? (T)B3[threadIdx.x] : (T)0;
    for (int x5 = 0; x5 < NN; x5++) {
#pragma unroll
        for (int x4 = 0; x4 < NN; x4++) {
            __syncthreads();                   // the previous chunk (and B slices) are consumed
#pragma unroll
            for (int u = 0; u < PF; u++) {
                const int e = threadIdx.x + 256 * u;
                if (e < N3) ech[e] = pf[u];
            }
            if (x4 == 0)
                for (int e = threadIdx.x; e < N3; e += 256) {
                    b1s[e] = (T)B1[(i64)x5 * N3 + e];
                    b2s[e] = (T)B2[(i64)x5 * N3 + e];
                }
            if (threadIdx.x < N2) b3s[threadIdx.x] = pb3;
            __syncthreads();
            {                                  // prefetch the next chunk in visiting order
                const int nx4 = x4 + 1 < NN ? x4 + 1 : 0, nx5 = x4 + 1 < NN ? x5 : x5 + 1;
                if (nx5 < NN) {
                    if (threadIdx.x < N2) pb3 = (T)B3[((i64)nx4 * N2 + nx5 * NN) * NN + threadIdx.x];
                    const T *nxt = slab + (i64)(nx4 * NN + nx5) * N3;
#pragma unroll
                    for (int u = 0; u < PF; u++) {
                        const int e = threadIdx.x + 256 * u;
                        pf[u] = e < N3 ? nxt[e] : (T)0;
                    }
                }
            }
            if (!c2w) {                        // C1[x4][x8][a] += E[x6,x7,x8] B1[(x5,x6,x7),a], K = (x6,x7)
                for (int ks = k0; ks < k1; ks++) {
                    const int kk = 4 * ks + lk;
                    const bool ok = li < NN && kk < N2;
                    const T a = ok ? ech[kk * NN + li] : (T)0;
                    const T b = ok ? b1s[kk * NN + li] : (T)0;
                    accA[x4] = MM::mma(a, b, accA[x4]);
                }
            } else {                           // C2[x4][x7][c] += E[x6,x7,x8] B2[(x5,x6,x8),c], K = (x6,x8)
                for (int ks = k0; ks < k1; ks++) {
                    const int kk = 4 * ks + lk;
                    const bool ok = li < NN && kk < N2;
                    const T a = ok ? ech[(kk / NN) * N2 + li * NN + kk % NN] : (T)0;
                    const T b = ok ? b2s[kk * NN + li] : (T)0;
                    accA[x4] = MM::mma(a, b, accA[x4]);
                }
            }
            // C3[(x7,x8)][b] += E[x6,x7,x8] B3[(x4,x5,x6),b], K = x6; row tiles 4u + wave
#pragma unroll
            for (int ks = 0; ks < KC; ks++) {
                const int x6 = 4 * ks + lk;
                const T b = x6 < NN && li < NN ? b3s[x6 * NN + li] : (T)0;
#pragma unroll
                for (int u = 0; u < TW; u++) {
                    const int r = 16 * (4 * u + wave) + li;
                    const T a = x6 < NN && r < N2 ? ech[x6 * N2 + r] : (T)0;
                    acc3[u] = MM::mma(a, b, acc3[u]);
                }
            }
        }
    }
    // C1 / C2: the upper halves (wavefronts 1 and 3) are added to the lower ones through LDS
#pragma unroll
    for (int x4 = 0; x4 < NN; x4++) {
        __syncthreads();
        if (wave & 1)
#pragma unroll
            for (int r = 0; r < 4; r++) part[wave >> 1][lane + 64 * r] = accA[x4][r];
        __syncthreads();
        if (!(wave & 1)) {                     // rows x8 (C1) / x7 (C2), columns a (C1) / c (C2)
            const int j = li;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = MM::row(lane, r);
                const T v = accA[x4][r] + part[wave >> 1][lane + 64 * r];
                if (i < NN && j < NN) {
                    if (!c2w)
                        C1[s + (i64)M * (obase + (i64)N3 * j + (i64)N2 * N2 * x4 + (i64)N2 * N3 * i)] = (double)v;
                    else
                        C2[s + (i64)M * (obase + (i64)N3 * i + (i64)N2 * N2 * x4 + (i64)N2 * N3 * j)] = (double)v;
                }
            }
        }
    }
    {                                          // C3: rows (a,c) = (x7,x8), columns b
        const int j = li;
#pragma unroll
        for (int u = 0; u < TW; u++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rr = 16 * (4 * u + wave) + MM::row(lane, r);
                if (rr < N2 && j < NN) {
                    const int a = rr / NN, c = rr % NN;
                    C3[s + (i64)M * (obase + (i64)N3 * a + (i64)N2 * N2 * j + (i64)N2 * N3 * c)] = (double)acc3[u][r];
                }
            }
    }
}

// pass B: workgroup g = (cp*n + ap)*n + bp; C4[(x7,x8)][b] = sum over chunks (x3,x5) and x6; the
// row tiles are spread over the four wavefronts
template <typename T, int NN, typename MM = Mfma<T>>
__global__ __launch_bounds__(256) void k_e4_pass_b(const T *__restrict__ E4, const double *__restrict__ B3, int M,
                                                   int s, double *C4) {
    constexpr int N2 = NN * NN, N3 = N2 * NN;
    constexpr int KC = (NN + 3) / 4;
    constexpr int T3 = (N2 + 15) / 16;
    constexpr int TW = (T3 + 3) / 4;           // tiles per wavefront
    typedef typename MM::V V;
    __shared__ T ech[N3];
    __shared__ T b3s[N2];
    const int g = blockIdx.x;
    const int bp = g % NN, ap = (g / NN) % NN, cp = g / N2;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const T *slab = E4 + (i64)(g / NN) * N3 * N3 + (i64)bp * N2 * N2;    // E4[cp,ap,0,bp,...]
    const i64 obase = ap + (i64)NN * bp + (i64)N2 * cp;
    V acc[TW];
#pragma unroll
    for (int u = 0; u < TW; u++) acc[u] = vzero<V>();
    constexpr int PF = (N3 + 255) / 256;       // next chunk prefetched into registers
    T pf[PF];
#pragma unroll
    for (int u = 0; u < PF; u++) {
        const int e = threadIdx.x + 256 * u;
        pf[u] = e < N3 ? slab[e] : (T)0;
    }
    for (int x3 = 0; x3 < NN; x3++)
        for (int x5 = 0; x5 < NN; x5++) {
            __syncthreads();
#pragma unroll
            for (int u = 0; u < PF; u++) {
                const int e = threadIdx.x + 256 * u;
                if (e < N3) ech[e] = pf[u];
            }
            for (int e = threadIdx.x; e < N2; e += 256) b3s[e] = (T)B3[((i64)x3 * N2 + x5 * NN) * NN + e];
            __syncthreads();
            {
                const int nx5 = x5 + 1 < NN ? x5 + 1 : 0, nx3 = x5 + 1 < NN ? x3 : x3 + 1;
                if (nx3 < NN) {
                    const T *nxt = slab + (i64)nx3 * N2 * N3 + (i64)nx5 * N3;
#pragma unroll
                    for (int u = 0; u < PF; u++) {
                        const int e = threadIdx.x + 256 * u;
                        pf[u] = e < N3 ? nxt[e] : (T)0;
                    }
                }
            }
#pragma unroll
            for (int ks = 0; ks < KC; ks++) {
                const int x6 = 4 * ks + lk;
                const T b = x6 < NN && li < NN ? b3s[x6 * NN + li] : (T)0;
#pragma unroll
                for (int u = 0; u < TW; u++) {
                    const int r = 16 * (wave + 4 * u) + li;
                    const T a = x6 < NN && r < N2 ? ech[x6 * N2 + r] : (T)0;
                    acc[u] = MM::mma(a, b, acc[u]);
                }
            }
        }
    const int j = li;
#pragma unroll
    for (int u = 0; u < TW; u++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int rr = 16 * (wave + 4 * u) + MM::row(lane, r);
            if (rr < N2 && j < NN) {
                const int a = rr / NN, c = rr % NN;
                C4[s + (i64)M * (obase + (i64)N3 * a + (i64)N2 * N2 * j + (i64)N2 * N3 * c)] = (double)acc[u][r];
            }
        }
}

ContractPlan *contract_plan_create(int n, int M) {
    ContractPlan *p = new ContractPlan();
    p->n = n; p->M = M;
    const i64 n4 = (i64)n * n * n * n;
    for (int c = 0; c < 3; c++)
        if (hipMalloc(&p->Bk[c], sizeof(double) * n4) != hipSuccess) p->Bk[c] = nullptr;
    return p;
}

void contract_plan_destroy(ContractPlan *p) {
    if (!p) return;
    for (int c = 0; c < 3; c++) (void)hipFree(p->Bk[c]);
    delete p;
}

template <typename T, int NN, typename MM>
static int run_passes(ContractPlan *p, const T *E4, int s, double **outs, hipStream_t st) {
    const unsigned nwg = NN * NN * NN;
    hipLaunchKernelGGL((k_e4_pass_a<T, NN, MM>), dim3(nwg), dim3(256), 0, st, E4, p->Bk[0], p->Bk[1], p->Bk[2], p->M,
                       s, outs[0], outs[1], outs[2]);
    hipLaunchKernelGGL((k_e4_pass_b<T, NN, MM>), dim3(nwg), dim3(256), 0, st, E4, p->Bk[2], p->M, s, outs[3]);
    NV2_LAUNCH_CHECK();
    return 0;
}

template <typename T, typename MM = Mfma<T>>
static int dispatch(ContractPlan *p, const T *E4, int s, double **outs, hipStream_t st) {
    switch (p->n) {
#define CASE(k) case k: return run_passes<T, k, MM>(p, E4, s, outs, st);
        CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13)
        CASE(14) CASE(15) CASE(16)
#undef CASE
        default: set_error("E4 contraction supports 2 <= nact <= 16 (got %d)", p->n); return -2;
    }
}

int launch_e4_contractions(ContractPlan *p, const void *E4, int prec, const double *atwo, int s, double *C1,
                           double *C2, double *C3, double *C4, hipStream_t st) {
    if (!p->Bk[0] || !p->Bk[1] || !p->Bk[2]) { set_error("E4 contraction: B matrices not allocated"); return -1; }
    const i64 n = p->n, n4 = n * n * n * n;
    hipLaunchKernelGGL(k_make_bk, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, p->n, atwo, p->Bk[0],
                       p->Bk[1], p->Bk[2]);
    NV2_LAUNCH_CHECK();
    double *outs[4] = {C1, C2, C3, C4};
    if (prec == 0) return dispatch<double>(p, (const double *)E4, s, outs, st);
    if (prec == 2) return dispatch<float, MfmaSplit3>(p, (const float *)E4, s, outs, st);
    return dispatch<float>(p, (const float *)E4, s, outs, st);
}

}  // namespace nv2

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
// tiles.  Pass A (two workgroups per (cp,ap,bp), one per half of x4, four wavefronts each)
// streams the contiguous sub-slab E4[cp,ap,bp,:,:,:,:,:] (n^5 values) once through LDS in chunks
// (x4,x5) of n^3 values and feeds C1, C2 and C3 from each chunk (the K ranges of C1 and C2 split
// over two wavefronts each, the C3 tiles over all four); pass B (one
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

// LDS layout of one E4 chunk E[x6][x7][x8] (n^3 values): each x7 row holds x8 at stride 1 in RS = 17
// slots, each x6 block is n rows (S6 = 17 n slots), and the blocks run on to a multiple of 4 with
// zero blocks.  Every slot the chunk does not fill is zeroed once at the start and never written
// again, so each MFMA operand is read at a per-lane base plus a compile-time offset (the offset field
// of ds_read), with no per-lane condition:
//   C1: A = E[x6,x7,x8 = lane%16] with K = (x6,x7) in order: slot 17 k + lane%16 (past the last k:
//       the zero blocks);
//   C2: A = E[x6, x7 = lane%16, x8] with K = (x6, x8) and x8 run to 16 (B zero and A a zero pad slot
//       for x8 >= n): slot S6 x6 + 17 (lane%16) + x8; 17 being odd, the 16 lanes of one K value hit
//       16 different bank pairs;
//   C3/C4: A = E[x6, x7 = tile, x8 = lane%16] with K = x6 run to a multiple of 4 (zero blocks).
// Rows or columns beyond n land in output rows and columns that are never stored.
template <int NN> struct ChunkLayout {
    static constexpr int RS = 17;
    static constexpr int S6 = NN * RS;
    static constexpr int KC = (NN + 3) / 4;
    static constexpr int KA = (NN * NN + 3) / 4;
    static constexpr int cmax(int a, int b) { return a > b ? a : b; }
    static constexpr int SIZE = cmax(cmax(4 * KC * S6, 4 * KA * RS + 16), cmax((NN - 1) * S6 + 15 * RS + 16, 512));
    __device__ static int at(int x6, int x7, int x8) { return x6 * S6 + x7 * RS + x8; }
};

// One wavefront's share of C1 for one chunk: K steps K0 .. K1-1 of K = (x6,x7) (n^2 values, padded to
// a multiple of 4 with zero rows of B1).  Fully unrolled, reads at fixed offsets from two bases.
template <typename T, int NN, typename MM, int K0, int K1>
__device__ __forceinline__ void kloop_c1(const T *ech, const T *bs, int li, int lk, typename MM::V &acc) {
    const T *pa = ech + lk * ChunkLayout<NN>::RS + li, *pb = bs + lk * NN + li;
#pragma unroll
    for (int ks = K0; ks < K1; ks++) acc = MM::mma(pa[4 * ks * ChunkLayout<NN>::RS], pb[4 * ks * NN], acc);
}
// C2 likewise, K = (x6, x8) with x8 in 0..15 (4 n steps; the B2 slice is staged in that order, rows
// x8 >= n zero)
template <typename T, int NN, typename MM, int K0, int K1>
__device__ __forceinline__ void kloop_c2(const T *ech, const T *bs, int li, int lk, typename MM::V &acc) {
    typedef ChunkLayout<NN> CL;
    const T *pa = ech + li * CL::RS + lk, *pb = bs + lk * NN + li;
#pragma unroll
    for (int ks = K0; ks < K1; ks++) acc = MM::mma(pa[(ks / 4) * CL::S6 + 4 * (ks % 4)], pb[4 * ks * NN], acc);
}

// C3 (pass A) / C4 (pass B) for one chunk: rows (a,c) = (x7, x8) as one tile per x7, columns b, K = x6
// (padded to a multiple of 4: zero blocks of the chunk, zero rows of B3).  Wavefront W owns the tiles
// x7 = W + 4u, u < (n - W + 3) / 4: a compile-time count, so every MFMA is unconditional.
template <int NN, int W> struct C3Tiles { static constexpr int value = (NN - W + 3) / 4; };
template <typename T, int NN, typename MM, int W, int TW>
__device__ __forceinline__ void kloop_c3(const T *ech, const T *b3s, int li, int lk, typename MM::V (&acc)[TW]) {
    typedef ChunkLayout<NN> CL;
    constexpr int TWW = C3Tiles<NN, W>::value;
    static_assert(TWW <= TW, "accumulator tiles");
    const T *pa = ech + lk * CL::S6 + W * CL::RS + li, *pb = b3s + lk * NN + li;
#pragma unroll
    for (int ks = 0; ks < CL::KC; ks++) {
        const T b = pb[4 * ks * NN];
#pragma unroll
        for (int u = 0; u < TWW; u++) acc[u] = MM::mma(pa[4 * ks * CL::S6 + 4 * u * CL::RS], b, acc[u]);
    }
}
// one chunk of pass A for wavefront W: its half of C1 (W = 0, 1) or C2 (W = 2, 3), and its C3 tiles
template <typename T, int NN, typename MM, int W, int TW>
__device__ __forceinline__ void chunk_a(const T *ech, const T *b1s, const T *b2s, const T *b3s, int li, int lk,
                                        typename MM::V &accA, typename MM::V (&acc3)[TW]) {
    constexpr int KA = ChunkLayout<NN>::KA, KH = (KA + 1) / 2, KA2 = 4 * NN;
    if (W == 0) kloop_c1<T, NN, MM, 0, KH>(ech, b1s, li, lk, accA);
    if (W == 1) kloop_c1<T, NN, MM, KH, KA>(ech, b1s, li, lk, accA);
    if (W == 2) kloop_c2<T, NN, MM, 0, KA2 / 2>(ech, b2s, li, lk, accA);
    if (W == 3) kloop_c2<T, NN, MM, KA2 / 2, KA2>(ech, b2s, li, lk, accA);
    kloop_c3<T, NN, MM, W, TW>(ech, b3s, li, lk, acc3);
}

// pass A: workgroup 2 g + h, g = (cp*n + ap)*n + bp, takes the chunks (x4,x5) of x4 group h (x4 in
// h G4 .. h G4 + G4 - 1), x5-outer, so the B1 and B2 slices of one x5 are staged in LDS once and serve
// its G4 chunks, and C1 and C2 keep one accumulator tile per x4 of the group in registers (G4 tiles
// instead of n: the register budget of two workgroups per CU).  Two workgroups per slab make 2 n^3
// equal workgroups, so the last round of the grid is a smaller fraction of it.  C3 sums over both
// groups: each workgroup adds its half to the zeroed output with an FP64 atomic add, and two addends
// onto zero give a+b = b+a whichever lands first.  Wavefronts 0 and 1 split the K range of C1, 2 and
// 3 that of C2 (their halves are added through LDS at the end), and all four share the C3 tiles (one
// per x7), so every wavefront issues about the same number of MFMAs per chunk.  Each chunk is read from
// HBM once.
// amdgpu_waves_per_eu(2): two workgroups per CU, so one's barriers and chunk stores overlap the
// other's MFMAs (the LDS footprint, 77 KB at n = 14 in FP64, allows two).
template <typename T, int NN, typename MM = Mfma<T>>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_e4_pass_a(const T *__restrict__ E4, const double *__restrict__ B1,
                                                   const double *__restrict__ B2, const double *__restrict__ B3,
                                                   int M, int s, double *C1, double *C2, double *C3, int g0) {
    typedef ChunkLayout<NN> CL;
    constexpr int N2 = NN * NN, N3 = N2 * NN;
    constexpr int KA = (N2 + 3) / 4;          // K steps of C1 / C2 per chunk (K = n^2)
    constexpr int KH = (KA + 1) / 2;          // K steps of the lower half
    constexpr int KC = (NN + 3) / 4;          // K steps of C3 per chunk (K = n)
    constexpr int TW = (NN + 3) / 4;          // C3 tiles per wavefront (one tile per x7)
    constexpr int G4 = (NN + 1) / 2;          // x4 values per group
    constexpr int KA2 = 4 * NN;               // K steps of C2 (x8 run to 16)
    constexpr int BS1 = 4 * KA * NN + 16;     // B1 slice with zero rows and tail
    constexpr int BS20 = 16 * NN * NN + 16;   // B2 slice in the (x6, x8 < 16) order
    constexpr int BS2 = BS20 > 512 ? BS20 : 512;   // (and room for two partial tiles)
    constexpr int B3S = 4 * KC * NN + 16;
    typedef typename MM::V V;
    __shared__ T ech[CL::SIZE], b1s[BS1], b2s[BS2], b3s[B3S];
    const int gb = g0 + blockIdx.x;
    const int g = gb >> 1, x4g = (gb & 1) * G4;    // this workgroup's x4 group: x4g .. x4g + g4 - 1
    const int g4 = NN - x4g < G4 ? NN - x4g : G4;
    const int bp = g % NN, ap = (g / NN) % NN, cp = g / N2;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: scalar branches
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const T *slab = E4 + (i64)g * N2 * N3;   // E4[cp,ap,bp,...]
    const i64 obase = ap + (i64)NN * bp + (i64)N2 * cp;
    for (int e = threadIdx.x; e < CL::SIZE; e += 256) ech[e] = (T)0;
    for (int e = threadIdx.x; e < BS1; e += 256) b1s[e] = (T)0;
    for (int e = threadIdx.x; e < BS2; e += 256) b2s[e] = (T)0;
    for (int e = threadIdx.x; e < B3S; e += 256) b3s[e] = (T)0;
    V acc3[TW];
#pragma unroll
    for (int u = 0; u < TW; u++) acc3[u] = vzero<V>();
    // the next E4 chunk is loaded into registers while the current one is multiplied
    constexpr int PF = (N3 + 255) / 256;
    T pf[PF];
#pragma unroll
    for (int u = 0; u < PF; u++) {
        const int e = threadIdx.x + 256 * u;
        pf[u] = e < N3 ? slab[(i64)x4g * NN * N3 + e] : (T)0;          // chunk (x4 = x4g, x5 = 0)
    }
    // and the B3 slice of the next chunk (N2 <= 256 values: one per thread)
    static_assert(N2 <= 256, "one B3 value per thread");
    T pb3 = threadIdx.x < N2 ? (T)B3[(i64)x4g * N2 * NN + threadIdx.x] : (T)0;
    const bool c2w = wave >= 2;
    {
        V accA[G4];
#pragma unroll
        for (int u = 0; u < G4; u++) accA[u] = vzero<V>();
        for (int x5 = 0; x5 < NN; x5++) {
#pragma unroll
            for (int u = 0; u < G4; u++) {
                if (u >= g4) break;
                const int x4 = x4g + u;
                __syncthreads();               // the previous chunk (and B slices) are consumed
#pragma unroll
                for (int v = 0; v < PF; v++) {
                    const int e = threadIdx.x + 256 * v;
                    if (e < N3) ech[CL::at(e / N2, (e / NN) % NN, e % NN)] = pf[v];
                }
                if (u == 0)
                    for (int e = threadIdx.x; e < N3; e += 256) {
                        b1s[e] = (T)B1[(i64)x5 * N3 + e];
                        b2s[((e / N2) * 16 + (e / NN) % NN) * NN + e % NN] = (T)B2[(i64)x5 * N3 + e];
                    }
                if (threadIdx.x < N2) b3s[threadIdx.x] = pb3;
                __syncthreads();
                {                              // prefetch the next chunk in visiting order
                    int nx4 = x4 + 1, nx5 = x5;
                    if (u + 1 >= g4) {         // end of the group's x4 run: next x5, or none
                        nx4 = x4g;
                        nx5 = x5 + 1;
                        if (nx5 == NN) nx4 = NN;
                    }
                    if (nx4 < NN) {
                        if (threadIdx.x < N2) pb3 = (T)B3[((i64)nx4 * N2 + nx5 * NN) * NN + threadIdx.x];
                        const T *nxt = slab + (i64)(nx4 * NN + nx5) * N3;
#pragma unroll
                        for (int v = 0; v < PF; v++) {
                            const int e = threadIdx.x + 256 * v;
                            pf[v] = e < N3 ? nxt[e] : (T)0;
                        }
                    }
                }
                // C1[x4][x8][a] += E[x6,x7,x8] B1[(x5,x6,x7),a]; C2[x4][x7][c] += E[x6,x7,x8] B2[(x5,x6,x8),c]
                // and C3[(x7,x8)][b] += E[x6,x7,x8] B3[(x4,x5,x6),b], K = x6
                switch (wave) {
                    case 0: chunk_a<T, NN, MM, 0, TW>(ech, b1s, b2s, b3s, li, lk, accA[u], acc3); break;
                    case 1: chunk_a<T, NN, MM, 1, TW>(ech, b1s, b2s, b3s, li, lk, accA[u], acc3); break;
                    case 2: chunk_a<T, NN, MM, 2, TW>(ech, b1s, b2s, b3s, li, lk, accA[u], acc3); break;
                    default: chunk_a<T, NN, MM, 3, TW>(ech, b1s, b2s, b3s, li, lk, accA[u], acc3); break;
                }
            }
        }
        // C1 / C2 of the group: the upper halves (wavefronts 1 and 3) are added to the lower ones through
        // the B2 slice buffer (its zero rows meet zero pad slots of the chunk, so what this leaves there
        // does not matter)
        T *part = b2s;
#pragma unroll
        for (int u = 0; u < G4; u++) {
            if (u >= g4) break;
            const int x4 = x4g + u;
            __syncthreads();
            if (wave & 1)
#pragma unroll
                for (int r = 0; r < 4; r++) part[256 * (wave >> 1) + lane + 64 * r] = accA[u][r];
            __syncthreads();
            if (!(wave & 1)) {                 // rows x8 (C1) / x7 (C2), columns a (C1) / c (C2)
                const int j = li;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = MM::row(lane, r);
                    const T v = accA[u][r] + part[256 * (wave >> 1) + lane + 64 * r];
                    if (i < NN && j < NN) {
                        if (!c2w)
                            C1[s + (i64)M * (obase + (i64)N3 * j + (i64)N2 * N2 * x4 + (i64)N2 * N3 * i)] = (double)v;
                        else
                            C2[s + (i64)M * (obase + (i64)N3 * i + (i64)N2 * N2 * x4 + (i64)N2 * N3 * j)] = (double)v;
                    }
                }
            }
        }
    }
    {                                          // C3: rows (a,c) = (x7 = tile, x8), columns b; the two x4
        const int j = li;                      // groups' halves are added to the zeroed output
#pragma unroll
        for (int u = 0; u < TW; u++) {
            const int a = wave + 4 * u;
            if (a >= NN) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int c = MM::row(lane, r);
                if (c < NN && j < NN)
                    unsafeAtomicAdd(&C3[s + (i64)M * (obase + (i64)N3 * a + (i64)N2 * N2 * j + (i64)N2 * N3 * c)],
                                    (double)acc3[u][r]);
            }
        }
    }
}

// pass B: workgroup g = (cp*n + ap)*n + bp; C4[(x7,x8)][b] = sum over chunks (x3,x5) and x6, one tile per
// x7, the tiles spread over the four wavefronts.  (Two workgroups per slab, as in pass A, made this
// streaming pass slower.)
template <typename T, int NN, typename MM = Mfma<T>>
__global__ __launch_bounds__(256) void k_e4_pass_b(const T *__restrict__ E4, const double *__restrict__ B3, int M,
                                                   int s, double *C4, int g0) {
    typedef ChunkLayout<NN> CL;
    constexpr int N2 = NN * NN, N3 = N2 * NN;
    constexpr int KC = (NN + 3) / 4;
    constexpr int TW = (NN + 3) / 4;           // tiles per wavefront
    constexpr int B3S = 4 * KC * NN + 16;
    typedef typename MM::V V;
    __shared__ T ech[CL::SIZE];
    __shared__ T b3s[B3S];
    const int g = g0 + blockIdx.x;
    const int bp = g % NN, ap = (g / NN) % NN, cp = g / N2;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: scalar branches
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const T *slab = E4 + (i64)(g / NN) * N3 * N3 + (i64)bp * N2 * N2;    // E4[cp,ap,0,bp,...]
    const i64 obase = ap + (i64)NN * bp + (i64)N2 * cp;
    for (int e = threadIdx.x; e < CL::SIZE; e += 256) ech[e] = (T)0;
    for (int e = threadIdx.x; e < B3S; e += 256) b3s[e] = (T)0;
    V acc[TW];
#pragma unroll
    for (int u = 0; u < TW; u++) acc[u] = vzero<V>();
    constexpr int PF = (N3 + 255) / 256;       // the next two chunks prefetched into registers
    constexpr int NQ = NN * NN;                // chunks q = x3 * NN + x5
    T pfa[PF], pfb[PF], pba, pbb;
    auto fetch = [&](int q, T (&pf)[PF], T &pb) {
        if (q >= NQ) return;
        const int x3 = q / NN, x5 = q % NN;
        if (threadIdx.x < N2) pb = (T)B3[((i64)x3 * N2 + x5 * NN) * NN + threadIdx.x];
        const T *src = slab + (i64)x3 * N2 * N3 + (i64)x5 * N3;
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int e = threadIdx.x + 256 * u;
            pf[u] = e < N3 ? src[e] : (T)0;
        }
    };
    pba = pbb = (T)0;
    fetch(0, pfa, pba);
    fetch(1, pfb, pbb);
    auto step = [&](int q, T (&pf)[PF], T &pb) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int e = threadIdx.x + 256 * u;
            if (e < N3) ech[CL::at(e / N2, (e / NN) % NN, e % NN)] = pf[u];
        }
        if (threadIdx.x < N2) b3s[threadIdx.x] = pb;
        __syncthreads();
        fetch(q + 2, pf, pb);
        switch (wave) {
            case 0: kloop_c3<T, NN, MM, 0, TW>(ech, b3s, li, lk, acc); break;
            case 1: kloop_c3<T, NN, MM, 1, TW>(ech, b3s, li, lk, acc); break;
            case 2: kloop_c3<T, NN, MM, 2, TW>(ech, b3s, li, lk, acc); break;
            default: kloop_c3<T, NN, MM, 3, TW>(ech, b3s, li, lk, acc); break;
        }
    };
    for (int q = 0; q < NQ; q += 2) {
        step(q, pfa, pba);
        if (q + 1 < NQ) step(q + 1, pfb, pbb);
    }
    const int j = li;
#pragma unroll
    for (int u = 0; u < TW; u++) {
        const int a = wave + 4 * u;
        if (a >= NN) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int c = MM::row(lane, r);
            if (c < NN && j < NN)
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

// C[s + M k] = 0 for k < cnt: one state's entries of a contraction output
__global__ void k_zero_state(double *C, int M, int s, i64 cnt) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < cnt) C[s + (i64)M * k] = 0.0;
}

template <typename T, int NN, typename MM>
static int run_passes(ContractPlan *p, const T *E4, int s, double **outs, hipStream_t st) {
    const unsigned nwg = NN * NN * NN;
    const int nch = stage1_chunks();
    const unsigned chunk = ((nwg + nch - 1) / nch + 7) / 8 * 8;
    const i64 n6 = (i64)nwg * nwg;
    hipLaunchKernelGGL(k_zero_state, dim3((unsigned)((n6 + 255) / 256)), dim3(256), 0, st, outs[2], p->M, s, n6);
    const unsigned nwa = 2 * nwg, chunka = 2 * chunk;    // pass A: two workgroups per slab
    for (unsigned g0 = 0; g0 < nwa; g0 += chunka)
        hipLaunchKernelGGL((k_e4_pass_a<T, NN, MM>), dim3(nwa - g0 < chunka ? nwa - g0 : chunka), dim3(256), 0, st, E4,
                           p->Bk[0], p->Bk[1], p->Bk[2], p->M, s, outs[0], outs[1], outs[2], (int)g0);
    for (unsigned g0 = 0; g0 < nwg; g0 += chunk)
        hipLaunchKernelGGL((k_e4_pass_b<T, NN, MM>), dim3(nwg - g0 < chunk ? nwg - g0 : chunk), dim3(256), 0, st, E4,
                           p->Bk[2], p->M, s, outs[3], (int)g0);
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

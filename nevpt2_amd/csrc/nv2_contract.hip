// The O(n^9) part of bamat: four contractions of E4 = <E E E E> with the active
// ERIs (koopro4QD.F90:4069-4094).  Each is a batched GEMM
//     out[t, m, col] = sum_k E4[atoff[t] + amoff[m] + akoff[k]] * Bk[k, col]
// with t = (cp,ap,bp) batches, m a pair of free E4 indices, k a triple of summed
// indices and col the single free ERI index (N = n <= 16).  Index tables turn
// the permuted tensor views into plain offsets.  Outputs land directly in
// amat's layout (state fastest).
#include "nv2_common.h"
#include "nv2_stage1.h"

namespace nv2 {

struct ContractPlan {
    int n, M;
    i64 *atoff[4], *amoff[4], *akoff[4], *otoff[4], *omoff[4], *ocoff[4];
    double *Bk[4];
};

static const int KC = 64;   // K chunk staged in LDS

template <typename TA, typename TC, int NN>
__global__ __launch_bounds__(256) void k_contract(const TA *__restrict__ A, const i64 *__restrict__ atoff,
                                                  const i64 *__restrict__ amoff, const i64 *__restrict__ akoff,
                                                  const double *__restrict__ Bm, int K, int Mr, i64 rows,
                                                  const i64 *__restrict__ otoff, const i64 *__restrict__ omoff,
                                                  const i64 *__restrict__ ocoff, int M, int s,
                                                  double *__restrict__ out) {
    __shared__ TC Bs[KC * NN];
    __shared__ i64 ks[KC];
    const i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = r < rows;
    const i64 t = valid ? r / Mr : 0, m = valid ? r % Mr : 0;
    const i64 base = valid ? atoff[t] + amoff[m] : 0;
    TC acc[NN];
#pragma unroll
    for (int c = 0; c < NN; c++) acc[c] = (TC)0;
    for (int k0 = 0; k0 < K; k0 += KC) {
        const int kc = min(KC, K - k0);
        for (int e = threadIdx.x; e < kc * NN; e += blockDim.x) Bs[e] = (TC)Bm[(i64)k0 * NN + e];
        for (int e = threadIdx.x; e < kc; e += blockDim.x) ks[e] = akoff[k0 + e];
        __syncthreads();
        if (valid) {
            for (int kk = 0; kk < kc; kk++) {
                const TC a = (TC)A[base + ks[kk]];
#pragma unroll
                for (int c = 0; c < NN; c++) acc[c] += a * Bs[kk * NN + c];
            }
        }
        __syncthreads();
    }
    if (valid) {
        const i64 ob = otoff[t] + omoff[m];
#pragma unroll
        for (int c = 0; c < NN; c++) out[s + (i64)M * (ob + ocoff[c])] = (double)acc[c];
    }
}

// B matrices from atwo (Fortran layout): B1[(d,j,i),a]=w(d,j,i,a) B2[(d,j,i),c]=w(d,j,c,i) B3[(i,d,j),b]=w(d,j,b,i)
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

ContractPlan *contract_plan_create(int n, int M) {
    ContractPlan *p = new ContractPlan();
    p->n = n; p->M = M;
    const i64 N = n, n2 = N * N, n3 = n2 * N;
    i64 s[9];
    s[8] = 1;
    for (int k = 7; k >= 1; k--) s[k] = s[k + 1] * N;    // s[k] = n^(8-k)
    std::vector<i64> at(n3), am(n2), ak(n3), ot(n3), om(n2), oc(N);
    auto up = [](std::vector<i64> &h) {
        i64 *d = nullptr;
        if (hipMalloc(&d, sizeof(i64) * h.size()) != hipSuccess) return (i64 *)nullptr;
        hipMemcpy(d, h.data(), sizeof(i64) * h.size(), hipMemcpyHostToDevice);
        return d;
    };
    for (int c = 0; c < 4; c++) {
        for (i64 R = 0; R < N; R++)
            for (i64 P = 0; P < N; P++)
                for (i64 Q = 0; Q < N; Q++) {
                    i64 t = (R * N + P) * N + Q;
                    at[t] = R * s[1] + P * s[2] + (c == 3 ? Q * s[4] : Q * s[3]);
                    ot[t] = P + N * Q + n2 * R;
                }
        for (i64 x = 0; x < N; x++)
            for (i64 y = 0; y < N; y++) {
                i64 m = x * N + y;
                switch (c) {
                    case 0: am[m] = x * s[4] + y * s[8]; om[m] = n2 * n2 * x + n2 * n3 * y; break;  // (b,c)
                    case 1: am[m] = x * s[4] + y * s[7]; om[m] = n2 * n2 * x + n3 * y; break;       // (b,a)
                    default: am[m] = x * s[7] + y * s[8]; om[m] = n3 * x + n2 * n3 * y; break;      // (a,c)
                }
            }
        for (i64 x0 = 0; x0 < N; x0++)
            for (i64 x1 = 0; x1 < N; x1++)
                for (i64 x2 = 0; x2 < N; x2++) {
                    i64 k = (x0 * N + x1) * N + x2;
                    switch (c) {
                        case 0: ak[k] = x0 * s[5] + x1 * s[6] + x2 * s[7]; break;   // (d,j,i)
                        case 1: ak[k] = x0 * s[5] + x1 * s[6] + x2 * s[8]; break;   // (d,j,i), i at x8
                        case 2: ak[k] = x0 * s[4] + x1 * s[5] + x2 * s[6]; break;   // (i,d,j), i at x4
                        default: ak[k] = x0 * s[3] + x1 * s[5] + x2 * s[6]; break;  // (i,d,j), i at x3
                    }
                }
        for (i64 col = 0; col < N; col++) {
            switch (c) {
                case 0: oc[col] = n3 * col; break;         // a
                case 1: oc[col] = n2 * n3 * col; break;    // c
                default: oc[col] = n2 * n2 * col; break;   // b
            }
        }
        p->atoff[c] = up(at); p->amoff[c] = up(am); p->akoff[c] = up(ak);
        p->otoff[c] = up(ot); p->omoff[c] = up(om); p->ocoff[c] = up(oc);
    }
    for (int c = 0; c < 3; c++) {
        if (hipMalloc(&p->Bk[c], sizeof(double) * n3 * N) != hipSuccess) p->Bk[c] = nullptr;
    }
    p->Bk[3] = p->Bk[2];
    return p;
}

void contract_plan_destroy(ContractPlan *p) {
    if (!p) return;
    for (int c = 0; c < 4; c++) {
        hipFree(p->atoff[c]); hipFree(p->amoff[c]); hipFree(p->akoff[c]);
        hipFree(p->otoff[c]); hipFree(p->omoff[c]); hipFree(p->ocoff[c]);
    }
    for (int c = 0; c < 3; c++) hipFree(p->Bk[c]);
    delete p;
}

template <typename TA, typename TC, int NN>
static int run4(ContractPlan *p, const TA *E4, int s, double **outs, hipStream_t st) {
    const i64 n = p->n, rows = n * n * n * n * n;
    const int bs = 256;
    for (int c = 0; c < 4; c++) {
        hipLaunchKernelGGL((k_contract<TA, TC, NN>), dim3((unsigned)((rows + bs - 1) / bs)), dim3(bs), 0, st, E4,
                           p->atoff[c], p->amoff[c], p->akoff[c], p->Bk[c], (int)(n * n * n), (int)(n * n), rows,
                           p->otoff[c], p->omoff[c], p->ocoff[c], p->M, s, outs[c]);
        NV2_LAUNCH_CHECK();
    }
    return 0;
}

template <typename TA, typename TC>
static int dispatch(ContractPlan *p, const TA *E4, int s, double **outs, hipStream_t st) {
    switch (p->n) {
#define CASE(k) case k: return run4<TA, TC, k>(p, E4, s, outs, st);
        CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13)
        CASE(14) CASE(15) CASE(16)
#undef CASE
        default: set_error("E4 contraction supports 2 <= nact <= 16 (got %d)", p->n); return -2;
    }
}

int launch_e4_contractions(ContractPlan *p, const void *E4, int prec, const double *atwo, int s, double *C1,
                           double *C2, double *C3, double *C4, hipStream_t st) {
    const i64 n = p->n, n4 = n * n * n * n;
    hipLaunchKernelGGL(k_make_bk, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, p->n, atwo, p->Bk[0],
                       p->Bk[1], p->Bk[2]);
    NV2_LAUNCH_CHECK();
    double *outs[4] = {C1, C2, C3, C4};
    if (prec == 0) return dispatch<double, double>(p, (const double *)E4, s, outs, st);
    return dispatch<float, float>(p, (const float *)E4, s, outs, st);
}

}  // namespace nv2

// CI vectors in the reference's file04 determinant form -> the string-ordered arrays of nv2_ctx_set_ci.
//
// file04 (koopro4QD.F90:147-176) lists determinant J as excitations of the closed-shell reference: holes
// trou(nd(J)+k) and particles part(nd(J)+k), k = 1..ne(J), spin-orbitals numbered 1..norb (one spin) and
// norb+1..2norb (the other); particle 2 norb + 1 is the fictitious orbital of an ion, i.e. the electron is
// removed (koopro4QD.F90:188-204).  The determinant is the product of the pairs a+_part a_trou in listed
// order applied to the reference, the convention whose relative signs Jdifgen computes by counting the
// crossings of the two lists (koopro4QD.F90:3763-3886).  Here the pairs are applied to an occupation bit
// string with the Jordan-Wigner sign of each operator (spin-orbitals in ascending order), which gives the
// determinant's phase relative to its ascending-order form; that form, restricted to the active orbitals
// (alpha = spin-orbitals 1..norb, beta = norb+1..2norb), is what the RDM builders index.  Core orbitals are
// closed and virtual ones empty in every determinant, which leaves one common sign for the whole vector.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/nevpt2_amd.h"

namespace nv2 {
void set_error(const char *fmt, ...);

namespace {

struct Bits {
    std::vector<uint64_t> w;
    explicit Bits(int64_t n) : w((size_t)((n + 63) / 64), 0) {}
    bool get(int64_t k) const { return (w[(size_t)(k >> 6)] >> (k & 63)) & 1u; }
    void flip(int64_t k) { w[(size_t)(k >> 6)] ^= (uint64_t)1 << (k & 63); }
    int below(int64_t k) const {        // occupied spin-orbitals with index < k
        int c = 0;
        for (int64_t q = 0; q < (k >> 6); q++) c += __builtin_popcountll(w[(size_t)q]);
        if (k & 63) c += __builtin_popcountll(w[(size_t)(k >> 6)] & (((uint64_t)1 << (k & 63)) - 1));
        return c;
    }
};

int64_t binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    int64_t r = 1;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

// rank of an ascending k-subset of range(n) in lexicographic order (itertools.combinations order)
int64_t lex_rank(const int *occ, int k, int n) {
    int64_t r = 0;
    int prev = -1;
    for (int i = 0; i < k; i++) {
        for (int v = prev + 1; v < occ[i]; v++) r += binom(n - v - 1, k - i - 1);
        prev = occ[i];
    }
    return r;
}

}  // namespace
}  // namespace nv2

using namespace nv2;

extern "C" int nv2_ci_from_file04(int64_t norb, int64_t ncore, int64_t nact, int64_t nocc, int64_t metat, int64_t ncf,
                                  const int16_t *ne, const int64_t *nd, const int16_t *trou, const int16_t *part,
                                  const double *c, double *ci, int64_t *nalpha, int64_t *nbeta) {
    if (norb < 1 || ncore < 0 || nact < 1 || nact > 30 || ncore + nact > norb || nocc < ncore ||
        nocc > ncore + nact || metat < 1 || ncf < 1 || !ne || !nd || !c || !nalpha || !nbeta) {
        set_error("nv2_ci_from_file04: inconsistent dimensions (norb=%lld ncore=%lld nact=%lld nocc=%lld)",
                  (long long)norb, (long long)ncore, (long long)nact, (long long)nocc);
        return -2;
    }
    const int64_t nso = 2 * norb;
    const int n = (int)nact;
    int na = -1, nb = -1;
    int64_t nsa = 0, nsb = 0;
    std::vector<int> oa(n), ob(n);
    for (int pass = 0; pass < 2; pass++) {        // pass 0: electron counts; pass 1: scatter
        for (int64_t J = 0; J < ncf; J++) {
            Bits d(nso);
            for (int64_t q = 0; q < nocc; q++) { d.flip(q); d.flip(norb + q); }
            int sign = 1;
            for (int k = 0; k < ne[J]; k++) {
                if (!trou || !part) { set_error("nv2_ci_from_file04: excitations without trou/part"); return -2; }
                const int64_t t = (int64_t)trou[nd[J] + k] - 1, p = (int64_t)part[nd[J] + k] - 1;
                if (t < 0 || t >= nso || !d.get(t)) {
                    set_error("nv2_ci_from_file04: determinant %lld: hole %lld is not occupied", (long long)J + 1,
                              (long long)t + 1);
                    return -2;
                }
                if (d.below(t) & 1) sign = -sign;  // a_t
                d.flip(t);
                if (p == nso) continue;            // fictitious orbital of an ion: the electron leaves
                if (p < 0 || p > nso || d.get(p)) {
                    set_error("nv2_ci_from_file04: determinant %lld: particle %lld is not empty", (long long)J + 1,
                              (long long)p + 1);
                    return -2;
                }
                if (d.below(p) & 1) sign = -sign;  // a+_p
                d.flip(p);
            }
            // core closed, virtual empty, then the active strings
            int ka = 0, kb = 0;
            for (int64_t q = 0; q < norb; q++) {
                const bool inact = q >= ncore && q < ncore + nact;
                const bool a = d.get(q), b = d.get(norb + q);
                if (!inact) {
                    if (a != (q < ncore) || b != (q < ncore)) {
                        set_error("nv2_ci_from_file04: determinant %lld changes the occupation of orbital %lld outside "
                                  "the active space", (long long)J + 1, (long long)q + 1);
                        return -2;
                    }
                    continue;
                }
                if (a) oa[ka++] = (int)(q - ncore);
                if (b) ob[kb++] = (int)(q - ncore);
            }
            if (pass == 0) {
                if (na < 0) { na = ka; nb = kb; }
                if (ka != na || kb != nb) {
                    set_error("nv2_ci_from_file04: determinant %lld has %d/%d active electrons per spin, not %d/%d",
                              (long long)J + 1, ka, kb, na, nb);
                    return -2;
                }
                continue;
            }
            const int64_t ia = lex_rank(oa.data(), na, n), ib = lex_rank(ob.data(), nb, n);
            for (int64_t s = 0; s < metat; s++)
                ci[(s * nsa + ia) * nsb + ib] = sign * c[J + ncf * s];
        }
        if (pass == 0) {
            *nalpha = na;
            *nbeta = nb;
            if (!ci) return 0;                 // size query: the caller allocates metat*C(nact,na)*C(nact,nb)
            nsa = binom(n, na);
            nsb = binom(n, nb);
            memset(ci, 0, sizeof(double) * (size_t)(metat * nsa * nsb));
        }
    }
    *nalpha = na;
    *nbeta = nb;
    return 0;
}

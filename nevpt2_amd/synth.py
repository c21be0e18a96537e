"""Seeded synthetic NEVPT2 inputs (CPU, numpy).

The NEVPT2 hot path consumes spin-free 1-4 particle reduced density matrices
(RDMs), Cholesky-factored two-electron integrals and a one-electron matrix.
The reference builds RDMs from a determinant CI vector (``rofour``/``ro3``...,
koopro4QD.F90:921-3014) and reads integrals from a Molcas ``ijkl.h5``; both are
outside the hot path.  This module generates the same *kind* of inputs from a
seed, in the reference's array conventions:

* RDM convention (pinned by the accessor closed forms, eexx_functions.F90:125-133):
  ``G_k(p1..pk; q1..qk) = sum_sigma <a+_{p1 s1} .. a+_{pk sk} a_{qk sk} .. a_{q1 s1}>``,
  i.e. creation index ``p_i`` is paired with annihilation index ``q_i``.
* Packed 4-RDM ``ro4(nwords, n, n, n, n)`` with the descending-sorted creation
  4-tuple addressed by ``lindice`` (indices_norm.F90:41-44, 64-72) and the
  annihilation partners permuted alongside (ord8, ord_utils.F90:71-129).
* Cholesky vectors ``B(L, norb(norb+1)/2)``, pair column ``indice(p,q)``
  (ijkl_utils.F90:257, 676-677); one-electron integrals packed the same way.

RDMs are built exactly (no sampling) from the CI vector by spin-block string
algebra: for ``a`` alpha and ``b`` beta annihilators the matrix
``W[(K_a,K_b),(Q_a,Q_b)] = <K_a K_b| B_{Q_b} A_{Q_a} |Psi>`` over sorted
annihilation tuples gives the spin block ``D = W_bra^T W_ket``; the spin-free
RDM is the sum of the antisymmetric expansions of the blocks over all spin
patterns (each pattern's reordering sign cancels between creators and
annihilators).
"""
from __future__ import annotations

from itertools import combinations, permutations
from math import comb

import numpy as np
import scipy.sparse as sp


# --------------------------------------------------------------------------
# determinant strings
# --------------------------------------------------------------------------
def strings(n: int, k: int):
    """All k-electron occupation strings of n orbitals (sorted tuples, lexicographic)."""
    if k < 0 or k > n:
        return []
    return list(combinations(range(n), k))


def annihilation_operator(n: int, nel: int, a: int) -> sp.csr_matrix:
    """Sparse map ``A[(K, Q), I]`` for ``a_{q_a} .. a_{q_1} |I>`` with Q sorted ascending.

    Rows are ``K * nQ + Q`` with K an (nel-a)-string; the entry is the fermionic
    sign when ``K = I minus Q`` and zero otherwise.
    """
    I_list = strings(n, nel)
    K_list = strings(n, nel - a)
    Q_list = strings(n, a)
    if not I_list or not K_list or not Q_list:
        return sp.csr_matrix((max(len(K_list), 1) * max(len(Q_list), 1), max(len(I_list), 1)))
    kidx = {s: i for i, s in enumerate(K_list)}
    nq = len(Q_list)
    rows, cols, vals = [], [], []
    for ii, I in enumerate(I_list):
        for qi, Q in enumerate(combinations(I, a)):
            # combinations(I, a) enumerates subsets of I in lexicographic order; map to global Q index
            sign = 1
            for j, q in enumerate(Q):
                cnt = I.index(q) - j          # occupied orbitals below q once q_1..q_{j-1} are gone
                if cnt & 1:
                    sign = -sign
            K = tuple(o for o in I if o not in Q)
            rows.append(kidx[K] * nq + _comb_rank(Q, n))
            cols.append(ii)
            vals.append(float(sign))
    return sp.csr_matrix((vals, (rows, cols)), shape=(len(K_list) * nq, len(I_list)))


_RANK_CACHE: dict = {}


def _comb_rank(Q, n):
    key = (n, len(Q))
    tab = _RANK_CACHE.get(key)
    if tab is None:
        tab = {s: i for i, s in enumerate(combinations(range(n), len(Q)))}
        _RANK_CACHE[key] = tab
    return tab[tuple(Q)]


def antisym_expander(n: int, a: int) -> np.ndarray:
    """U[p_1..p_a (flattened, C order), Q] = sign of the sort p -> Q, zero on repeats."""
    Q_list = strings(n, a)
    U = np.zeros((n ** a, max(len(Q_list), 1)))
    if a == 0:
        U[0, 0] = 1.0
        return U
    for qi, Q in enumerate(Q_list):
        for perm in permutations(range(a)):
            p = tuple(Q[i] for i in perm)
            # parity of perm
            inv = sum(1 for i in range(a) for j in range(i + 1, a) if perm[i] > perm[j])
            flat = 0
            for x in p:
                flat = flat * n + x
            U[flat, qi] = -1.0 if inv & 1 else 1.0
    return U


# --------------------------------------------------------------------------
# CI vectors
# --------------------------------------------------------------------------
def random_ci(n: int, nel: int, seed: int, nroots: int = 1, keep: int | None = None) -> np.ndarray:
    """Seeded N(0,1) CI vectors over all Sz=0 determinants of CAS(nel, n).

    Returns ``C[root, I_alpha, I_beta]``, normalised; with ``keep`` only the
    ``keep`` largest-|c| determinants survive before renormalisation; several
    roots are Gram-Schmidt orthonormalised in root order.
    """
    if nel % 2:
        raise ValueError("Sz=0 generator needs an even electron count")
    na = comb(n, nel // 2)
    rng = np.random.Generator(np.random.Philox(seed))
    C = rng.standard_normal((nroots, na, na))
    if keep is not None and keep < na * na:
        for r in range(nroots):
            flat = C[r].ravel()
            cut = np.argpartition(np.abs(flat), flat.size - keep)[: flat.size - keep]
            flat[cut] = 0.0
    for r in range(nroots):
        v = C[r].ravel()
        for s in range(r):
            u = C[s].ravel()
            v -= (u @ v) * u
        v /= np.linalg.norm(v)
    return C


# --------------------------------------------------------------------------
# spin-free RDMs
# --------------------------------------------------------------------------
class _Blocks:
    """Cache of string annihilation operators for one CAS(nel, n), Sz=0."""

    def __init__(self, n: int, nel: int):
        self.n, self.nel = n, nel
        self.nalpha = nel // 2
        self._A = {}
        self._U = {}

    def A(self, a):
        if a not in self._A:
            self._A[a] = annihilation_operator(self.n, self.nalpha, a)
        return self._A[a]

    def U(self, a):
        if a not in self._U:
            self._U[a] = antisym_expander(self.n, a)
        return self._U[a]

    def W(self, C, a, b):
        """W[(K_a K_b), (Q_a Q_b)] for one CI vector C[I_a, I_b]."""
        n, nal = self.n, self.nalpha
        if a > nal or b > nal:
            return None
        Aa, Ab = self.A(a), self.A(b)
        nKa, nQa = comb(n, nal - a), comb(n, a)
        nKb, nQb = comb(n, nal - b), comb(n, b)
        X = Aa @ C                      # (Ka Qa, Ib)
        X = (Ab @ X.T).T                # (Ka Qa, Kb Qb)
        X = np.asarray(X).reshape(nKa, nQa, nKb, nQb).transpose(0, 2, 1, 3)
        return X.reshape(nKa * nKb, nQa * nQb)

    def block(self, Cbra, Cket, a, b):
        """Antisymmetric spin block T_ab[p_a.. p'_b.. ; q_a.. q'_b..] as an n^k x n^k matrix."""
        Wb = self.W(Cbra, a, b)
        if Wb is None:
            return None
        Wk = Wb if Cket is Cbra else self.W(Cket, a, b)
        D = Wb.T @ Wk
        Ua, Ub = self.U(a), self.U(b)
        UU = np.kron(Ua, Ub)            # n^(a+b) x (nQa nQb)
        return UU @ D @ UU.T


def spinfree_rdm(k: int, Cbra: np.ndarray, Cket: np.ndarray | None = None, blocks: _Blocks | None = None,
                 n: int | None = None, nel: int | None = None) -> np.ndarray:
    """Spin-free k-RDM ``<bra| a+_{p1}..a+_{pk} a_{qk}..a_{q1} |ket>`` as an array [p1..pk, q1..qk]."""
    if blocks is None:
        blocks = _Blocks(n, nel)
    n = blocks.n
    if Cket is None:
        Cket = Cbra
    out = np.zeros((n,) * (2 * k))
    for a in range(k + 1):
        b = k - a
        T = blocks.block(Cbra, Cket, a, b)
        if T is None:
            continue
        T = T.reshape((n,) * (2 * k))
        for S in combinations(range(k), a):
            Sc = tuple(i for i in range(k) if i not in S)
            order = S + Sc                       # T's creator slots hold positions S then Sc
            # T axis t (creator) corresponds to position order[t]; inverse permutation for transpose
            axes = [0] * k
            for t, pos in enumerate(order):
                axes[pos] = t
            axes = axes + [k + x for x in axes]
            out += T.transpose(axes)
    return out


# --------------------------------------------------------------------------
# packed 4-RDM and the reference's index helpers (indices_norm.F90, ord_utils.F90)
# --------------------------------------------------------------------------
def nwords(n: int) -> int:
    return n * (n + 1) * (n + 2) * (n + 3) // 24


def lindice(i1, i2, i3, i4):
    """1-based packed index of a descending 4-tuple (indices_norm.F90:31-44)."""
    return (i1 - 1) * i1 * (i1 + 1) * (i1 + 2) // 24 + (i2 - 1) * i2 * (i2 + 1) // 6 + (i3 - 1) * i3 // 2 + i4


def pack_ro4(G4: np.ndarray) -> np.ndarray:
    """Full Gamma4[p1..p4, q1..q4] (0-based) -> ro4[nwords, n, n, n, n] (Fortran layout, 1 state)."""
    n = G4.shape[0]
    ro4 = np.zeros((nwords(n), n, n, n, n), order="F")
    for i1 in range(1, n + 1):
        for i2 in range(1, i1 + 1):
            for i3 in range(1, i2 + 1):
                for i4 in range(1, i3 + 1):
                    ro4[lindice(i1, i2, i3, i4) - 1] = G4[i1 - 1, i2 - 1, i3 - 1, i4 - 1]
    return ro4


# --------------------------------------------------------------------------
# integrals
# --------------------------------------------------------------------------
def tri_index(p, q):
    """0-based position of the pair (p,q) (0-based orbitals) in the reference's packed triangle."""
    hi, lo = max(p, q) + 1, min(p, q) + 1
    return hi * (hi - 1) // 2 + lo - 1


def random_integrals(ncore: int, nact: int, nvirt: int, seed: int, rank_factor: int = 3):
    """Seeded Cholesky vectors and one-electron matrix.

    ``(pq|rs) = sum_L B^L_pq B^L_rs`` with rank ``L = rank_factor * norb`` and symmetric
    ``B_pq ~ N(0,1) exp(-|p-q|/8) / sqrt(norb)``; ``h`` = symmetric N(0, 0.3) plus diagonal
    orbital energies (inactive U[-20,-0.5], active U[-0.5,0.5], virtual U[0.1,5]).
    Returns ``(B[L, npair] Fortran-ordered, oneint[npair], eps[norb])``.
    """
    norb = ncore + nact + nvirt
    L = rank_factor * norb
    rng = np.random.Generator(np.random.Philox(seed + 7919))
    p = np.arange(norb)
    damp = np.exp(-np.abs(p[:, None] - p[None, :]) / 8.0) / np.sqrt(norb)
    npair = norb * (norb + 1) // 2
    B = np.zeros((L, npair), order="F")
    hi, lo = np.tril_indices(norb)           # hi >= lo, row-major over hi: matches indice ordering
    cols = hi * (hi + 1) // 2 + lo
    for l0 in range(0, L, 64):
        l1 = min(L, l0 + 64)
        X = rng.standard_normal((l1 - l0, npair))
        B[l0:l1, cols] = X * damp[hi, lo][None, :]
    h = rng.normal(0.0, 0.3, size=(norb, norb))
    h = 0.5 * (h + h.T)
    eps = np.concatenate([rng.uniform(-20.0, -0.5, ncore), rng.uniform(-0.5, 0.5, nact),
                          rng.uniform(0.1, 5.0, nvirt)])
    h[p, p] += eps
    oneint = np.zeros(npair)
    oneint[cols] = h[hi, lo]
    return B, oneint, eps


def eri_full(B: np.ndarray, norb: int, orbs=None) -> np.ndarray:
    """Dense (pq|rs) over a subset of orbitals (0-based), from packed Cholesky vectors."""
    if orbs is None:
        orbs = np.arange(norb)
    orbs = np.asarray(orbs)
    m = len(orbs)
    idx = np.empty((m, m), dtype=np.int64)
    for i, p in enumerate(orbs):
        for j, q in enumerate(orbs):
            idx[i, j] = tri_index(p, q)
    Bs = B[:, idx.ravel()].reshape(B.shape[0], m, m)
    return np.einsum("lpq,lrs->pqrs", Bs, Bs, optimize=True)


# --------------------------------------------------------------------------
# one-call problem generator
# --------------------------------------------------------------------------
class Problem:
    """A seeded single- or multi-root NEVPT2 input set in the reference's layouts.

    Attributes (Fortran-ordered numpy arrays, state index first except ro4):
      ncore, nact, nvirt, norb, nele, metat
      ro4 (nwords,n,n,n,n,metat), daaa (M,n^6), taa (M,n^4), dal (M,n,n)
      B (L, npair), oneint (npair,)
      transition RDMs (ncoppie, ...) when metat > 1: daaaoff, taaoff, daloff
    """

    def __init__(self, nact, ncore, nvirt, seed=1, nroots=1, keep=None, rank_factor=3, nele=None, host_rdms=True,
                 ci=None):
        """host_rdms=False keeps only the CI vectors ``C`` (the RDMs are then built on the GPU from them,
        ``Evaluator.for_problem(P, rdms="gpu")``); ro4 and the lower RDMs are None.  ``ci`` (C[root, I_alpha,
        I_beta], orthonormal) replaces the seeded random CI vectors, e.g. by eigenvectors of the active
        Hamiltonian; the integrals still come from ``seed``."""
        from . import rdmtrace
        self.nact, self.ncore, self.nvirt = nact, ncore, nvirt
        self.norb = nact + ncore + nvirt
        self.nele = nact if nele is None else nele
        self.metat = nroots
        self.seed = seed
        n, M = nact, nroots
        if ci is None:
            self.C = random_ci(n, self.nele, seed, nroots=nroots, keep=keep)
        else:
            self.C = np.asarray(ci, dtype=np.float64).reshape(nroots, comb(n, self.nele // 2), -1)
        self.nalpha = self.nbeta = self.nele // 2
        self.ncoppie = M * (M - 1) // 2
        self.B, self.oneint, self.eps0 = random_integrals(ncore, nact, nvirt, seed, rank_factor)
        self.nchovec = self.B.shape[0]
        if not host_rdms:
            self.ro4 = self.daaa = self.taa = self.dal = None
            self.daaaoff = self.taaoff = self.daloff = None
            return
        blk = _Blocks(n, self.nele)
        self.ro4 = np.zeros((nwords(n), n, n, n, n, M), order="F")
        for s in range(M):
            G4 = spinfree_rdm(4, self.C[s], blocks=blk)
            self.ro4[..., s] = pack_ro4(G4)
            del G4
        self.daaa, self.taa, self.dal = rdmtrace.traces_from_ro4(self.ro4, n, self.nele)
        if self.ncoppie:
            P = self.ncoppie
            self.daaaoff = np.zeros((P,) + (n,) * 6, order="F")
            ij = 0
            for i in range(M):
                for j in range(i):
                    # pair (i, j), i > j, in the reference's ijcou order (koopro4QD.F90:879-881);
                    # bra = state i, ket = state j: the class drivers pair daoff(ijcou,p,q) of
                    # (istate > jstate) with <istate|E_pq|jstate> (e.g. qdnevpt.F:1912-1918, 3395-3406)
                    self.daaaoff[ij] = spinfree_rdm(3, self.C[i], self.C[j], blocks=blk)
                    ij += 1
            self.taaoff = rdmtrace.bro2(self.daaaoff, n, self.nele)
            self.daloff = rdmtrace.bro1(self.taaoff, n, self.nele)

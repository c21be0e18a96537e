"""Fock-space brute force: NEVPT2 and the Koopmans matrices from explicit operators.

Oracle (test infrastructure only).  Nothing here uses an RDM, an accessor
closed form or any routine of ``oracle.stage1``/``oracle.stage2``: states are
vectors over occupation-number bit strings, operators are sparse matrices
(Jordan-Wigner), and every quantity is computed from its definition.  It pins
the restatements in two ways:

* ``KoopmansBrute``: each Koopmans matrix of koop_matrices.F90 and
  koopro4QD.F90:3974-4459 as a spin-summed expectation value
  ``<Psi| O_bra^+ [H_act, O_ket] |Psi>`` of the active Hamiltonian, for any
  (not necessarily eigen-) state Psi of the active space.
* ``nevpt2``: SC- and PC-NEVPT2 per perturber class for a CASCI eigenfunction
  Psi_0 of the Dyall Hamiltonian, from the first-order functions themselves
  (Angeli, Cimiraglia, Malrieu, JCP 117, 9138 (2002), cited at
  qdnevpt_main.F90:29-41): the SC function of an external label set is the
  projection of H|Psi_0> onto the determinants with that core/virtual
  occupation; the PC space is spanned by the spin-free excitations of that
  class with those external labels applied to Psi_0, orthonormalised with the
  same linear-dependence threshold and diagonalised under H_D.

Conventions: spin-orbital k < norb is alpha orbital k, k >= norb is beta
orbital k - norb; bit k of a state index is its occupation; a determinant is
a+_{k1} a+_{k2} ... |0> with k ascending, so a_k carries the sign
(-1)^(occupied spin-orbitals below k).  E_pq = sum_sigma a+_{p sigma} a_{q sigma}.
Orbitals 0..nc-1 are core, nc..nc+n-1 active, the rest virtual.
"""
from __future__ import annotations

from itertools import combinations

import numpy as np
import scipy.sparse as sp

CLASS_NAMES = ["(0)", "(+1)", "(-1)", "(+2)", "(-2)", "(+1)'", "(-1)'", "(0)'"]


def _popcount(x):
    x = np.asarray(x, dtype=np.int64)
    c = np.zeros_like(x)
    while np.any(x):
        c += x & 1
        x = x >> 1
    return c


class FockSpace:
    """Sparse second-quantised operators over the states ``basis`` (bit strings of 2*norb spin-orbitals).

    ``basis`` defaults to the whole Fock space; operators that change the particle number are only
    available on the whole space.  Number- and Ms-conserving operators (E_pq) can be built on any
    subset closed under them (``sector``).
    """

    def __init__(self, norb: int, basis=None):
        self.norb = norb
        self.nso = 2 * norb
        if basis is None:
            basis = np.arange(1 << self.nso, dtype=np.int64)
        self.basis = np.asarray(basis, dtype=np.int64)
        self.dim = len(self.basis)
        assert np.all(np.diff(self.basis) > 0)
        self.index = {int(b): i for i, b in enumerate(self.basis)}
        self._ex = {}

    @classmethod
    def sector(cls, norb: int, nalpha: int, nbeta: int):
        """All determinants with nalpha alpha and nbeta beta electrons."""
        out = []
        for sa in combinations(range(norb), nalpha):
            ba = sum(1 << p for p in sa)
            for sb in combinations(range(norb), nbeta):
                out.append(ba | sum(1 << (norb + p) for p in sb))
        return cls(norb, np.array(sorted(out), dtype=np.int64))

    # ------------------------------------------------------------------ elementary operators
    def ann(self, k: int) -> sp.csr_matrix:
        """a_k on the whole Fock space."""
        assert self.dim == 1 << self.nso
        src = self.basis[(self.basis >> k) & 1 == 1]
        dst = src & ~(1 << k)
        sign = np.where(_popcount(src & ((1 << k) - 1)) & 1, -1.0, 1.0)
        return sp.csr_matrix((sign, (dst, src)), shape=(self.dim, self.dim))

    def cre(self, k: int) -> sp.csr_matrix:
        return self.ann(k).T.tocsr()

    def ex(self, k: int, l: int) -> sp.csr_matrix:
        """a+_k a_l (spin-orbitals) on ``basis``."""
        key = (k, l)
        if key in self._ex:
            return self._ex[key]
        b = self.basis
        has_l = (b >> l) & 1 == 1
        src = b[has_l]
        mid = src & ~(1 << l)
        ok = (mid >> k) & 1 == 0
        src, mid = src[ok], mid[ok]
        dst = mid | (1 << k)
        par = _popcount(src & ((1 << l) - 1)) + _popcount(mid & ((1 << k) - 1))
        sign = np.where(par & 1, -1.0, 1.0)
        rows = np.searchsorted(self.basis, dst)
        cols = np.searchsorted(self.basis, src)
        m = sp.csr_matrix((sign, (rows, cols)), shape=(self.dim, self.dim))
        self._ex[key] = m
        return m

    def E(self, p: int, q: int) -> sp.csr_matrix:
        """Spin-free E_pq = a+_{p alpha} a_{q alpha} + a+_{p beta} a_{q beta}."""
        n = self.norb
        return (self.ex(p, q) + self.ex(n + p, n + q)).tocsr()

    def hamiltonian(self, orbs, k1: np.ndarray, g: np.ndarray) -> sp.csr_matrix:
        """sum_pq k1_pq E_pq + 1/2 sum_pqrs g_pqrs E_pq E_rs over the orbitals ``orbs`` (indices into
        k1/g are positions in ``orbs``).  With k1 = h - 1/2 sum_c g_pccq this is the usual
        sum h E + 1/2 sum g (E E - delta E)."""
        orbs = list(orbs)
        m = len(orbs)
        Es = [[self.E(orbs[p], orbs[q]) for q in range(m)] for p in range(m)]
        H = sp.csr_matrix((self.dim, self.dim))
        for p in range(m):
            for q in range(m):
                X = k1[p, q] * sp.identity(self.dim, format="csr")
                for r in range(m):
                    for s in range(m):
                        if g[p, q, r, s] != 0.0:
                            X = X + (0.5 * g[p, q, r, s]) * Es[r][s]
                H = H + Es[p][q] @ X
        return H.tocsr()


def h_eff_normal(h: np.ndarray, g: np.ndarray) -> np.ndarray:
    """k1_pq = h_pq - 1/2 sum_c g_pccq, so that sum h E + 1/2 sum g (E E - delta E) = sum k1 E + 1/2 sum g E E."""
    return h - 0.5 * np.einsum("pccq->pq", g)


# ---------------------------------------------------------------------------- inputs
def dense_integrals(P):
    """h (norb x norb) and (pq|rs) from a synth.Problem's packed one-electron integrals and Cholesky vectors."""
    norb = P.norb
    p = np.arange(norb)
    hi = np.maximum(p[:, None], p[None, :])
    lo = np.minimum(p[:, None], p[None, :])
    tri = hi * (hi + 1) // 2 + lo
    h = np.asarray(P.oneint)[tri]
    Bf = np.asarray(P.B)[:, tri]
    g = np.einsum("lpq,lrs->pqrs", Bf, Bf, optimize=True)
    return h, g


def active_hamiltonian(P, h=None, g=None):
    """Dyall active Hamiltonian integrals: h_eff_ab = h_ab + sum_core [2 (ab|jj) - (aj|jb)], (ab|cd)."""
    if h is None:
        h, g = dense_integrals(P)
    nc, n = P.ncore, P.nact
    A = slice(nc, nc + n)
    C = np.arange(nc)
    heff = h[A, A] + 2.0 * np.einsum("abjj->ab", g[A, A][:, :, C][:, :, :, C]) \
        - np.einsum("ajjb->ab", g[A][:, C][:, :, C][:, :, :, A])
    return heff, g[A, A, A, A].copy()


def spin_squared(C: np.ndarray, n: int, na: int, nb: int) -> float:
    """<S^2> of a CI vector C[I_alpha, I_beta] (synth's string order), S^2 = S_- S_+ + S_z (S_z + 1)."""
    fs = FockSpace(n)
    psi = np.zeros(fs.dim)
    for ia, a in enumerate(combinations(range(n), na)):
        ba = sum(1 << p for p in a)
        for ib, b in enumerate(combinations(range(n), nb)):
            psi[ba | sum(1 << (n + p) for p in b)] = C[ia, ib]
    a = [fs.ann(k) for k in range(2 * n)]
    Sp = sum(a[p].T @ a[n + p] for p in range(n))
    sz = 0.5 * (na - nb)
    v = Sp @ psi
    return float(v @ v) + sz * (sz + 1.0) * float(psi @ psi)


def casci_ground_state(P, nalpha=None, nbeta=None, nroots=None, spin=None):
    """Lowest eigenvector of the active Hamiltonian in the (nalpha, nbeta) sector, as C[I_alpha, I_beta]
    in synth's string order (combinations of range(n), lexicographic; alpha bit p, beta bit n+p).
    Returns (C, E, gap to the next root); with ``nroots`` the lowest nroots vectors C[root, ...] and
    their energies; ``spin`` keeps only roots with <S^2> = spin (spin + 1)."""
    n = P.nact
    na = P.nele // 2 if nalpha is None else nalpha
    nb = P.nele - na if nbeta is None else nbeta
    heff, gaa = active_hamiltonian(P)
    fs = FockSpace.sector(n, na, nb)
    H = fs.hamiltonian(range(n), h_eff_normal(heff, gaa), gaa).toarray()
    w, V = np.linalg.eigh(H)
    sa = list(combinations(range(n), na))
    sb = list(combinations(range(n), nb))
    C = np.zeros((V.shape[1], len(sa), len(sb)))
    for ia, a in enumerate(sa):
        ba = sum(1 << p for p in a)
        for ib, b in enumerate(sb):
            C[:, ia, ib] = V[fs.index[ba | sum(1 << (n + p) for p in b)], :]
    if spin is not None:
        keep = [r for r in range(len(w)) if abs(spin_squared(C[r], n, na, nb) - spin * (spin + 1.0)) < 1e-8]
        C, w = C[keep], w[keep]
    if nroots is not None:
        return C[:nroots], w[:nroots]
    return C[0], float(w[0]), float(w[1] - w[0])


# ---------------------------------------------------------------------------- Koopmans matrices
class KoopmansBrute:
    """Spin-summed <Psi| O_bra^+ [H_act, O_ket] |Psi> on the active Fock space of one CI vector.

    ``C`` is the CI vector in synth's layout; ``heff``/``gaa`` the active Hamiltonian.  ``H`` below is
    sum heff E + 1/2 sum g (E E - delta E), the active part of the Dyall Hamiltonian.
    """

    def __init__(self, C: np.ndarray, n: int, nalpha: int, nbeta: int, heff: np.ndarray, gaa: np.ndarray):
        self.n = n
        fs = FockSpace(n)
        self.fs = fs
        self.a = [fs.ann(k) for k in range(2 * n)]
        self.c = [m.T.tocsr() for m in self.a]
        psi = np.zeros(fs.dim)
        for ia, sa in enumerate(combinations(range(n), nalpha)):
            ba = sum(1 << p for p in sa)
            for ib, sb in enumerate(combinations(range(n), nbeta)):
                psi[ba | sum(1 << (n + p) for p in sb)] = C[ia, ib]
        self.psi = psi
        self.E = [[(self.c[p] @ self.a[q] + self.c[n + p] @ self.a[n + q]).tocsr() for q in range(n)]
                  for p in range(n)]
        k1 = h_eff_normal(heff, gaa)
        H = sp.csr_matrix((fs.dim, fs.dim))
        for p in range(n):
            for q in range(n):
                X = k1[p, q] * sp.identity(fs.dim, format="csr")
                for r in range(n):
                    for s in range(n):
                        X = X + (0.5 * gaa[p, q, r, s]) * self.E[r][s]
                H = H + self.E[p][q] @ X
        self.H = H.tocsr()
        self.Hpsi = self.H @ psi

    def comm_ket(self, ops):
        """[H, O] |Psi> for O = product of sparse operators (rightmost acts first)."""
        v = self.psi
        for op in reversed(ops):
            v = op @ v
        Ov = v
        HOv = self.H @ Ov
        w = self.Hpsi
        for op in reversed(ops):
            w = op @ w
        return HOv - w

    def bra(self, ops):
        """O^+ acting to the left on <Psi|, i.e. the ket O|Psi> (then <bra| = its transpose)."""
        v = self.psi
        for op in reversed(ops):
            v = op @ v
        return v


# ---------------------------------------------------------------------------- full-space NEVPT2
class FullSpace:
    """One N-electron state Psi_0 = |core closed> x |active CI> over core + active + virtual orbitals,
    with the Dyall partition H = H_D + V in the orbitals that diagonalise the core-core and
    virtual-virtual blocks of the Fock operator of Psi_0."""

    def __init__(self, P, Cs, thr=1.0e-6):
        nc, n, nv, norb = P.ncore, P.nact, P.nvirt, P.norb
        self.nc, self.n, self.nv, self.norb, self.thr = nc, n, nv, norb, thr
        na_act = P.nele // 2
        nb_act = P.nele - na_act
        fs = FockSpace.sector(norb, nc + na_act, nc + nb_act)
        self.fs = fs
        core = (1 << nc) - 1
        sa = list(combinations(range(n), na_act))
        sb = list(combinations(range(n), nb_act))
        self.psis = []
        for C in Cs:
            psi = np.zeros(fs.dim)
            for ia, a in enumerate(sa):
                ba = core | (sum(1 << p for p in a) << nc)
                for ib, b in enumerate(sb):
                    bb = (core | (sum(1 << p for p in b) << nc)) << norb
                    psi[fs.index[ba | bb]] = C[ia, ib]
            self.psis.append(psi)
        b = fs.basis
        self.occ = np.stack([((b >> p) & 1) + ((b >> (norb + p)) & 1) for p in range(norb)], axis=1)
        self.h0, self.g0 = dense_integrals(P)
        self.Es = {}
        self.frozen = ()
        self.triangle = "upper"
        # the reference's PC basis order for (+2) and (-2) (pairs (b, a) with b slow) and classes whose K
        # it stores with kets as rows ((+1)' and (-1)': coef(abc, a'b'c') = amat(a',b',c',a,b,c))
        self.variant = {3: 1, 4: 1}
        self.transposed = {5: True, 6: True}

    def E(self, p, q):
        key = (p, q)
        if key not in self.Es:
            self.Es[key] = self.fs.E(p, q)
        return self.Es[key]

    def one_rdm_active(self, psi):
        n, nc = self.n, self.nc
        return np.array([[psi @ (self.E(nc + a, nc + b) @ psi) for b in range(n)] for a in range(n)])

    def setup_state(self, s):
        """Fock operator of state s, its canonical core and virtual orbitals, the rotated integrals,
        H|Psi_s>, the Dyall Hamiltonian and E_0 = <Psi_s|H_D|Psi_s>."""
        nc, n, nv, norb = self.nc, self.n, self.nv, self.norb
        psi = self.psis[s]
        h, g = self.h0, self.g0
        gam = self.one_rdm_active(psi)
        D = np.zeros((norb, norb))
        D[:nc, :nc] = 2.0 * np.eye(nc)
        D[nc:nc + n, nc:nc + n] = gam
        F = h + np.einsum("pqkl,kl->pq", g, D) - 0.5 * np.einsum("pklq,kl->pq", g, D)
        U = np.eye(norb)
        eps = np.diag(F).copy()
        for blk in (slice(0, nc), slice(nc + n, norb)):
            if blk.stop > blk.start:
                w, V = np.linalg.eigh(F[blk, blk])
                U[blk, blk] = V
                eps[blk] = w
        hr = U.T @ h @ U
        gr = np.einsum("pqrs,pa,qb,rc,sd->abcd", g, U, U, U, U, optimize=True)
        # H|Psi> in the rotated orbitals.  Rotating core among core and virtual among virtual orbitals
        # leaves Psi_0 (closed core, empty virtuals) unchanged up to a sign, so the vector is reused.
        self.hr, self.gr = hr, gr
        Hpsi = self.apply_H(psi)
        # Dyall Hamiltonian: orbital energies on core and virtual occupations, the active Hamiltonian
        # with the closed-core effective one-electron operator on the active orbitals.
        A = np.arange(nc, nc + n)
        Cc = np.arange(nc)
        heff = hr[np.ix_(A, A)] + 2.0 * np.einsum("abjj->ab", gr[np.ix_(A, A, Cc, Cc)]) \
            - np.einsum("ajjb->ab", gr[np.ix_(A, Cc, Cc, A)])
        gaa = gr[np.ix_(A, A, A, A)]
        Hact = self.fs.hamiltonian(A, h_eff_normal(heff, gaa), gaa)
        ext = np.zeros(self.fs.dim)
        for p in list(range(nc)) + list(range(nc + n, norb)):
            ext += eps[p] * self.occ[:, p]
        self.HD = (Hact + sp.diags(ext)).tocsr()
        self.E0 = float(psi @ (self.HD @ psi))
        self.residual = float(np.linalg.norm(self.HD @ psi - self.E0 * psi))
        self.psi, self.Hpsi, self.eps, self.gam, self.U = psi, Hpsi, eps, gam, U
        # for a state that is not an eigenfunction of H_D the excitation operators' commutators need
        # H and the excitations applied to H_D|Psi_0> as well: [H_D, O]|Psi> = H_D O|Psi> - O H_D|Psi>
        self.HDpsi = self.HD @ psi
        self.H_HDpsi = self.apply_H(self.HDpsi)
        self.state = s
        return self

    def apply_H(self, x):
        """The full Hamiltonian sum k1 E + 1/2 sum g E E (rotated orbitals) applied to x."""
        norb, hr, gr = self.norb, self.hr, self.gr
        k1 = h_eff_normal(hr, gr)
        v = {(r, q): self.E(r, q) @ x for r in range(norb) for q in range(norb)}
        out = np.zeros(self.fs.dim)
        for p in range(norb):
            for q in range(norb):
                w = k1[p, q] * x
                for r in range(norb):
                    for t in range(norb):
                        if gr[p, q, r, t] != 0.0:
                            w = w + 0.5 * gr[p, q, r, t] * v[(r, t)]
                out += self.E(p, q) @ w
        return out

    # ------------------------------------------------------------------ class definitions
    def sector_mask(self, holes, parts):
        """Determinants whose core orbitals miss the electrons ``holes`` (a multiset of core orbitals) and
        whose virtual orbitals hold ``parts`` (a multiset of virtual orbitals), every other core orbital
        doubly occupied and every other virtual orbital empty."""
        nc, n, norb = self.nc, self.n, self.norb
        want = np.zeros(norb, dtype=np.int64)
        want[:nc] = 2
        for i in holes:
            want[i] -= 1
        for r in parts:
            want[r] += 1
        ext = np.r_[np.arange(nc), np.arange(nc + n, norb)]
        return np.all(self.occ[:, ext] == want[ext][None, :], axis=1)

    def labels(self, k):
        """External label sets of class k; frozen core orbitals (the reference's .FROZEN) are never holes."""
        nc, n, norb = self.nc, self.n, self.norb
        core = [i for i in range(nc) if i not in self.frozen]
        virt = range(nc + n, norb)
        if k == 0:
            return [((i, j), (r, s)) for i in core for j in core if i <= j for r in virt for s in virt if r <= s]
        if k == 1:
            return [((i, j), (r,)) for i in core for j in core if i <= j for r in virt]
        if k == 2:
            return [((i,), (r, s)) for i in core for r in virt for s in virt if r <= s]
        if k == 3:
            return [((i, j), ()) for i in core for j in core if i <= j]
        if k == 4:
            return [((), (r, s)) for r in virt for s in virt if r <= s]
        if k == 5:
            return [((i,), ()) for i in core]
        if k == 6:
            return [((), (r,)) for r in virt]
        return [((i,), (r,)) for i in core for r in virt]

    def pc_functions(self, k, holes, parts, psi=None):
        """The spin-free excitations of class k with these external labels applied to Psi_0 (or psi).

        Returns (functions, blocks): the functions in the order in which the reference indexes its
        PC basis, and the index ranges of the external spin-coupling types whose blocks share one
        active-space matrix (classes (+1) and (-1): the reference solves the active problem once and
        couples the spin types afterwards)."""
        nc, n = self.nc, self.n
        act = range(nc, nc + n)
        psi = self.psi if psi is None else psi
        V = self.variant.get(k, 0)
        if k == 0:
            (i, j), (r, s) = holes, parts
            ops = [[(r, i), (s, j)], [(s, i), (r, j)]]
        elif k == 1:
            (i, j), (r,) = holes, parts
            ops = [[(r, i), (a, j)] for a in act] + ([[(r, j), (a, i)] for a in act] if i != j else [])
        elif k == 2:
            (i,), (r, s) = holes, parts
            ops = [[(r, i), (s, a)] for a in act] + ([[(s, i), (r, a)] for a in act] if r != s else [])
        elif k == 3:
            (i, j) = holes
            if i == j:
                ops = [[(a, i), (b, i)] for a in act for b in act if a <= b]
            elif V == 0:
                ops = [[(a, i), (b, j)] for a in act for b in act]
            else:
                ops = [[(b, i), (a, j)] for a in act for b in act]
        elif k == 4:
            (r, s) = parts
            if r == s:
                ops = [[(r, a), (r, b)] for a in act for b in act if a <= b]
            elif V == 0:
                ops = [[(r, a), (s, b)] for a in act for b in act]
            else:
                ops = [[(r, b), (s, a)] for a in act for b in act]
        elif k == 5:
            (i,) = holes
            ops = [[(b, i), (a, c)] for a in act for b in act for c in act]
        elif k == 6:
            (r,) = parts
            ops = [[(r, b), (a, c)] for a in act for b in act for c in act]
        else:
            (i,), (r,) = holes, parts
            ops = [[(r, i), (a, b)] for a in act for b in act] + [[(a, i), (r, b)] for a in act for b in act]
        out = []
        for prod in ops:
            v = psi
            for p, q in reversed(prod):
                v = self.E(p, q) @ v
            out.append(v)
        m = len(ops)
        blocks = [range(0, n), range(n, m)] if k in (1, 2) and m == 2 * n else [range(0, m)]
        return np.array(out), blocks

    def symmetrise(self, k, Kd, blocks):
        """The symmetric matrix the reference diagonalises: LAPACK reads the upper triangle of its K
        (dsymm 'L','U'), whose rows are bras in its basis order, or kets where it stores K transposed."""
        if self.triangle == "sym":
            return 0.5 * (Kd + Kd.T)
        Kr = Kd.T if self.transposed.get(k, False) else Kd
        out = np.empty_like(Kr)
        for b1 in blocks:
            for b2 in blocks:
                X = Kr[np.ix_(b1, b2)]
                out[np.ix_(b1, b2)] = np.triu(X) + np.triu(X, 1).T
        return out

    def class_energies(self, k):
        """(norm_sc, e_sc, norm_pc, e_pc, min kept / max dropped metric eigenvalue) of class k."""
        e_sc = n_sc = e_pc = n_pc = 0.0
        kept_min, drop_max = np.inf, 0.0
        HD, E0 = self.HD, self.E0
        for holes, parts in self.labels(k):
            mask = self.sector_mask(holes, parts)
            v = np.where(mask, self.Hpsi, 0.0)
            N = float(v @ v)
            if N > 1e-28:
                # <O^+ [H_D, O]>/N with O the part of H that maps the closed-core, empty-virtual space to
                # this sector: O|x> = P_sector H|x> for every x in that space, here x = H_D|Psi_0>
                d = (float(v @ (HD @ v)) - float(v @ np.where(mask, self.H_HDpsi, 0.0))) / N
                e_sc -= N / d
                n_sc += N / d ** 2
            Phi, blocks = self.pc_functions(k, holes, parts)
            S = Phi @ Phi.T
            PhiH, _ = self.pc_functions(k, holes, parts, self.HDpsi)
            Kd = Phi @ (HD @ Phi.T) - Phi @ PhiH.T          # <O_a^+ [H_D, O_b]>, bra a, ket b
            Kd = self.symmetrise(k, Kd, blocks)
            r = Phi @ self.Hpsi
            w, Uw = np.linalg.eigh(S)
            keep = w >= self.thr
            if keep.any():
                kept_min = min(kept_min, float(w[keep].min()))
            if (~keep).any():
                drop_max = max(drop_max, float(np.abs(w[~keep]).max()))
            X = Uw[:, keep] / np.sqrt(w[keep])[None, :]
            lam, V = np.linalg.eigh(X.T @ Kd @ X)
            y = V.T @ (X.T @ r)
            e_pc -= float(np.sum(y * y / lam))
            n_pc += float(np.sum(y * y / lam ** 2))
        return n_sc, e_sc, n_pc, e_pc, kept_min, drop_max

    def class_rho1(self, k):
        """First-order density block of PC class k in {5, 6, 7}: <Psi_0|E_pq|Psi_1> with Psi_1 the class's
        PC first-order function, -sum_mu y_mu / lam_mu Phi_mu over the H_D eigenfunctions Phi_mu of each
        external label's span.  (0)': [i, r] with E_ir; (+1)': [i, d] with E_id; (-1)': [d, r] with E_dr
        (canonical core and virtual orbitals, rows and columns as the .NATORB blocks)."""
        nc, n, nv = self.nc, self.n, self.nv
        out = {5: np.zeros((nc, n)), 6: np.zeros((n, nv)), 7: np.zeros((nc, nv))}[k]
        HD, psi = self.HD, self.psi
        for holes, parts in self.labels(k):
            Phi, blocks = self.pc_functions(k, holes, parts)
            S = Phi @ Phi.T
            PhiH, _ = self.pc_functions(k, holes, parts, self.HDpsi)
            Kd = self.symmetrise(k, Phi @ (HD @ Phi.T) - Phi @ PhiH.T, blocks)
            w, Uw = np.linalg.eigh(S)
            keep = w >= self.thr
            X = Uw[:, keep] / np.sqrt(w[keep])[None, :]
            lam, V = np.linalg.eigh(X.T @ Kd @ X)
            T = V.T @ X.T                                     # Phi_mu = sum_a T[mu, a] Phi_a
            y = T @ (Phi @ self.Hpsi)
            psi1 = -((y / lam) @ T) @ Phi
            if k == 7:
                (i,), (r,) = holes, parts
                out[i, r - nc - n] = float(psi1 @ (self.E(r, i) @ psi))
            elif k == 5:
                (i,) = holes
                for d in range(n):
                    out[i, d] = float(psi1 @ (self.E(nc + d, i) @ psi))
            else:
                (r,) = parts
                for d in range(n):
                    out[d, r - nc - n] = float(psi1 @ (self.E(r, nc + d) @ psi))
        return out

    def class_couplings(self, k):
        """QD-NEVPT2 couplings of class k into column s of Heff (this state's perturbers):
        SC  sum_l <P_l H Psi_j | P_l H Psi_s> / (E_0 - E_l), E_l from state s's SC function;
        PC  sum_mu <H Psi_j|Phi_mu><Phi_mu|H Psi_s> / (E_0 - E_mu) over state s's PC eigenfunctions.
        Returns (sc[j], pc[j]) for every state j (entry s holds the diagonal energies)."""
        M = len(self.psis)
        Hj = [self.Hpsi if j == self.state else self.apply_H(self.psis[j]) for j in range(M)]
        sc, pc = np.zeros(M), np.zeros(M)
        HD = self.HD
        for holes, parts in self.labels(k):
            mask = self.sector_mask(holes, parts)
            v = np.where(mask, self.Hpsi, 0.0)
            N = float(v @ v)
            if N > 1e-28:
                d = (float(v @ (HD @ v)) - float(v @ np.where(mask, self.H_HDpsi, 0.0))) / N
                for j in range(M):
                    sc[j] -= float(np.where(mask, Hj[j], 0.0) @ v) / d
            Phi, blocks = self.pc_functions(k, holes, parts)
            S = Phi @ Phi.T
            PhiH, _ = self.pc_functions(k, holes, parts, self.HDpsi)
            Kd = self.symmetrise(k, Phi @ (HD @ Phi.T) - Phi @ PhiH.T, blocks)
            w, Uw = np.linalg.eigh(S)
            keep = w >= self.thr
            X = Uw[:, keep] / np.sqrt(w[keep])[None, :]
            lam, V = np.linalg.eigh(X.T @ Kd @ X)
            T = V.T @ X.T
            ys = T @ (Phi @ self.Hpsi)
            for j in range(M):
                pc[j] -= float(np.sum((T @ (Phi @ Hj[j])) * ys / lam))
        return sc, pc


def nevpt2(P, s=0, thr=1.0e-6, classes=range(8), frozen=()):
    """Per-class (norm_sc, e_sc, norm_pc, e_pc) of state s, in the reference's class order, plus the
    state's E_0 residual |H_D Psi - E_0 Psi| and the metric eigenvalue gap per class.  ``frozen``:
    canonical core orbitals (0-based, ascending energy) excluded as holes."""
    fsp = FullSpace(P, [P.C[t] for t in range(P.metat)], thr).setup_state(s)
    fsp.frozen = tuple(frozen)
    out = np.zeros((8, 4))
    gaps = {}
    for k in classes:
        r = fsp.class_energies(k)
        out[k] = r[:4]
        gaps[k] = r[4:]
    return out, fsp.residual, gaps


def qd_heff(P, thr=1.0e-6, classes=range(8)):
    """Second-order QD-NEVPT2 Heff without E_0 (column i = perturbers of state i), SC and PC, and the
    per-class contributions off[k] = (SC, PC), as M x M matrices."""
    M = P.metat
    fsp = FullSpace(P, [P.C[t] for t in range(M)], thr)
    Hsc, Hpc = np.zeros((M, M)), np.zeros((M, M))
    off = {}
    for i in range(M):
        fsp.setup_state(i)
        for k in classes:
            sc, pc = fsp.class_couplings(k)
            off.setdefault(k, (np.zeros((M, M)), np.zeros((M, M))))
            off[k][0][:, i] = sc
            off[k][1][:, i] = pc
            Hsc[:, i] += sc
            Hpc[:, i] += pc
    return Hsc, Hpc, off

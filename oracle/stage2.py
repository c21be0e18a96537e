"""Stage-2 oracle: perturber-class energies of SC- and PC-NEVPT2 (single state).

Oracle (test infrastructure only).  Restates qdnevpt.F for one state of a
C1 (nsym=1) Cholesky run, 0-based orbitals: core 0..nc-1, active nc..nc+n-1,
virtual nc+n..norb-1.  Loops over the outer external index follow the
reference; inner index sums are numpy reductions.  Accessors are the stage-2
copies with swapped arguments (qdnevpt.F:6448-6519), built as tensors here.

Parity: every class's SC and PC energy and norm equals SC/PC-NEVPT2 computed
from the first-order functions themselves by Fock-space brute force
(oracle/fock.py, tests/test_oracle_brute.py): for CASCI eigenfunctions at
1e-11 relative, and for random CI vectors in commutator form with LAPACK's
upper triangle of the non-symmetric Koopmans matrix taken in the reference's
basis order (within 1e-10 relative plus 1e-10 times the class norm).  Class
(0) also equals the closed-shell MP2 pair formula (tests/test_stage2_oracle.py).
The reference Fortran itself is not buildable here (hdf5_utils is external).
"""
from __future__ import annotations

import numpy as np

THR_LINDEP = 1.0e-6     # thr_lindep default (nevpt2_cfg.F90)

# timing samples (bench cpu_baseline only): outer external loops run on a fraction of their range
_FRAC = None
_SCALES = []


def _outer(N):
    if _FRAC is None or N == 0:
        return range(N)
    k = max(1, int(np.ceil(_FRAC * N)))
    _SCALES.append(N / k)
    return range(min(N, k))


# ---------------------------------------------------------------------------- helpers
def tri_idx(norb):
    """0-based position of pair (p,q) in the packed triangle, as an (norb,norb) table."""
    p = np.arange(norb)
    hi = np.maximum(p[:, None], p[None, :])
    lo = np.minimum(p[:, None], p[None, :])
    return hi * (hi + 1) // 2 + lo


def unpack_oneint(oneint, norb):
    return np.asarray(oneint)[tri_idx(norb)]


def full_B(B, norb):
    """Cholesky vectors as a dense (L, norb, norb) symmetric array."""
    return np.asarray(B)[:, tri_idx(norb)]


def pc_solve(S, K, thr=THR_LINDEP):
    """Metric orthogonalisation + diagonalisation of the PC Koopmans matrix.

    Restates the common block of every PC class (e.g. qdnevpt.F:1974-2037,
    caux0p :4744-4758): dsyev('V','U') of the metric, drop eigenvalues below
    thr (negatives too), caux = U/sqrt(w), K' = caux^T K caux with K taken from
    its upper triangle (dsymm 'L','U'), dsyev('V','U') of K', coef = caux V.
    Returns (coef[N, nnew], w[nnew], nnew, metric eigenvalues).
    """
    w1, U = np.linalg.eigh(S, UPLO="U")
    keep = w1 >= thr
    caux = U[:, keep] / np.sqrt(w1[keep])[None, :]
    Ks = np.triu(K) + np.triu(K, 1).T
    Kp = caux.T @ (Ks @ caux)
    w2, V = np.linalg.eigh(Kp, UPLO="U")
    return caux @ V, w2, int(keep.sum()), w1


FROZEN_EPS = -1.0e30

# planted mistakes for the tests that show the .NATORB pins catch them (None in every real use)
_RHO_SLIP = None


class Ctx:
    """Per-state quantities shared by the classes.

    ``frozen`` lists canonical core orbitals (0-based, ascending orbital energy) that the reference's
    .FROZEN keyword keeps out of every class's external core loops (``if (zgel(ii)) goto``,
    qdnevpt.F:733-745, e.g. :1831-1840).  They stay in the Fock operator; here their orbital energy is
    set to FROZEN_EPS, which gives every term with a hole in them a denominator of ~1e30, so each
    such term vanishes to rounding -- the skip, without touching each class's loops."""

    def __init__(self, P, koop, s=0, frozen=()):
        nc, n, nv, norb = P.ncore, P.nact, P.nvirt, P.norb
        self.nc, self.n, self.nv, self.norb = nc, n, nv, norb
        self.s = s
        self.da = P.dal[s]
        self.daa = P.taa[s]
        self.ro3 = P.daaa[s]
        self.koop = {k: v[s] for k, v in koop.items() if k not in ("aj", "atwo")}
        h = unpack_oneint(P.oneint, norb)
        Bf = full_B(P.B, norb)
        self.h, self.Bf = h, Bf
        self.fock(P)
        self.transform()
        for i in frozen:
            if not 0 <= i < nc:
                raise ValueError(f"frozen orbital {i} is not a core orbital (ncore={nc})")
            self.eps[i] = FROZEN_EPS

    # ------------------------------------------------------------ Fock + canonicalisation (qdnevpt.F:369-784)
    def fock(self, P):
        nc, n, nv = self.nc, self.n, self.nv
        C, A, V = slice(0, nc), slice(nc, nc + n), slice(nc + n, self.norb)
        Bf, h, da = self.Bf, self.h, self.da
        Jcore = np.einsum("lkk->l", Bf[:, C, C])
        Jact = np.einsum("lab,ba->l", Bf[:, A, A], da)
        eps = np.zeros(self.norb)
        if nc:
            Fc = h[C, C] + 2.0 * np.einsum("lij,l->ij", Bf[:, C, C], Jcore) \
                - np.einsum("lik,lkj->ij", Bf[:, C, C], Bf[:, C, C]) \
                + np.einsum("lij,l->ij", Bf[:, C, C], Jact) \
                - 0.5 * np.einsum("lib,laj,ba->ij", Bf[:, C, A], Bf[:, A, C], da, optimize=True)
            w, U = np.linalg.eigh(Fc, UPLO="L")
            eps[C] = w
            self.ccore = U
        else:
            self.ccore = np.zeros((0, 0))
        Fv = h[V, V].copy()
        if nc:
            Fv += 2.0 * np.einsum("lrs,l->rs", Bf[:, V, V], Jcore) - np.einsum("lrj,ljs->rs", Bf[:, V, C], Bf[:, C, V])
        Fv += np.einsum("lrs,l->rs", Bf[:, V, V], Jact) \
            - 0.5 * np.einsum("lrb,las,ba->rs", Bf[:, V, A], Bf[:, A, V], da, optimize=True)
        w, U = np.linalg.eigh(Fv, UPLO="L")
        eps[V] = w
        self.cvirt = U
        self.eps = eps

    # ------------------------------------------------------------ cholesky_transform (qdnevpt.F:7599-7717)
    def transform(self):
        nc, n = self.nc, self.n
        C, A, V = slice(0, nc), slice(nc, nc + n), slice(nc + n, self.norb)
        Bf, Cc, Cv = self.Bf, self.ccore, self.cvirt
        self.Bcc = np.einsum("lkm,ki,mj->lij", Bf[:, C, C], Cc, Cc, optimize=True)
        self.Bcv = np.einsum("lkm,ki,mr->lir", Bf[:, C, V], Cc, Cv, optimize=True)
        self.Bca = np.einsum("lka,ki->lia", Bf[:, C, A], Cc, optimize=True)
        self.Bav = np.einsum("lam,mr->lar", Bf[:, A, V], Cv, optimize=True)
        self.Baa = Bf[:, A, A]
        # one-electron integrals in the canonical bases (newint_ir / newint_ar / newint_ia)
        self.h_cv = Cc.T @ self.h[C, V] @ Cv
        self.h_av = self.h[A, V] @ Cv
        self.h_ca = Cc.T @ self.h[C, A]

    # ------------------------------------------------------------ stage-2 accessor tensors (qdnevpt.F:6448-6519)
    def ee2(self):
        n, I = self.n, np.eye(self.n)
        return self.daa.transpose(2, 0, 3, 1) + np.einsum("bc,da->abcd", I, self.da)    # taa(b,d,a,c)+d(b,c)dal(d,a)

    def ee2tr(self):
        I = np.eye(self.n)
        return -self.ee2().transpose(0, 1, 3, 2) + 2.0 * np.einsum("cd,ba->abcd", I, self.da)

    def eee2(self):
        I = np.eye(self.n)
        d3, d2, d1 = self.ro3, self.daa, self.da
        E = d3.transpose(3, 0, 4, 1, 5, 2).copy()        # daaa(b,d,f,a,c,e)
        E += np.einsum("ed,bfac->abcdef", I, d2)
        E += np.einsum("eb,fdac->abcdef", I, d2)
        E += np.einsum("bc,dfae->abcdef", I, d2)
        E += np.einsum("bc,ed,fa->abcdef", I, I, d1)
        return E

    def eee2t(self, E3=None, E2=None):
        I = np.eye(self.n)
        E3 = self.eee2() if E3 is None else E3
        E2 = self.ee2() if E2 is None else E2
        return -E3.transpose(0, 1, 3, 2, 4, 5) + 2.0 * np.einsum("cd,abef->abcdef", I, E2)


# ---------------------------------------------------------------------------- classes
def v0(c: Ctx):
    """Class (0), SC = PC (qdnevpt.F:1661-1773)."""
    nc, nv = c.nc, c.nv
    e = c.eps
    ec, ev = e[:nc], e[nc + c.n:]
    E = P = 0.0
    for ii in _outer(nc):
        for ij in range(ii, nc):
            # traint(r,s) = (ij r | s ii)
            T = np.einsum("lr,ls->rs", c.Bcv[:, ij, :], c.Bcv[:, ii, :])
            ab1 = T
            ab2 = T.T
            zij = ii == ij
            iu = np.triu_indices(nv)
            a1, a2 = ab1[iu], ab2[iu]
            zrs = iu[0] == iu[1]
            if zij:
                dint = np.where(zrs, a1, np.sqrt(2.0) * a1)
            else:
                dint = np.where(zrs, np.sqrt(2.0) * a1, 2.0 * np.sqrt(a1 ** 2 + a2 ** 2 - a1 * a2))
            co = -dint / (ev[iu[0]] + ev[iu[1]] - ec[ii] - ec[ij])
            E += float(np.sum(dint * co))
            P += float(np.sum(co * co))
    return E, P, E, P


def v1k(c: Ctx, thr=THR_LINDEP):
    """Class (+1): SC (qdnevpt.F:1775-1934) and PC (:1974-2037, e2p1cont :3076-3254)."""
    nc, n, nv = c.nc, c.n, c.nv
    e = c.eps
    ec, ev = e[:nc], e[nc + n:]
    gkp1a = c.koop["1-EAKO"]
    ro = 2.0 * np.eye(n) - c.da                      # ro(a,a') = 2 delta - da(a,a')
    Esc = Psc = 0.0
    # PC setup
    Kpc = gkp1a * 0.5
    coef, w, nnew, _ = pc_solve(2.0 * np.eye(n) - c.da, Kpc, thr)
    s_pc = ro.T @ coef                               # s(a,mu) = sum_a' ro(a',a) coef(a',mu)
    Epc = Ppc = 0.0
    for ii in _outer(nc):
        for ij in range(ii, nc):
            Tj = np.einsum("la,lr->ar", c.Bca[:, ii, :], c.Bcv[:, ij, :])     # (a ii | r ij)
            Tk = np.einsum("la,lr->ar", c.Bca[:, ij, :], c.Bcv[:, ii, :])     # (a ij | r ii)
            if ii == ij:
                cint = 0.5 * np.einsum("ar,br->rab", Tj, Tj)
            else:
                cint = np.einsum("ar,br->rab", Tj, Tj) + np.einsum("ar,br->rab", Tk, Tk) \
                    - 0.5 * np.einsum("ar,br->rab", Tj, Tk) - 0.5 * np.einsum("ar,br->rab", Tk, Tj)
            dn = 2.0 * np.einsum("ab,rab->r", ro, cint)
            ek = np.einsum("ab,rab->r", gkp1a, cint)
            ok = np.abs(dn) > 0.0
            ekn = np.where(ok, ek / np.where(ok, dn, 1.0), 0.0)
            den = ec[ii] + ec[ij] - ev - ekn
            co = np.where(ok, dn / np.where(ok, den, 1.0), 0.0)
            Esc += float(co.sum())
            Psc += float(np.where(ok, co / np.where(ok, den, 1.0), 0.0).sum())
            # PC
            deno = w[None, :] + ev[:, None] - ec[ii] - ec[ij]          # (r, mu)
            if ii != ij:
                rj = (Tj + Tk).T @ s_pc                                 # (r, mu)
                rjp = (Tj - Tk).T @ s_pc
                contri = (0.5 * rj ** 2 + 1.5 * rjp ** 2) / deno
            else:
                rj = Tj.T @ s_pc
                contri = rj ** 2 / deno
            Epc -= float(contri.sum())
            Ppc += float((contri / deno).sum())
    return Esc, Psc, Epc, Ppc


def vm1k(c: Ctx, thr=THR_LINDEP):
    """Class (-1): SC (qdnevpt.F:2095-2262) and PC (:2297-2308, e2m1cont :2400-2561)."""
    nc, n, nv = c.nc, c.n, c.nv
    e = c.eps
    ec, ev = e[:nc], e[nc + n:]
    gkm1a = c.koop["1-IPKO"]
    coef, w, nnew, _ = pc_solve(c.da.copy(), gkm1a * 0.5, thr)
    s_pc = c.da.T @ coef                               # s(a,mu) = sum_a' da(a',a) coef(a',mu)
    Esc = Psc = Epc = Ppc = 0.0
    for ir in _outer(nv):
        for is_ in range(ir, nv):
            Tj = np.einsum("li,la->ia", c.Bcv[:, :, ir], c.Bav[:, :, is_])     # (r i | s a)
            Tk = np.einsum("li,la->ia", c.Bcv[:, :, is_], c.Bav[:, :, ir])     # (s i | r a)
            if ir == is_:
                cint = 0.5 * np.einsum("ia,ib->iab", Tj, Tj)
            else:
                cint = np.einsum("ia,ib->iab", Tj, Tj) + np.einsum("ia,ib->iab", Tk, Tk) \
                    - 0.5 * np.einsum("ia,ib->iab", Tj, Tk) - 0.5 * np.einsum("ia,ib->iab", Tk, Tj)
            dn = 2.0 * np.einsum("ab,iab->i", c.da, cint)
            ek = np.einsum("ab,iab->i", gkm1a, cint)
            ok = np.abs(dn) >= 1e-14
            ekn = np.where(ok, -ek / np.where(ok, dn, 1.0), 0.0)
            den = ekn + ec - ev[ir] - ev[is_]
            co = np.where(ok, dn / np.where(ok, den, 1.0), 0.0)
            Esc += float(co.sum())
            Psc += float(np.where(ok, co / np.where(ok, den, 1.0), 0.0).sum())
            deno = w[None, :] + ev[ir] + ev[is_] - ec[:, None]         # (i, mu)
            if ir != is_:
                r1 = (Tj + Tk) @ s_pc
                r2 = (Tj - Tk) @ s_pc
                contri = (0.5 * r1 ** 2 + 1.5 * r2 ** 2) / deno
            else:
                r1 = Tj @ s_pc
                contri = r1 ** 2 / deno
            Epc -= float(contri.sum())
            Ppc += float((contri / deno).sum())
    return Esc, Psc, Epc, Ppc


def fillcm2(X):
    n = X.shape[0]
    return X.reshape(n * n, n * n).copy()            # S[(i,j),(k,l)] = X(i,j,k,l), ij = i*n+j


def fillcm2p(X):
    n = X.shape[0]
    iu = [(i, j) for i in range(n) for j in range(i, n)]
    m = len(iu)
    S = np.empty((m, m))
    for p, (i, j) in enumerate(iu):
        for q, (k, l) in enumerate(iu):
            S[p, q] = X[i, j, k, l] + X[i, j, l, k]
    return S, iu


def v2mod(c: Ctx, thr=THR_LINDEP):
    """Class (+2): SC (qdnevpt.F:3256-3445), PC e2p2cont/e2p2contp (:2804-3072)."""
    nc, n = c.nc, c.n
    e = c.eps
    ec = e[:nc]
    daat, gkp2aa = c.koop["2-HRDM"], c.koop["2-EAKO"]
    Esc = Psc = 0.0
    for ir in _outer(nc):
        for is_ in range(ir, nc):
            T = np.einsum("la,lb->ab", c.Bca[:, ir, :], c.Bca[:, is_, :])   # (i a | j b)
            cint = np.einsum("CD,cd->CDcd", T, T)                            # dint1(c',d')*dint2(c,d)
            dum = daat.copy()                                                # daat(c',d',c,d)
            dumk = gkp2aa.copy()
            if ir == is_:
                cint = cint * 0.25
                dum = dum + daat.transpose(0, 1, 3, 2)                      # daat(c',d',d,c)
                dumk = dumk + gkp2aa.transpose(1, 0, 2, 3)                  # gkp2aa(d',c',c,d)
            dn = float(np.sum(cint * dum))
            dk = float(np.sum(cint * dumk))
            if abs(dn) >= 1e-14:
                den = ec[ir] + ec[is_] - dk / dn
                co = dn / den
                Esc += co
                Psc += co / den
    # PC, i != j
    coef, w, nnew, _ = pc_solve(fillcm2(daat), fillcm2(gkp2aa), thr)
    s_pc = fillcm2(daat).T @ coef                     # s(ab,mu) = sum_a'b' daat(a',b',a,b) coef(a'b',mu)
    Epc = Ppc = 0.0
    for i in _outer(nc):
        for j in range(i + 1, nc):
            T = np.einsum("la,lb->ab", c.Bca[:, i, :], c.Bca[:, j, :])     # (i a | j b)
            v = T.T.reshape(-1)                                              # traint(b,a) at ab = a*n+b
            aij = v @ s_pc
            deno = w - ec[i] - ec[j]
            contri = aij ** 2 / deno
            Epc -= float(contri.sum())
            Ppc += float((contri / deno).sum())
    # PC, i == j
    Sp, iu = fillcm2p(daat)
    Kp, _ = fillcm2p(gkp2aa)
    coef, w, nnew, _ = pc_solve(Sp, Kp, thr)
    Mh = np.array([[0.5 * (daat[ap, bp, a, b] + daat[ap, bp, b, a]) for (a, b) in iu] for (ap, bp) in iu])
    s_pc = Mh.T @ coef
    for i in _outer(nc):
        T = np.einsum("la,lb->ab", c.Bca[:, i, :], c.Bca[:, i, :])
        v = np.array([T[a, a] if a == b else 2.0 * T[b, a] for (a, b) in iu])
        aij = v @ s_pc
        deno = w - 2.0 * ec[i]
        contri = aij ** 2 / deno
        Epc -= float(contri.sum())
        Ppc += float((contri / deno).sum())
    return Esc, Psc, Epc, Ppc


def vm2mod(c: Ctx, thr=THR_LINDEP):
    """Class (-2): SC (qdnevpt.F:3685-3849), PC e2m2cont/e2m2contp (:2563-2802)."""
    nc, n, nv = c.nc, c.n, c.nv
    ev = c.eps[nc + n:]
    daa, gkm2aa = c.daa, c.koop["2-IPKO"]
    Bvr = c.Bav.transpose(0, 2, 1)                    # [l, r, a] = B'(r a)
    Esc = Psc = 0.0
    for ir in _outer(nv):
        for is_ in range(ir, nv):
            D = np.einsum("la,lb->ab", Bvr[:, is_, :], Bvr[:, ir, :])     # dint1(a,b) = (s a | r b)
            cint = np.einsum("cd,CD->CDcd", D, D)
            dum = daa.copy()
            dumk = gkm2aa.copy()
            if ir == is_:
                cint = cint * 0.25
                dum = dum + daa.transpose(0, 1, 3, 2)
                dumk = dumk + gkm2aa.transpose(1, 0, 2, 3)
            dn = float(np.sum(cint * dum))
            dk = float(np.sum(cint * dumk))
            if abs(dn) >= 1e-14:
                den = -dk / dn - ev[ir] - ev[is_]
                co = dn / den
                Esc += co
                Psc += co / den
    coef, w, nnew, _ = pc_solve(fillcm2(daa), fillcm2(gkm2aa), thr)
    s_pc = fillcm2(daa).T @ coef
    Epc = Ppc = 0.0
    for ir in _outer(nv):
        for is_ in range(ir + 1, nv):
            D = np.einsum("la,lb->ab", Bvr[:, ir, :], Bvr[:, is_, :])     # (r a | s b)
            v = D.T.reshape(-1)                                              # dint1(b,a) at ab=a*n+b
            rs = v @ s_pc
            deno = w + ev[ir] + ev[is_]
            contri = rs ** 2 / deno
            Epc -= float(contri.sum())
            Ppc += float((contri / deno).sum())
    Sp, iu = fillcm2p(daa)
    Kp, _ = fillcm2p(gkm2aa)
    coef, w, nnew, _ = pc_solve(Sp, Kp, thr)
    Mh = np.array([[0.5 * (daa[ap, bp, a, b] + daa[ap, bp, b, a]) for (a, b) in iu] for (ap, bp) in iu])
    s_pc = Mh.T @ coef
    for ir in _outer(nv):
        D = np.einsum("la,lb->ab", Bvr[:, ir, :], Bvr[:, ir, :])
        v = np.array([D[a, b] * (1.0 if a == b else 2.0) for (a, b) in iu])
        rs = v @ s_pc
        deno = w + 2.0 * ev[ir]
        contri = rs ** 2 / deno
        Epc -= float(contri.sum())
        Ppc += float((contri / deno).sum())
    return Esc, Psc, Epc, Ppc


def v0pmod(c: Ctx, thr=THR_LINDEP, rho=None):
    """Class (0)': SC (qdnevpt.F:4326-4611), PC fillc0p/e20pcont (:4195-4264, :4771-4999).
    With a dict ``rho``, also the core-virtual block of the first-order density matrix that e20pcont
    accumulates for .NATORB (:4823-4861, :4971-4976): rho['cv'][i, r] = -sum_mu c_ir,mu vaux_mu with
    vaux_mu = sum_ab da(a,b) (2 coef(ab,mu) - coef(ab+n^2,mu))."""
    nc, n, nv = c.nc, c.n, c.nv
    ec, ev = c.eps[:nc], c.eps[nc + n:]
    da, daa = c.da, c.daa
    ga, gd, gf = c.koop["gk0pa"], c.koop["gk0pd"], c.koop["gk0pf"]
    E2 = c.ee2()
    I = np.eye(n)
    # heffd(k,r) = h'(k,r) + sum_j 2 (jj|kr) - (jk|jr)
    heff = c.h_cv + 2.0 * np.einsum("ljj,lkr->kr", c.Bcc, c.Bcv) - np.einsum("ljk,ljr->kr", c.Bcc, c.Bcv)
    # dum1[a',b',a,b] = ee2(b',a',a,b); dum2 = ee2tr(b',b,a',a) + delta(a,b) da(b',a')
    E2tr = c.ee2tr()
    X1 = np.einsum("BAab->ABab", E2)                                         # index order (a',b',a,b)
    X2 = np.einsum("BbAa->ABab", E2tr) + np.einsum("ab,BA->ABab", I, da)
    G1 = np.einsum("BAab->ABab", ga)
    G2 = np.einsum("BAab->ABab", gd)
    Esc = Psc = 0.0
    for ii in _outer(nc):
        Dj = np.einsum("lr,lab->rab", c.Bcv[:, ii, :], c.Baa)                 # (i r | a b)
        Dk = np.einsum("la,lrb->rab", c.Bca[:, ii, :], c.Bav.transpose(0, 2, 1)[:, :, :])  # (i a | r b)
        for ir in range(nv):
            dj, dk = Dj[ir], Dk[ir]
            dint1 = 2.0 * np.einsum("AB,ab->ABab", dj, dj) - np.einsum("AB,ab->ABab", dj, dk) \
                - np.einsum("AB,ab->ABab", dk, dj)
            dint2 = np.einsum("AB,ab->ABab", dk, dk)
            dint11 = np.einsum("AB,ab->ABab", dj, dj) - 0.5 * np.einsum("AB,ab->ABab", dj, dk) \
                - 0.5 * np.einsum("ab,AB->ABab", dj, dk)
            dn = float(np.sum(dint1 * X1 + dint2 * X2))
            dkoo = float(np.sum(dint11 * G1 + dint2 * G2))
            d1 = 2.0 * dj - dk
            dummy = float(np.sum(da.T * d1))                                   # sum da(b,a) d1(a,b)
            dummyk = float(np.sum(gf * d1 * 0.5))
            hh = heff[ii, ir]
            dn = dn + 2.0 * hh * dummy + 2.0 * hh ** 2
            dkoo = dkoo + hh * dummyk
            if abs(dn) > 1e-14:
                den = ec[ii] - ev[ir] - dkoo / dn
                co = dn / den
                Esc += co
                Psc += co / den
    # PC: 2n^2 metric (fillc0p)
    nn = n * n
    S = np.zeros((2 * nn, 2 * nn))
    K = np.zeros((2 * nn, 2 * nn))
    Vm = np.einsum("BAab->abAB", E2).reshape(nn, nn)        # V[ab,a'b'] = ee2(b',a',a,b)
    Ga = np.einsum("baAB->abAB", ga).reshape(nn, nn)        # ga(b,a,a',b')
    Gd = np.einsum("baAB->abAB", gd).reshape(nn, nn)
    S[:nn, :nn] = 2.0 * Vm; K[:nn, :nn] = Ga
    S[:nn, nn:] = -Vm; K[:nn, nn:] = -0.5 * Ga
    S[nn:, :nn] = -Vm; K[nn:, :nn] = -0.5 * Ga
    V2 = -np.einsum("BabA->abAB", daa) + 2.0 * np.einsum("aA,Bb->abAB", I, da)   # -daa(b',a,b,a') + 2 d(a,a') da(b',b)
    S[nn:, nn:] = V2.reshape(nn, nn); K[nn:, nn:] = Gd
    coef, w, nnew, _ = pc_solve(S, K, thr)
    # e20pcont: s = V coef[:nn], sp = V coef[nn:], spp = V2' coef[nn:]
    V2p = (-np.einsum("AbaB->abAB", daa) + 2.0 * np.einsum("Aa,bB->abAB", I, da)).reshape(nn, nn)
    sm = Vm @ coef[:nn]
    sp = Vm @ coef[nn:]
    spp = V2p @ coef[nn:]
    daT = da.T.reshape(-1)                                   # da(b,a) at ab = a*n+b
    if rho is not None:
        vaux = (2.0 * coef[:nn] - coef[nn:]).T @ da.reshape(-1)       # da(a,b) at ab = a*n+b
        if _RHO_SLIP == "v0p_one_coef":
            vaux = (coef[:nn] - coef[nn:]).T @ da.reshape(-1)
        rho["cv"] = np.zeros((nc, nv))
    Epc = Ppc = 0.0
    for i in _outer(nc):
        Dj = np.einsum("lr,lab->rab", c.Bcv[:, i, :], c.Baa).reshape(nv, nn)
        Dk = np.einsum("la,lrb->rab", c.Bca[:, i, :], c.Bav.transpose(0, 2, 1)).reshape(nv, nn)
        hh = heff[i]                                         # (r)
        rimu = (2.0 * Dj - Dk) @ sm + 2.0 * hh[:, None] * (daT @ coef[:nn])[None, :]
        rimup = -Dj @ sp + Dk @ spp - hh[:, None] * (daT @ coef[nn:])[None, :]
        deno = w[None, :] + ev[:, None] - ec[i]
        co = rimu + rimup
        cir = co / deno
        Epc -= float(np.sum(co * cir))
        Ppc += float(np.sum(cir ** 2))
        if rho is not None:
            rho["cv"][i] = -(cir @ vaux)
    return Esc, Psc, Epc, Ppc


def _pm1_prime(c: Ctx, plus: bool, thr=THR_LINDEP, rho=None):
    """Classes (+1)' (vpupE, qdnevpt.F:5437-5867) and (-1)' (vmupE, :5869-6445).
    With a dict ``rho``, also the block of the first-order density matrix that e21pcont (core-active,
    :5052-5055, :5113-5130, :5187-5192) or e2m1pcont (active-virtual, :5266-5269, :5326-5341, :5407-5412)
    accumulates for .NATORB: rho[x, d] = -sum_mu co_mu vaux(d, mu), vaux(d, mu) = sum_abc coef(mu, abc)
    ee2tl(d,b,a,c) for (+1)' and ee2(d,b,a,c) for (-1)'; stored as rho['ca'][i, d] and rho['av'][d, r]."""
    nc, n, nv = c.nc, c.n, c.nv
    ec, ev = c.eps[:nc], c.eps[nc + n:]
    da = c.da
    E2 = c.ee2()
    E3 = c.eee2()
    if plus:
        amat, bmat, cmat, dmat = (c.koop[k] for k in ("amatEA", "bmatEA", "cmatEA", "dmatEA"))
        X2 = c.ee2tr()                                        # ee2tr(c',a',b',a)
        X3 = c.eee2t(E3, E2)                                  # eee2t(c',a',b',b,a,c)
        ro1 = 2.0 * np.eye(n) - da                            # (2 delta - da(a',a))
        # heffd(a,i) = h'(i,a) + sum_j 2 (jj|ia) - (ji|ja)
        heff = (c.h_ca + 2.0 * np.einsum("ljj,lia->ia", c.Bcc, c.Bca) - np.einsum("lji,lja->ia", c.Bcc, c.Bca)).T
        # dint1(a,b,c) = (a b | c i): [ext, a, b, c]
        D = np.einsum("lab,lic->iabc", c.Baa, c.Bca)
        ext_e = ec
    else:
        amat, bmat, cmat, dmat = (c.koop[k] for k in ("amatIP", "bmatIP", "cmatIP", "dmatIP"))
        X2 = E2
        X3 = E3
        ro1 = da
        Bvr = c.Bav                                          # [l, a, r]
        heff = c.h_av + 2.0 * np.einsum("lar,ljj->ar", Bvr, c.Bcc) - np.einsum("lja,ljr->ar", c.Bca, c.Bcv) \
            - np.einsum("lab,lbr->ar", c.Baa, Bvr)
        D = np.einsum("lab,lcr->rabc", c.Baa, Bvr)            # (a b | c r)
        ext_e = ev
    next_ = D.shape[0]
    # SC: v[a'b'c'] = dint1(a',c',b') ; X3 block M[a'b'c', abc] = X3(c',a',b',b,a,c)
    M = np.einsum("CABbac->ABCabc", X3).reshape(n ** 3, n ** 3)
    Am = np.ascontiguousarray(amat).reshape(n ** 3, n ** 3)   # Am[a'b'c', abc] = amat(a',b',c',a,b,c)
    Q2 = np.einsum("CABa->ABCa", X2).reshape(n ** 3, n)       # X2(c',a',b',a)
    BC = (np.einsum("ABCa->ABCa", bmat) + np.einsum("aABC->ABCa", cmat)).reshape(n ** 3, n)
    Esc = Psc = 0.0
    for x in _outer(next_):
        v = np.einsum("acb->abc", D[x]).reshape(-1)          # v[a'b'c'] = dint1(a',c',b')
        h = heff[:, x]
        dumm = float(h @ ro1 @ h)
        dumm += 2.0 * float(v @ Q2 @ h) + float(v @ M @ v)
        dumk = float(h @ dmat @ h) + float(v @ BC @ h) + float(v @ Am @ v)
        if dumm != 0.0:
            if plus:
                den = ext_e[x] - dumk / dumm
            else:
                den = -dumk / dumm - ext_e[x]
            co = dumm / den
            Esc += co
            Psc += co / den
    # PC: n^3 metric
    Spc = np.einsum("CABbac->abcABC", X3).reshape(n ** 3, n ** 3)    # s(abc,a'b'c') = X3(c',a',b',b,a,c)
    Kpc = np.einsum("ABCabc->abcABC", amat).reshape(n ** 3, n ** 3)  # coef(abc,a'b'c') = amat(a',b',c',a,b,c)
    coef, w, nnew, _ = pc_solve(Spc, Kpc, thr)
    coefT = coef.T                                                   # coefT[mu, abc]
    Sb = coefT @ M                                                   # S[mu, abc] = sum coefT[mu,a'b'c'] X3(c',a',b',b,a,c)
    sp = coefT @ Q2                                                  # sp[mu, a]
    if rho is not None:
        I = np.eye(n)
        if plus:       # T[abc, d] = ee2tl(d,b,a,c) = -ee2(b,d,a,c) + 2 delta(d,b) da(c,a)
            T = -np.einsum("bdac->abcd", E2) + 2.0 * np.einsum("db,ca->abcd", I, da)
            if _RHO_SLIP == "plus_as_minus":
                T = np.einsum("dbac->abcd", E2)
        else:          # T[abc, d] = ee2(d,b,a,c)
            T = np.einsum("dbac->abcd", E2)
            if _RHO_SLIP == "minus_bc_swapped":
                T = np.einsum("dcab->abcd", E2)
        vaux = coefT @ T.reshape(n ** 3, n)                          # vaux[mu, d]
        blk = np.zeros((next_, n))
    Epc = Ppc = 0.0
    for x in _outer(next_):
        v = np.einsum("acb->abc", D[x]).reshape(-1)                  # dint1(a,c,b) at abc
        rmu = sp @ heff[:, x] + Sb @ v
        deno = w - ext_e[x] if plus else w + ext_e[x]
        co = rmu / deno
        Epc -= float(np.sum(rmu * co))
        Ppc += float(np.sum(co ** 2))
        if rho is not None:
            blk[x] = -(co @ vaux)
    if rho is not None:
        if plus:
            rho["ca"] = blk
        else:
            rho["av"] = blk.T.copy()
    return Esc, Psc, Epc, Ppc


def vpupE(c, thr=THR_LINDEP, rho=None):
    return _pm1_prime(c, True, thr, rho)


def vmupE(c, thr=THR_LINDEP, rho=None):
    return _pm1_prime(c, False, thr, rho)


def rho1(P, koop, s=0, thr=THR_LINDEP, with_4rdm=True, frozen=()):
    """The .NATORB density matrix of state s (qdnevpt.F:1240-1303): core block 2, active block the 1-RDM,
    core-virtual / active-virtual / core-active blocks the first-order terms of the PC classes (0)',
    (-1)' and (+1)' (zero without the 4-RDM classes), symmetric, virtual block 0; in the canonical core and
    virtual orbitals of state s.  Returns (rho, rho_mo, occ): rho_mo = U rho U^T back in the input orbitals
    (U = diag(ccore, 1, cvirt)), occ the eigenvalues of rho in ascending order (the dsyev print of
    qdnevpt.F:1414-1423)."""
    c = Ctx(P, koop, s, frozen)
    nc, n, nv, norb = c.nc, c.n, c.nv, c.norb
    blocks = {}
    v0pmod(c, thr, rho=blocks)
    if with_4rdm:
        vpupE(c, thr, rho=blocks)
        vmupE(c, thr, rho=blocks)
    A, V = slice(nc, nc + n), slice(nc + n, norb)
    R = np.zeros((norb, norb))
    R[:nc, :nc] = 2.0 * np.eye(nc)
    R[A, A] = c.da
    for key, (r0, c0) in (("cv", (slice(0, nc), V)), ("av", (A, V)), ("ca", (slice(0, nc), A))):
        if key in blocks:
            R[r0, c0] = blocks[key]
            R[c0, r0] = blocks[key].T
    U = np.zeros((norb, norb))
    U[:nc, :nc] = c.ccore
    U[A, A] = np.eye(n)
    U[V, V] = c.cvirt
    return R, U @ R @ U.T, np.linalg.eigvalsh(R)


def order_occupations(w):
    """order_eigvalvec (qdnevpt.F:1484-1519): selection sort by descending |w|, the later of equal moduli
    first (its test is >=)."""
    w = list(w)
    m = len(w)
    for i in range(m - 1):
        emin, imin = w[i], i
        for j in range(i + 1, m):
            if abs(w[j]) >= abs(emin):
                emin, imin = w[j], j
        if imin != i:
            w[imin] = w[i]
            w[i] = emin
    return np.array(w)


CLASS_NAMES = ["(0)", "(+1)", "(-1)", "(+2)", "(-2)", "(+1)'", "(-1)'", "(0)'"]


def state_energies(P, koop, s=0, thr=THR_LINDEP, with_4rdm=True, frozen=()):
    """Per-class (norm_sc, e_sc, norm_pc, e_pc) for state s, in the reference's class order
    (0),(+1),(-1),(+2),(-2),(+1)',(-1)',(0)' (qdnevpt.F:1019-1045); ``frozen``: see Ctx."""
    c = Ctx(P, koop, s, frozen)
    out = np.zeros((8, 4))
    for k, fn in ((0, v0), (1, v1k), (2, vm1k), (3, v2mod), (4, vm2mod), (7, v0pmod)):
        e_sc, p_sc, e_pc, p_pc = fn(c, thr) if fn is not v0 else fn(c)
        out[k] = (p_sc, e_sc, p_pc, e_pc)
    if with_4rdm:
        for k, fn in ((6, vmupE), (5, vpupE)):
            e_sc, p_sc, e_pc, p_pc = fn(c, thr)
            out[k] = (p_sc, e_sc, p_pc, e_pc)
    return out, c


def state_energies_timed(P, koop, s=0, frac=0.05, thr=THR_LINDEP):
    """Timing sample for the CPU baseline: each class's outer external loops run on
    ``frac`` of their range; each class's time is scaled by the largest range/sample
    ratio it used.  Returns (table of the partial sums, per-class seconds, total seconds)."""
    import time
    global _FRAC
    t0 = time.perf_counter()
    c = Ctx(P, koop, s)
    t_ctx = time.perf_counter() - t0
    out = np.zeros((8, 4))
    secs = {}
    _FRAC = frac
    try:
        for k, fn in ((0, v0), (1, v1k), (2, vm1k), (3, v2mod), (4, vm2mod), (7, v0pmod), (6, vmupE), (5, vpupE)):
            _SCALES.clear()
            t0 = time.perf_counter()
            e_sc, p_sc, e_pc, p_pc = fn(c) if fn is v0 else fn(c, thr)
            dt = time.perf_counter() - t0
            secs[CLASS_NAMES[k]] = dt * (max(_SCALES) if _SCALES else 1.0)
            out[k] = (p_sc, e_sc, p_pc, e_pc)
    finally:
        _FRAC = None
    return out, secs, t_ctx + sum(secs.values())

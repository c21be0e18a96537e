"""QD-NEVPT2 oracle: off-diagonal class terms and the effective Hamiltonian.

Oracle (test infrastructure only).  For each state i (``istate``) and every
other state j the reference adds <j|...|i>-type couplings to the effective
Hamiltonian, e2mp(j,i) (SC) and e2en(j,i) (PC), inside every class driver
(qdnevpt.F:1775-6445).  They reuse state i's orbitals, integrals, Koopmans
matrices and denominators and replace i's RDMs by the transition RDMs of the
pair (daoff/daaoff/ro3off, stored once per pair I>J at index
ijcou = (I-1)(I-2)/2 + J, qdnevpt.F:1912).  Every access below is transcribed
from the reference with its own index order, including which pair
orientation it uses for j<i and j>i; constant (norm) terms are absent there,
as in the reference.  The Heff post-processing restates qdnevpt.F:1046-1204
(E0 on the diagonal, dgeev, ascending sort) and hermitian_heff
(qdnevpt.F:8089-8184).

Diagonal entries come from oracle.stage2 unchanged; this module only adds the
off-diagonal sums.  Parity: every class's couplings equal
<H Psi_j|P (E_0 - H_D)^-1 P|H Psi_i> over state i's SC function and PC
eigenfunctions, by Fock-space brute force (oracle/fock.py), for CASCI
eigenfunctions and random CI vectors (tests/test_oracle_brute.py); the
hermitisation is pinned by its algebra (tests/test_qd_oracle.py).
"""
from __future__ import annotations

import numpy as np

from . import stage2 as S2
from .stage2 import THR_LINDEP, pc_solve, fillcm2, fillcm2p


def pair(i, j):
    """0-based position of the unordered state pair (i, j), i != j (qdnevpt.F:1912)."""
    hi, lo = max(i, j), min(i, j)
    return hi * (hi - 1) // 2 + lo


# ---------------------------------------------------------------------------- stage-2 accessors on any RDM set
def t_ee2(d2, d1):
    """ee2(a,b,c,d) = taa(b,d,a,c) + d(b,c) dal(d,a) (qdnevpt.F:6462-6469)."""
    n = d1.shape[0]
    return d2.transpose(2, 0, 3, 1) + np.einsum("bc,da->abcd", np.eye(n), d1)


def t_ee2tr(d2, d1):
    """ee2tr(a,b,c,d) = -ee2(a,b,d,c) + 2 d(c,d) dal(b,a) (qdnevpt.F:6490-6497)."""
    n = d1.shape[0]
    return -t_ee2(d2, d1).transpose(0, 1, 3, 2) + 2.0 * np.einsum("cd,ba->abcd", np.eye(n), d1)


def t_ee2tl(d2, d1):
    """ee2tl(a,b,c,d) = -ee2(b,a,c,d) + 2 d(a,b) dal(d,c) (qdnevpt.F:6499-6506)."""
    n = d1.shape[0]
    return -t_ee2(d2, d1).transpose(1, 0, 2, 3) + 2.0 * np.einsum("ab,dc->abcd", np.eye(n), d1)


def t_eee2(d3, d2, d1):
    """eee2(a..f) = daaa(b,d,f,a,c,e) + delta terms (qdnevpt.F:6448-6460)."""
    n = d1.shape[0]
    I = np.eye(n)
    E = d3.transpose(3, 0, 4, 1, 5, 2).copy()
    E += np.einsum("ed,bfac->abcdef", I, d2)
    E += np.einsum("eb,fdac->abcdef", I, d2)
    E += np.einsum("bc,dfae->abcdef", I, d2)
    E += np.einsum("bc,ed,fa->abcdef", I, I, d1)
    return E


def t_eee2t(d3, d2, d1):
    """eee2t(a..f) = -eee2(a,b,d,c,e,f) + 2 d(c,d) ee2(a,b,e,f) (qdnevpt.F:6480-6488)."""
    n = d1.shape[0]
    return -t_eee2(d3, d2, d1).transpose(0, 1, 3, 2, 4, 5) + 2.0 * np.einsum("cd,abef->abcdef", np.eye(n),
                                                                             t_ee2(d2, d1))


def t_ro2toff(d2, d1):
    """ro2toff(ap,bp,a,b) (qdnevpt.F:6508-6518): two-hole transition density without constants."""
    n = d1.shape[0]
    I = np.eye(n)
    R = d2.copy()
    R -= 2.0 * np.einsum("Aa,Bb->ABab", I, d1)
    R -= 2.0 * np.einsum("Bb,Aa->ABab", I, d1)
    R += np.einsum("Ab,Ba->ABab", I, d1)
    R += np.einsum("Ba,Ab->ABab", I, d1)
    return R


class Trans:
    """Transition RDMs of a synth.Problem by unordered pair, with derived tensors cached."""

    def __init__(self, P):
        self.M = P.metat
        self.d1 = P.daloff
        self.d2 = P.taaoff
        self.d3 = P.daaaoff
        self._c = {}

    def get(self, name, p):
        key = (name, p)
        if key not in self._c:
            d1, d2, d3 = self.d1[p], self.d2[p], self.d3[p]
            f = {"ee2": lambda: t_ee2(d2, d1), "ee2tr": lambda: t_ee2tr(d2, d1), "ee2tl": lambda: t_ee2tl(d2, d1),
                 "eee2": lambda: t_eee2(d3, d2, d1), "eee2t": lambda: t_eee2t(d3, d2, d1),
                 "ro2toff": lambda: t_ro2toff(d2, d1)}[name]
            self._c[key] = f()
        return self._c[key]


def _others(c, T):
    return [j for j in range(T.M) if j != c.s]


# ---------------------------------------------------------------------------- (+1)
def v1k_off(c, T, thr=THR_LINDEP):
    """Off-diagonal V(+1): SC qdnevpt.F:1909-1943, PC e2p1cont :3106-3239."""
    nc, n, i = c.nc, c.n, c.s
    ec, ev = c.eps[:nc], c.eps[nc + n:]
    gkp1a = c.koop["1-EAKO"]
    ro = 2.0 * np.eye(n) - c.da
    js = _others(c, T)
    sc, pc = np.zeros(T.M), np.zeros(T.M)
    # SC: ro(ia,iap) = -daoff(p,iap,ia) for j<i, -daoff(p,ia,iap) for j>i
    roj = {j: (-T.d1[pair(i, j)].T if j < i else -T.d1[pair(i, j)]) for j in js}
    coef, w, nnew, _ = pc_solve(2.0 * np.eye(n) - c.da, gkp1a * 0.5, thr)
    s_pc = ro.T @ coef
    # PC: s(a,mu,j) = sum_ap ro(ap,a) coef(ap,mu), ro = -daoff(p,a,ap) (j<i), -daoff(p,ap,a) (j>i)
    s_j = {j: ((-T.d1[pair(i, j)].T) if j < i else (-T.d1[pair(i, j)])).T @ coef for j in js}
    for ii in range(nc):
        for ij in range(ii, nc):
            Tj = np.einsum("la,lr->ar", c.Bca[:, ii, :], c.Bcv[:, ij, :])
            Tk = np.einsum("la,lr->ar", c.Bca[:, ij, :], c.Bcv[:, ii, :])
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
            for j in js:
                dnj = 2.0 * np.einsum("ab,rab->r", roj[j], cint)
                sc[j] += float(np.where(ok, dnj / np.where(ok, den, 1.0), 0.0).sum())
            deno = w[None, :] + ev[:, None] - ec[ii] - ec[ij]
            if ii != ij:
                rj, rjp = (Tj + Tk).T @ s_pc, (Tj - Tk).T @ s_pc
                for j in js:
                    rjj, rjpj = (Tj + Tk).T @ s_j[j], (Tj - Tk).T @ s_j[j]
                    pc[j] -= float(((0.5 * rjj * rj + 1.5 * rjpj * rjp) / deno).sum())
            else:
                rj = Tj.T @ s_pc
                for j in js:
                    pc[j] -= float((Tj.T @ s_j[j] * rj / deno).sum())
    return sc, pc


# ---------------------------------------------------------------------------- (-1)
def vm1k_off(c, T, thr=THR_LINDEP):
    """Off-diagonal V(-1): SC qdnevpt.F:2227-2259, PC e2m1cont :2434-2552."""
    nc, n, nv, i = c.nc, c.n, c.nv, c.s
    ec, ev = c.eps[:nc], c.eps[nc + n:]
    gkm1a = c.koop["1-IPKO"]
    js = _others(c, T)
    sc, pc = np.zeros(T.M), np.zeros(T.M)
    # SC: ro(ia,iap) = daoff(p,ia,iap) (j<i), daoff(p,iap,ia) (j>i)
    roj = {j: (T.d1[pair(i, j)] if j < i else T.d1[pair(i, j)].T) for j in js}
    coef, w, nnew, _ = pc_solve(c.da.copy(), gkm1a * 0.5, thr)
    s_pc = c.da.T @ coef
    # PC: s(a,mu,j) = sum_ap ro(ap,a) coef(ap,mu), ro = daoff(p,ap,a) (j<i), daoff(p,a,ap) (j>i)
    s_j = {j: (T.d1[pair(i, j)] if j < i else T.d1[pair(i, j)].T).T @ coef for j in js}
    for ir in range(nv):
        for is_ in range(ir, nv):
            Tj = np.einsum("li,la->ia", c.Bcv[:, :, ir], c.Bav[:, :, is_])
            Tk = np.einsum("li,la->ia", c.Bcv[:, :, is_], c.Bav[:, :, ir])
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
            for j in js:
                dnj = 2.0 * np.einsum("ab,iab->i", roj[j], cint)
                sc[j] += float(np.where(ok, dnj / np.where(ok, den, 1.0), 0.0).sum())
            deno = w[None, :] + ev[ir] + ev[is_] - ec[:, None]
            if ir != is_:
                r1, r2 = (Tj + Tk) @ s_pc, (Tj - Tk) @ s_pc
                for j in js:
                    r1j, r2j = (Tj + Tk) @ s_j[j], (Tj - Tk) @ s_j[j]
                    pc[j] -= float(((0.5 * r1j * r1 + 1.5 * r2j * r2) / deno).sum())
            else:
                r1 = Tj @ s_pc
                for j in js:
                    pc[j] -= float((Tj @ s_j[j] * r1 / deno).sum())
    return sc, pc


# ---------------------------------------------------------------------------- (+2)
def v2mod_off(c, T, thr=THR_LINDEP):
    """Off-diagonal V(+2): SC qdnevpt.F:3401-3416, PC e2p2cont :2854-2926 and e2p2contp :2986-3066."""
    nc, n, i = c.nc, c.n, c.s
    ec = c.eps[:nc]
    daat, gkp2aa = c.koop["2-HRDM"], c.koop["2-EAKO"]
    js = _others(c, T)
    sc, pc = np.zeros(T.M), np.zeros(T.M)
    RT = {j: T.get("ro2toff", pair(i, j)) for j in js}
    for ir in range(nc):
        for is_ in range(ir, nc):
            Tm = np.einsum("la,lb->ab", c.Bca[:, ir, :], c.Bca[:, is_, :])
            cint = np.einsum("CD,cd->CDcd", Tm, Tm)
            dum = daat.copy()
            dumk = gkp2aa.copy()
            if ir == is_:
                cint = cint * 0.25
                dum = dum + daat.transpose(0, 1, 3, 2)
                dumk = dumk + gkp2aa.transpose(1, 0, 2, 3)
            dn = float(np.sum(cint * dum))
            dk = float(np.sum(cint * dumk))
            if abs(dn) < 1e-14:
                continue
            den = ec[ir] + ec[is_] - dk / dn
            for j in js:
                R = RT[j]
                if j < i:   # ro2toff(c',d',c,d) [+ ro2toff(d',c',c,d)]
                    X = R.copy()
                    if ir == is_:
                        X = X + np.einsum("DCcd->CDcd", R)
                else:       # ro2toff(c,d,c',d') [+ ro2toff(c,d,d',c')]
                    X = np.einsum("cdCD->CDcd", R)
                    if ir == is_:
                        X = X + np.einsum("cdDC->CDcd", R)
                sc[j] += float(np.sum(cint * X)) / den
    # PC, i != j: s(ab,mu,j) = sum_a'b' daatoff(b,a,b',a') coef (j<i); daatoff(b',a',b,a) (j>i)
    coef, w, nnew, _ = pc_solve(fillcm2(daat), fillcm2(gkp2aa), thr)
    s_pc = fillcm2(daat).T @ coef
    Y = {j: (np.einsum("baBA->ABab", RT[j]) if j < i else np.einsum("BAba->ABab", RT[j])).reshape(n * n, n * n)
         for j in js}
    s_j = {j: Y[j].T @ coef for j in js}
    for a in range(nc):
        for b in range(a + 1, nc):
            Tm = np.einsum("la,lb->ab", c.Bca[:, a, :], c.Bca[:, b, :])
            v = Tm.T.reshape(-1)
            aij = v @ s_pc
            deno = w - ec[a] - ec[b]
            for j in js:
                pc[j] -= float(((v @ s_j[j]) * aij / deno).sum())
    # PC, i == j (upper pairs)
    Sp, iu = fillcm2p(daat)
    Kp, _ = fillcm2p(gkp2aa)
    coef, w, nnew, _ = pc_solve(Sp, Kp, thr)
    Mh = np.array([[0.5 * (daat[ap, bp, a, b] + daat[ap, bp, b, a]) for (a, b) in iu] for (ap, bp) in iu])
    s_pc = Mh.T @ coef
    s_j = {}
    for j in js:
        R = RT[j]
        if j < i:   # 0.5 (daatoff(b,a,b',a') + daatoff(b,a,a',b'))
            Yj = np.array([[0.5 * (R[b, a, bp, ap] + R[b, a, ap, bp]) for (a, b) in iu] for (ap, bp) in iu])
        else:       # 0.5 (daatoff(b',a',b,a) + daatoff(a',b',b,a))
            Yj = np.array([[0.5 * (R[bp, ap, b, a] + R[ap, bp, b, a]) for (a, b) in iu] for (ap, bp) in iu])
        s_j[j] = Yj.T @ coef
    for x in range(nc):
        Tm = np.einsum("la,lb->ab", c.Bca[:, x, :], c.Bca[:, x, :])
        v = np.array([Tm[a, a] if a == b else 2.0 * Tm[b, a] for (a, b) in iu])
        aij = v @ s_pc
        deno = w - 2.0 * ec[x]
        for j in js:
            pc[j] -= float(((v @ s_j[j]) * aij / deno).sum())
    return sc, pc


# ---------------------------------------------------------------------------- (-2)
def vm2mod_off(c, T, thr=THR_LINDEP):
    """Off-diagonal V(-2): SC qdnevpt.F:3806-3822, PC e2m2cont :2613-2675 and e2m2contp :2737-2796."""
    nc, n, nv, i = c.nc, c.n, c.nv, c.s
    ev = c.eps[nc + n:]
    daa, gkm2aa = c.daa, c.koop["2-IPKO"]
    Bvr = c.Bav.transpose(0, 2, 1)
    js = _others(c, T)
    sc, pc = np.zeros(T.M), np.zeros(T.M)
    D2 = {j: T.d2[pair(i, j)] for j in js}
    for ir in range(nv):
        for is_ in range(ir, nv):
            D = np.einsum("la,lb->ab", Bvr[:, is_, :], Bvr[:, ir, :])
            cint = np.einsum("cd,CD->CDcd", D, D)
            dum = daa.copy()
            dumk = gkm2aa.copy()
            if ir == is_:
                cint = cint * 0.25
                dum = dum + daa.transpose(0, 1, 3, 2)
                dumk = dumk + gkm2aa.transpose(1, 0, 2, 3)
            dn = float(np.sum(cint * dum))
            dk = float(np.sum(cint * dumk))
            if abs(dn) < 1e-14:
                continue
            den = -dk / dn - ev[ir] - ev[is_]
            for j in js:
                R = D2[j]
                if j < i:   # daaoff(c,d,c',d') [+ daaoff(c,d,d',c')]
                    X = np.einsum("cdCD->CDcd", R)
                    if ir == is_:
                        X = X + np.einsum("cdDC->CDcd", R)
                else:       # daaoff(c',d',c,d) [+ daaoff(d',c',c,d)]
                    X = R.copy()
                    if ir == is_:
                        X = X + np.einsum("DCcd->CDcd", R)
                sc[j] += float(np.sum(cint * X)) / den
    coef, w, nnew, _ = pc_solve(fillcm2(daa), fillcm2(gkm2aa), thr)
    s_pc = fillcm2(daa).T @ coef
    # s(ab,mu,j): daaoff(b',a',b,a) (j<i), daaoff(b,a,b',a') (j>i)
    Y = {j: (np.einsum("BAba->ABab", D2[j]) if j < i else np.einsum("baBA->ABab", D2[j])).reshape(n * n, n * n)
         for j in js}
    s_j = {j: Y[j].T @ coef for j in js}
    for ir in range(nv):
        for is_ in range(ir + 1, nv):
            D = np.einsum("la,lb->ab", Bvr[:, ir, :], Bvr[:, is_, :])
            v = D.T.reshape(-1)
            rs = v @ s_pc
            deno = w + ev[ir] + ev[is_]
            for j in js:
                pc[j] -= float(((v @ s_j[j]) * rs / deno).sum())
    Sp, iu = fillcm2p(daa)
    Kp, _ = fillcm2p(gkm2aa)
    coef, w, nnew, _ = pc_solve(Sp, Kp, thr)
    Mh = np.array([[0.5 * (daa[ap, bp, a, b] + daa[ap, bp, b, a]) for (a, b) in iu] for (ap, bp) in iu])
    s_pc = Mh.T @ coef
    s_j = {}
    for j in js:
        R = D2[j]
        if j < i:   # 0.5 (daaoff(b',a',b,a) + daaoff(a',b',b,a))
            Yj = np.array([[0.5 * (R[bp, ap, b, a] + R[ap, bp, b, a]) for (a, b) in iu] for (ap, bp) in iu])
        else:       # 0.5 (daaoff(b,a,a',b') + daaoff(b,a,b',a'))
            Yj = np.array([[0.5 * (R[b, a, ap, bp] + R[b, a, bp, ap]) for (a, b) in iu] for (ap, bp) in iu])
        s_j[j] = Yj.T @ coef
    for ir in range(nv):
        D = np.einsum("la,lb->ab", Bvr[:, ir, :], Bvr[:, ir, :])
        v = np.array([D[a, b] * (1.0 if a == b else 2.0) for (a, b) in iu])
        rs = v @ s_pc
        deno = w + 2.0 * ev[ir]
        for j in js:
            pc[j] -= float(((v @ s_j[j]) * rs / deno).sum())
    return sc, pc


# ---------------------------------------------------------------------------- (0)'
def v0pmod_off(c, T, thr=THR_LINDEP):
    """Off-diagonal V(0)': SC qdnevpt.F:4520-4593 (same index order for j<i and j>i),
    PC e20pcont :4862-4982."""
    nc, n, nv, i = c.nc, c.n, c.nv, c.s
    ec, ev = c.eps[:nc], c.eps[nc + n:]
    da, daa = c.da, c.daa
    ga, gd, gf = c.koop["gk0pa"], c.koop["gk0pd"], c.koop["gk0pf"]
    I = np.eye(n)
    js = _others(c, T)
    sc, pc = np.zeros(T.M), np.zeros(T.M)
    heff = c.h_cv + 2.0 * np.einsum("ljj,lkr->kr", c.Bcc, c.Bcv) - np.einsum("ljk,ljr->kr", c.Bcc, c.Bcv)
    E2, E2tr = c.ee2(), c.ee2tr()
    X1 = np.einsum("BAab->ABab", E2)
    X2 = np.einsum("BbAa->ABab", E2tr) + np.einsum("ab,BA->ABab", I, da)
    G1 = np.einsum("BAab->ABab", ga)
    G2 = np.einsum("BAab->ABab", gd)
    # off: dum1 = ee2(b',a',a,b), dum2 = ee2tr(b',b,a',a) + d(a,b) daoff(a',b'); dummy: daoff(b,a)+daoff(a,b)
    X1j, X2j, DSj = {}, {}, {}
    for j in js:
        p = pair(i, j)
        X1j[j] = np.einsum("BAab->ABab", T.get("ee2", p))
        X2j[j] = np.einsum("BbAa->ABab", T.get("ee2tr", p)) + np.einsum("ab,AB->ABab", I, T.d1[p])
        DSj[j] = T.d1[p] + T.d1[p].T
    for ii in range(nc):
        Dj = np.einsum("lr,lab->rab", c.Bcv[:, ii, :], c.Baa)
        Dk = np.einsum("la,lrb->rab", c.Bca[:, ii, :], c.Bav.transpose(0, 2, 1))
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
            dummy = float(np.sum(da.T * d1))
            dummyk = float(np.sum(gf * d1 * 0.5))
            hh = heff[ii, ir]
            dn = dn + 2.0 * hh * dummy + 2.0 * hh ** 2
            dkoo = dkoo + hh * dummyk
            if abs(dn) <= 1e-14:
                continue
            den = ec[ii] - ev[ir] - dkoo / dn
            for j in js:
                dnj = float(np.sum(dint1 * X1j[j] + dint2 * X2j[j])) + hh * float(np.sum(DSj[j] * d1))
                sc[j] += dnj / den
    # PC
    nn = n * n
    S = np.zeros((2 * nn, 2 * nn))
    K = np.zeros((2 * nn, 2 * nn))
    Vm = np.einsum("BAab->abAB", E2).reshape(nn, nn)
    Ga = np.einsum("baAB->abAB", ga).reshape(nn, nn)
    Gd = np.einsum("baAB->abAB", gd).reshape(nn, nn)
    S[:nn, :nn] = 2.0 * Vm; K[:nn, :nn] = Ga
    S[:nn, nn:] = -Vm; K[:nn, nn:] = -0.5 * Ga
    S[nn:, :nn] = -Vm; K[nn:, :nn] = -0.5 * Ga
    V2 = -np.einsum("BabA->abAB", daa) + 2.0 * np.einsum("aA,Bb->abAB", I, da)
    S[nn:, nn:] = V2.reshape(nn, nn); K[nn:, nn:] = Gd
    coef, w, nnew, _ = pc_solve(S, K, thr)
    V2p = (-np.einsum("AbaB->abAB", daa) + 2.0 * np.einsum("Aa,bB->abAB", I, da)).reshape(nn, nn)
    sm, sp, spp = Vm @ coef[:nn], Vm @ coef[nn:], V2p @ coef[nn:]
    daT = da.T.reshape(-1)
    mats = {}
    for j in js:
        p = pair(i, j)
        E2p, D2p, D1p = T.get("ee2", p), T.d2[p], T.d1[p]
        if j < i:   # val = ee2(b,a,a',b'), val2 = -daaoff(a,b',a',b) + 2 d(a',a) daoff(b',b), ro = daoff(b,a)
            Vj = np.einsum("baAB->abAB", E2p).reshape(nn, nn)
            V2j = (-np.einsum("aBAb->abAB", D2p) + 2.0 * np.einsum("Aa,Bb->abAB", I, D1p)).reshape(nn, nn)
            roj = D1p.T.reshape(-1)
        else:       # val = ee2(b',a',a,b), val2 = -daaoff(a',b,a,b') + 2 d(a',a) daoff(b,b'), ro = daoff(a,b)
            Vj = np.einsum("BAab->abAB", E2p).reshape(nn, nn)
            V2j = (-np.einsum("AbaB->abAB", D2p) + 2.0 * np.einsum("Aa,bB->abAB", I, D1p)).reshape(nn, nn)
            roj = D1p.reshape(-1)
        mats[j] = (Vj @ coef[:nn], Vj @ coef[nn:], V2j @ coef[nn:], roj @ coef[:nn], roj @ coef[nn:])
    for x in range(nc):
        Dj = np.einsum("lr,lab->rab", c.Bcv[:, x, :], c.Baa).reshape(nv, nn)
        Dk = np.einsum("la,lrb->rab", c.Bca[:, x, :], c.Bav.transpose(0, 2, 1)).reshape(nv, nn)
        hh = heff[x]
        rimu = (2.0 * Dj - Dk) @ sm + 2.0 * hh[:, None] * (daT @ coef[:nn])[None, :]
        rimup = -Dj @ sp + Dk @ spp - hh[:, None] * (daT @ coef[nn:])[None, :]
        deno = w[None, :] + ev[:, None] - ec[x]
        co_i = rimu + rimup
        for j in js:
            smj, spj, sppj, r0, r1 = mats[j]
            rj = (2.0 * Dj - Dk) @ smj + 2.0 * hh[:, None] * r0[None, :]
            rjp = -Dj @ spj + Dk @ sppj - hh[:, None] * r1[None, :]
            pc[j] -= float(np.sum((rj + rjp) * co_i / deno))
    return sc, pc


# ---------------------------------------------------------------------------- (+1)' and (-1)'
def pm1_prime_off(c, T, plus, thr=THR_LINDEP):
    """Off-diagonal V(+1)' (vpupE SC qdnevpt.F:5602-5668, e21pcont :5061-5101, 5157-5198) and
    V(-1)' (vmupE SC :6059-6152, e2m1pcont :5276-5314, 5375-5419)."""
    nc, n, i = c.nc, c.n, c.s
    ec, ev = c.eps[:nc], c.eps[nc + n:]
    da = c.da
    E2, E3 = c.ee2(), c.eee2()
    n3 = n ** 3
    if plus:
        amat, bmat, cmat, dmat = (c.koop[k] for k in ("amatEA", "bmatEA", "cmatEA", "dmatEA"))
        X2, X3 = c.ee2tr(), c.eee2t(E3, E2)
        ro1 = 2.0 * np.eye(n) - da
        heff = (c.h_ca + 2.0 * np.einsum("ljj,lia->ia", c.Bcc, c.Bca) - np.einsum("lji,lja->ia", c.Bcc, c.Bca)).T
        D = np.einsum("lab,lic->iabc", c.Baa, c.Bca)
        ext_e = ec
    else:
        amat, bmat, cmat, dmat = (c.koop[k] for k in ("amatIP", "bmatIP", "cmatIP", "dmatIP"))
        X2, X3 = E2, E3
        ro1 = da
        Bvr = c.Bav
        heff = c.h_av + 2.0 * np.einsum("lar,ljj->ar", Bvr, c.Bcc) - np.einsum("lja,ljr->ar", c.Bca, c.Bcv) \
            - np.einsum("lab,lbr->ar", c.Baa, Bvr)
        D = np.einsum("lab,lcr->rabc", c.Baa, Bvr)
        ext_e = ev
    next_ = D.shape[0]
    js = _others(c, T)
    sc, pc = np.zeros(T.M), np.zeros(T.M)
    M = np.einsum("CABbac->ABCabc", X3).reshape(n3, n3)
    Am = np.ascontiguousarray(amat).reshape(n3, n3)
    Q2 = np.einsum("CABa->ABCa", X2).reshape(n3, n)
    BC = (bmat + np.einsum("aABC->ABCa", cmat)).reshape(n3, n)
    # SC off-diagonal matrices (index a' b' c' rows, a / abc columns)
    R1, Q2j, Mj = {}, {}, {}
    for j in js:
        p = pair(i, j)
        D1p = T.d1[p]
        if plus:
            R1[j] = -D1p if j < i else -D1p.T                 # -daoff(a',a) (j<i), -daoff(a,a') (j>i)
            Q2j[j] = (np.einsum("CABa->ABCa", T.get("ee2tr", p))
                      + np.einsum("aBAC->ABCa", T.get("ee2tl", p))).reshape(n3, n)
            E3tp = T.get("eee2t", p)
            Mj[j] = (np.einsum("CABbac->ABCabc", E3tp) if j < i else np.einsum("cabBAC->ABCabc", E3tp)).reshape(n3, n3)
        else:
            R1[j] = D1p.T if j < i else D1p                   # daoff(a,a') (j<i), daoff(a',a) (j>i)
            E2p = T.get("ee2", p)
            Q2j[j] = (np.einsum("CABa->ABCa", E2p) + np.einsum("aBAC->ABCa", E2p)).reshape(n3, n)
            E3p = T.get("eee2", p)
            Mj[j] = (np.einsum("CABbac->ABCabc", E3p) if j < i else np.einsum("cabBAC->ABCabc", E3p)).reshape(n3, n3)
    for x in range(next_):
        v = np.einsum("acb->abc", D[x]).reshape(-1)
        h = heff[:, x]
        dumm = float(h @ ro1 @ h) + 2.0 * float(v @ Q2 @ h) + float(v @ M @ v)
        dumk = float(h @ dmat @ h) + float(v @ BC @ h) + float(v @ Am @ v)
        if dumm == 0.0:
            continue
        den = (ext_e[x] - dumk / dumm) if plus else (-dumk / dumm - ext_e[x])
        for j in js:
            dj = float(h @ R1[j] @ h) + float(v @ Q2j[j] @ h) + float(v @ Mj[j] @ v)
            sc[j] += dj / den
    # PC
    Spc = np.einsum("CABbac->abcABC", X3).reshape(n3, n3)
    Kpc = np.einsum("ABCabc->abcABC", amat).reshape(n3, n3)
    coef, w, nnew, _ = pc_solve(Spc, Kpc, thr)
    coefT = coef.T
    Sb, sp = coefT @ M, coefT @ Q2
    Sbj, spj = {}, {}
    for j in js:
        p = pair(i, j)
        if plus:   # s: eee2t(c,a,b,b',a',c') (j<i) / eee2t(c',a',b',b,a,c) (j>i); sp: ee2tl(a,b',a',c') / ee2tr(c',a',b',a)
            E3tp = T.get("eee2t", p)
            Mp = np.einsum("cabBAC->ABCabc", E3tp) if j < i else np.einsum("CABbac->ABCabc", E3tp)
            Qp = np.einsum("aBAC->ABCa", T.get("ee2tl", p)) if j < i else np.einsum("CABa->ABCa", T.get("ee2tr", p))
        else:      # s: eee2(c,a,b,b',a',c') (j<i) / eee2(c',a',b',b,a,c) (j>i); sp: ee2(a,b',a',c') / ee2(c',a',b',a)
            E3p, E2p = T.get("eee2", p), T.get("ee2", p)
            Mp = np.einsum("cabBAC->ABCabc", E3p) if j < i else np.einsum("CABbac->ABCabc", E3p)
            Qp = np.einsum("aBAC->ABCa", E2p) if j < i else np.einsum("CABa->ABCa", E2p)
        Sbj[j] = coefT @ Mp.reshape(n3, n3)
        spj[j] = coefT @ Qp.reshape(n3, n)
    for x in range(next_):
        v = np.einsum("acb->abc", D[x]).reshape(-1)
        rmu = sp @ heff[:, x] + Sb @ v
        deno = w - ext_e[x] if plus else w + ext_e[x]
        for j in js:
            rj = spj[j] @ heff[:, x] + Sbj[j] @ v
            pc[j] -= float(np.sum(rmu * rj / deno))
    return sc, pc


# ---------------------------------------------------------------------------- Heff
CLASS_OFF = [(1, v1k_off), (2, vm1k_off), (3, v2mod_off), (4, vm2mod_off), (7, v0pmod_off),
             (6, lambda c, T, thr: pm1_prime_off(c, T, False, thr)),
             (5, lambda c, T, thr: pm1_prime_off(c, T, True, thr))]


def heff_matrices(P, koop, thr=THR_LINDEP, with_4rdm=True, frozen=()):
    """Second-order Heff (SC, PC) without E0, e2mp/e2en of qdnevpt.F (column i = perturbers of state i),
    the per-state class tables, and the per-class off-diagonal contributions off[k] (SC, PC)."""
    M = P.metat
    T = Trans(P)
    Hsc, Hpc = np.zeros((M, M)), np.zeros((M, M))
    tables = []
    off = {}
    for i in range(M):
        tab, c = S2.state_energies(P, koop, i, thr, with_4rdm, frozen)
        tables.append(tab)
        Hsc[i, i] = tab[:, 1].sum()
        Hpc[i, i] = tab[:, 3].sum()
        for k, fn in CLASS_OFF:
            if k in (5, 6) and not with_4rdm:
                continue
            sc, pc = fn(c, T, thr)
            off.setdefault(k, (np.zeros((M, M)), np.zeros((M, M))))
            off[k][0][:, i] += sc
            off[k][1][:, i] += pc
            Hsc[:, i] += sc
            Hpc[:, i] += pc
    return Hsc, Hpc, np.array(tables), off


def diagonalize_heff(H, e0):
    """qdnevpt.F:1056-1112 + hermitian_heff (:8089-8184): returns (sorted eigenvalues, Hermitian Heff~)."""
    A = np.array(H, dtype=float) + np.diag(e0)
    w, vr = np.linalg.eig(A)
    if np.abs(np.imag(w)).max(initial=0.0) > 0.0:
        raise ValueError("complex Heff eigenvalues")
    w, vr = np.real(w), np.real(vr)
    order = np.argsort(w, kind="stable")
    w, vr = w[order], vr[:, order]
    S = vr.T @ vr
    sw, U = np.linalg.eigh(S, UPLO="U")
    Sm12 = (U / np.sqrt(sw)) @ U.T
    C = vr @ Sm12
    Ht = C @ np.diag(w) @ C.T
    e = np.linalg.eigvalsh(Ht, UPLO="U")
    return e, Ht

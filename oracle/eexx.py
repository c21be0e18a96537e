"""Accessor closed forms: <E E E E>, <E E E>, <E E> and their hole variants.

Oracle (test infrastructure only).  Indices are 0-based; RDM arrays carry the
state index first (``daaa[s, p1,p2,p3, q1,q2,q3]``), ``ro4`` is the packed
reference layout ``ro4[word, i5,i6,i7,i8, s]``.

Scalar forms follow eexx_functions.F90 line by line; ``s2_*`` follow the
stage-2 copies with swapped arguments (qdnevpt.F:6448-6519).  ``T_*`` build the
same quantities as full tensors over all index values (used by the vectorised
builders).  Parity: scalar forms pinned against the reference's own compiled
accessor modules (oracle/ref) and against second-quantisation brute force
(oracle/brute); tensor forms pinned against the scalar forms.
"""
from __future__ import annotations

import numpy as np


# ----------------------------------------------------------------------------
# packed 4-RDM gather: ord8 (ord_utils.F90:71-129) + lindice (indices_norm.F90:41-44)
# ----------------------------------------------------------------------------
def norm4(t):
    s = sorted(t, reverse=True)
    s = [x + 1 for x in s]
    return (s[0] - 1) * s[0] * (s[0] + 1) * (s[0] + 2) // 24 + (s[1] - 1) * s[1] * (s[1] + 1) // 6 + \
        (s[2] - 1) * s[2] // 2 + s[3]


def ord8(c, a):
    """Descending selection sort of the creation 4-tuple with partners alongside (ord_utils.F90:71-129)."""
    if norm4(c) >= norm4(a):
        nv1, nv2 = list(c), list(a)
    else:
        nv1, nv2 = list(a), list(c)
    for i in range(4):
        imax, jmax = nv1[i], i
        for j in range(i + 1, 4):
            if nv1[j] > imax:
                imax, jmax = nv1[j], j
        if jmax != i:
            nv1[i], nv1[jmax] = nv1[jmax], nv1[i]
            nv2[i], nv2[jmax] = nv2[jmax], nv2[i]
    return nv1, nv2


def lindice0(i1, i2, i3, i4):
    """0-based packed word of a descending 0-based 4-tuple."""
    a, b, c, d = i1 + 1, i2 + 1, i3 + 1, i4 + 1
    return (a - 1) * a * (a + 1) * (a + 2) // 24 + (b - 1) * b * (b + 1) // 6 + (c - 1) * c // 2 + d - 1


def g4(ro4, s, c, a):
    nv1, nv2 = ord8(c, a)
    return ro4[lindice0(*nv1), nv2[0], nv2[1], nv2[2], nv2[3], s]


# ----------------------------------------------------------------------------
# stage-1 accessors (eexx_functions.F90)
# ----------------------------------------------------------------------------
def eeee(a, b, c, d, e, f, g, h, d3, d2, d1, ro4, s):
    """<E_ab E_cd E_ef E_gh> (eexx_functions.F90:25-61)."""
    v = g4(ro4, s, (a, c, e, g), (b, d, f, h))
    if b == c: v += d3[s, a, e, g, d, f, h]
    if b == e: v += d3[s, a, c, g, f, d, h]
    if b == g: v += d3[s, a, c, e, h, d, f]
    if d == e: v += d3[s, a, g, c, b, h, f]
    if d == g: v += d3[s, a, c, e, b, h, f]
    if f == g: v += d3[s, a, e, c, b, h, d]
    if b == c:
        if d == e: v += d2[s, a, g, f, h]
        if d == g: v += d2[s, a, e, h, f]
        if f == g: v += d2[s, a, e, d, h]
    if b == e and d == g: v += d2[s, a, c, f, h]
    if b == e and f == g: v += d2[s, a, c, h, d]
    if b == g and d == e: v += d2[s, a, c, h, f]
    if d == e and f == g: v += d2[s, a, c, b, h]
    if b == c and d == e and f == g: v += d1[s, a, h]
    return v


def eeeet(a, b, c, d, e, f, g, h, d3, d2, d1, ro4, s):
    """eexx_functions.F90:64-76."""
    v = -eeee(a, b, d, c, e, f, g, h, d3, d2, d1, ro4, s)
    if c == d: v += 2.0 * eee2(a, b, e, f, g, h, d3, d2, d1, s)
    return v


def eee2(a, b, c, d, e, f, daaa, taa, dal, s):
    """<E_ab E_cd E_ef> (eexx_functions.F90:79-94)."""
    v = daaa[s, a, c, e, b, d, f]
    if e == d: v += taa[s, a, c, b, f]
    if e == b: v += taa[s, a, c, f, d]
    if b == c: v += taa[s, a, e, d, f]
    if b == c and e == d: v += dal[s, a, f]
    return v


def eee2t(a, b, c, d, e, f, daaa, taa, dal, s):
    """eexx_functions.F90:97-108."""
    v = -eee2(a, b, d, c, e, f, daaa, taa, dal, s)
    if c == d: v += 2.0 * ee2(a, b, e, f, taa, dal, s)
    return v


def eee2tl(a, b, c, d, e, f, daaa, taa, dal, s):
    """eexx_functions.F90:111-122."""
    v = -eee2(b, a, c, d, e, f, daaa, taa, dal, s)
    if a == b: v += 2.0 * ee2(c, d, e, f, taa, dal, s)
    return v


def ee2(a, b, c, d, taa, dal, s):
    """<E_ab E_cd> (eexx_functions.F90:125-135)."""
    v = taa[s, a, c, b, d]
    if b == c: v += dal[s, a, d]
    return v


def ee2t(a, b, c, d, taa, dal, s):
    """eexx_functions.F90:138-148."""
    v = -ee2(b, a, c, d, taa, dal, s)
    if a == b: v += 2.0 * dal[s, c, d]
    return v


def ee2tr(a, b, c, d, taa, dal, s):
    """eexx_functions.F90:151-161."""
    v = -ee2(a, b, d, c, taa, dal, s)
    if c == d: v += 2.0 * dal[s, a, b]
    return v


def ro3t(a, b, c, ap, bp, cp, d3, d2, d1, s):
    """Three-hole density (eexx_functions.F90:164-207)."""
    v = -d3[s, ap, bp, cp, a, b, c]
    if ap == a: v += 2.0 * d2[s, bp, cp, b, c]
    if bp == b: v += 2.0 * d2[s, ap, cp, a, c]
    if cp == c: v += 2.0 * d2[s, ap, bp, a, b]
    if bp == c: v -= d2[s, ap, cp, a, b]
    if ap == c: v -= d2[s, cp, bp, a, b]
    if cp == b: v -= d2[s, bp, ap, c, a]
    if ap == b: v -= d2[s, bp, cp, a, c]
    if cp == a: v -= d2[s, ap, bp, c, b]
    if bp == a: v -= d2[s, ap, cp, b, c]
    if cp == c and bp == b: v -= 4.0 * d1[s, ap, a]
    if cp == b and bp == c: v += 2.0 * d1[s, ap, a]
    if cp == c and ap == a: v -= 4.0 * d1[s, bp, b]
    if cp == a and ap == c: v += 2.0 * d1[s, bp, b]
    if ap == a and bp == b: v -= 4.0 * d1[s, cp, c]
    if ap == b and bp == a: v += 2.0 * d1[s, cp, c]
    if cp == c and ap == b: v += 2.0 * d1[s, bp, a]
    if cp == b and ap == c: v -= d1[s, bp, a]
    if cp == c and bp == a: v += 2.0 * d1[s, ap, b]
    if cp == a and bp == c: v -= d1[s, ap, b]
    if bp == b and ap == c: v += 2.0 * d1[s, cp, a]
    if bp == c and ap == b: v -= d1[s, cp, a]
    if bp == b and cp == a: v += 2.0 * d1[s, ap, c]
    if bp == a and cp == b: v -= d1[s, ap, c]
    if ap == a and cp == b: v += 2.0 * d1[s, bp, c]
    if ap == b and cp == a: v -= d1[s, bp, c]
    if ap == a and bp == c: v += 2.0 * d1[s, cp, b]
    if ap == c and bp == a: v -= d1[s, cp, b]
    if ap == a and bp == b and cp == c: v += 8.0
    if ap == b and bp == a and cp == c: v -= 4.0
    if ap == a and bp == c and cp == b: v -= 4.0
    if ap == c and bp == b and cp == a: v -= 4.0
    if ap == b and bp == c and cp == a: v += 2.0
    if ap == c and bp == a and cp == b: v += 2.0
    return v


def ro2t(ap, bp, a, b, d2, d1, s):
    """Two-hole density (eexx_functions.F90:210-225)."""
    v = d2[s, a, b, ap, bp]
    if ap == a: v -= 2.0 * d1[s, b, bp]
    if bp == b: v -= 2.0 * d1[s, a, ap]
    if ap == b: v += d1[s, a, bp]
    if bp == a: v += d1[s, b, ap]
    if ap == a and bp == b: v += 4.0
    if ap == b and bp == a: v -= 2.0
    return v


# ----------------------------------------------------------------------------
# stage-2 copies with swapped arguments (qdnevpt.F:6448-6519)
# ----------------------------------------------------------------------------
def s2_eee2(a, b, c, d, e, f, daaa, taa, dal, s):
    """qdnevpt.F:6450-6460."""
    v = daaa[s, b, d, f, a, c, e]
    if e == d: v += taa[s, b, f, a, c]
    if e == b: v += taa[s, f, d, a, c]
    if b == c: v += taa[s, d, f, a, e]
    if b == c and e == d: v += dal[s, f, a]
    return v


def s2_ee2(a, b, c, d, taa, dal, s):
    """qdnevpt.F:6462-6469."""
    v = taa[s, b, d, a, c]
    if b == c: v += dal[s, d, a]
    return v


def s2_ee2_dypc(a, b, c, d, taa, dal, s):
    """qdnevpt.F:6471-6478."""
    v = taa[s, a, c, b, d]
    if b == c: v += dal[s, a, d]
    return v


def s2_eee2t(a, b, c, d, e, f, daaa, taa, dal, s):
    """qdnevpt.F:6480-6488."""
    v = -s2_eee2(a, b, d, c, e, f, daaa, taa, dal, s)
    if c == d: v += 2.0 * s2_ee2(a, b, e, f, taa, dal, s)
    return v


def s2_ee2tr(a, b, c, d, taa, dal, s):
    """qdnevpt.F:6490-6497."""
    v = -s2_ee2(a, b, d, c, taa, dal, s)
    if c == d: v += 2.0 * dal[s, b, a]
    return v


def s2_ee2tl(a, b, c, d, taa, dal, s):
    """qdnevpt.F:6499-6506."""
    v = -s2_ee2(b, a, c, d, taa, dal, s)
    if a == b: v += 2.0 * dal[s, d, c]
    return v


def s2_ro2toff(ap, bp, a, b, d2, d1, ij):
    """qdnevpt.F:6508-6518 (transition two-hole density; no constant terms)."""
    v = d2[ij, ap, bp, a, b]
    if ap == a: v -= 2.0 * d1[ij, bp, b]
    if bp == b: v -= 2.0 * d1[ij, ap, a]
    if ap == b: v += d1[ij, bp, a]
    if bp == a: v += d1[ij, ap, b]
    return v


# ----------------------------------------------------------------------------
# full tensors (one state); d3 = daaa[s], d2 = taa[s], d1 = dal[s]
# ----------------------------------------------------------------------------
def _delta(n):
    return np.eye(n)


def T_ee2(d2, d1):
    """E2[a,b,c,d] = ee2(a,b,c,d)."""
    n = d1.shape[0]
    I = _delta(n)
    return d2.transpose(0, 2, 1, 3) + np.einsum("bc,ad->abcd", I, d1)


def T_ee2t(d2, d1):
    n = d1.shape[0]
    I = _delta(n)
    E2 = T_ee2(d2, d1)
    return -E2.transpose(1, 0, 2, 3) + 2.0 * np.einsum("ab,cd->abcd", I, d1)


def T_ee2tr(d2, d1):
    n = d1.shape[0]
    I = _delta(n)
    E2 = T_ee2(d2, d1)
    return -E2.transpose(0, 1, 3, 2) + 2.0 * np.einsum("cd,ab->abcd", I, d1)


def T_eee2(d3, d2, d1):
    """E3[a,b,c,d,e,f] = eee2(a..f) = daaa(a,c,e,b,d,f) + delta terms."""
    n = d1.shape[0]
    I = _delta(n)
    E3 = d3.transpose(0, 3, 1, 4, 2, 5).copy()
    E3 += np.einsum("ed,acbf->abcdef", I, d2)
    E3 += np.einsum("eb,acfd->abcdef", I, d2)
    E3 += np.einsum("bc,aedf->abcdef", I, d2)
    E3 += np.einsum("bc,ed,af->abcdef", I, I, d1)
    return E3


def T_eee2t(d3, d2, d1, E3=None, E2=None):
    n = d1.shape[0]
    I = _delta(n)
    if E3 is None: E3 = T_eee2(d3, d2, d1)
    if E2 is None: E2 = T_ee2(d2, d1)
    return -E3.transpose(0, 1, 3, 2, 4, 5) + 2.0 * np.einsum("cd,abef->abcdef", I, E2)


def T_eee2tl(d3, d2, d1, E3=None, E2=None):
    n = d1.shape[0]
    I = _delta(n)
    if E3 is None: E3 = T_eee2(d3, d2, d1)
    if E2 is None: E2 = T_ee2(d2, d1)
    return -E3.transpose(1, 0, 2, 3, 4, 5) + 2.0 * np.einsum("ab,cdef->abcdef", I, E2)


def T_eeee(G4u, d3, d2, d1):
    """E4[a,b,c,d,e,f,g,h] = eeee(a..h); G4u is the unpacked ro4 gather G4u[c1..c4, a1..a4]."""
    n = d1.shape[0]
    I = _delta(n)
    E4 = G4u.transpose(0, 4, 1, 5, 2, 6, 3, 7).copy()          # [a,b,c,d,e,f,g,h] <- G(a,c,e,g; b,d,f,h)
    E4 += np.einsum("bc,aegdfh->abcdefgh", I, d3, optimize=True)
    E4 += np.einsum("be,acgfdh->abcdefgh", I, d3, optimize=True)
    E4 += np.einsum("bg,acehdf->abcdefgh", I, d3, optimize=True)
    E4 += np.einsum("de,agcbhf->abcdefgh", I, d3, optimize=True)
    E4 += np.einsum("dg,acebhf->abcdefgh", I, d3, optimize=True)
    E4 += np.einsum("fg,aecbhd->abcdefgh", I, d3, optimize=True)
    E4 += np.einsum("bc,de,agfh->abcdefgh", I, I, d2, optimize=True)
    E4 += np.einsum("bc,dg,aehf->abcdefgh", I, I, d2, optimize=True)
    E4 += np.einsum("bc,fg,aedh->abcdefgh", I, I, d2, optimize=True)
    E4 += np.einsum("be,dg,acfh->abcdefgh", I, I, d2, optimize=True)
    E4 += np.einsum("be,fg,achd->abcdefgh", I, I, d2, optimize=True)
    E4 += np.einsum("bg,de,achf->abcdefgh", I, I, d2, optimize=True)
    E4 += np.einsum("de,fg,acbh->abcdefgh", I, I, d2, optimize=True)
    E4 += np.einsum("bc,de,fg,ah->abcdefgh", I, I, I, d1, optimize=True)
    return E4


def T_ro2t(d2, d1):
    """R2[ap,bp,a,b] = ro2t(ap,bp,a,b)."""
    n = d1.shape[0]
    I = _delta(n)
    R = d2.transpose(2, 3, 0, 1).copy()                           # d2(a,b,ap,bp)
    R -= 2.0 * np.einsum("Pa,bQ->PQab", I, d1)
    R -= 2.0 * np.einsum("Qb,aP->PQab", I, d1)
    R += np.einsum("Pb,aQ->PQab", I, d1)
    R += np.einsum("Qa,bP->PQab", I, d1)
    R += 4.0 * np.einsum("Pa,Qb->PQab", I, I)
    R -= 2.0 * np.einsum("Pb,Qa->PQab", I, I)
    return R


def T_ro3t(d3, d2, d1):
    """R3[a,b,c,ap,bp,cp] = ro3t(a,b,c,ap,bp,cp) (eexx_functions.F90:164-207)."""
    n = d1.shape[0]
    I = _delta(n)
    e = lambda *args: np.einsum(*args, optimize=True)
    # letters: a b c = a,b,c ; A B C = ap,bp,cp ; output 'abcABC'
    R = -d3.transpose(3, 4, 5, 0, 1, 2).copy()                   # d3(ap,bp,cp,a,b,c)
    R += 2.0 * e("Aa,BCbc->abcABC", I, d2)
    R += 2.0 * e("Bb,ACac->abcABC", I, d2)
    R += 2.0 * e("Cc,ABab->abcABC", I, d2)
    R -= e("Bc,ACab->abcABC", I, d2)
    R -= e("Ac,CBab->abcABC", I, d2)
    R -= e("Cb,BAca->abcABC", I, d2)
    R -= e("Ab,BCac->abcABC", I, d2)
    R -= e("Ca,ABcb->abcABC", I, d2)
    R -= e("Ba,ACbc->abcABC", I, d2)
    R -= 4.0 * e("Cc,Bb,Aa->abcABC", I, I, d1)
    R += 2.0 * e("Cb,Bc,Aa->abcABC", I, I, d1)
    R -= 4.0 * e("Cc,Aa,Bb->abcABC", I, I, d1)
    R += 2.0 * e("Ca,Ac,Bb->abcABC", I, I, d1)
    R -= 4.0 * e("Aa,Bb,Cc->abcABC", I, I, d1)
    R += 2.0 * e("Ab,Ba,Cc->abcABC", I, I, d1)
    R += 2.0 * e("Cc,Ab,Ba->abcABC", I, I, d1)
    R -= e("Cb,Ac,Ba->abcABC", I, I, d1)
    R += 2.0 * e("Cc,Ba,Ab->abcABC", I, I, d1)
    R -= e("Ca,Bc,Ab->abcABC", I, I, d1)
    R += 2.0 * e("Bb,Ac,Ca->abcABC", I, I, d1)
    R -= e("Bc,Ab,Ca->abcABC", I, I, d1)
    R += 2.0 * e("Bb,Ca,Ac->abcABC", I, I, d1)
    R -= e("Ba,Cb,Ac->abcABC", I, I, d1)
    R += 2.0 * e("Aa,Cb,Bc->abcABC", I, I, d1)
    R -= e("Ab,Ca,Bc->abcABC", I, I, d1)
    R += 2.0 * e("Aa,Bc,Cb->abcABC", I, I, d1)
    R -= e("Ac,Ba,Cb->abcABC", I, I, d1)
    R += 8.0 * e("Aa,Bb,Cc->abcABC", I, I, I)
    R -= 4.0 * e("Ab,Ba,Cc->abcABC", I, I, I)
    R -= 4.0 * e("Aa,Bc,Cb->abcABC", I, I, I)
    R -= 4.0 * e("Ac,Bb,Ca->abcABC", I, I, I)
    R += 2.0 * e("Ab,Bc,Ca->abcABC", I, I, I)
    R += 2.0 * e("Ac,Ba,Cb->abcABC", I, I, I)
    return R

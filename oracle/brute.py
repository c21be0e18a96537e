"""Second-quantisation brute force on tiny active spaces.

Oracle (test infrastructure only).  Builds |Psi> as a dict over spin-orbital
bit strings (alpha orbital p -> bit p, beta -> bit n+p) and applies spin-free
excitation operators E_pq = sum_sigma a+_{p sigma} a_{q sigma} directly.  It
pins, independently of any closed form:

* the RDM convention used by the generator and the accessors
  (``G_k(p..;q..) = <a+_p1 .. a+_pk a_qk .. a_q1>``, eexx_functions.F90:125-133),
* the accessors' closed forms: ee2 = <E_ab E_cd>, eee2 = <E_ab E_cd E_ef>,
  eeee = <E_ab E_cd E_ef E_gh> (eexx_functions.F90:25-135).
"""
from __future__ import annotations

from itertools import combinations

import numpy as np


def ci_to_dict(C: np.ndarray, n: int, nalpha: int) -> dict:
    """C[I_alpha, I_beta] (strings in lexicographic combination order) -> {bits: coeff}."""
    strs = list(combinations(range(n), nalpha))
    psi = {}
    for ia, sa in enumerate(strs):
        ba = sum(1 << p for p in sa)
        for ib, sb in enumerate(strs):
            bb = sum(1 << (n + p) for p in sb)
            c = C[ia, ib]
            if c != 0.0:
                psi[ba | bb] = c
    return psi


def _apply_ex(psi: dict, p: int, q: int) -> dict:
    """a+_p a_q on spin-orbitals."""
    out = {}
    for det, c in psi.items():
        if not (det >> q) & 1:
            continue
        s = bin(det & ((1 << q) - 1)).count("1")
        d1 = det & ~(1 << q)
        if (d1 >> p) & 1:
            continue
        s += bin(d1 & ((1 << p) - 1)).count("1")
        d2 = d1 | (1 << p)
        out[d2] = out.get(d2, 0.0) + (-c if s & 1 else c)
    return out


def apply_E(psi: dict, p: int, q: int, n: int) -> dict:
    out = _apply_ex(psi, p, q)
    for det, c in _apply_ex(psi, n + p, n + q).items():
        out[det] = out.get(det, 0.0) + c
    return out


def dot(a: dict, b: dict) -> float:
    return sum(c * b.get(d, 0.0) for d, c in a.items())


def expect_E(psi: dict, pairs, n: int, bra: dict | None = None) -> float:
    """<bra| E_{p1 q1} E_{p2 q2} ... |psi> (rightmost operator applied first)."""
    v = psi
    for p, q in reversed(pairs):
        v = apply_E(v, p, q, n)
    return dot(psi if bra is None else bra, v)

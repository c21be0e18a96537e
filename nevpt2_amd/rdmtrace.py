"""Packed 4-RDM helpers and partial traces for input preparation (CPU, numpy).

The reference stores the spin-free 4-RDM packed over the descending-sorted
creation 4-tuple (``ro4(nwords,n,n,n,n,metat)``, koopro4QD.F90:263) and reads an
element through ``ord8`` + ``lindice`` (eexx_functions.F90:38-43).  The lower
RDMs are its partial traces ``bro3``/``bro2``/``bro1`` (koopro4QD.F90:4881-5049).
These vectorised CPU versions generate stage-2 inputs for seeded problems; the
GPU library computes the same quantities on the device.
"""
from __future__ import annotations

from itertools import product

import numpy as np


def _sorted_tables(n: int):
    """For every 4-tuple (0-based, C-order flat index): ord8's descending sort.

    Returns (norm[n^4], perm[n^4,4]) where ``perm`` is the position order that
    ord8's selection sort produces (ties keep the earlier position), and
    ``norm`` is the reference's packed index (1-based lindice of the sorted tuple).
    """
    tup = np.array(list(product(range(1, n + 1), repeat=4)), dtype=np.int64)   # 1-based
    N = tup.shape[0]
    perm = np.tile(np.arange(4), (N, 1))
    vals = tup.copy()
    for i in range(4):
        # selection: first position j >= i holding the maximum of vals[:, i:]
        sub = vals[:, i:]
        jrel = np.argmax(sub, axis=1)            # argmax returns the first maximum
        j = i + jrel
        rows = np.arange(N)
        vi = vals[rows, i].copy(); vj = vals[rows, j].copy()
        vals[rows, i] = vj; vals[rows, j] = vi
        pi = perm[rows, i].copy(); pj = perm[rows, j].copy()
        perm[rows, i] = pj; perm[rows, j] = pi
    s = vals
    norm = (s[:, 0] - 1) * s[:, 0] * (s[:, 0] + 1) * (s[:, 0] + 2) // 24 + (s[:, 1] - 1) * s[:, 1] * (s[:, 1] + 1) // 6 \
        + (s[:, 2] - 1) * s[:, 2] // 2 + s[:, 3]
    return norm, perm


def unpack_ro4(ro4s: np.ndarray) -> np.ndarray:
    """One state's packed ro4 (nwords,n,n,n,n) -> full G4[p1..p4,q1..q4] by the ord8/lindice gather."""
    n = ro4s.shape[1]
    norm, perm = _sorted_tables(n)
    N = n ** 4
    q = np.array(list(product(range(n), repeat=4)), dtype=np.int64)          # 0-based tuples
    # packed source as [word, flat(i5,i6,i7,i8)] with the partner tuple in C order
    Rc = np.ascontiguousarray(np.asarray(ro4s).reshape(ro4s.shape[0], N))
    out = np.empty((N, N))
    w = n ** np.arange(3, -1, -1)
    for p in range(N):
        # case A: norm(p) >= norm(q): creation tuple p, partners q permuted by p's sort
        qp = (q[:, perm[p]] * w).sum(axis=1)
        valA = Rc[norm[p] - 1, qp]
        # case B: norm(q) > norm(p): creation tuple q, partners p permuted by q's sort
        ptup = q[p]
        pq = (ptup[perm] * w).sum(axis=1)            # p permuted by each q's sort
        valB = Rc[norm - 1, pq]
        out[p] = np.where(norm[p] >= norm, valA, valB)
    return out.reshape((n,) * 8)


def bro3(G4u: np.ndarray, n: int, nele: int) -> np.ndarray:
    """Gamma3(i,j,k;ii,jj,kk) = sum_l G4(i,j,k,l;ii,jj,kk,l)/(N-3) (koopro4QD.F90:4881-4909)."""
    return np.einsum("ijklmnol->ijkmno", G4u) / (nele - 3)


def bro2(daaa: np.ndarray, n: int, nele: int) -> np.ndarray:
    """taa(s,i,j,ii,jj) = sum_l daaa(s,i,j,l,ii,jj,l)/(N-2) (koopro4QD.F90:4975-4999)."""
    return np.asfortranarray(np.einsum("sijlmnl->sijmn", daaa) / (nele - 2))


def bro1(taa: np.ndarray, n: int, nele: int) -> np.ndarray:
    """dal(s,i,j) = sum_l taa(s,i,l,j,l)/(N-1) (koopro4QD.F90:5028-5046)."""
    return np.asfortranarray(np.einsum("siljl->sij", taa) / (nele - 1))


def traces_from_ro4(ro4: np.ndarray, n: int, nele: int):
    """(daaa, taa, dal) with the state index first, Fortran ordered, from ro4(...,metat)."""
    M = ro4.shape[-1]
    daaa = np.zeros((M,) + (n,) * 6, order="F")
    for s in range(M):
        daaa[s] = bro3(unpack_ro4(ro4[..., s]), n, nele)
    taa = bro2(daaa, n, nele)
    dal = bro1(taa, n, nele)
    return daaa, taa, dal

"""Stage-1 host interface: the Koopmans-matrix builder of koopro4QD.

Mirrors the reference's stage-1 operator API (koopro4QD.F90, koop_matrices.F90)
with the same argument meaning and array layouts; every routine runs on the GPU
through libnevpt2_amd.so.  Arrays are numpy, Fortran-ordered, state index first.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L


def _n4(n):
    return n * (n + 1) * (n + 2) * (n + 3) // 24


def bro(ro4, nact, metat, nele):
    """bro3/bro2/bro1 (koopro4QD.F90:4881-5049): (daaa, taa, dal) from the packed 4-RDM."""
    n, M = nact, metat
    ro4 = L.fin(ro4)
    if ro4.shape != (_n4(n), n, n, n, n, M):
        raise ValueError(f"ro4 shape {ro4.shape} != {(_n4(n), n, n, n, n, M)}")
    daaa, taa, dal = L.zeros(M, *(n,) * 6), L.zeros(M, *(n,) * 4), L.zeros(M, n, n)
    L.check(L.lib().nv2_bro(n, M, nele, L.ptr(ro4), L.ptr(daaa), L.ptr(taa), L.ptr(dal)), "nv2_bro")
    return daaa, taa, dal


def get_aj_atwo(ncore, nact, norb, cholesky, oneint):
    """get_aj_atwo (koopro4QD.F90:3900): Dyall h (norb triangle) and active ERIs."""
    B = L.fin(cholesky)
    one = np.ascontiguousarray(oneint, dtype=np.float64)
    aj = np.zeros(norb * (norb + 1) // 2)
    atwo = L.zeros(nact, nact, nact, nact)
    L.check(L.lib().nv2_get_aj_atwo(ncore, nact, norb, B.shape[0], L.ptr(B), L.ptr(one), L.ptr(aj), L.ptr(atwo)),
            "nv2_get_aj_atwo")
    return aj, atwo


def koop_matrices(metat, nact, ncore, norb, atwo, daaa, taa, dal, aj):
    """koop_matrices_ctl + bbmat/bcmat/bdmat: every Koopmans matrix except the A-matrices.

    Returns a dict keyed by the reference's nevpt.h5 dataset names
    (koop_matrices.F90:53-157, koopro4QD.F90:383-418).
    """
    n, M = nact, metat
    s2, s4 = (M, n, n), (M, n, n, n, n)
    names = ["1-IPKO", "1-EAKO", "2-HRDM", "2-IPKO", "2-EAKO", "gk0pa", "gk0pd", "gk0pf",
             "bmatIP", "bmatEA", "cmatIP", "cmatEA", "dmatIP", "dmatEA"]
    shapes = [s2, s2, s4, s4, s4, s4, s4, s2, s4, s4, s4, s4, s2, s2]
    outs = [L.zeros(*s) for s in shapes]
    aj = np.ascontiguousarray(aj, dtype=np.float64)
    rc = L.lib().nv2_koop_matrices(M, n, ncore, norb, L.ptr(L.fin(atwo)), L.ptr(L.fin(daaa)), L.ptr(L.fin(taa)),
                                   L.ptr(L.fin(dal)), L.ptr(aj), *[L.ptr(o) for o in outs])
    L.check(rc, "nv2_koop_matrices")
    return dict(zip(names, outs))


def bamat(atwo, ro4, daaa, taa, dal, nact, ncore, aj, norb, metat, prec=L.PREC_FP64):
    """bamat (koopro4QD.F90:3974): (amat, atmat) = (amatIP, amatEA), shape (metat, n^6)."""
    n, M = nact, metat
    amat, atmat = L.zeros(M, *(n,) * 6), L.zeros(M, *(n,) * 6)
    aj = np.ascontiguousarray(aj, dtype=np.float64)
    rc = L.lib().nv2_bamat(n, ncore, norb, M, L.ptr(L.fin(atwo)), L.ptr(L.fin(ro4)), L.ptr(L.fin(daaa)),
                           L.ptr(L.fin(taa)), L.ptr(L.fin(dal)), L.ptr(aj), L.ptr(amat), L.ptr(atmat), int(prec))
    L.check(rc, "nv2_bamat")
    return amat, atmat

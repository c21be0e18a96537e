"""Stage-2 host interface: second-order energies of qdnevpt.

Mirrors the reference's per-state class evaluation (qdnevpt.F:858-993) and its
result table (qdnevpt.F:1019-1045); all arithmetic runs on the GPU through
libnevpt2_amd.so.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L

CLASS_NAMES = ["(0)", "(+1)", "(-1)", "(+2)", "(-2)", "(+1)'", "(-1)'", "(0)'"]
THR_LINDEP = 1.0e-6    # .LINDEP default (nevpt2_cfg.F90)

_KOOP_ORDER = ["1-IPKO", "1-EAKO", "2-HRDM", "2-IPKO", "2-EAKO", "gk0pa", "gk0pd", "gk0pf",
               "amatIP", "bmatIP", "cmatIP", "dmatIP", "amatEA", "bmatEA", "cmatEA", "dmatEA"]


def state_energies(ncore, nact, nvirt, metat, cholesky, oneint, dal, taa, daaa, koop, istate,
                   thr_lindep=THR_LINDEP, no_4rdm=False):
    """Per-class (norm_sc, e_sc, norm_pc, e_pc) of one state (0-based), shape (8, 4).

    ``koop`` maps the nevpt.h5 dataset names to the stage-1 matrices.
    """
    B = L.fin(cholesky)
    one = np.ascontiguousarray(oneint, dtype=np.float64)
    arrs = []
    for k in _KOOP_ORDER:
        if no_4rdm and k[1:4] == "mat":
            arrs.append(None)
        else:
            arrs.append(L.fin(koop[k]))
    ptrs = [L.ptr(a) if a is not None else None for a in arrs]
    out = np.zeros(32)
    rc = L.lib().nv2_qdnevpt_state(ncore, nact, nvirt, metat, B.shape[0], L.ptr(B), L.ptr(one), L.ptr(L.fin(dal)),
                                   L.ptr(L.fin(taa)), L.ptr(L.fin(daaa)), *ptrs, istate, float(thr_lindep),
                                   L.ptr(out))
    L.check(rc, "nv2_qdnevpt_state")
    return out.reshape(8, 4)


def format_table(tab, state=1):
    """The reference's per-state class table (format 1x,A,4f20.15; qdnevpt.F:1019-1045)."""
    lines = [f" state number:{state:4d}", " -----------------",
             " class        norm(sc)           energy(sc)           norm(pc)           energy(pc)"]
    labels = ["(0)  ", "(+1) ", "(-1) ", "(+2) ", "(-2) ", "(+1)'", "(-1)'", "(0)' "]
    for k, lab in enumerate(labels):
        lines.append(" " + lab + "".join(f"{v:20.15f}" for v in tab[k]))
    lines.append(" " + "-" * 85)
    tot = tab.sum(axis=0)
    lines.append(" Total" + "".join(f"{v:20.15f}" for v in tot))
    return "\n".join(lines)


def qd_heff(heff2, e0):
    """Hermitised zero+second-order Heff and its eigenvalues (the QD-NEVPT2 energies) on the GPU."""
    H = L.fin(np.asarray(heff2, dtype=np.float64))
    M = H.shape[0]
    e = np.ascontiguousarray(e0, dtype=np.float64)
    out = np.zeros((M, M), order="F")
    w = np.zeros(M)
    L.check(L.lib().nv2_qd_heff(M, L.ptr(H), L.ptr(e), L.ptr(out), L.ptr(w)), "nv2_qd_heff")
    return w, out

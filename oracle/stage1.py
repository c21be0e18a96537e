"""Stage-1 oracle entry points (ctypes over the C restatement in oracle/c).

Oracle (test infrastructure only).  Every function returns Fortran-ordered
numpy arrays in the reference layouts (state index first).  The C routines
restate koop_matrices.F90 and koopro4QD.F90:3900-5049 line by line.

Parity: every Koopmans matrix (1-/2-IPKO, 1-/2-EAKO, 2-HRDM, gk0pa/gk0pd/gk0pf,
A/B/C/D, IP and EA) equals its operator definition, a spin-summed
<Psi| O^+ [H_act, O'] |Psi> computed by Fock-space brute force (oracle/fock.py),
element by element at 1e-12 relative on random CI vectors; the accessors equal
the reference's own compiled eexx_functions.F90 (oracle/ref) bit for bit.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

i64 = C.c_int64
dp = C.POINTER(C.c_double)
lp = C.POINTER(C.c_int64)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "lib", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
        _LIB = C.CDLL(path)
        for name in ("orc_eeee", "orc_eeeet"):
            getattr(_LIB, name).restype = C.c_double
            getattr(_LIB, name).argtypes = [i64, i64, dp, dp, dp, dp, i64, lp]
        for name in ("orc_eee2", "orc_eee2t", "orc_eee2tl", "orc_ro3t"):
            getattr(_LIB, name).restype = C.c_double
            getattr(_LIB, name).argtypes = [i64, i64, dp, dp, dp, i64, lp]
    return _LIB


def ref():
    """The reference's own accessor modules (oracle/_ref), or None when not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_eexx.so")
        if not os.path.exists(path):
            return None
        _REF = C.CDLL(path)
        _REF.ref_init.argtypes = [i64]
        _REF.ref_nwords.restype = i64
        for name in ("ref_eeee", "ref_eeeet"):
            getattr(_REF, name).restype = C.c_double
            getattr(_REF, name).argtypes = [lp, dp, dp, dp, dp, i64, i64, i64]
        _REF.ref_eee2.restype = C.c_double
        _REF.ref_eee2.argtypes = [i64, lp, dp, dp, dp, i64, i64, i64]
        _REF.ref_ee2.restype = C.c_double
        _REF.ref_ee2.argtypes = [i64, lp, dp, dp, i64, i64, i64]
    return _REF


def P(a):
    assert a.flags.f_contiguous or a.flags.c_contiguous and a.ndim <= 1
    return a.ctypes.data_as(dp)


def F(a):
    return np.asfortranarray(a, dtype=np.float64)


def _z(*shape):
    return np.zeros(shape, order="F")


def bro3(ro4, n, M, nele):
    out = _z(M, *(n,) * 6)
    lib().orc_bro3(i64(n), i64(M), i64(nele), P(F(ro4)), P(out))
    return out


def bro2(daaa, n, M, nele):
    out = _z(M, *(n,) * 4)
    lib().orc_bro2(i64(n), i64(M), i64(nele), P(F(daaa)), P(out))
    return out


def bro1(taa, n, M, nele):
    out = _z(M, n, n)
    lib().orc_bro1(i64(n), i64(M), i64(nele), P(F(taa)), P(out))
    return out


def get_aj_atwo(ncore, nact, norb, B, oneint):
    L = B.shape[0]
    aj = np.zeros(norb * (norb + 1) // 2)
    atwo = _z(nact, nact, nact, nact)
    lib().orc_get_aj_atwo(i64(ncore), i64(nact), i64(norb), i64(L), P(F(B)), P(np.ascontiguousarray(oneint)),
                          P(aj), P(atwo))
    return aj, atwo


def koopE(n, ncore, M, atwo, taa, dal, aj):
    ipa, eaa = _z(M, n, n), _z(M, n, n)
    lib().orc_koopE(i64(n), i64(ncore), i64(M), P(F(atwo)), P(F(taa)), P(F(dal)), P(np.ascontiguousarray(aj)),
                    P(ipa), P(eaa))
    return ipa, eaa


def hrdm(n, M, taa, dal):
    out = _z(M, n, n, n, n)
    lib().orc_hrdm(i64(n), i64(M), P(F(taa)), P(F(dal)), P(out))
    return out


def koop2E(n, ncore, M, atwo, daaa, taa, dal, aj):
    ip, ea = _z(M, n, n, n, n), _z(M, n, n, n, n)
    lib().orc_koop2E(i64(n), i64(ncore), i64(M), P(F(atwo)), P(F(daaa)), P(F(taa)), P(F(dal)),
                     P(np.ascontiguousarray(aj)), P(ip), P(ea))
    return ip, ea


def koopman0pE(n, ncore, M, atwo, daaa, taa, dal, aj):
    ga, gd, gf = _z(M, n, n, n, n), _z(M, n, n, n, n), _z(M, n, n)
    lib().orc_koopman0pE(i64(n), i64(ncore), i64(M), P(F(atwo)), P(F(daaa)), P(F(taa)), P(F(dal)),
                         P(np.ascontiguousarray(aj)), P(ga), P(gd), P(gf))
    return ga, gd, gf


def bamat(n, ncore, M, atwo, ro4, daaa, taa, dal, aj, pair_limit=-1):
    """A-matrices; pair_limit >= 0 computes only the first (a,c) outer pairs (timing samples)."""
    am, at = _z(M, *(n,) * 6), _z(M, *(n,) * 6)
    lib().orc_bamat(i64(n), i64(ncore), i64(M), P(F(atwo)), P(F(ro4)), P(F(daaa)), P(F(taa)), P(F(dal)),
                    P(np.ascontiguousarray(aj)), P(am), P(at), i64(pair_limit))
    return am, at


def bbmat(n, ncore, M, atwo, daaa, taa, dal, aj):
    b, bt = _z(M, n, n, n, n), _z(M, n, n, n, n)
    lib().orc_bbmat(i64(n), i64(ncore), i64(M), P(F(atwo)), P(F(daaa)), P(F(taa)), P(F(dal)),
                    P(np.ascontiguousarray(aj)), P(b), P(bt))
    return b, bt


def bcmat(n, ncore, M, atwo, daaa, taa, dal, aj):
    c, ct = _z(M, n, n, n, n), _z(M, n, n, n, n)
    lib().orc_bcmat(i64(n), i64(ncore), i64(M), P(F(atwo)), P(F(daaa)), P(F(taa)), P(F(dal)),
                    P(np.ascontiguousarray(aj)), P(c), P(ct))
    return c, ct


def bdmat(n, ncore, M, atwo, taa, dal, aj):
    d, dt = _z(M, n, n), _z(M, n, n)
    lib().orc_bdmat(i64(n), i64(ncore), i64(M), P(F(atwo)), P(F(taa)), P(F(dal)),
                    P(np.ascontiguousarray(aj)), P(d), P(dt))
    return d, dt


def koopmans_all(prob):
    """All stage-1 outputs for a synth.Problem, keyed by the nevpt.h5 dataset names."""
    n, nc, M, nel = prob.nact, prob.ncore, prob.metat, prob.nele
    aj, atwo = get_aj_atwo(nc, n, prob.norb, prob.B, prob.oneint)
    daaa, taa, dal = prob.daaa, prob.taa, prob.dal
    out = {"aj": aj, "atwo": atwo}
    out["1-IPKO"], out["1-EAKO"] = koopE(n, nc, M, atwo, taa, dal, aj)
    out["2-HRDM"] = hrdm(n, M, taa, dal)
    out["2-IPKO"], out["2-EAKO"] = koop2E(n, nc, M, atwo, daaa, taa, dal, aj)
    out["gk0pa"], out["gk0pd"], out["gk0pf"] = koopman0pE(n, nc, M, atwo, daaa, taa, dal, aj)
    out["amatIP"], out["amatEA"] = bamat(n, nc, M, atwo, prob.ro4, daaa, taa, dal, aj)
    out["bmatIP"], out["bmatEA"] = bbmat(n, nc, M, atwo, daaa, taa, dal, aj)
    out["cmatIP"], out["cmatEA"] = bcmat(n, nc, M, atwo, daaa, taa, dal, aj)
    out["dmatIP"], out["dmatEA"] = bdmat(n, nc, M, atwo, taa, dal, aj)
    return out

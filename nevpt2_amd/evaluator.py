"""Device-resident NEVPT2 evaluation (stage 1 + stage 2 in one call).

Replaces the reference's two-executable pipeline koopro4QD.x -> nevpt.h5 ->
QDnevpt2.x (koopro4QD_main.F90, qdnevpt_main.F90) for the hot path: the packed
4-RDM and the Cholesky integrals are uploaded once and every intermediate stays
in HBM.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

PHASES = ["partial traces", "Koopmans <=3-RDM + B/C/D", "E4 build", "E4 contractions", "bamat rest", "stage 2"]


class Evaluator:
    def __init__(self, ncore, nact, nvirt, metat, nele, nchovec):
        self.nc, self.n, self.nv, self.M, self.nele, self.L = ncore, nact, nvirt, metat, nele, nchovec
        self.norb = ncore + nact + nvirt
        h = L.lib().nv2_ctx_create(ncore, nact, nvirt, metat, nele, nchovec)
        if not h:
            raise L.NV2Error(f"nv2_ctx_create failed: {L.last_error()}")
        self._h = h

    @classmethod
    def for_problem(cls, P, rdms="host"):
        """Context for a synthetic problem; rdms="gpu" builds the RDMs on the device from P.C."""
        ev = cls(P.ncore, P.nact, P.nvirt, P.metat, P.nele, P.B.shape[0])
        ev.set_integrals(P.B, P.oneint)
        if rdms == "gpu":
            ev.set_ci(P.C, P.nele // 2, P.nele - P.nele // 2)
            return ev
        ev.set_ro4(P.ro4)
        if P.metat > 1:
            ev.set_trdm(P.daaaoff, P.taaoff, P.daloff)
        return ev

    def set_ci(self, C, nalpha, nbeta):
        """Builds ro4 (and, for several states, the transition RDMs) on the GPU from CI vectors
        C[root, I_alpha, I_beta] over lexicographic strings."""
        Cc = np.ascontiguousarray(C, dtype=np.float64)
        L.check(L.lib().nv2_ctx_set_ci(self._h, int(nalpha), int(nbeta), L.ptr(Cc)), "nv2_ctx_set_ci")

    def set_trdm(self, trdm3, trdm2, trdm1):
        """Transition RDMs (3-TRDM, 2-TRDM, 1-TRDM) of the state pairs, for QD-NEVPT2."""
        L.check(L.lib().nv2_ctx_set_trdm(self._h, L.ptr(L.fin(trdm3)), L.ptr(L.fin(trdm2)), L.ptr(L.fin(trdm1))),
                "nv2_ctx_set_trdm")

    def heff(self):
        """Second-order effective Hamiltonians (SC, PC) without E0; element [j, i] = e2mp(j,i)."""
        M = self.M
        sc = np.zeros(M * M)
        pc = np.zeros(M * M)
        L.check(L.lib().nv2_ctx_get(self._h, b"Heff-SC", L.ptr(sc), sc.size), "nv2_ctx_get(Heff-SC)")
        L.check(L.lib().nv2_ctx_get(self._h, b"Heff-PC", L.ptr(pc), pc.size), "nv2_ctx_get(Heff-PC)")
        return sc.reshape(M, M, order="F"), pc.reshape(M, M, order="F")

    def qd_terms(self):
        """Per-class QD couplings: array [i, k, 0|1, j] = SC|PC contribution of class k to Heff(j, i)."""
        M = self.M
        q = np.zeros(16 * M * M)
        L.check(L.lib().nv2_ctx_get(self._h, b"QD", L.ptr(q), q.size), "nv2_ctx_get(QD)")
        return q.reshape(M, 8, 2, M)

    def pc_metric(self):
        """Per state and class: (smallest metric eigenvalue, kept, Cholesky path flag), shape (metat, 8, 3)."""
        a = np.zeros(24 * self.M)
        L.check(L.lib().nv2_ctx_get(self._h, b"PC-metric", L.ptr(a), a.size), "nv2_ctx_get(PC-metric)")
        return a.reshape(self.M, 8, 3)

    def pc_timings(self):
        """Per state: (device ms of the (+-1)' K' reduction, its order, its batch, form: 1 tridiagonal,
        2 band), shape (metat, 4)."""
        a = np.zeros(4 * self.M)
        L.check(L.lib().nv2_ctx_get(self._h, b"PC-timings", L.ptr(a), a.size), "nv2_ctx_get(PC-timings)")
        return a.reshape(self.M, 4)

    def pc_graphs(self):
        """hipGraph launches so far of the batched eigensolver call sites (worker-0 metrics, worker-0 K',
        Fock blocks): the first call with a problem list tunes and solves, the second captures, later
        ones replay."""
        a = np.zeros(3)
        L.check(L.lib().nv2_ctx_get(self._h, b"PC-graphs", L.ptr(a), 3), "nv2_ctx_get(PC-graphs)")
        return a

    def set_option(self, name, value):
        """Run options: "no4rdm" (the reference's .NO4RDM) skips the A-D matrices and the (+-1)' classes;
        "natorb" (.NATORB) also forms each state's first-order density matrix (see natorb())."""
        L.check(L.lib().nv2_ctx_set_option(self._h, name.encode(), int(value)), f"nv2_ctx_set_option({name})")

    def natorb(self):
        """After a run with the "natorb" option: per state the first-order density matrix in its canonical
        orbitals, the same matrix in the input orbitals, and its eigenvalues in ascending order; shapes
        (metat, norb, norb), (metat, norb, norb), (metat, norb)."""
        no, M = self.norb, self.M
        out = []
        for name, cnt in (("rho1", no * no), ("rho1-mo", no * no), ("natocc", no)):
            a = np.zeros(cnt * M)
            L.check(L.lib().nv2_ctx_get(self._h, name.encode(), L.ptr(a), a.size), f"nv2_ctx_get({name})")
            out.append(a.reshape(M, no, no).transpose(0, 2, 1).copy() if cnt == no * no else a.reshape(M, no))
        return tuple(out)

    def set_frozen(self, orbitals):
        """The reference's .FROZEN: 1-based canonical core orbital numbers that are never holes."""
        a = np.ascontiguousarray(orbitals, dtype=np.int64)
        L.check(L.lib().nv2_ctx_set_frozen(self._h, int(a.size), a.ctypes.data_as(C.POINTER(C.c_int64))),
                "nv2_ctx_set_frozen")

    def set_integrals(self, cholesky, oneint):
        B = L.fin(cholesky)
        one = np.ascontiguousarray(oneint, dtype=np.float64)
        L.check(L.lib().nv2_ctx_set_integrals(self._h, L.ptr(B), L.ptr(one)), "nv2_ctx_set_integrals")

    def set_ro4(self, ro4):
        L.check(L.lib().nv2_ctx_set_ro4(self._h, L.ptr(L.fin(ro4))), "nv2_ctx_set_ro4")

    def run(self, prec=L.PREC_FP64, thr_lindep=1.0e-6):
        """Per-state class table, shape (metat, 8, 4): (norm_sc, e_sc, norm_pc, e_pc)."""
        out = np.zeros(32 * self.M)
        L.check(L.lib().nv2_ctx_run(self._h, int(prec), float(thr_lindep), L.ptr(out)), "nv2_ctx_run")
        return out.reshape(self.M, 8, 4)

    def timings(self):
        ms = np.zeros(6)
        L.check(L.lib().nv2_ctx_timings(self._h, L.ptr(ms)), "nv2_ctx_timings")
        return dict(zip(PHASES, ms))

    def get(self, name, shape):
        a = np.zeros(shape, order="F")
        L.check(L.lib().nv2_ctx_get(self._h, name.encode(), L.ptr(a), a.size), f"nv2_ctx_get({name})")
        return a

    def close(self):
        if self._h:
            L.lib().nv2_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ci_rdm(n, nalpha, nbeta, k, bra, ket=None):
    """Spin-free k-RDM (or transition k-RDM) from CI vectors on the GPU, as an array [p1..pk, q1..qk]."""
    b = np.ascontiguousarray(bra, dtype=np.float64)
    kt = None if ket is None else np.ascontiguousarray(ket, dtype=np.float64)
    out = np.zeros(n ** (2 * k))
    L.check(L.lib().nv2_ci_rdm(int(n), int(nalpha), int(nbeta), int(k), L.ptr(b), None if kt is None else L.ptr(kt),
                               L.ptr(out)), "nv2_ci_rdm")
    return out.reshape((n,) * (2 * k), order="F")

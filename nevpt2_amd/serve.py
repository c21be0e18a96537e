"""Request loop: concurrent NEVPT2 evaluations on one GPU (nv2_server_* in the C ABI).

Requests are admitted first-in first-out when one of the server's slots is free
and their device memory fits the HBM budget; admitted requests run concurrently,
so a request that arrives while others are on the GPU overlaps with them.
Contexts stay resident per problem shape and are reused by later requests.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class _Request(C.Structure):
    _fields_ = [("ncore", C.c_int64), ("nact", C.c_int64), ("nvirt", C.c_int64), ("metat", C.c_int64),
                ("nele", C.c_int64), ("nchovec", C.c_int64),
                ("cholesky", L.dp), ("oneint", L.dp), ("ro4", L.dp),
                ("trdm3", L.dp), ("trdm2", L.dp), ("trdm1", L.dp),
                ("prec", C.c_int32), ("thr_lindep", C.c_double), ("out", L.dp), ("heff", L.dp)]


class Server:
    def __init__(self, slots=4, hbm_budget_bytes=0):
        h = L.lib().nv2_server_create(int(slots), int(hbm_budget_bytes))
        if not h:
            raise L.NV2Error(f"nv2_server_create failed: {L.last_error()}")
        self._h = h
        self._keep = {}          # request id -> arrays the library reads or writes until wait()

    def submit(self, P, prec=L.PREC_FP64, thr_lindep=1.0e-6):
        """Host-data request for a synthetic problem P (nevpt2_amd.synth.Problem); returns its id."""
        M = P.metat
        arrs = {"cholesky": L.fin(P.B), "oneint": np.ascontiguousarray(P.oneint, dtype=np.float64),
                "ro4": L.fin(P.ro4), "out": np.zeros(32 * M)}
        qd = M > 1 and getattr(P, "daaaoff", None) is not None
        if qd:
            arrs.update(trdm3=L.fin(P.daaaoff), trdm2=L.fin(P.taaoff), trdm1=L.fin(P.daloff), heff=np.zeros(2 * M * M))
        null = L.dp()
        r = _Request(P.ncore, P.nact, P.nvirt, M, P.nele, P.B.shape[0],
                     L.ptr(arrs["cholesky"]), L.ptr(arrs["oneint"]), L.ptr(arrs["ro4"]),
                     L.ptr(arrs["trdm3"]) if qd else null, L.ptr(arrs["trdm2"]) if qd else null,
                     L.ptr(arrs["trdm1"]) if qd else null, int(prec), float(thr_lindep),
                     L.ptr(arrs["out"]), L.ptr(arrs["heff"]) if qd else null)
        rid = L.lib().nv2_server_submit(self._h, C.byref(r))
        if rid <= 0:
            raise L.NV2Error(f"nv2_server_submit: {L.last_error()}")
        arrs["_req"] = r
        arrs["_M"] = M
        self._keep[rid] = arrs
        return rid

    def submit_ctx(self, ev, prec=L.PREC_FP64, thr_lindep=1.0e-6):
        """Run an Evaluator whose inputs are resident; returns the request id."""
        out = np.zeros(32 * ev.M)
        rid = L.lib().nv2_server_submit_ctx(self._h, ev._h, int(prec), float(thr_lindep), L.ptr(out))
        if rid <= 0:
            raise L.NV2Error(f"nv2_server_submit_ctx: {L.last_error()}")
        self._keep[rid] = {"out": out, "_M": ev.M}
        return rid

    def wait(self, rid):
        """Blocks until the request is done: (class table (metat, 8, 4), Heff or None, latency ms, queued ms)."""
        lat = np.zeros(1)
        qms = np.zeros(1)
        rc = L.lib().nv2_server_wait(self._h, int(rid), L.ptr(lat), L.ptr(qms))
        arrs = self._keep.pop(rid)
        if rc != 0:
            raise L.NV2Error(f"request {rid} failed: {L.last_error()}")
        M = arrs["_M"]
        heff = None
        if "heff" in arrs:
            h = arrs["heff"]
            heff = (h[:M * M].reshape(M, M, order="F"), h[M * M:].reshape(M, M, order="F"))
        return arrs["out"].reshape(M, 8, 4), heff, float(lat[0]), float(qms[0])

    def close(self):
        if self._h:
            L.lib().nv2_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def ctx_bytes(ncore, nact, nvirt, metat, nchovec):
    """Device bytes one evaluation context of this shape holds (the server's admission estimate)."""
    return int(L.lib().nv2_ctx_bytes(ncore, nact, nvirt, metat, nchovec))

"""nv2_koop_matrices_noalloc (device arrays in and out, static device scratch, no allocation) equals
nv2_koop_matrices on the same inputs, on two calls in a row, the second with a smaller active space
(the shared scratch keeps the first call's larger contents; the second must not read them)."""
import ctypes as C

import numpy as np
import pytest

from nevpt2_amd import synth, koopro, _lib as L

NAMES = ["1-IPKO", "1-EAKO", "2-HRDM", "2-IPKO", "2-EAKO", "gk0pa", "gk0pd", "gk0pf",
         "bmatIP", "bmatEA", "cmatIP", "cmatEA", "dmatIP", "dmatEA"]
H2D, D2H = 1, 2


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


class Dev:
    """Device copies for the test (the test allocates; the entry point under test does not)."""

    def __init__(self, hip):
        self.hip, self.bufs = hip, []

    def put(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        self.bufs.append(p)
        return p

    def get(self, p, shape):
        a = np.zeros(shape, order="F")
        assert self.hip.hipMemcpy(a.ctypes.data, p, a.nbytes, D2H) == 0
        return a

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry_point(gpu):
    hip = _hip()
    for n, nc, nv, seed, M, nele in [(6, 3, 4, 2, 2, 6), (4, 2, 3, 5, 1, 4)]:
        print(f"[noalloc] nact={n} ncore={nc} nvirt={nv} roots={M} nele={nele}", flush=True)
        P = synth.Problem(n, nc, nv, seed=seed, nroots=M, nele=nele)
        aj, atwo = koopro.get_aj_atwo(nc, n, P.norb, P.B, P.oneint)
        ref = koopro.koop_matrices(M, n, nc, P.norb, atwo, P.daaa, P.taa, P.dal, aj)
        d = Dev(hip)
        try:
            ins = [d.put(x) for x in (atwo, P.daaa, P.taa, P.dal, aj)]
            outs = [d.put(np.full(ref[k].shape, np.nan, order="F")) for k in NAMES]
            rc = L.lib().nv2_koop_matrices_noalloc(M, n, nc, P.norb, *ins, *outs, None)
            L.check(rc, "nv2_koop_matrices_noalloc")
            assert hip.hipDeviceSynchronize() == 0
            for k, p in zip(NAMES, outs):
                got = d.get(p, ref[k].shape)
                scale = max(1.0, float(np.abs(ref[k]).max()))
                err = float(np.abs(got - ref[k]).max())
                print(f"  {k:7s} max|noalloc - nv2_koop_matrices| = {err:.2e} (scale {scale:.2e})")
                assert err <= 1e-12 * scale, (k, err)
        finally:
            d.free()

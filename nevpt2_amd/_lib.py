"""ctypes binding of libnevpt2_amd.so (the C ABI in include/nevpt2_amd.h).

The library is built in-tree by ``python -m nevpt2_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libnevpt2_amd.so")

i64 = C.c_int64
i32 = C.c_int32
dp = C.POINTER(C.c_double)

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16X3 = 2

_lib = None


class NV2Error(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise NV2Error(f"{LIBPATH} is not built; run `python -m nevpt2_amd.build`")
        _lib = C.CDLL(LIBPATH)
        _declare(_lib)
    return _lib


def _declare(L):
    L.nv2_last_error.argtypes = [C.c_char_p, i64]
    L.nv2_device_count.restype = C.c_int
    L.nv2_bro.argtypes = [i64, i64, i64, dp, dp, dp, dp]
    L.nv2_get_aj_atwo.argtypes = [i64, i64, i64, i64, dp, dp, dp, dp]
    L.nv2_koop_matrices.argtypes = [i64, i64, i64, i64] + [dp] * 19
    L.nv2_koop_matrices_noalloc.argtypes = [i64, i64, i64, i64] + [C.c_void_p] * 19 + [C.c_void_p]
    L.nv2_bamat.argtypes = [i64, i64, i64, i64] + [dp] * 8 + [i32]
    L.nv2_qdnevpt_state.argtypes = [i64] * 5 + [dp] * 21 + [i64, C.c_double, dp]
    L.nv2_ctx_create.restype = C.c_void_p
    L.nv2_ctx_create.argtypes = [i64] * 6
    L.nv2_ctx_set_integrals.argtypes = [C.c_void_p, dp, dp]
    L.nv2_ctx_set_ro4.argtypes = [C.c_void_p, dp]
    L.nv2_ctx_set_trdm.argtypes = [C.c_void_p, dp, dp, dp]
    L.nv2_ctx_set_ci.argtypes = [C.c_void_p, i64, i64, dp]
    L.nv2_ci_rdm.argtypes = [i64, i64, i64, i32, dp, dp, dp]
    L.nv2_qd_heff.argtypes = [i64, dp, dp, dp, dp]
    L.nv2_ctx_run.argtypes = [C.c_void_p, i32, C.c_double, dp]
    L.nv2_ctx_timings.argtypes = [C.c_void_p, dp]
    L.nv2_ctx_set_option.argtypes = [C.c_void_p, C.c_char_p, i64]
    L.nv2_ctx_set_frozen.argtypes = [C.c_void_p, i64, C.POINTER(C.c_int64)]
    L.nv2_ci_from_file04.argtypes = [i64] * 6 + [C.c_void_p] * 6 + [C.POINTER(C.c_int64)] * 2
    L.nv2_ctx_get.argtypes = [C.c_void_p, C.c_char_p, dp, i64]
    L.nv2_ctx_destroy.argtypes = [C.c_void_p]
    L.nv2_ctx_destroy.restype = None
    L.nv2_server_create.restype = C.c_void_p
    L.nv2_server_create.argtypes = [i32, i64]
    L.nv2_server_submit.restype = i64
    L.nv2_server_submit.argtypes = [C.c_void_p, C.c_void_p]
    L.nv2_server_submit_ctx.restype = i64
    L.nv2_server_submit_ctx.argtypes = [C.c_void_p, C.c_void_p, i32, C.c_double, dp]
    L.nv2_server_wait.argtypes = [C.c_void_p, i64, dp, dp]
    L.nv2_server_destroy.argtypes = [C.c_void_p]
    L.nv2_server_destroy.restype = None
    L.nv2_ctx_bytes.restype = i64
    L.nv2_syev_batched.argtypes = [i64, C.POINTER(C.c_int64), C.POINTER(dp), C.POINTER(dp)]
    L.nv2_ctx_bytes.argtypes = [i64] * 5


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().nv2_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str):
    if rc != 0:
        raise NV2Error(f"{what} failed (rc={rc}): {last_error()}")


def ptr(a: np.ndarray):
    if not (a.flags.f_contiguous or a.flags.c_contiguous) or a.dtype != np.float64:
        raise ValueError("arrays passed to libnevpt2_amd must be contiguous float64")
    return a.ctypes.data_as(dp)


def fin(a) -> np.ndarray:
    """Input array as Fortran-contiguous float64 (the reference layout)."""
    return np.asfortranarray(a, dtype=np.float64)


def syev_batched(mats):
    """Eigen-decomposition of every symmetric matrix of ``mats`` (upper triangles read) in one
    batched GPU solve; returns [(w ascending, V columns)], as numpy.linalg.eigh does."""
    n = len(mats)
    As = [np.array(m, dtype=np.float64, order="F", copy=True) for m in mats]
    ws = [np.zeros(a.shape[0]) for a in As]
    N = (C.c_int64 * n)(*[a.shape[0] for a in As])
    Ap = (dp * n)(*[ptr(a) for a in As])
    wp = (dp * n)(*[ptr(w) for w in ws])
    check(lib().nv2_syev_batched(n, N, Ap, wp), "nv2_syev_batched")
    return list(zip(ws, As))


def zeros(*shape) -> np.ndarray:
    return np.zeros(shape, order="F")

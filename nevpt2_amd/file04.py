"""The reference's file04: CI vectors as excitation lists of a closed-shell reference determinant.

One Fortran unformatted sequential record (koopro4QD.F90:147-176, read again at qdnevpt.F:215-232):

    NORB, NOCA, NOCB, METAT, NCF, II, I          int64 (the reference builds with -fdefault-integer-8)
    (NE(J) int16, ND(J) int64), J = 1..NCF       excitations of determinant J and their offset
    (TROU(J) int16, PART(J) int16), J = 1..II    hole / particle spin-orbitals (1..norb, norb+1..2norb)
    C(J, state), J fastest                       I = NCF * METAT values
    E(state), EMP(state)                         state energies
    EPS(NORB) twice, NNROT, (x, NNROT), MMETAT, (ZHEFF, MMETAT) logical*1, two logical*1, FACTOR, ESCF

The record is framed by 4-byte length markers (negative for the subrecords of a record above 2 GiB).
``to_ci`` turns the excitation lists into the string-ordered CI arrays of nv2_ctx_set_ci through the
library (nv2_ci_from_file04).
"""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np

from . import _lib as L

_MAXSUB = 2 ** 31 - 9


def _read_record(f) -> bytes:
    out = bytearray()
    while True:
        hdr = f.read(4)
        if len(hdr) < 4:
            raise EOFError("file04: truncated record marker")
        n = struct.unpack("<i", hdr)[0]
        body = f.read(abs(n))
        f.read(4)
        out += body
        if n >= 0:
            return bytes(out)


def _write_record(f, payload: bytes):
    if len(payload) <= _MAXSUB:
        f.write(struct.pack("<i", len(payload)))
        f.write(payload)
        f.write(struct.pack("<i", len(payload)))
        return
    off = 0
    while off < len(payload):
        chunk = payload[off:off + _MAXSUB]
        last = off + len(chunk) >= len(payload)
        m = len(chunk) if last else -len(chunk)
        f.write(struct.pack("<i", m))
        f.write(chunk)
        f.write(struct.pack("<i", m))
        off += len(chunk)


def read(path):
    """The file04 record as a dict: norb, noca, nocb, metat, ncf, ne, nd, trou, part, c (ncf, metat), e."""
    with open(path, "rb") as f:
        rec = _read_record(f)
    norb, noca, nocb, metat, ncf, ii, i = struct.unpack_from("<7q", rec, 0)
    off = 56
    pair = np.frombuffer(rec, dtype=np.dtype([("ne", "<i2"), ("nd", "<i8")]), count=ncf, offset=off)
    off += 10 * ncf
    tp = np.frombuffer(rec, dtype=np.dtype([("t", "<i2"), ("p", "<i2")]), count=ii, offset=off)
    off += 4 * ii
    c = np.frombuffer(rec, dtype="<f8", count=ncf * metat, offset=off).reshape(metat, ncf).T.copy()
    off += 8 * ncf * metat
    e = np.frombuffer(rec, dtype="<f8", count=metat, offset=off).copy()
    return {"norb": norb, "noca": noca, "nocb": nocb, "metat": metat, "ncf": ncf, "nexc": ii, "ncoef": i,
            "ne": pair["ne"].astype(np.int16), "nd": pair["nd"].astype(np.int64),
            "trou": tp["t"].astype(np.int16), "part": tp["p"].astype(np.int16), "c": c, "e": e}


def write(path, norb, noca, ne, nd, trou, part, c, e):
    """Writes a file04 record (the fields after E are zeros / false, FACTOR = 1)."""
    c = np.asarray(c, dtype=np.float64)
    ncf, metat = c.shape
    ii = len(trou)
    buf = bytearray(struct.pack("<7q", norb, noca, noca, metat, ncf, ii, ncf * metat))
    pair = np.zeros(ncf, dtype=np.dtype([("ne", "<i2"), ("nd", "<i8")]))
    pair["ne"], pair["nd"] = ne, nd
    buf += pair.tobytes()
    tp = np.zeros(ii, dtype=np.dtype([("t", "<i2"), ("p", "<i2")]))
    tp["t"], tp["p"] = trou, part
    buf += tp.tobytes()
    buf += np.asfortranarray(c).tobytes(order="F")
    buf += np.asarray(e, dtype="<f8").tobytes() + np.zeros(metat).tobytes()      # E, EMP
    buf += np.zeros(2 * norb).tobytes()                                           # EPS twice
    buf += struct.pack("<q", 0) + struct.pack("<q", 0)                            # NNROT, MMETAT
    buf += struct.pack("<??", False, False) + struct.pack("<dd", 1.0, 0.0)        # zz, zz, FACTOR, ESCF
    with open(path, "wb") as f:
        _write_record(f, bytes(buf))


def to_ci(f04, ncore, nact):
    """(C[root, I_alpha, I_beta], nalpha, nbeta) for nv2_ctx_set_ci, through nv2_ci_from_file04."""
    lib = L.lib()
    ne = np.ascontiguousarray(f04["ne"], dtype=np.int16)
    nd = np.ascontiguousarray(f04["nd"], dtype=np.int64)
    tr = np.ascontiguousarray(f04["trou"], dtype=np.int16)
    pa = np.ascontiguousarray(f04["part"], dtype=np.int16)
    c = np.asfortranarray(f04["c"], dtype=np.float64)
    na, nb = C.c_int64(), C.c_int64()
    args = [f04["norb"], ncore, nact, f04["noca"], f04["metat"], f04["ncf"], ne.ctypes.data, nd.ctypes.data,
            tr.ctypes.data if tr.size else None, pa.ctypes.data if pa.size else None, c.ctypes.data]
    L.check(lib.nv2_ci_from_file04(*args, None, C.byref(na), C.byref(nb)), "nv2_ci_from_file04")
    from math import comb
    out = np.zeros((f04["metat"], comb(nact, na.value), comb(nact, nb.value)))
    L.check(lib.nv2_ci_from_file04(*args, out.ctypes.data, C.byref(na), C.byref(nb)), "nv2_ci_from_file04")
    return out, na.value, nb.value

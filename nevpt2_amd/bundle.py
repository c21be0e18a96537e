"""Input bundle of the Fortran host: one stream file with the packed 4-RDM,
Cholesky vectors, one-electron integrals and reference energies.

It replaces, for the hot path, the reference's file04 + ijkl.h5 + *.RasOrb.N
inputs (koopro4QD.F90:147-176, ijkl_utils.F90:210-263, info_state_energy.F90:41-107).
Layout (little endian): 8 bytes "NV2BUNDL", int64[8] = (version, ncore, nact,
nvirt, metat, nele, nchovec, 0), then float64 e0(metat), oneint(npair),
cholesky(nchovec*npair) [column-major], ro4(nwords*nact^4*metat) [column-major].
Version 2 (metat > 1) appends the transition RDMs of the P = metat(metat-1)/2
state pairs, 3-TRDM(P,n^6), 2-TRDM(P,n^4), 1-TRDM(P,n^2), column-major
(the nevpt.h5 datasets of koopro4QD.F90:473-477).  Version 3 stops after the Cholesky vectors: the
integrals and reference energies only, for a host run whose CI vectors come from the reference's file04.
"""
from __future__ import annotations

import numpy as np


def write_bundle(path, P, e0=None, rdms=True):
    M = P.metat
    e0 = np.zeros(M) if e0 is None else np.asarray(e0, dtype=np.float64)
    version = (2 if M > 1 else 1) if rdms else 3
    hdr = np.array([version, P.ncore, P.nact, P.nvirt, M, P.nele, P.B.shape[0], 0], dtype="<i8")
    with open(path, "wb") as f:
        f.write(b"NV2BUNDL")
        f.write(hdr.tobytes())
        f.write(e0.astype("<f8").tobytes())
        f.write(np.ascontiguousarray(P.oneint, dtype="<f8").tobytes())
        f.write(np.asarray(P.B, dtype="<f8").tobytes(order="F"))
        if not rdms:
            return
        f.write(np.asarray(P.ro4, dtype="<f8").tobytes(order="F"))
        if M > 1:
            for a in (P.daaaoff, P.taaoff, P.daloff):
                f.write(np.asarray(a, dtype="<f8").tobytes(order="F"))


def read_table(stdout: str):
    """Parse the host's per-state class tables back into arrays (metat, 8, 4)."""
    tabs, cur = [], None
    for line in stdout.splitlines():
        if line.startswith(" state number:"):
            cur = []
            tabs.append(cur)
        elif cur is not None and line[:6].strip() in ("(0)", "(+1)", "(-1)", "(+2)", "(-2)", "(+1)'", "(-1)'", "(0)'"):
            vals = line[6:]
            cur.append([float(vals[i:i + 20]) for i in range(0, 80, 20)])
    return np.array(tabs)


def read_qd_summary(stdout: str):
    """Parse the host's QD summary lines ' @ state  E0  SC  PC' into an array (metat, 3)."""
    rows = []
    for line in stdout.splitlines():
        if line.startswith(" @"):
            parts = line[2:].split()
            rows.append([float(x) for x in parts[1:4]])
    return np.array(rows)

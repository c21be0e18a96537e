"""The reference's file04 (CI vectors as excitation lists, koopro4QD.F90:147-176) on the CPU: the record
reader and writer, the conversion to string-ordered CI arrays (nv2_ci_from_file04), and its phase
convention pinned by the reference's own determinant algebra restated in oracle/cipsi.py (Jdifgen,
koopro4QD.F90:3763-3886): the 1-RDM built from Jdifgen's matrix elements on the excitation lists equals
the 1-RDM of the converted vector, with the holes and particles of every determinant in shuffled order."""
import os

import numpy as np
import pytest

from nevpt2_amd import synth, file04, _lib
from oracle import cipsi

CASES = [(4, 2, 3, 5, 4, 1), (5, 1, 2, 3, 4, 2), (6, 3, 2, 7, 6, 1)]   # nact, ncore, nvirt, seed, nele, roots


@pytest.mark.parametrize("case", CASES, ids=lambda c: "cas%d_%de_%droot" % (c[0], c[4], c[5]))
def test_file04_roundtrip_and_jdifgen_phases(tmp_path, case):
    n, nc, nv, seed, nele, M = case
    P = synth.Problem(n, nc, nv, seed=seed, nele=nele, nroots=M, host_rdms=False)
    f = cipsi.lists_from_ci(P.C, n, nc, P.norb, nele // 2, nele // 2, seed=seed)
    path = tmp_path / "file04"
    file04.write(path, P.norb, f["noca"], f["ne"], f["nd"], f["trou"], f["part"], f["c"], -1.0 - np.arange(M))
    g = file04.read(path)
    assert g["ncf"] == f["ncf"] and np.array_equal(g["trou"], f["trou"]) and np.allclose(g["e"], -1.0 - np.arange(M))
    Cback, na, nb = file04.to_ci(g, nc, n)
    assert (na, nb) == (nele // 2, nele // 2)
    assert np.array_equal(Cback, P.C)
    for s in range(M):
        gj = cipsi.one_rdm(g, nc, n, s)
        gs = synth.spinfree_rdm(1, P.C[s], n=n, nel=nele)
        assert np.abs(gj - gs).max() < 1e-13
        assert np.abs(gs - np.diag(np.diag(gs))).max() > 1e-2        # off-diagonal elements are tested


def test_planted_phase_convention_error_is_caught():
    """Lists phased as a+ .. a+ a .. a (creators first, annihilators in listed order) instead of the pair
    product: Jdifgen's matrix elements then give a 1-RDM that is not the vector's, so the pin above
    distinguishes the two conventions."""
    n, nc, nv, seed, nele = 4, 2, 3, 5, 4
    P = synth.Problem(n, nc, nv, seed=seed, nele=nele, host_rdms=False)
    f = cipsi.lists_from_ci(P.C, n, nc, P.norb, 2, 2, seed=seed, grouped=True)
    gj = cipsi.one_rdm(f, nc, n, 0)
    gs = synth.spinfree_rdm(1, P.C[0], n=n, nel=nele)
    assert np.abs(gj - gs).max() > 1e-3


def test_rejects_excitations_outside_the_active_space():
    f = {"norb": 6, "noca": 3, "metat": 1, "ncf": 2, "ne": np.array([0, 1], dtype=np.int16),
         "nd": np.array([0, 0], dtype=np.int64), "trou": np.array([1], dtype=np.int16),
         "part": np.array([4], dtype=np.int16), "c": np.array([[0.8], [0.6]])}
    with pytest.raises(_lib.NV2Error, match="outside the active space"):
        file04.to_ci(f, 1, 4)      # hole in core orbital 1

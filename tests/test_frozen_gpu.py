"""Frozen core orbitals (.FROZEN, qdnevpt.F:733-745) on the GPU against the oracle: single state, QD with
two roots, and through the Fortran host's input reader."""
import os
import subprocess

import numpy as np
import pytest

from nevpt2_amd import synth, bundle
from nevpt2_amd.evaluator import Evaluator
from oracle import stage1 as S1, stage2 as S2, qd as Q
from conftest import energy_tol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
def test_frozen_single_state(gpu):
    print("[frozen] CAS(6,6) nc=4 nv=8, frozen core orbitals 1 and 3", flush=True)
    P = synth.Problem(6, 4, 8, seed=2)
    koop = S1.koopmans_all(P)
    ref, _ = S2.state_energies(P, koop, 0, frozen=(0, 2))
    ev = Evaluator.for_problem(P)
    try:
        ev.set_frozen([1, 3])
        got = ev.run()[0]
        ev.set_frozen([])
        unfrozen = ev.run()[0]
    finally:
        ev.close()
    for k in range(8):
        for col, ncol in ((1, 0), (3, 2)):
            d = abs(got[k, col] - ref[k, col])
            print(f"  class {k} col {col} |dE| {d:.2e} (norm {ref[k, ncol]:.2e})")
            assert d <= energy_tol(ref[k, ncol])
    full, _ = S2.state_energies(P, koop, 0)
    assert np.abs(unfrozen[:, [1, 3]] - full[:, [1, 3]]).max() <= energy_tol(full[:, [0, 2]]).max()


@pytest.mark.gpu
def test_frozen_qd_two_roots(gpu):
    print("[frozen] QD CAS(4,4) nc=3 nv=3, 2 roots, frozen core orbital 2", flush=True)
    P = synth.Problem(4, 3, 3, seed=1, nroots=2)
    koop = S1.koopmans_all(P)
    Hsc, Hpc, _, _ = Q.heff_matrices(P, koop, frozen=(1,))
    ev = Evaluator.for_problem(P)
    try:
        ev.set_frozen([2])
        ev.run()
        hsc, hpc = ev.heff()
    finally:
        ev.close()
    print(f"  max|dHeff| SC {np.abs(hsc - Hsc).max():.2e} PC {np.abs(hpc - Hpc).max():.2e}")
    # as in the unfrozen QD test: relative to the largest Heff element
    assert np.abs(hsc - Hsc).max() <= 1e-9 * max(1.0, np.abs(Hsc).max())
    assert np.abs(hpc - Hpc).max() <= 1e-9 * max(1.0, np.abs(Hpc).max())


@pytest.mark.gpu
def test_host_frozen_keyword(gpu, tmp_path):
    exe = os.path.join(ROOT, "nevpt2_amd", "nevpt2_amd.x")
    print("[fortran host] .FROZEN CAS(4,4) nc=3 nv=4", flush=True)
    P = synth.Problem(4, 3, 4, seed=3)
    bundle.write_bundle(tmp_path / "in.bin", P, e0=[-1.5])
    inp = ("*MODEL\n.CHOLES\n*KOOPRO\n.STATES\n1\n.NACTEL\n4\n.NO ZOR\n*QDNEVP\n.LINDEP\n1.0d-6\n"
           ".FROZEN\n2\n1 2\n.BUNDLE\nin.bin\n*END OF\n")
    (tmp_path / "input.nevpt").write_text(inp)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Frozen orbitals" in r.stdout
    tab = bundle.read_table(r.stdout)[0]
    koop = S1.koopmans_all(P)
    ref, _ = S2.state_energies(P, koop, 0, frozen=(0, 1))
    assert (np.abs(tab[:, 1] - ref[:, 1]) <= energy_tol(ref[:, 0])).all()
    assert (np.abs(tab[:, 3] - ref[:, 3]) <= energy_tol(ref[:, 2])).all()

"""The Fortran host (input.nevpt reader + ISO_C_BINDING calls into the GPU library)."""
import os
import subprocess

import numpy as np
import pytest

from nevpt2_amd import synth, bundle
from oracle import stage1 as S1, stage2 as S2, qd as Q
from conftest import energy_tol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "nevpt2_amd", "nevpt2_amd.x")

INPUT = """*MODEL
.CHOLES
{model}*KOOPRO
.STATES
{nstates}
.NACTEL
{nele}
.NO ZOR
*QDNEVP
.LINDEP
1.0d-6
.BUNDLE
{bundle}
*END OF
"""


def _need_exe():
    if not os.path.exists(EXE):
        pytest.fail("nevpt2_amd.x not built (run __graft_entry__.build())")


def test_host_rejects_unknown_keyword(tmp_path):
    _need_exe()
    (tmp_path / "input.nevpt").write_text("*QDNEVP\n.BOGUS \n*END OF\n")
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert "illegal keyword .BOGUS" in r.stdout + r.stderr


@pytest.mark.gpu
def test_host_end_to_end(gpu, tmp_path):
    _need_exe()
    print("[fortran host] CAS(4,4) nc=2 nv=5", flush=True)
    P = synth.Problem(4, 2, 5, seed=3)
    bundle.write_bundle(tmp_path / "in.bin", P, e0=[-1.5])
    (tmp_path / "input.nevpt").write_text(INPUT.format(nstates=1, nele=P.nele, bundle="in.bin", model=""))
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    tab = bundle.read_table(r.stdout)
    koop = S1.koopmans_all(P)
    ref, _ = S2.state_energies(P, koop, 0)
    # the printed table carries 15 decimals; bound as in conftest.energy_tol
    assert (np.abs(tab[0][:, 1] - ref[:, 1]) <= energy_tol(ref[:, 0])).all()
    assert (np.abs(tab[0][:, 3] - ref[:, 3]) <= energy_tol(ref[:, 2])).all()
    assert " @   1" in r.stdout


@pytest.mark.gpu
def test_host_qd_two_roots(gpu, tmp_path):
    """Two roots: the host reads the transition RDMs, prints Heff and the QD summary,
    whose energies are the eigenvalues of the hermitised Heff (qdnevpt.F:1046-1230)."""
    _need_exe()
    print("[fortran host] QD CAS(4,4) nc=2 nv=3, 2 roots", flush=True)
    P = synth.Problem(4, 2, 3, seed=1, nroots=2)     # a seed whose SC and PC Heff have real spectra
    e0 = np.array([-1.5, -1.2])
    bundle.write_bundle(tmp_path / "in.bin", P, e0=e0)
    (tmp_path / "input.nevpt").write_text(INPUT.format(nstates=2, nele=P.nele, bundle="in.bin", model=""))
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "QD-NEVPT2: FINAL RESULTS" in r.stdout
    assert "Heff~ (Hermitian) --> SC" in r.stdout
    summ = bundle.read_qd_summary(r.stdout)
    koop = S1.koopmans_all(P)
    Hsc, Hpc, _, _ = Q.heff_matrices(P, koop)
    esc, _ = Q.diagonalize_heff(Hsc, e0)
    epc, _ = Q.diagonalize_heff(Hpc, e0)
    # the summary prints 7 decimals
    assert np.abs(summ[:, 0] - e0).max() < 1e-6
    assert np.abs(summ[:, 1] - esc).max() < 2e-6 * max(1.0, np.abs(esc).max())
    assert np.abs(summ[:, 2] - epc).max() < 2e-6 * max(1.0, np.abs(epc).max())


@pytest.mark.gpu
def test_host_no4rdm_skips_the_prime_classes(gpu, tmp_path):
    """.NO4RDM: the (+1)' and (-1)' classes are never formed (qdnevpt.F:949-991), so they are absent
    from the table, from Heff and from the summary energies alike."""
    _need_exe()
    print("[fortran host] .NO4RDM QD CAS(4,4) nc=2 nv=3, 2 roots", flush=True)
    P = synth.Problem(4, 2, 3, seed=1, nroots=2)
    e0 = np.array([-1.5, -1.2])
    bundle.write_bundle(tmp_path / "in.bin", P, e0=e0)
    (tmp_path / "input.nevpt").write_text(INPUT.format(nstates=2, nele=P.nele, bundle="in.bin", model=".NO4RDM\n"))
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "V(-1p) terms will not be calculated" in r.stdout
    tabs = bundle.read_table(r.stdout)
    koop = S1.koopmans_all(P)
    Hsc, Hpc, tables, _ = Q.heff_matrices(P, koop, with_4rdm=False)
    for s in range(2):
        assert np.all(tabs[s][[5, 6]] == 0.0)
        assert (np.abs(tabs[s][:, 1] - tables[s][:, 1]) <= energy_tol(tables[s][:, 0])).all()
    summ = bundle.read_qd_summary(r.stdout)
    esc, _ = Q.diagonalize_heff(Hsc, e0)
    epc, _ = Q.diagonalize_heff(Hpc, e0)
    assert np.abs(summ[:, 1] - esc).max() < 2e-6 * max(1.0, np.abs(esc).max())
    assert np.abs(summ[:, 2] - epc).max() < 2e-6 * max(1.0, np.abs(epc).max())


FILE04_INPUT = """*MODEL
.CHOLES
*KOOPRO
.STATES
{nstates}
.NACTEL
{nele}
.FILE04
file04
*QDNEVP
.LINDEP
1.0d-6
.BUNDLE
ints.bin
.FILE04
file04
*END OF
"""


@pytest.mark.gpu
@pytest.mark.parametrize("roots", [1, 2])
def test_host_reads_file04(gpu, tmp_path, roots):
    """CI vectors from the reference's file04 (excitation lists, shuffled hole/particle order), integrals
    from an integrals-only bundle: the host converts the lists, builds the RDMs (and transition RDMs) on the
    GPU and reproduces the oracle's table and QD summary; E0 comes from file04."""
    from nevpt2_amd import file04
    from oracle import cipsi
    _need_exe()
    print(f"[fortran host] file04 CAS(4,4) nc=2 nv=3, {roots} root(s)", flush=True)
    P = synth.Problem(4, 2, 3, seed=1, nroots=roots)
    e0 = np.array([-1.5, -1.2])[:roots]
    f = cipsi.lists_from_ci(P.C, P.nact, P.ncore, P.norb, 2, 2, seed=7)
    file04.write(tmp_path / "file04", P.norb, f["noca"], f["ne"], f["nd"], f["trou"], f["part"], f["c"], e0)
    bundle.write_bundle(tmp_path / "ints.bin", P, e0=np.zeros(roots), rdms=False)
    (tmp_path / "input.nevpt").write_text(FILE04_INPUT.format(nstates=roots, nele=P.nele))
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FILE04 informations" in r.stdout
    tabs = bundle.read_table(r.stdout)
    koop = S1.koopmans_all(P)
    for s in range(roots):
        ref, _ = S2.state_energies(P, koop, s)
        assert (np.abs(tabs[s][:, 1] - ref[:, 1]) <= energy_tol(ref[:, 0])).all()
        assert (np.abs(tabs[s][:, 3] - ref[:, 3]) <= energy_tol(ref[:, 2])).all()
    if roots > 1:
        summ = bundle.read_qd_summary(r.stdout)
        Hsc, Hpc, _, _ = Q.heff_matrices(P, koop)
        esc, _ = Q.diagonalize_heff(Hsc, e0)
        assert np.abs(summ[:, 0] - e0).max() < 1e-6
        assert np.abs(summ[:, 1] - esc).max() < 2e-6 * max(1.0, np.abs(esc).max())

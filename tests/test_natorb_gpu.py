""".NATORB (qdnevpt.F:1240-1467) on the GPU against the oracle: the first-order density matrix of each state
(core-virtual from (0)', active-virtual from (-1)', core-active from (+1)') and its eigenvalues, on every path
its blocks take: the tridiagonal form of the (+-1)' K' (n < 8), the band form summed on the chain's stream
(one state, n >= 8) or after it (several states), the metric eigen path (thr_lindep = 0.1), frozen core
orbitals and .NO4RDM.

The canonical core and virtual orbitals are eigenvectors whose signs the solvers choose, so the matrix is
compared back in the input orbitals (rho1-mo, U rho U^T: the signs cancel) and elementwise in modulus in
the canonical orbitals.  Like the class energies, an element is a sum of r_mu v_mu / d_mu over the PC
functions, so its round-off follows the class's norm: the bound is 1e-9 times max(1, max|rho|)."""
import os
import subprocess

import numpy as np
import pytest

from nevpt2_amd import synth, bundle
from nevpt2_amd.evaluator import Evaluator
from oracle import stage1 as S1, stage2 as S2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# name: (nact, ncore, nvirt, seed, nele, nroots, thr_lindep, frozen (0-based), 4-RDM classes)
CASES = {
    "trd": (4, 2, 3, 13, 4, 1, 1e-6, (), True),
    "trd_qd": (4, 3, 3, 1, 4, 2, 1e-6, (), True),
    "band_chain": (8, 2, 3, 3, 8, 1, 1e-6, (), True),
    "band_qd": (8, 2, 2, 5, 8, 2, 1e-6, (), True),
    "eigen_thr": (5, 2, 3, 7, 6, 1, 0.1, (), True),
    "frozen": (4, 3, 3, 11, 4, 1, 1e-6, (0, 2), True),
    "no4rdm": (4, 2, 3, 13, 4, 1, 1e-6, (), False),
}


def _check(name, got, ref):
    can, mo, occ = got
    Rc, Rmo, w = ref
    scale = max(1.0, float(np.abs(Rmo).max()))
    d_mo = float(np.abs(mo - Rmo).max())
    d_abs = float(np.abs(np.abs(can) - np.abs(Rc)).max())
    d_occ = float(np.abs(occ - w).max())
    print(f"  {name}: max|rho| {np.abs(Rmo - np.diag(np.diag(Rmo))).max():.3e}  |d rho_mo| {d_mo:.2e}  "
          f"|d |rho|| {d_abs:.2e}  |d occ| {d_occ:.2e}", flush=True)
    tol = 1e-9 * scale
    assert d_mo <= tol and d_abs <= tol and d_occ <= tol, (name, d_mo, d_abs, d_occ, tol)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_natorb_density(gpu, name):
    n, nc, nv, seed, nele, M, thr, frozen, with4 = CASES[name]
    print(f"[natorb] {name}: CAS({nele},{n}) nc={nc} nv={nv} roots={M} thr={thr} frozen={frozen} 4rdm={with4}",
          flush=True)
    P = synth.Problem(n, nc, nv, seed=seed, nele=nele, nroots=M)
    koop = S1.koopmans_all(P)
    ev = Evaluator.for_problem(P)
    try:
        ev.set_option("natorb", 1)
        if not with4:
            ev.set_option("no4rdm", 1)
        if frozen:
            ev.set_frozen([i + 1 for i in frozen])
        ev.run(thr_lindep=thr)
        can, mo, occ = ev.natorb()
    finally:
        ev.close()
    for s in range(M):
        ref = S2.rho1(P, koop, s, thr=thr, with_4rdm=with4, frozen=frozen)
        _check(f"{name} state {s}", (can[s], mo[s], occ[s]), ref)
        if not with4:        # without the (+-1)' classes their blocks stay zero
            A = slice(nc, nc + n)
            assert np.abs(can[s][:nc, A]).max() == 0.0 and np.abs(can[s][A, nc + n:]).max() == 0.0


@pytest.mark.gpu
def test_natorb_off_leaves_energies(gpu):
    """The option adds work after the class sums and changes no energy."""
    print("[natorb] CAS(4,4) nc=2 nv=3: energies with and without .NATORB", flush=True)
    P = synth.Problem(4, 2, 3, seed=13, nele=4)
    ev = Evaluator.for_problem(P)
    try:
        a = ev.run()
        ev.set_option("natorb", 1)
        b = ev.run()
        ev.set_option("natorb", 0)
        c = ev.run()
    finally:
        ev.close()
    assert np.allclose(a, b, rtol=1e-13, atol=1e-13) and np.allclose(a, c, rtol=1e-13, atol=1e-13)


@pytest.mark.gpu
def test_host_natorb_keyword(gpu, tmp_path):
    """The Fortran host's .NATORB output: the density matrix and the occupation numbers as the reference
    prints them (qdnevpt.F:1297-1460), against the oracle."""
    exe = os.path.join(ROOT, "nevpt2_amd", "nevpt2_amd.x")
    print("[fortran host] .NATORB CAS(4,4) nc=2 nv=3", flush=True)
    P = synth.Problem(4, 2, 3, seed=13, nele=4)
    bundle.write_bundle(tmp_path / "in.bin", P, e0=[-1.5])
    inp = ("*MODEL\n.CHOLES\n*KOOPRO\n.STATES\n1\n.NACTEL\n4\n.NO ZOR\n*QDNEVP\n.LINDEP\n1.0d-6\n.NATORB\n"
           ".BUNDLE\nin.bin\n*END OF\n")
    (tmp_path / "input.nevpt").write_text(inp)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    assert "Density matrix 0   for state" in out and "Natural orbital occupation numbers for state" in out
    koop = S1.koopmans_all(P)
    Rc, _, w = S2.rho1(P, koop, 0)
    lines = out.splitlines()
    k = next(i for i, ln in enumerate(lines) if "orbital occupations" in ln)
    vals = []
    for ln in lines[k + 1:]:
        if "----" in ln:
            break
        vals += [float(x) for x in ln.split()]
    assert len(vals) == P.norb
    assert np.abs(np.array(vals) - w).max() <= 1e-7            # printed with 8 decimals
    ordered = S2.order_occupations(w)
    nc, n = P.ncore, P.nact
    sums = {"core": ordered[:nc].sum(), "active": ordered[nc:nc + n].sum(), "virtual": ordered[nc + n:].sum()}
    for key, v in sums.items():
        ln = next(x for x in lines if f"{key:7s} occupation numbers" in x)
        assert abs(float(ln.split(":")[1]) - v) <= 1e-5
    tot = next(x for x in lines if "Total sum of occupation numbers" in x)
    assert abs(float(tot.split("=")[1]) - w.sum()) <= 1e-5

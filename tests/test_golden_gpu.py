"""GPU parity at the workload sizes: the default HIP path (RDMs built on the GPU from the seeded CI
vector, every Koopmans matrix, the eight classes) against golden fixtures written by the oracle
(tests/golden/make_fixtures.py), and size-independent properties at CAS(14,14) and CAS(12,12)x6.

Fixtures: CAS(6,6) nc=10 nv=40 and CAS(10,10) nc=20 nv=200 (the two single-state configurations the
reference path and the FP64 drop-in are quoted on; the latter runs the split-K GEMMs and the band form
of K' at order 1000), CAS(8,8) (order 512, the smallest band-form case) and CAS(9,9) with 8 electrons.
Per class the observed |dE| and the class norm are printed.  The bar is 1e-9 Eh; a class whose norm
exceeds 500 gets 2e-12 Eh per unit of norm on top, because E = -sum r^2/d and norm = sum r^2/d^2 make a
relative round-off x in a small denominator d move E by about norm * d * x.
"""
import json
import os

import numpy as np
import pytest

from nevpt2_amd import synth, qdnevpt
from nevpt2_amd.evaluator import Evaluator
from conftest import energy_tol

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = ["c1", "cas8", "cas9", "c2"]
NAMES = ["(0)", "(+1)", "(-1)", "(+2)", "(-2)", "(+1)'", "(-1)'", "(0)'"]


def class_tol(norm):
    return float(energy_tol(norm))


def load(name):
    with open(os.path.join(HERE, "golden", f"{name}.json")) as f:
        return json.load(f)


def check_table(got, ref, label):
    """Per-class energies within class_tol(norm), norms within 1e-9 relative; prints what it saw."""
    worst = 0.0
    for k in range(8):
        for col, ncol, form in ((1, 0, "SC"), (3, 2, "PC")):
            d = abs(got[k, col] - ref[k, col])
            worst = max(worst, d)
            print(f"  {label} {NAMES[k]:6s} {form} E {ref[k, col]: .12f} norm {ref[k, ncol]:.3e} |dE| {d:.2e}")
            assert d <= class_tol(ref[k, ncol]), (label, NAMES[k], form, d)
            dn = abs(got[k, ncol] - ref[k, ncol])
            assert dn <= 1e-9 * max(1.0, abs(ref[k, ncol])) + 2.0 * abs(ref[k, ncol]) * class_tol(ref[k, ncol]) \
                / max(abs(ref[k, col]), 1e-300), (label, NAMES[k], form, "norm", dn)
    print(f"  {label}: largest |dE| {worst:.2e} Eh", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("name", GOLDEN)
def test_golden_fixture(gpu, name):
    fx = load(name)
    n, nc, nv, seed, nele = fx["nact"], fx["ncore"], fx["nvirt"], fx["seed"], fx["nele"]
    print(f"[golden {name}] nact={n} ncore={nc} nvirt={nv} seed={seed} nele={nele}", flush=True)
    P = synth.Problem(n, nc, nv, seed=seed, nele=nele, host_rdms=False)
    ev = Evaluator.for_problem(P, rdms="gpu")
    try:
        got = ev.run()[0]
        check_table(got, np.array(fx["table"]), name)
        for k, rec in fx["koopmans"].items():
            a = ev.get(k, tuple([1] + rec["shape"]))[0]
            scale = max(1.0, float(np.sqrt(rec["sumsq"] / a.size)) * 10.0, float(np.abs(a).max()))
            vals = np.array([a[tuple(p)] for p in rec["positions"]])
            dv = float(np.abs(vals - np.array(rec["values"])).max())
            ds = abs(float(a.sum()) - rec["sum"])
            dq = abs(float((a * a).sum()) - rec["sumsq"])
            print(f"  {name} {k:7s} sampled max|d| {dv:.2e} sum d {ds:.2e} sumsq rel d "
                  f"{dq / max(rec['sumsq'], 1e-300):.2e}")
            assert dv <= 1e-11 * scale, (k, dv, scale)
            assert ds <= 1e-11 * scale * a.size ** 0.5, (k, ds)
            assert dq <= 1e-10 * max(rec["sumsq"], 1e-300), (k, dq)
    finally:
        ev.close()


@pytest.mark.gpu
def test_cas14_pc_paths_agree(gpu):
    """CAS(14,14) nc=30 nv=400, CI truncated to 2e5 determinants (the precision-form configuration):
    the (+-1)' PC sums through the band form of K' (default), the tridiagonal form and the metric's
    eigen path are exact rewritings of one sum and agree.  The K' form touches only the (+-1)' classes
    (the others are bit-identical between band and tridiagonal); the metric path of every other PC class
    switches with it (the factored metric by default, its eigen-decomposition with NV2_METRIC_EIGEN), so
    their PC energies agree to the same bound, and their SC energies are bit-identical."""
    print("[cas14] nact=14 ncore=30 nvirt=400 keep=200000 seed=1", flush=True)
    P = synth.Problem(14, 30, 400, seed=1, keep=200000, host_rdms=False)
    ev = Evaluator.for_problem(P, rdms="gpu")
    saved = {k: os.environ.pop(k, None) for k in ("NV2_K_TRIDIAG", "NV2_K_BAND", "NV2_METRIC_EIGEN")}
    try:
        base = ev.run()[0]
        forms = {"band": ev.pc_timings()[0, 3]}
        os.environ["NV2_K_TRIDIAG"] = "1"
        tri = ev.run()[0]
        forms["tridiagonal"] = ev.pc_timings()[0, 3]
        os.environ.pop("NV2_K_TRIDIAG")
        os.environ["NV2_METRIC_EIGEN"] = "1"
        eig = ev.run()[0]
        metric = ev.pc_metric()[0]
    finally:
        for k, v in saved.items():
            os.environ.pop(k, None)
            if v is not None:
                os.environ[k] = v
        ev.close()
    assert forms == {"band": 2.0, "tridiagonal": 1.0}, forms
    assert np.all(metric[[5, 6], 2] == 0.0), metric      # the eigen path really ran for (+-1)'
    assert np.isfinite(base).all() and (base[:, [0, 2]] >= 0).all()
    for other, label in ((tri, "tridiagonal"), (eig, "eigen")):
        for k in range(8):
            if k in (5, 6):
                d = abs(other[k, 3] - base[k, 3])
                # the three forms round differently; the sum is amplified by the class norm
                tol = 1e-9 * max(1.0, abs(base[k, 3])) + 2e-12 * base[k, 2]
                print(f"  band vs {label} {NAMES[k]} PC E {base[k, 3]: .10f} norm {base[k, 2]:.3e} |dE| {d:.2e}")
                assert d <= tol, (label, NAMES[k], d, tol)
                assert other[k, 1] == base[k, 1]
            elif label == "tridiagonal" or k == 0:
                assert np.array_equal(other[k], base[k]), (label, NAMES[k])
            else:
                d = abs(other[k, 3] - base[k, 3])
                tol = 1e-9 * max(1.0, abs(base[k, 3])) + 2e-12 * base[k, 2]
                print(f"  factored vs eigen metric {NAMES[k]} PC E {base[k, 3]: .10f} norm {base[k, 2]:.3e} |dE| {d:.2e}")
                assert d <= tol, (label, NAMES[k], d, tol)
                assert other[k, 1] == base[k, 1] and other[k, 0] == base[k, 0]


@pytest.mark.gpu
def test_cas12_six_roots_qd_properties(gpu):
    """CAS(12,12)x6 nc=25 nv=300 (the QD configuration): the Heff diagonal is each state's class total,
    every state's class table equals a single-state run of that state alone, and the hermitised
    effective Hamiltonian is symmetric with eigenvalues equal to those of the raw Heff + E0."""
    M = 6
    print(f"[cas12 x{M}] nact=12 ncore=25 nvirt=300 seed=1", flush=True)
    P = synth.Problem(12, 25, 300, seed=1, nroots=M, host_rdms=False)
    ev = Evaluator.for_problem(P, rdms="gpu")
    try:
        tab = ev.run()
        hsc, hpc = ev.heff()
    finally:
        ev.close()
    for i in range(M):
        assert abs(hsc[i, i] - tab[i, :, 1].sum()) <= 1e-12 * max(1.0, abs(hsc[i, i]))
        assert abs(hpc[i, i] - tab[i, :, 3].sum()) <= 1e-12 * max(1.0, abs(hpc[i, i]))
    # one state alone (the first and the last): same class table
    for s in (0, M - 1):
        Ps = synth.Problem(12, 25, 300, seed=1, nroots=1, host_rdms=False, ci=P.C[s:s + 1])
        ev1 = Evaluator.for_problem(Ps, rdms="gpu")
        try:
            alone = ev1.run()[0]
        finally:
            ev1.close()
        for k in range(8):
            for col, ncol in ((1, 0), (3, 2)):
                d = abs(alone[k, col] - tab[s, k, col])
                print(f"  state {s} {NAMES[k]} col {col} |dE| alone vs in QD {d:.2e} (norm {tab[s, k, ncol]:.2e})")
                assert d <= class_tol(tab[s, k, ncol]), (s, NAMES[k], col, d)
    e0 = np.linspace(-100.0, -99.5, M)
    for H in (hsc, hpc):
        w, Ht = qdnevpt.qd_heff(H, e0)
        assert np.abs(Ht - Ht.T).max() <= 1e-10 * np.abs(Ht).max()
        raw = np.sort(np.linalg.eigvals(H + np.diag(e0)).real)
        assert np.abs(np.sort(w) - raw).max() <= 1e-9 * max(1.0, np.abs(raw).max())

"""The batched eigensolver (nv2_syev_batched: one launch per step for every matrix of a batch of
mixed orders) against LAPACK (numpy.linalg.eigh) on the cases the PC classes produce: random
spectra on both sides of the LDS/global switch (order 120), exact degeneracies, a rank-deficient
metric (a cluster of zero eigenvalues), a decoupled padding block, tight clusters (one of 100
vectors, and the spectrum of a metric close to 0.5 I, which is one cluster), graded
eigenvalues, a diagonal and a zero matrix, matrices of norm 1e-200 and 1e150, and order 1."""
import os

import numpy as np
import pytest

from nevpt2_amd import _lib as L
from nevpt2_amd import synth
from nevpt2_amd.evaluator import Evaluator


def _sym(rng, n):
    a = rng.standard_normal((n, n))
    return a + a.T


def _with_spectrum(rng, lam):
    q, _ = np.linalg.qr(rng.standard_normal((len(lam), len(lam))))
    return (q * lam) @ q.T


def _cases():
    rng = np.random.default_rng(11)
    cases = [(f"random{n}", _sym(rng, n)) for n in (1, 2, 3, 14, 64, 105, 119, 120, 121, 122, 196, 392, 500)]
    cases.append(("degenerate x12", _with_spectrum(rng, np.r_[np.full(12, 2.0), np.linspace(-1, 1, 188)])))
    x = rng.standard_normal((150, 40))
    cases.append(("rank 40 of 150", x @ x.T))
    pad = np.zeros((130, 130))
    pad[:100, :100] = _sym(rng, 100)
    pad[100:, 100:] = 7.0 * np.eye(30)
    cases.append(("padding block", pad))
    cases.append(("tight cluster", _with_spectrum(rng, np.r_[1 + 1e-10 * np.arange(6), np.linspace(-3, 3, 194)])))
    cases.append(("cluster of 100", _with_spectrum(rng, np.r_[0.5 + 1e-6 * np.arange(100), np.linspace(1, 3, 200)])))
    cases.append(("0.5 I + 1e-3 A", 0.5 * np.eye(392) + 1e-3 * _sym(rng, 392) / np.sqrt(392)))
    cases.append(("graded", _with_spectrum(rng, 10.0 ** -np.arange(14.0))))
    cases.append(("diagonal", np.diag(np.arange(40.0))))
    cases.append(("zero", np.zeros((9, 9))))
    cases.append(("scaled 1e-200", 1e-200 * _sym(rng, 30)))
    cases.append(("scaled 1e+150", 1e150 * _sym(rng, 30)))
    upper_only = np.triu(_sym(rng, 60))        # lower triangle zero: only the upper one may be read
    cases.append(("upper only", upper_only))
    return cases


@pytest.mark.gpu
def test_syev_batched_matches_lapack(gpu):
    cases = _cases()
    print(f"[syevb] {len(cases)} matrices, orders {[a.shape[0] for _, a in cases]}", flush=True)
    got = L.syev_batched([a for _, a in cases])
    for (name, a), (w, v) in zip(cases, got):
        s = np.triu(a) + np.triu(a, 1).T
        wr = np.linalg.eigvalsh(s)
        scale = max(float(np.abs(wr).max()), 1e-200)     # the zero matrix: an absolute bound
        n = s.shape[0]
        ew = float(np.abs(w - wr).max()) / scale
        res = float(np.abs(s @ v - v * w).max()) / scale
        orth = float(np.abs(v.T @ v - np.eye(n)).max())
        print(f"  {name:16s} N={n:4d} |w - w_lapack| {ew:.1e}  |AV - VW| {res:.1e}  |V'V - I| {orth:.1e}")
        assert np.all(np.diff(w) >= 0), name
        assert ew <= 1e-13 * max(1.0, np.sqrt(n)), (name, ew)
        assert res <= 1e-12, (name, res)
        assert orth <= 1e-12, (name, orth)


@pytest.mark.gpu
def test_syev_batched_rejects_orders_above_its_limit(gpu):
    with pytest.raises(L.NV2Error, match="outside"):
        L.syev_batched([np.eye(1025)])


@pytest.mark.gpu
def test_graph_replays_match_direct_launches(gpu):
    """Inside an evaluation the batched solves run through plans: the first call with a problem list
    picks the launch shapes by timing and solves, the second captures its launches in a hipGraph, later
    calls replay it.  Every replay, with other evaluations and eager work between them, gives the same
    class table bit for bit as direct launches (NV2_SYEVB_GRAPH=0) on the same inputs."""
    print("[syevb graphs] nact=6 ncore=3 nvirt=6 seed=4 (and a CAS(5,5) context between replays)", flush=True)
    P = synth.Problem(6, 3, 6, seed=4, nele=6)
    Q = synth.Problem(5, 2, 4, seed=7, nele=4)
    ev, other = Evaluator.for_problem(P), Evaluator.for_problem(Q)
    try:
        os.environ["NV2_SYEVB_GRAPH"] = "0"
        ref = ev.run()
        ref_other = other.run()
        os.environ.pop("NV2_SYEVB_GRAPH")
        outs, launches = [], []
        for _ in range(5):
            outs.append(ev.run())
            assert np.array_equal(other.run(), ref_other)
            launches.append(ev.pc_graphs().copy())
    finally:
        os.environ.pop("NV2_SYEVB_GRAPH", None)
        ev.close()
        other.close()
    print("  graph launches after each run (metrics, K', Fock):", [l.tolist() for l in launches])
    for k, o in enumerate(outs):
        assert np.array_equal(o, ref), k
    # the K' and Fock call sites replay from the third run with graphs on (the direct run was the first)
    assert launches[-1][1] >= 3 and launches[-1][2] >= 3, launches

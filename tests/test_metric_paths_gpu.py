"""The PC sums through the Cholesky-factored metric equal those through its eigen-decomposition
(the (+-1)' classes and, through the batched solves, every other PC class).

When no metric eigenvalue falls below thr_lindep the reference's thresholded inverse is S^-1, and
the PC sums are invariant under caux -> caux O (O orthogonal); the library then factors S = U^T U
and uses U^-1 in place of caux.  NV2_METRIC_EIGEN=1 forces the eigen-decomposition path on the
same resident inputs; both must agree to relative 1e-9 (the Cholesky path and the eigen path
differ by rounding amplified by the metric's condition number).
"""
import os

import numpy as np
import pytest

from nevpt2_amd import synth
from nevpt2_amd.evaluator import Evaluator


@pytest.mark.gpu
@pytest.mark.parametrize("n,nele,nc,nv,seed,keep", [(8, 8, 3, 6, 5, None), (7, 6, 2, 5, 3, None), (8, 8, 3, 6, 6, 600)])
def test_cholesky_metric_matches_eigen_path(gpu, n, nele, nc, nv, seed, keep):
    print(f"[metric paths] nact={n} nele={nele} ncore={nc} nvirt={nv} seed={seed} keep={keep}", flush=True)
    P = synth.Problem(n, nc, nv, seed=seed, keep=keep, nele=nele, host_rdms=False)
    ev = Evaluator.for_problem(P, rdms="gpu")
    try:
        os.environ.pop("NV2_METRIC_EIGEN", None)
        t_chol = ev.run(0)
        m_chol = ev.pc_metric()
        os.environ["NV2_METRIC_EIGEN"] = "1"
        t_eig = ev.run(0)
        m_eig = ev.pc_metric()
    finally:
        os.environ.pop("NV2_METRIC_EIGEN", None)
        ev.close()
    print("(+1)'/(-1)' metric: Cholesky flags", m_chol[0, 5:7, 2], "kept", m_eig[0, 5:7, 1],
          "min eigenvalue", m_eig[0, 5:7, 0], flush=True)
    assert np.all(m_eig[0, :, 2] == 0)
    # the factored path is taken exactly when the eigen path keeps every eigenvalue of both metrics
    # (the (+1)' and (-1)' problems are solved as one batch)
    N = n ** 3
    both = all(m_eig[0, c, 1] == N for c in (5, 6))
    for c in (5, 6):
        assert (m_chol[0, c, 2] == 1) == both
    assert np.allclose(t_chol, t_eig, rtol=1e-9, atol=1e-10)


@pytest.mark.gpu
def test_metric_below_threshold_takes_eigen_path(gpu):
    """With thr_lindep above the smallest metric eigenvalue the factorisation of S - thr*I fails and
    the eigen path (with its thresholding) runs; results equal the forced eigen path."""
    n, nele, nc, nv, seed = 7, 6, 2, 5, 3
    print(f"[metric threshold] nact={n} nele={nele} seed={seed} thr=0.1", flush=True)
    P = synth.Problem(n, nc, nv, seed=seed, nele=nele, host_rdms=False)
    ev = Evaluator.for_problem(P, rdms="gpu")
    try:
        os.environ.pop("NV2_METRIC_EIGEN", None)
        t_def = ev.run(0, thr_lindep=0.1)
        m_def = ev.pc_metric()
        os.environ["NV2_METRIC_EIGEN"] = "1"
        t_eig = ev.run(0, thr_lindep=0.1)
        m_eig = ev.pc_metric()
    finally:
        os.environ.pop("NV2_METRIC_EIGEN", None)
        ev.close()
    print("kept", m_eig[0, 5:7, 1], "min eigenvalue", m_eig[0, 5:7, 0], flush=True)
    assert min(m_eig[0, 5:7, 0]) < 0.1 and np.all(m_def[0, 5:7, 2] == 0)
    assert np.all(m_def[0, 5:7, 1] == m_eig[0, 5:7, 1]) and m_eig[0, 5:7, 1].min() < n ** 3
    # (+-1)': the eigen path both times; the other PC classes may keep their factored metric (their
    # eigenvalues can all lie above 0.1), which rounds differently
    assert np.allclose(t_def[:, 5:7], t_eig[:, 5:7], rtol=1e-12, atol=1e-12)
    assert np.allclose(t_def, t_eig, rtol=1e-9, atol=1e-10)

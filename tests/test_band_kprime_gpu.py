"""The (+-1)' PC sums through the band form of K' equal those through its tridiagonal form.

With the factored metric the library reduces K' = Q B Q^T to half-bandwidth 64 (CholeskyQR2
panels, two-sided block-reflector updates) and solves (B + s_x) u = Q^T y_x by banded elimination
with partial pivoting, one workgroup per external orbital; NV2_K_TRIDIAG=1 reduces K' to
tridiagonal form instead (one Householder column per launch, tridiagonal solves).  NV2_K_BAND=1
forces the band form at these small orders (by default it starts at N = 512).  Both are exact
rewritings of sum_mu r_mu^2 / (w_mu + s_x) and must agree to relative 1e-10 on every class, the
QD couplings included.  The orders cover a final Householder block of exactly 64 rows (N = 512)
and ragged ones (N = 216, 343).  Smaller active spaces (nact 4, 5 at these fillings) have
singular metrics and take the eigen path, which has no band form.
"""
import os

import numpy as np
import pytest

from nevpt2_amd import synth
from nevpt2_amd.evaluator import Evaluator

CASES = [  # (nact, nele, ncore, nvirt, seed, roots): both (+-1)' metrics positive definite above thr
    (6, 6, 2, 5, 2, 2),
    (7, 6, 2, 5, 3, 1),
    (7, 8, 2, 5, 3, 1),
    (8, 8, 3, 6, 5, 1),
]


@pytest.mark.gpu
@pytest.mark.parametrize("n,nele,nc,nv,seed,M", CASES)
def test_band_form_matches_tridiagonal_form(gpu, n, nele, nc, nv, seed, M):
    print(f"[K' band] nact={n} nele={nele} ncore={nc} nvirt={nv} seed={seed} roots={M}", flush=True)
    P = synth.Problem(n, nc, nv, seed=seed, nele=nele, nroots=M, host_rdms=False)
    ev = Evaluator.for_problem(P, rdms="gpu")
    try:
        os.environ.pop("NV2_K_TRIDIAG", None)
        os.environ["NV2_K_BAND"] = "1"
        t_band = ev.run()
        q_band = ev.heff() if M > 1 else None
        f_band = ev.pc_timings()
        chol = ev.pc_metric()[:, 5:7, 2]
        os.environ.pop("NV2_K_BAND", None)
        os.environ["NV2_K_TRIDIAG"] = "1"
        t_trd = ev.run()
        q_trd = ev.heff() if M > 1 else None
        f_trd = ev.pc_timings()
    finally:
        os.environ.pop("NV2_K_TRIDIAG", None)
        os.environ.pop("NV2_K_BAND", None)
        ev.close()
    print("forms", f_band[:, 3], f_trd[:, 3], "Cholesky", chol.ravel(), flush=True)
    for s in range(M):
        if chol[s].all():
            assert f_band[s, 3] == 2 and f_band[s, 1] == n ** 3
            assert f_trd[s, 3] == 1
    assert np.allclose(t_band, t_trd, rtol=1e-10, atol=1e-11)
    if M > 1:
        for a, b in zip(q_band, q_trd):
            assert np.allclose(a, b, rtol=1e-10, atol=1e-11)
    assert chol.all(), "the Cholesky path (and with it the band form) was not taken"


@pytest.mark.gpu
def test_metric_eigenvalue_test_paths_agree(gpu):
    """The Cholesky path needs every eigenvalue of the (+-1)' metrics above thr.  The default test
    reads trace(S^-1) = ||U^-1||_F^2 < 1/thr off the inverse factor the path forms anyway; when that
    bound is not met the library factors S - thr*I (the exact test).  NV2_CHOL_EXACT=1 sends every
    metric through the exact test: both must take the Cholesky path and give the same energies."""
    n, nele, nc, nv, seed = 7, 6, 2, 5, 3
    print(f"[metric test] nact={n} nele={nele} ncore={nc} nvirt={nv} seed={seed}", flush=True)
    P = synth.Problem(n, nc, nv, seed=seed, nele=nele, nroots=1, host_rdms=False)
    ev = Evaluator.for_problem(P, rdms="gpu")
    try:
        os.environ.pop("NV2_CHOL_EXACT", None)
        t_trace = ev.run()
        c_trace = ev.pc_metric()[:, 5:7, 2]
        os.environ["NV2_CHOL_EXACT"] = "1"
        t_exact = ev.run()
        c_exact = ev.pc_metric()[:, 5:7, 2]
    finally:
        os.environ.pop("NV2_CHOL_EXACT", None)
        ev.close()
    assert c_trace.all() and c_exact.all(), (c_trace, c_exact)
    assert np.allclose(t_trace, t_exact, rtol=1e-13, atol=1e-14)

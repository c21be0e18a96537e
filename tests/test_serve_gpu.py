"""Request loop on the GPU: every result is the same alone and in a mix of concurrent requests.

Requests of different shapes (CAS(4,4), CAS(6,6) with two roots and QD couplings,
CAS(5,5) with nele < nact) are first evaluated one at a time through their own
Evaluator, then submitted together to a server whose slots run them concurrently;
the class tables and effective Hamiltonians must agree to 1e-10 Eh (the same kernels
and solver calls run in both cases, so they agree to round-off).
"""
import numpy as np
import pytest

from nevpt2_amd import synth
from nevpt2_amd._lib import NV2Error
from nevpt2_amd.evaluator import Evaluator
from nevpt2_amd.serve import Server, ctx_bytes

SHAPES = [  # (nact, ncore, nvirt, seed, roots, active electrons)
    (4, 2, 3, 5, 1, 4),
    (6, 3, 5, 2, 2, 6),
    (5, 2, 4, 3, 1, 4),
]


@pytest.fixture(scope="module")
def problems(gpu):
    out = []
    for n, nc, nv, seed, M, nele in SHAPES:
        print(f"[serve case] nact={n} ncore={nc} nvirt={nv} seed={seed} roots={M} nele={nele}", flush=True)
        P = synth.Problem(n, nc, nv, seed=seed, nroots=M, nele=nele)
        ev = Evaluator.for_problem(P)
        tab = ev.run(0)
        heff = ev.heff() if M > 1 else None
        ev.close()
        out.append((P, tab, heff))
    return out


def _same(a, b, tol=1e-10):
    assert np.max(np.abs(np.asarray(a) - np.asarray(b))) <= tol


@pytest.mark.gpu
def test_mixed_requests_match_alone(problems):
    with Server(slots=3) as srv:
        ids = [(k, srv.submit(P)) for rep in range(2) for k, (P, _, _) in enumerate(problems)]
        for k, rid in ids:
            tab, heff, lat, qms = srv.wait(rid)
            _same(tab, problems[k][1])
            if problems[k][2] is not None:
                _same(heff[0], problems[k][2][0])
                _same(heff[1], problems[k][2][1])
            assert lat >= qms >= 0.0


@pytest.mark.gpu
def test_budget_reuses_resident_contexts(problems):
    P, ref, _ = problems[0]
    one = ctx_bytes(P.ncore, P.nact, P.nvirt, P.metat, P.B.shape[0])
    # room for one context: the four requests run one after the other on the pooled context
    with Server(slots=2, hbm_budget_bytes=int(1.5 * one)) as srv:
        ids = [srv.submit(P) for _ in range(4)]
        for rid in ids:
            tab, _, _, _ = srv.wait(rid)
            _same(tab, ref)


@pytest.mark.gpu
def test_request_above_budget_fails(problems):
    P = problems[1][0]
    with Server(slots=1, hbm_budget_bytes=4096) as srv:
        rid = srv.submit(P)
        with pytest.raises(NV2Error, match="above the HBM budget"):
            srv.wait(rid)


@pytest.mark.gpu
def test_resident_contexts_run_concurrently(problems):
    P, ref, _ = problems[2]
    evs = [Evaluator.for_problem(P) for _ in range(3)]
    try:
        with Server(slots=3) as srv:
            ids = [srv.submit_ctx(ev) for ev in evs for _ in range(1)]
            for rid in ids:
                tab, _, _, _ = srv.wait(rid)
                _same(tab, ref)
    finally:
        for ev in evs:
            ev.close()

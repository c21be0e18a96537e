"""Spin-free RDMs built on the GPU from CI vectors (nv2_ci_rdm, nv2_ctx_set_ci).

The host string-algebra builder of the synthetic inputs (nevpt2_amd.synth.spinfree_rdm) is
pinned to second-quantisation brute force in tests/test_oracle_pins.py.  The GPU builder must
reproduce it to 1e-12 (FP64; the same products summed in another order) for k = 1..4, for
transition RDMs, for truncated CI vectors and for nele < nact; an open-shell vector
(nalpha != nbeta), which the host builder does not take, is checked against brute force
directly.  Built into a context, the RDMs must give the same packed ro4 and transition RDMs
(1e-12) and the same class energies (relative 1e-9) as the host-built inputs.
"""
from itertools import combinations, product

import numpy as np
import pytest

from nevpt2_amd import synth
from nevpt2_amd.evaluator import Evaluator, ci_rdm
from oracle import brute

CASES = [  # (nact, nele, seed, keep: largest-|c| determinants kept)
    (4, 4, 11, None),
    (6, 6, 3, None),
    (6, 6, 4, 120),
    (5, 4, 3, None),
]


@pytest.mark.gpu
@pytest.mark.parametrize("n,nele,seed,keep", CASES)
def test_ci_rdm_matches_host_builder(gpu, n, nele, seed, keep):
    print(f"[ci rdm] nact={n} nele={nele} seed={seed} keep={keep}", flush=True)
    C = synth.random_ci(n, nele, seed, nroots=2, keep=keep)
    blk = synth._Blocks(n, nele)
    na = nele // 2
    for k in (1, 2, 3, 4):
        ref = synth.spinfree_rdm(k, C[0], blocks=blk)
        got = ci_rdm(n, na, na, k, C[0])
        assert np.max(np.abs(got - ref)) < 1e-12, k
    for k in (1, 2, 3):        # transition RDMs <C1| .. |C0>
        ref = synth.spinfree_rdm(k, C[1], C[0], blocks=blk)
        got = ci_rdm(n, na, na, k, C[1], C[0])
        assert np.max(np.abs(got - ref)) < 1e-12, k


def _psi(C, n, na, nb):
    psi = {}
    for ia, sa in enumerate(combinations(range(n), na)):
        ba = sum(1 << p for p in sa)
        for ib, sb in enumerate(combinations(range(n), nb)):
            if C[ia, ib] != 0.0:
                psi[ba | sum(1 << (n + p) for p in sb)] = C[ia, ib]
    return psi


@pytest.mark.gpu
def test_ci_rdm_open_shell_brute(gpu):
    n, na, nb = 4, 2, 1
    rng = np.random.default_rng(9)
    C = rng.standard_normal((6, 4))
    C /= np.linalg.norm(C)
    psi = _psi(C, n, na, nb)
    g1 = ci_rdm(n, na, nb, 1, C)
    g2 = ci_rdm(n, na, nb, 2, C)
    assert abs(np.trace(g1) - 3.0) < 1e-12
    for p, q in product(range(n), repeat=2):
        assert abs(g1[p, q] - brute.expect_E(psi, [(p, q)], n)) < 1e-12
    # <a+p1 a+p2 a_q2 a_q1> = <E_p1q1 E_p2q2> - delta(q1,p2) <E_p1q2>
    for p1, p2, q1, q2 in product(range(n), repeat=4):
        val = brute.expect_E(psi, [(p1, q1), (p2, q2)], n)
        if q1 == p2:
            val -= brute.expect_E(psi, [(p1, q2)], n)
        assert abs(g2[p1, p2, q1, q2] - val) < 1e-12, (p1, p2, q1, q2)


@pytest.mark.gpu
@pytest.mark.parametrize("n,nc,nv,M,seed,nele", [(6, 3, 5, 2, 2, 6), (5, 2, 4, 1, 3, 4), (4, 2, 3, 3, 7, 4)])
def test_context_from_ci_matches_host_inputs(gpu, n, nc, nv, M, seed, nele):
    print(f"[ci ctx] nact={n} ncore={nc} nvirt={nv} roots={M} seed={seed} nele={nele}", flush=True)
    P = synth.Problem(n, nc, nv, seed=seed, nroots=M, nele=nele)
    evh = Evaluator.for_problem(P)
    evg = Evaluator.for_problem(P, rdms="gpu")
    try:
        ro4 = evg.get("ro4", P.ro4.shape)
        assert np.max(np.abs(ro4 - P.ro4)) < 1e-12
        if M > 1:
            for name, ref in (("3-TRDM", P.daaaoff), ("2-TRDM", P.taaoff), ("1-TRDM", P.daloff)):
                got = evg.get(name, ref.shape)
                assert np.max(np.abs(got - ref)) < 1e-12, name
        # the inputs agree to rounding (1e-16); a class whose denominators nearly vanish (PC energies of
        # 1e5 Eh occur on these random vectors) amplifies that, so the bound is relative
        th, tg = evh.run(0), evg.run(0)
        assert np.allclose(th, tg, rtol=1e-9, atol=1e-9)
        if M > 1:
            for a, b in zip(evh.heff(), evg.heff()):
                assert np.allclose(a, b, rtol=1e-9, atol=1e-9)
    finally:
        evh.close()
        evg.close()

"""CPU pins of the stage-2 oracle."""
import numpy as np

from nevpt2_amd import synth
from oracle import stage1 as S1, stage2 as S2


def test_class0_is_mp2_pair_energy():
    """Class (0) (core pair -> virtual pair) equals the closed-shell MP2 formula
    -sum_ijrs (ir|js)[2(ir|js)-(is|jr)]/(e_r+e_s-e_i-e_j) in the canonical orbitals."""
    P = synth.Problem(4, 3, 7, seed=4)
    koop = S1.koopmans_all(P)
    out, c = S2.state_energies(P, koop, 0)
    nc, n = P.ncore, P.nact
    K = np.einsum("lir,ljs->irjs", c.Bcv, c.Bcv)
    e = c.eps
    ec, ev = e[:nc], e[nc + n:]
    D = ev[None, :, None, None] + ev[None, None, None, :] - ec[:, None, None, None] - ec[None, None, :, None]
    mp2 = -np.sum(K * (2 * K - K.transpose(0, 3, 2, 1)) / D)
    assert abs(out[0, 1] - mp2) < 1e-13
    assert out[0, 1] == out[0, 3]


def test_canonical_orbitals_diagonalise_fock():
    P = synth.Problem(4, 3, 7, seed=4)
    koop = S1.koopmans_all(P)
    _, c = S2.state_energies(P, koop, 0, with_4rdm=False)
    # rebuilding the transformed core-core Cholesky vectors from the canonical orbitals is orthogonal
    assert np.allclose(c.ccore.T @ c.ccore, np.eye(P.ncore), atol=1e-12)
    assert np.allclose(c.cvirt.T @ c.cvirt, np.eye(P.nvirt), atol=1e-12)


def test_pc_solve_limits():
    """With a unit metric and a symmetric K, pc_solve reduces to a plain eigen-decomposition."""
    rng = np.random.default_rng(0)
    K = rng.standard_normal((6, 6)); K = K + K.T
    coef, w, nnew, w1 = S2.pc_solve(np.eye(6), K)
    assert nnew == 6
    assert np.allclose(np.sort(w), np.linalg.eigvalsh(K))
    # metric eigenvalues below the threshold are dropped, negative ones too
    S = np.diag([1.0, 2.0, 1e-9, -0.5, 3.0, 4.0])
    coef, w, nnew, w1 = S2.pc_solve(S, K)
    assert nnew == 4

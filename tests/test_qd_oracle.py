"""QD oracle checks on the CPU: Heff assembly and the hermitisation are pinned by
mathematics (eigenvalue invariance, symmetry), the per-pair index orientation
by the transition-RDM bra/ket identity."""
import numpy as np

from nevpt2_amd import synth
from oracle import stage1 as S1, stage2 as S2, qd as Q


def _problem():
    P = synth.Problem(4, 2, 3, seed=5, nroots=3)
    return P, S1.koopmans_all(P)


def test_heff_diagonal_is_single_state_total():
    P, koop = _problem()
    Hsc, Hpc, tabs, off = Q.heff_matrices(P, koop)
    for i in range(P.metat):
        tab, _ = S2.state_energies(P, koop, i)
        assert abs(Hsc[i, i] - tab[:, 1].sum()) < 1e-12
        assert abs(Hpc[i, i] - tab[:, 3].sum()) < 1e-12
    # class (0) has no couplings; every other class does
    assert 0 not in off
    assert all(np.abs(a).max() > 0 for a, _ in off.values())


def test_hermitisation_keeps_the_spectrum():
    """Heff~ = C' diag(w) C'^T is symmetric with the eigenvalues of the non-Hermitian Heff
    (hermitian_heff, qdnevpt.F:8089-8184)."""
    rng = np.random.default_rng(3)
    H = np.diag([-1.0, -0.5, 0.2, 0.9]) + 0.05 * rng.standard_normal((4, 4))
    e, Ht = Q.diagonalize_heff(H, np.zeros(4))
    assert np.allclose(Ht, Ht.T, atol=1e-13)
    assert np.allclose(np.sort(np.linalg.eigvals(H).real), e, atol=1e-12)
    assert np.allclose(np.linalg.eigvalsh(Ht), e, atol=1e-12)
    # a symmetric Heff is returned unchanged (C' is then orthogonal with C' = V)
    Hs = 0.5 * (H + H.T)
    _, Ht2 = Q.diagonalize_heff(Hs, np.zeros(4))
    assert np.allclose(Ht2, Hs, atol=1e-12)


def test_transition_rdm_orientation():
    """The reference stores one pair I>J; the other orientation is the adjoint, which for
    spin-free RDMs swaps the creation and annihilation index groups.  Both orientations
    of the generator agree with that identity, which the off-diagonal terms rely on."""
    P = synth.Problem(4, 2, 3, seed=5, nroots=2)
    blk = synth._Blocks(4, P.nele)
    g1 = synth.spinfree_rdm(1, P.C[1], P.C[0], blocks=blk)
    g1t = synth.spinfree_rdm(1, P.C[0], P.C[1], blocks=blk)
    assert np.allclose(g1, g1t.T, atol=1e-13)
    g2 = synth.spinfree_rdm(2, P.C[1], P.C[0], blocks=blk)
    g2t = synth.spinfree_rdm(2, P.C[0], P.C[1], blocks=blk)
    assert np.allclose(g2, g2t.transpose(2, 3, 0, 1), atol=1e-13)
    # stored arrays are the partial traces of the pair's 3-RDM, like the diagonal ones
    assert np.allclose(P.daloff[0], g1, atol=1e-12)

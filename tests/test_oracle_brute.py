"""Pins of the oracle by Fock-space brute force (oracle/fock.py), independent of any RDM formula.

1. Every Koopmans matrix of stage 1 (koop_matrices.F90, koopro4QD.F90:3974-4459) equals a
   spin-summed commutator expectation value <Psi| O_bra^+ [H_act, O_ket] |Psi> of the active
   Dyall Hamiltonian, for random (non-eigen) CI vectors: element by element at 1e-12 relative.
2. For a CASCI eigenfunction the reference's own self-checks hold: bmat(i,j,ii,jj) = cmat(jj,i,j,ii)
   (fdiff, koopro4QD.F90:5053-5112) and gk0pf = 0 (fdiff_F, koop_matrices.F90:591-619).
3. The eight classes' SC and PC energies and norms of oracle/stage2.py equal SC/PC-NEVPT2 computed
   from the first-order functions themselves: for eigenfunctions (where NEVPT2 is defined without
   conventions) and for random CI vectors (commutator form; LAPACK's upper triangle of the
   non-symmetric Koopmans matrix taken in the reference's basis order).
4. The QD couplings of oracle/qd.py equal <H Psi_j|P (E_0 - H_D)^-1 P|H Psi_i> over state i's SC
   function / PC eigenfunctions.
5. Plausible slips (a transposed A-matrix, a B/C mix-up, a 1e-6 error in a Koopmans matrix) fail (3).
6. The .NATORB density blocks of oracle/stage2.py (qdnevpt.F:4823-4988, 5052-5203, 5266-5424) equal
   <Psi_0|E_pq|Psi_1> with Psi_1 the PC first-order function of classes (0)', (+1)' and (-1)', and slips in
   their density vectors fail.
"""
import numpy as np
import pytest

from nevpt2_amd import synth
from oracle import fock, qd, stage1 as S1, stage2 as S2

# (nact, ncore, nvirt, seed, active electrons)
SHAPES = [(4, 2, 2, 11, 4), (4, 1, 2, 5, 6), (5, 1, 2, 7, 6), (4, 2, 3, 13, 4), (6, 1, 1, 8, 6)]


def _eigen_problem(nact, nc, nv, seed, nele, nroots=1):
    P0 = synth.Problem(nact, nc, nv, seed=seed, nele=nele, nroots=nroots)
    if nroots == 1:
        C, _, gap = fock.casci_ground_state(P0)
        assert gap > 1e-3
        C = C[None]
    else:
        C, _ = fock.casci_ground_state(P0, nroots=nroots, spin=0)
        assert len(C) == nroots
    return synth.Problem(nact, nc, nv, seed=seed, nele=nele, nroots=nroots, ci=C)


def _koop_brute(P, s=0):
    """Every stage-1 Koopmans matrix of state s from its operator definition."""
    n, na = P.nact, P.nele // 2
    heff, gaa = fock.active_hamiltonian(P)
    K = fock.KoopmansBrute(P.C[s], n, na, P.nele - na, heff, gaa)
    a, c, E = K.a, K.c, K.E
    sp = (0, n)
    out = {}
    # one-index: 1-IPKO(ap,a) = 2 sum_s <a+_{ap s} [H, a_{a s}]>, 1-EAKO with a+ <-> a; D = 1/2 of them with aj
    bra_h = [[K.bra([a[s + x]]) for s in sp] for x in range(n)]
    bra_p = [[K.bra([c[s + x]]) for s in sp] for x in range(n)]
    com_h = [[K.comm_ket([a[s + x]]) for s in sp] for x in range(n)]
    com_p = [[K.comm_ket([c[s + x]]) for s in sp] for x in range(n)]
    one = lambda B, Cm: np.array([[sum(B[x][t] @ Cm[y][t] for t in range(2)) for y in range(n)] for x in range(n)])
    out["1-IPKO"] = 2.0 * one(bra_h, com_h)
    out["1-EAKO"] = 2.0 * one(bra_p, com_p)
    out["dmatIP"] = one(bra_h, com_h)
    out["dmatEA"] = one(bra_p, com_p)
    # gk0pf(a,b) = 2 <[H, E_ab]>
    out["gk0pf"] = 2.0 * np.array([[K.psi @ K.comm_ket([E[x][y]]) for y in range(n)] for x in range(n)])
    # two-index
    T = {k: np.zeros((n,) * 4) for k in ("hh", "pp", "EE", "x1", "hrdm")}
    for i0 in range(n):
        for i1 in range(n):
            for i2 in range(n):
                for i3 in range(n):
                    idx = (i0, i1, i2, i3)
                    hh = pp = x1 = hrdm = 0.0
                    for s in sp:
                        for t in sp:
                            # 2-IPKO(ap,bp,a,b) = sum <a+_{ap s} a+_{bp t} [H, a_{b t} a_{a s}]>
                            hh += K.bra([a[t + i1], a[s + i0]]) @ K.comm_ket([a[t + i3], a[s + i2]])
                            # 2-EAKO(ap,bp,a,b) = sum <a_{ap s} a_{bp t} [H, a+_{b t} a+_{a s}]>
                            pp += K.bra([c[t + i1], c[s + i0]]) @ K.comm_ket([c[t + i3], c[s + i2]])
                            # 2-HRDM(ap,bp,a,b) = sum <a_{ap s} a_{bp t} a+_{b t} a+_{a s}>
                            hrdm += K.bra([c[t + i1], c[s + i0]]) @ K.bra([c[t + i3], c[s + i2]])
                            # gk0pd(ap,bp,a,b) = sum <(a+_{bp s} a_{ap t})^+ [H, a+_{a s} a_{b t}]>
                            x1 += K.bra([c[s + i1], a[t + i0]]) @ K.comm_ket([c[s + i2], a[t + i3]])
                    T["hh"][idx], T["pp"][idx], T["x1"][idx], T["hrdm"][idx] = hh, pp, x1, hrdm
                    # gk0pa(ap,bp,a,b) = 2 <E_{ap bp} [H, E_ab]>
                    T["EE"][idx] = 2.0 * (K.bra([E[i1][i0]]) @ K.comm_ket([E[i2][i3]]))
    out["2-IPKO"], out["2-EAKO"], out["2-HRDM"] = T["hh"], T["pp"], T["hrdm"]
    out["gk0pa"], out["gk0pd"] = T["EE"], T["x1"]
    # three-index functions a_{x s} E_{yz} (IP) and a+_{x s} E_{yz} (EA)
    for tag, op in (("IP", a), ("EA", c)):
        m = n ** 3
        Bv = np.zeros((m, 2, K.fs.dim))
        Cv = np.zeros((m, 2, K.fs.dim))
        for f, (x, y, z) in enumerate(np.ndindex(n, n, n)):
            for t, s in enumerate(sp):
                Bv[f, t] = K.bra([op[s + x], E[y][z]])
                Cv[f, t] = K.comm_ket([op[s + x], E[y][z]])
        B1 = np.array([[K.bra([op[s + x]]) for s in sp] for x in range(n)])       # (n, 2, dim)
        C1 = np.array([[K.comm_ket([op[s + x]]) for s in sp] for x in range(n)])
        T33 = np.einsum("ftd,gtd->fg", Bv, Cv).reshape((n,) * 6)                  # [x y z, x' y' z']
        T31 = np.einsum("ftd,gtd->fg", Bv, C1).reshape((n,) * 4)                  # [x y z, w]
        T13 = np.einsum("ftd,gtd->fg", B1, Cv).reshape((n,) * 4)                  # [w, x y z]
        # amat(ap,bp,cp,a,b,c) = sum_s <E_{cp ap} a+_{bp s} [H, a_{b s} E_{a c}]>
        out["amat" + tag] = T33.transpose(1, 0, 2, 4, 3, 5)
        # bmat(ap,bp,cp,a) = sum_s <E_{cp ap} a+_{bp s} [H, a_{a s}]>
        out["bmat" + tag] = T31.transpose(1, 0, 2, 3)
        # cmat(ap,a,b,c) = sum_s <a+_{ap s} [H, a_{b s} E_{a c}]>
        out["cmat" + tag] = T13.transpose(0, 2, 1, 3)
    return out


KOOP_CASES = [(4, 2, 2, 11, 4), (4, 1, 2, 5, 6), (5, 1, 1, 3, 4)]


@pytest.mark.parametrize("shape", KOOP_CASES, ids=lambda s: "cas%d_%de" % (s[0], s[4]))
def test_koopmans_are_commutator_expectations(shape):
    """Random CI vectors: each Koopmans matrix equals its operator definition element by element."""
    n, nc, nv, seed, nele = shape
    P = synth.Problem(n, nc, nv, seed=seed, nele=nele)
    koop = S1.koopmans_all(P)
    brute = _koop_brute(P)
    for name, ref in brute.items():
        got = koop[name][0]
        err = float(np.abs(got - ref).max())
        scale = max(1.0, float(np.abs(ref).max()))
        print(f"{name}: max|oracle - brute| = {err:.2e} (scale {scale:.2e})")
        assert err <= 1e-12 * scale, name


def test_reference_self_checks_hold_for_an_eigenfunction():
    """fdiff: bmat(i,j,ii,jj) = cmat(jj,i,j,ii) (also btmat/ctmat); fdiff_F: gk0pf = 0; and the
    Koopmans matrices of an eigenfunction are symmetric in the sense K(bra,ket) = K(ket,bra)."""
    P = _eigen_problem(4, 2, 2, 11, 4)
    koop = S1.koopmans_all(P)
    for b, c in (("bmatIP", "cmatIP"), ("bmatEA", "cmatEA")):
        B, C = koop[b][0], koop[c][0]
        d = float(np.abs(B - C.transpose(1, 2, 3, 0)).max())
        assert d < 1e-10 * max(1.0, float(np.abs(B).max())), (b, d)
    assert float(np.abs(koop["gk0pf"][0]).max()) < 1e-10
    n = 4
    for name in ("amatIP", "amatEA"):
        A = koop[name][0].reshape(n ** 3, n ** 3)
        # amat(ap,bp,cp,a,b,c): bra function (bp; ap,cp), ket (b; a,c) -> symmetric under bra <-> ket
        assert float(np.abs(A - A.T).max()) < 1e-10 * float(np.abs(A).max()), name
    # a random CI vector breaks them (so the checks have teeth)
    Pr = synth.Problem(4, 2, 2, seed=11)
    kr = S1.koopmans_all(Pr)
    assert float(np.abs(kr["bmatIP"][0] - kr["cmatIP"][0].transpose(1, 2, 3, 0)).max()) > 1e-3
    assert float(np.abs(kr["gk0pf"][0]).max()) > 1e-3


def _compare(ref, bf, gaps=None, rel=1e-10, norm_rel=1e-10):
    for k in range(8):
        for col, ncol in ((1, 0), (3, 2)):
            d = abs(ref[k, col] - bf[k, col])
            tol = rel * max(1.0, abs(bf[k, col])) + norm_rel * abs(bf[k, ncol])
            print(f"{fock.CLASS_NAMES[k]:6s} {'SC' if col == 1 else 'PC'} E {bf[k, col]: .10e} "
                  f"norm {bf[k, ncol]:.3e} |dE| {d:.2e}")
            assert d <= tol, (fock.CLASS_NAMES[k], col, d, tol)
            # norm = sum r^2/d^2 and E = -sum r^2/d: a relative error x in the smallest denominator moves
            # the norm by about twice the relative amount it moves the energy
            dn = abs(ref[k, ncol] - bf[k, ncol])
            ntol = 1e-9 * max(1.0, abs(bf[k, ncol])) + 2.0 * abs(bf[k, ncol]) * tol / max(abs(bf[k, col]), 1e-300)
            assert dn <= ntol, (fock.CLASS_NAMES[k], "norm", dn, ntol)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "cas%d_nc%d_nv%d_%de" % (s[0], s[1], s[2], s[4]))
def test_classes_equal_first_order_functions_eigenfunction(shape):
    """CASCI eigenfunction: the oracle's eight classes (SC and PC) equal NEVPT2 from the perturbers."""
    P = _eigen_problem(*shape)
    koop = S1.koopmans_all(P)
    ref, _ = S2.state_energies(P, koop, 0)
    bf, residual, gaps = fock.nevpt2(P)
    assert residual < 1e-10
    for k, (kept_min, drop_max) in gaps.items():
        # the linear-dependence decisions are unambiguous: no metric eigenvalue near the threshold
        assert kept_min > 1e-5 and drop_max < 1e-10, (k, kept_min, drop_max)
    _compare(ref, bf, rel=1e-11, norm_rel=0.0)


@pytest.mark.parametrize("shape", SHAPES[:4], ids=lambda s: "cas%d_nc%d_nv%d_%de" % (s[0], s[1], s[2], s[4]))
def test_classes_equal_commutator_form_random_ci(shape):
    """Random CI vectors (the inputs of the GPU parity tests): SC from <O^+[H_D,O]>/<O^+O> with O the
    part of H reaching each sector, PC from the commutator matrix symmetrised by its upper triangle in
    the reference's basis.  Near-intruder classes (norm >> 1) amplify round-off by their norm."""
    P = synth.Problem(*shape[:3], seed=shape[3], nele=shape[4])
    koop = S1.koopmans_all(P)
    ref, _ = S2.state_energies(P, koop, 0)
    fsp = fock.FullSpace(P, [P.C[0]]).setup_state(0)
    bf = np.array([fsp.class_energies(k)[:4] for k in range(8)])
    _compare(ref, bf, rel=1e-10, norm_rel=1e-10)


@pytest.mark.parametrize("eig", [True, False], ids=["eigen", "random"])
def test_qd_couplings_equal_first_order_functions(eig):
    """QD-NEVPT2: every class's off-diagonal Heff entries of oracle/qd.py."""
    shape = (4, 2, 2, 11, 4)
    M = 2 if eig else 3
    P = _eigen_problem(*shape, nroots=M) if eig else synth.Problem(*shape[:3], seed=shape[3], nele=4, nroots=M)
    koop = S1.koopmans_all(P)
    Hsc, Hpc, tabs, off = qd.heff_matrices(P, koop)
    bsc, bpc, boff = fock.qd_heff(P)
    offd = ~np.eye(M, dtype=bool)
    for k, (osc, opc) in off.items():
        for name, o, b, t in (("SC", osc, boff[k][0], 0), ("PC", opc, boff[k][1], 2)):
            # a coupling into column i is a sum like class i's energy: bound its error by that class's norm
            norm_i = np.array([tabs[i][k, t] for i in range(M)])[None, :]
            d = np.abs(o - b)[offd]
            tol = (1e-11 * np.maximum(1.0, np.abs(b)) + 1e-10 * norm_i * np.ones((M, 1)))[offd]
            print(f"{fock.CLASS_NAMES[k]:6s} {name} max|d| {d.max():.2e} scale {np.abs(b)[offd].max():.2e}")
            assert (d <= tol).all(), (fock.CLASS_NAMES[k], name, d, tol)
    assert float(np.abs(bsc).max()) > 1e-4 and float(np.abs(bpc[offd]).max()) > 1e-5


@pytest.mark.parametrize("slip", ["amat_transposed", "bc_swapped", "eako_scaled"])
def test_plausible_slips_fail(slip):
    """Mistakes of the kind a restatement makes are caught by the brute-force class energies."""
    P = _eigen_problem(4, 2, 2, 11, 4)
    koop = S1.koopmans_all(P)
    bf, _, _ = fock.nevpt2(P)
    k2 = dict(koop)
    n = 4
    if slip == "amat_transposed":
        A = koop["amatIP"][0].reshape(n ** 3, n ** 3)
        # perturb the symmetric eigenfunction matrix into a non-symmetric one, then transpose it
        k2["amatIP"] = np.asfortranarray(koop["amatIP"].copy())
        k2["amatIP"][0] = (A + 1e-3 * np.triu(A, 1)).T.reshape((n,) * 6)
        cls = 6
    elif slip == "bc_swapped":
        k2["bmatIP"] = np.asfortranarray(koop["cmatIP"].transpose(0, 2, 3, 4, 1) * 1.001)
        cls = 6
    else:
        k2["1-EAKO"] = koop["1-EAKO"] * (1.0 + 1e-6)
        cls = 1
    ref, _ = S2.state_energies(P, k2, 0)
    d = max(abs(ref[cls, 1] - bf[cls, 1]), abs(ref[cls, 3] - bf[cls, 3]))
    assert d > 1e-9, (slip, d)


@pytest.mark.parametrize("frozen", [(0,), (0, 2)], ids=["one", "two"])
def test_frozen_core_orbitals_are_never_holes(frozen):
    """.FROZEN (qdnevpt.F:733-745): the oracle's orbital-energy device for frozen core orbitals equals the
    reference's skip, i.e. NEVPT2 from first-order functions whose external labels exclude those holes."""
    P = _eigen_problem(4, 3, 2, 11, 4)
    koop = S1.koopmans_all(P)
    ref, _ = S2.state_energies(P, koop, 0, frozen=frozen)
    bf, _, _ = fock.nevpt2(P, frozen=frozen)
    full, _ = S2.state_energies(P, koop, 0)
    _compare(ref, bf, rel=1e-11, norm_rel=0.0)
    assert abs(full[:, 1].sum() - ref[:, 1].sum()) > 1e-3       # freezing changes the energy


def _brute_rho1(P, s=0, frozen=()):
    """The .NATORB density matrix from the first-order functions, back in the input orbitals."""
    fsp = fock.FullSpace(P, [P.C[t] for t in range(P.metat)]).setup_state(s)
    fsp.frozen = tuple(frozen)
    nc, n, norb = P.ncore, P.nact, P.norb
    A, V = slice(nc, nc + n), slice(nc + n, norb)
    B = np.zeros((norb, norb))
    B[:nc, :nc] = 2.0 * np.eye(nc)
    B[A, A] = fsp.gam
    for k, (r0, c0) in ((7, (slice(0, nc), V)), (5, (slice(0, nc), A)), (6, (A, V))):
        blk = fsp.class_rho1(k)
        B[r0, c0] = blk
        B[c0, r0] = blk.T
    return fsp.U @ B @ fsp.U.T


@pytest.mark.parametrize("case", ["eigen_a", "eigen_b", "random", "frozen"])
def test_natorb_density_equals_first_order_functions(case):
    """rho_pq = <Psi_0|E_pq|Psi_1> in every first-order block, for the input orbitals (the canonical
    orbitals' signs cancel in U rho U^T)."""
    frozen = ()
    if case == "eigen_a":
        P = _eigen_problem(4, 2, 2, 11, 4)
    elif case == "eigen_b":
        P = _eigen_problem(5, 1, 2, 7, 6)
    elif case == "random":
        P = synth.Problem(4, 2, 3, seed=13, nele=4)
    else:
        P, frozen = _eigen_problem(4, 3, 2, 11, 4), (1,)
    koop = S1.koopmans_all(P)
    _, Rmo, occ = S2.rho1(P, koop, 0, frozen=frozen)
    Bmo = _brute_rho1(P, 0, frozen)
    off = np.abs(Rmo - np.diag(np.diag(Rmo))).max()
    d = float(np.abs(Rmo - Bmo).max())
    print(f"{case}: max|rho offdiag| {off:.3e}  |d| {d:.2e}")
    assert off > 1e-2
    assert d <= (1e-12 if case != "random" else 1e-9) * max(1.0, float(np.abs(Bmo).max()))
    assert np.abs(occ - np.linalg.eigvalsh(Bmo)).max() <= 1e-9 * max(1.0, float(np.abs(Bmo).max()))


@pytest.mark.parametrize("slip", ["v0p_one_coef", "plus_as_minus", "minus_bc_swapped"])
def test_natorb_slips_fail(slip, monkeypatch):
    P = _eigen_problem(4, 2, 2, 11, 4)
    koop = S1.koopmans_all(P)
    Bmo = _brute_rho1(P)
    monkeypatch.setattr(S2, "_RHO_SLIP", slip)
    _, Rmo, _ = S2.rho1(P, koop, 0)
    d = float(np.abs(Rmo - Bmo).max())
    assert d > 1e-6, (slip, d)


def test_order_occupations_is_order_eigvalvec():
    """order_eigvalvec (qdnevpt.F:1484-1519) sorts by descending modulus and, with >=, lets the later of
    two equal moduli go first."""
    got = S2.order_occupations([0.1, -1.9, 1.9, 0.5])
    assert list(got) == [1.9, -1.9, 0.5, 0.1]

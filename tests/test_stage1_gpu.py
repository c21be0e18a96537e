"""GPU parity of stage 1 (Koopmans builds) against the C oracle, through the C ABI."""
import numpy as np
import pytest

from nevpt2_amd import synth, koopro
from oracle import stage1 as S

CASES = [  # (nact, ncore, nvirt, seed, nroots)
    (4, 2, 3, 11, 1),
    (6, 3, 5, 1, 1),
    (6, 2, 4, 5, 2),
]


def _close(got, ref, rel=1e-11, name=""):
    scale = max(1.0, float(np.abs(ref).max()))
    err = float(np.abs(got - ref).max())
    assert err <= rel * scale, f"{name}: max|diff|={err:.3e} scale={scale:.3e}"


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"cas{c[0]}_m{c[4]}")
def case(request):
    n, nc, nv, seed, M = request.param
    print(f"[stage1 case] nact={n} ncore={nc} nvirt={nv} seed={seed} roots={M}", flush=True)
    P = synth.Problem(n, nc, nv, seed=seed, nroots=M)
    ref = S.koopmans_all(P)
    return P, ref


@pytest.mark.gpu
def test_bro(gpu, case):
    P, _ = case
    d3, d2, d1 = koopro.bro(P.ro4, P.nact, P.metat, P.nele)
    r3 = S.bro3(P.ro4, P.nact, P.metat, P.nele)
    r2 = S.bro2(r3, P.nact, P.metat, P.nele)
    r1 = S.bro1(r2, P.nact, P.metat, P.nele)
    _close(d3, r3, name="daaa"); _close(d2, r2, name="taa"); _close(d1, r1, name="dal")


@pytest.mark.gpu
def test_aj_atwo(gpu, case):
    P, ref = case
    aj, atwo = koopro.get_aj_atwo(P.ncore, P.nact, P.norb, P.B, P.oneint)
    _close(aj, ref["aj"], name="aj"); _close(atwo, ref["atwo"], name="atwo")


@pytest.mark.gpu
def test_koop_matrices(gpu, case):
    P, ref = case
    out = koopro.koop_matrices(P.metat, P.nact, P.ncore, P.norb, ref["atwo"], P.daaa, P.taa, P.dal, ref["aj"])
    for k, v in out.items():
        _close(v, ref[k], name=k)


@pytest.mark.gpu
def test_bamat_fp64(gpu, case):
    P, ref = case
    am, at = koopro.bamat(ref["atwo"], P.ro4, P.daaa, P.taa, P.dal, P.nact, P.ncore, ref["aj"], P.norb, P.metat)
    _close(am, ref["amatIP"], name="amatIP"); _close(at, ref["amatEA"], name="amatEA")


@pytest.mark.gpu
def test_bamat_fp32(gpu, case):
    P, ref = case
    am, at = koopro.bamat(ref["atwo"], P.ro4, P.daaa, P.taa, P.dal, P.nact, P.ncore, ref["aj"], P.norb, P.metat,
                          prec=1)
    _close(am, ref["amatIP"], rel=2e-5, name="amatIP fp32"); _close(at, ref["amatEA"], rel=2e-5, name="amatEA fp32")


@pytest.mark.gpu
def test_bamat_bf16x3(gpu, case):
    """Split-BF16x3 contractions (three exact bf16 terms per FP32 operand, six products, FP32 accumulation)
    are as accurate as the FP32 form: FP32 E4 storage and FP32 accumulation set the error, not the split."""
    P, ref = case
    am, at = koopro.bamat(ref["atwo"], P.ro4, P.daaa, P.taa, P.dal, P.nact, P.ncore, ref["aj"], P.norb, P.metat,
                          prec=2)
    a32, t32 = koopro.bamat(ref["atwo"], P.ro4, P.daaa, P.taa, P.dal, P.nact, P.ncore, ref["aj"], P.norb, P.metat,
                            prec=1)
    for got, g32, r, nm in ((am, a32, ref["amatIP"], "amatIP"), (at, t32, ref["amatEA"], "amatEA")):
        scale = float(np.abs(r).max())
        e16, e32 = float(np.abs(got - r).max()), float(np.abs(g32 - r).max())
        print(f"  {nm}: max|bf16x3 - fp64| = {e16:.2e}, max|fp32 - fp64| = {e32:.2e} (scale {scale:.2e})")
        _close(got, r, rel=2e-5, name=nm + " bf16x3")
        assert e16 <= 4.0 * e32 + 1e-12 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("n", [5, 7, 8, 9])
def test_bamat_sizes(gpu, n):
    """The 4-RDM contraction passes at sizes whose paddings differ: n^2 not a multiple of 4 (odd n) pads
    the K range of C1 with zero rows, n not a multiple of 4 pads the x6 blocks of C3/C4, and n = 8 pads
    neither."""
    print(f"[stage1 bamat] nact={n} ncore=1 nvirt=2 nele={n - n % 2}", flush=True)
    P = synth.Problem(n, 1, 2, seed=3, nele=n - n % 2)
    ref = S.koopmans_all(P)
    for prec, rel in ((0, 1e-11), (1, 2e-5), (2, 2e-5)):
        am, at = koopro.bamat(ref["atwo"], P.ro4, P.daaa, P.taa, P.dal, P.nact, P.ncore, ref["aj"], P.norb,
                              P.metat, prec=prec)
        _close(am, ref["amatIP"], rel=rel, name=f"amatIP prec {prec}")
        _close(at, ref["amatEA"], rel=rel, name=f"amatEA prec {prec}")

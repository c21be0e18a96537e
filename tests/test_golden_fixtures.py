"""The committed golden fixtures are what the oracle computes from their recipes (CPU)."""
import json
import os

import numpy as np
import pytest

from nevpt2_amd import synth
from oracle import stage1 as S1, stage2 as S2

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name", ["c1", "cas8"])
def test_fixture_matches_oracle(name):
    with open(os.path.join(HERE, "golden", f"{name}.json")) as f:
        fx = json.load(f)
    P = synth.Problem(fx["nact"], fx["ncore"], fx["nvirt"], seed=fx["seed"], nele=fx["nele"])
    koop = S1.koopmans_all(P)
    tab, _ = S2.state_energies(P, koop, 0)
    assert np.allclose(tab, np.array(fx["table"]), rtol=1e-12, atol=1e-14)
    for k, rec in fx["koopmans"].items():
        a = koop[k][0]
        assert list(a.shape) == rec["shape"]
        assert np.allclose([a[tuple(p)] for p in rec["positions"]], rec["values"], rtol=1e-12, atol=1e-13)


def test_fixture_files_complete():
    for name in ("c1", "c2", "cas8", "cas9"):
        with open(os.path.join(HERE, "golden", f"{name}.json")) as f:
            fx = json.load(f)
        assert len(fx["table"]) == 8 and all(len(r) == 4 for r in fx["table"])
        assert np.isfinite(np.array(fx["table"])).all()
        assert len(fx["koopmans"]) == 16

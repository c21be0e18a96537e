import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def energy_tol(norm):
    """Per-class bound on |E_gpu - E_oracle| (Eh) for the FP64 path: 1e-9 Eh.

    A class energy is -sum_mu r_mu^2 / d_mu and its norm is sum_mu r_mu^2 / d_mu^2, so a round-off
    delta in a small denominator d moves the energy by about norm * delta.  Random CI vectors (not
    eigenfunctions of the Dyall Hamiltonian) give near-intruder classes with norms up to 1e6; only a
    class whose norm exceeds 500 gets 1e-11 Eh per unit of norm on top, i.e. denominators good to
    1e-11 Eh.  Observed: CAS(4,4) (+1)' PC, norm 520, |dE| 2.6e-9 Eh when the GPU builds the RDMs from
    the CI vector and the oracle on the host (5e-12 Eh per unit of norm); CAS(8,8) (-1)' PC, norm 2.6e5,
    |dE| 1.9e-9 Eh; every class with a norm below 500 stays below 2e-10 Eh.
    """
    import numpy as np
    norm = np.abs(norm)
    return 1e-9 + np.where(norm > 500.0, 1e-11 * norm, 0.0)


def has_gpu():
    try:
        from nevpt2_amd import _lib
        return _lib.lib().nv2_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not has_gpu():
        pytest.fail("GPU test run without a visible GPU or without libnevpt2_amd.so")
    return True

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def energy_tol(norm):
    """Per-class bound on |E_gpu - E_oracle| (Eh) for the FP64 path.

    Both sides are FP64 but sum and diagonalise in different orders.  A class
    energy is sum_mu r_mu^2 / d_mu and its norm is sum_mu r_mu^2 / d_mu^2, so a
    round-off shift delta in the denominators moves the energy by about
    norm * delta.  Random CI vectors (not eigenfunctions of the Dyall
    Hamiltonian) give norms up to 1e6 (near-zero denominators, intruders), and
    denominators taken through a metric with eigenvalues near the linear-
    dependence threshold carry round-off of order 1e-12 Eh.  Classes with
    norm <= 1 are held to 1e-9 Eh + 2e-12 Eh.
    """
    import numpy as np
    return 1e-9 + 2e-12 * np.abs(norm)


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

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def h2mech():
    from openhurricane_amd.chemistry import Mechanism

    return Mechanism.h2air()


@pytest.fixture(scope="session")
def h2oracle(h2mech):
    from oracle import Oracle

    return Oracle(h2mech)


@pytest.fixture(scope="session")
def c1opts():
    """Integrator settings of the H2-air workloads: rtol 1e-6, atol 1e-10 for
    the Newton test and for the step-error test."""
    from openhurricane_amd import _lib

    return _lib.default_opts(atol=1e-10, rtol=1e-6, atol_inc=1e-10, rtol_inc=1e-6)

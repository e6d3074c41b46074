import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU run (oracle at full size)")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle backend (test infrastructure)."""
    import oracle
    return oracle.backend()


@pytest.fixture(scope="session")
def gpu():
    """The GPU library backend; fails loudly when the HIP library or the GPU is missing."""
    import vortexlab_amd as vl
    be = vl.backend()
    n = be.lib.vl_device_count()
    if n < 1:
        raise RuntimeError("no HIP device visible: GPU tests need an MI355X")
    return be

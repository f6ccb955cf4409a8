import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "tenet.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def ctx():
    """One device context for the whole GPU session (device 0)."""
    import tenet_hip

    c = tenet_hip.Context(0)
    yield c
    c.close()

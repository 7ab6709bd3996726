import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library's kernels)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hip():
    """The HIP product; a GPU test fails (never skips) when it cannot load."""
    import casphy_amd
    from casphy_amd import _lib
    _lib.load()
    _lib.check(_lib.load().casphy_set_device(0))
    return casphy_amd

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU-oracle runs")


@pytest.fixture(scope="session")
def fglib():
    """The HIP library, on a machine with a GPU (fails, never skips, when
    the library or the GPU is missing)."""
    from fluidgrid_amd import lib
    L = lib.load()
    n = lib.gpu_count()
    assert n >= 1, "GPU tests need a visible GPU"
    return lib


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle

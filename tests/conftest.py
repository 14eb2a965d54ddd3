import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP kernels")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def ctx():
    """One device context for every GPU test (a HIP failure here is an error,
    never a skip: the GPU tests must run the native kernels)."""
    from gatlas_amd.device import Context
    c = Context(0)
    print("\n[gpu] " + c.info(), flush=True)
    yield c
    c.close()

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU statistics test")


@pytest.fixture(scope="session")
def ctx():
    """One library context on GPU 0 for the whole GPU session."""
    import mcml_gpu_amd
    c = mcml_gpu_amd.Context(0)
    yield c
    c.close()

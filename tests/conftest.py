import os
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
# GEMM autotuning decisions of this test process go to a file of its own
# (a fresh, empty cache), not to the user's cache.
os.environ.setdefault("MPS2_TUNE_CACHE", os.path.join(tempfile.gettempdir(), f"mps2_tune_test_{os.getpid()}.txt"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def ctx():
    from mps2_amd import engine
    n = engine.device_count()
    if n < 1:
        pytest.fail("GPU test collected on a host without a visible GPU")
    c = engine.Context(0)
    yield c
    c.close()

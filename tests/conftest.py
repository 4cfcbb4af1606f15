import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def hip():
    """One HIP backend for the session (the GPU tests run in one process)."""
    import lettuce_simulate_physics_amd as L
    be, why = L.make_hip_backend(0)
    if be is None:
        pytest.fail(f"HIP backend unavailable on a GPU test run: {why}")
    yield be
    be.close()

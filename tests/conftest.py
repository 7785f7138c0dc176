import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP engine")
    config.addinivalue_line("markers", "slow: long-running")


def _has_gpu():
    try:
        from rewl_simulator_amd import engine
        return engine.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not _has_gpu():
        pytest.fail("gpu-marked test ran without a visible GPU or without the built HIP engine")
    return True

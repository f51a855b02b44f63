import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def hip_device():
    """Skip unless a HIP device and the built library are present."""
    from aspartik_amd import _lib
    try:
        n = _lib.device_count()
    except RuntimeError as e:  # library not built: a real failure on a GPU box
        pytest.fail(str(e))
    if n == 0:
        pytest.skip("no HIP device")
    return 0

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (run on the MI355X box)")


@pytest.fixture(scope="session")
def gpu():
    """The library with a device; GPU tests fail (never skip) without one."""
    from pskel_amd import _lib

    n = _lib.device_count()
    assert n >= 1, "no HIP device visible: GPU tests must run on a GPU machine"
    return _lib.lib()

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


def gpu_available() -> bool:
    try:
        from pamtra_amd import _lib
        return _lib.library().pam_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_lib():
    from pamtra_amd import _lib
    lib = _lib.library()
    assert lib.pam_device_count() > 0, "no HIP device visible"
    return lib

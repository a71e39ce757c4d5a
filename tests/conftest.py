import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests through the C ABI")


def gpu_available():
    try:
        from dxmclib_amd import backend
        dev = backend.Device(0)
        dev.close()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def device():
    from dxmclib_amd import backend
    dev = backend.Device(0)
    yield dev
    dev.close()

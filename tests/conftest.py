import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vexelray-sim-fluid_amd")
for path in (ROOT, PKG):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: a longer CPU case")


def gpu_available() -> bool:
    try:
        from vexhip import _native
        return _native.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly when a GPU test runs without the HIP library or a device."""
    from vexhip import _native
    n = _native.device_count()
    if n < 1:
        raise RuntimeError("no HIP device visible to libvexelhip.so")
    return n

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gpu-elmm_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    """A HIP device is visible: hipGetDeviceCount of the HIP runtime that the
    library links (no torch involved)."""
    import ctypes
    for name in ("libamdhip64.so", "libamdhip64.so.7", "libamdhip64.so.6", "/opt/rocm/lib/libamdhip64.so"):
        try:
            hip = ctypes.CDLL(name)
        except OSError:
            continue
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    return False


@pytest.fixture(scope="session")
def require_gpu():
    """Skips a GPU test without a device; with ELMM_REQUIRE_GPU=1 a missing
    device fails it instead."""
    if not gpu_available():
        if os.environ.get("ELMM_REQUIRE_GPU") == "1":
            pytest.fail("ELMM_REQUIRE_GPU=1 but no HIP device is visible")
        pytest.skip("no GPU")

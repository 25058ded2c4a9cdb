import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP engine)")


def gpu_available() -> bool:
    """HIP device count through the runtime library (no torch import)."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        try:
            hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        except OSError:
            return False
    n = ctypes.c_int(0)
    return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0


@pytest.fixture(scope="session")
def require_gpu():
    if os.environ.get("CHAMP_ENGINE_LIB", "").endswith("libchamp_emu.so"):
        return  # host emulation build of the kernels
    if not gpu_available():
        pytest.fail("GPU test collected on a machine without a GPU")

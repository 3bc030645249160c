import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


# KDN_EMULATION names the bounds-checked CPU build of the library
# (tests/cpu_build): the GPU tests then run their calls on the CPU under
# AddressSanitizer, with the bounds-checked oracle, to find out-of-bounds
# accesses of kernels and launches. Never set on a GPU.
EMULATION = os.environ.get("KDN_EMULATION")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    if EMULATION:
        import ctypes

        import kodanuki_amd as k
        from tests import oracle_py
        k._bind(ctypes.CDLL(os.path.abspath(EMULATION)))
        oracle_py.PATH = os.path.join(os.path.dirname(os.path.abspath(EMULATION)), "liboracle_asan.so")


def has_gpu():
    """A HIP device node is present (the tests open the device through the
    library itself: the process then runs on the HIP runtime the library
    was built against, with no second runtime beside it)."""
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def device():
    """One device context for the whole GPU session."""
    import kodanuki_amd as k
    if not (EMULATION or has_gpu()):
        pytest.skip("no GPU")
    dev = k.Device(0)
    yield dev
    dev.close()

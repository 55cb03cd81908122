import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu_available():
    from suntans_amd import native
    try:
        native.dropin_lib()
    except native.NativeLibraryMissing as e:
        pytest.fail(str(e))
    import ctypes
    L = ctypes.CDLL(os.path.join(native.LIBDIR, "libsunhip.so"))
    n = L.sunhip_device_count()
    if n < 1:
        pytest.fail("no HIP device visible; GPU tests need an MI355X")
    return n

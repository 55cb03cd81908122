"""The C-ABI libraries load without a GPU and export every function their
headers declare; without a device the ABI reports an error instead of
crashing, and the drop-ins fail loudly (no CPU fallback)."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

from suntans_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, re.M)
    return sorted(set(n for n in names if n not in ("if", "return")))


@pytest.mark.parametrize("header,lib", [("sunhip.h", "libsunhip.so"), ("suntans_hip.h", "libsuntans.so"),
                                        ("sunstate.h", "libsunstate.so")])
def test_exports_every_declared_symbol(header, lib):
    names = declared(header)
    assert len(names) > 5
    L = C.CDLL(os.path.join(native.LIBDIR, lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_signatures_present():
    L = native.dropin_lib()
    for n in ("CGSolveQ", "CGSolve", "OperatorQ", "Preconditioner", "OperatorH", "HorizontalFaceScalars",
              "UpdateScalars"):
        assert hasattr(L, n)


def test_abi_reports_missing_device_without_crashing():
    from suntans_amd import device
    if device.device_count() > 0:
        pytest.skip("a GPU is visible")
    from suntans_amd import cases
    st = cases.tiny(0)
    with pytest.raises(RuntimeError, match="device"):
        device.DeviceGrid(st)


def test_dropin_fails_loudly_without_device():
    from suntans_amd import device
    if device.device_count() > 0:
        pytest.skip("a GPU is visible")
    code = ("import sys; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
            "from suntans_amd import cases\nimport _impls\n"
            "st = cases.tiny(0)\n_impls.dropin().cgsolveq(st)\nprint('returned')\n") % (ROOT, os.path.join(ROOT, "tests"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "returned" not in r.stdout
    assert "CGSolveQ" in r.stderr and "device" in r.stderr.lower()

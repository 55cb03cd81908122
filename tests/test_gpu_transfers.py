"""Transfers between the reference's flat host layout and the device fields
(include/sunhip.h: sunhip_upload_part, sunhip_download_parts,
sunhip_download_wait, sunhip_pinned).

Copies run on the grid's own copy stream through two staging buffers taken in
turn, and two downloads in parts may be in flight.  Checked here, on the
device ABI directly:
* an upload in parts then a download in parts give back every value bit for
  bit, for cell, w-type and edge fields, with two downloads in flight and the
  waits taken part by part in order;
* a third download in flight, an upload, a whole download or an upload in
  parts with downloads in flight are refused with an error, and the state is
  usable again once the downloads are waited for;
* a field uploaded while the previous field's scatter may still run, and read
  by a kernel right after, holds the new values (the copy of field i+1 shares
  no staging buffer with the scatter of field i).
"""
import ctypes as C

import numpy as np
import pytest

from suntans_amd import cases, device, native

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def _abi():
    L = device.lib()
    L.sunhip_upload_part.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_int64, C.c_int64, C.c_int]
    L.sunhip_download_parts.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_int]
    L.sunhip_download_wait.argtypes = [C.c_void_p, C.c_int64]
    L.sunhip_pinned.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.POINTER(_dp)]
    return L


@pytest.fixture(scope="module")
def grid(gpu_available):
    st = cases.tiny(11, nx=23, ny=9, Nkmax=13)
    dg = device.DeviceGrid(st)
    yield st, dg
    dg.close()


def _lengths(dg):
    z = dg.sizes()
    return {"s": z["Nflat"], "wnew": z["NflatW"], "u": z["NflatE"]}


def _pinned(L, dg, slot, n):
    p = _dp()
    assert L.sunhip_pinned(dg.h, slot, n, C.byref(p)) == 0
    return np.ctypeslib.as_array(p, shape=(n,))


def _upload_parts(L, dg, name, buf, parts):
    n = len(buf)
    for k in range(parts):
        a, b = n * k // parts, n * (k + 1) // parts
        rc = L.sunhip_upload_part(dg.h, name.encode(), native.dptr(buf), a, b - a, int(k == parts - 1))
        assert rc == 0, device.lib().sunhip_error()


def test_parts_round_trip_two_in_flight(grid):
    st, dg = grid
    L = _abi()
    lens = _lengths(dg)
    rng = np.random.default_rng(5)
    vals = {f: rng.standard_normal(n) for f, n in lens.items()}
    for f, parts in (("s", 3), ("wnew", 1), ("u", 5)):
        _upload_parts(L, dg, f, np.ascontiguousarray(vals[f]), parts)
    m = max(lens.values())
    order = [("s", 4), ("u", 3), ("wnew", 2), ("s", 1)]
    slots = [_pinned(L, dg, 0, m), _pinned(L, dg, 1, m)]
    for s in slots:
        s[:] = np.nan
    # two in flight: issue 0 and 1, then wait 0 part by part, issue 2, ...
    issued = 0
    for i, (f, parts) in enumerate(order):
        while issued < len(order) and issued < i + 2:
            g, p = order[issued]
            slots[issued & 1][:] = np.nan
            assert L.sunhip_download_parts(dg.h, g.encode(), native.dptr(slots[issued & 1]), p) == 0
            issued += 1
        n = lens[f]
        for upto in (1, n // 3, n // 2, n + 7):
            assert L.sunhip_download_wait(dg.h, upto) == 0
            k = min(upto, n)
            assert np.array_equal(slots[i & 1][:k], vals[f][:k]), (f, upto)
    assert L.sunhip_download_wait(dg.h, 1 << 40) == 0  # nothing in flight: no-op


def test_refusals_with_downloads_in_flight(grid):
    st, dg = grid
    L = _abi()
    lens = _lengths(dg)
    a = np.ascontiguousarray(np.arange(lens["s"], dtype=np.float64))
    dg.upload("s", a)
    out = [np.zeros(lens["s"]), np.zeros(lens["s"]), np.zeros(lens["s"])]
    assert L.sunhip_download_parts(dg.h, b"s", native.dptr(out[0]), 2) == 0
    assert L.sunhip_download_parts(dg.h, b"s", native.dptr(out[1]), 2) == 0
    assert L.sunhip_download_parts(dg.h, b"s", native.dptr(out[2]), 2) != 0  # a third
    assert L.sunhip_upload(dg.h, b"s", native.dptr(a)) != 0
    assert L.sunhip_download(dg.h, b"s", native.dptr(out[2])) != 0
    assert L.sunhip_upload_part(dg.h, b"s", native.dptr(a), 0, 5, 0) != 0
    assert L.sunhip_download_wait(dg.h, 1 << 40) == 0
    assert np.array_equal(out[0], a)
    assert L.sunhip_download_wait(dg.h, 1 << 40) == 0
    assert np.array_equal(out[1], a)
    b = a[::-1].copy()
    dg.upload("s", b)
    assert np.array_equal(dg.download("s", lens["s"]), b)


def test_back_to_back_uploads_then_kernel(grid):
    """Ten uploads in a row alternate the staging buffers while scatters may
    still run; a device copy right after reads the last values of each."""
    st, dg = grid
    lens = _lengths(dg)
    rng = np.random.default_rng(9)
    last = {}
    for r in range(10):
        f = ("s", "q")[r & 1]
        v = rng.standard_normal(lens["s"])
        dg.upload(f, v)
        last[f] = v
    dg.copy("opx", "s")
    dg.copy("opy", "q")
    assert np.array_equal(dg.download("opx", lens["s"]), last["s"])
    assert np.array_equal(dg.download("opy", lens["s"]), last["q"])

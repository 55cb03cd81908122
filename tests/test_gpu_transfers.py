"""Transfers between the reference's flat host layout and the device fields
(include/sunhip.h: sunhip_upload_part, sunhip_download_parts,
sunhip_download_wait, sunhip_pinned).

Copies run on the grid's own copy streams, one per direction, through
staging buffers taken in turn (two per direction), and two downloads in parts
may be in flight.  Checked here, on the device ABI directly:
* an upload in parts then a download in parts give back every value bit for
  bit, for cell, w-type and edge fields, with two downloads in flight and the
  waits taken part by part in order;
* a third download in flight and a whole download with downloads in flight
  are refused with an error, and the state is usable again once the
  downloads are waited for; uploads with downloads in flight go through and
  leave the downloads intact;
* a field uploaded while the previous field's scatter may still run, and read
  by a kernel right after, holds the new values (the copy of field i+1 shares
  no staging buffer with the scatter of field i);
* the scalar step in two calls (sunhip_update_scalars_faces, then the column
  update's inputs uploaded and the face values downloaded in flight, then
  sunhip_update_scalars_column) gives sunhip_update_scalars bit for bit.
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
    assert L.sunhip_download(dg.h, b"s", native.dptr(out[2])) != 0
    c = np.ascontiguousarray(-a)
    dg.upload("q", c)  # another field, uploaded with both downloads in flight
    assert L.sunhip_download_wait(dg.h, 1 << 40) == 0
    assert np.array_equal(out[0], a)
    assert L.sunhip_download_wait(dg.h, 1 << 40) == 0
    assert np.array_equal(out[1], a)
    assert np.array_equal(dg.download("q", lens["s"]), c)
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


@pytest.mark.parametrize("tvd", [1, 4])
def test_scalar_step_in_two_calls(gpu_available, tvd):
    st = cases.lock_exchange(seed=7)
    cases.timestep_inputs(st, np.random.default_rng(31), nonlinear=2)
    dg = device.DeviceGrid(st)
    L = _abi()
    L.sunhip_update_scalars_faces.argtypes = [C.c_void_p, C.POINTER(device.ScalarParams), C.c_char_p]
    L.sunhip_update_scalars_column.argtypes = [C.c_void_p, C.POINTER(device.ScalarParams), C.c_char_p]
    th, dt, ka = st.get_prop("theta"), st.get_prop("dt"), st.get_prop("kappa_s")
    column_inputs = {"wnew": st.get("wnew"), "wtmp2": st.get("wtmp2"), "dzz": st.get("dzz"),
                     "dzzold": st.get("dzzold"), "kappa_tv": st.get("kappa_tv")}

    def inputs():
        for f, v in (("s", st.get("s")), ("u", st.get("u")), ("utmp2", st.get("utmp2")), ("dzf", st.get("dzf"))):
            dg.upload(f, v)

    inputs()
    for f, v in column_inputs.items():
        dg.upload(f, v)
    dg.upload("active", st.active().astype(np.float64))
    dg.update_scalars(th, dt, ka, tvd, step=3)
    z = dg.sizes()
    want = {f: dg.download(f, z["NflatE" if f.startswith("Sf") else "Nflat"]) for f in ("s", "Cn", "SfHp", "SfHm")}
    # again in two calls, the column inputs uploaded (with other values first,
    # then the right ones) while the face values download
    for f, v in column_inputs.items():
        dg.upload(f, np.zeros_like(v))
    inputs()
    S = device.ScalarParams(th, dt, ka, tvd, 3, 0, 0)
    assert L.sunhip_update_scalars_faces(dg.h, C.byref(S), b"s") == 0
    out = [np.zeros(z["NflatE"]), np.zeros(z["NflatE"])]
    assert L.sunhip_download_parts(dg.h, b"SfHp", native.dptr(out[0]), 3) == 0
    assert L.sunhip_download_parts(dg.h, b"SfHm", native.dptr(out[1]), 2) == 0
    for f, v in column_inputs.items():
        dg.upload(f, v)
    assert L.sunhip_update_scalars_column(dg.h, C.byref(S), b"s") == 0
    assert L.sunhip_download_wait(dg.h, 1 << 40) == 0
    assert L.sunhip_download_wait(dg.h, 1 << 40) == 0
    assert np.array_equal(out[0], want["SfHp"]) and np.array_equal(out[1], want["SfHm"])
    assert np.array_equal(dg.download("s", z["Nflat"]), want["s"])
    assert np.array_equal(dg.download("Cn", z["Nflat"]), want["Cn"])
    S0 = device.ScalarParams(th, dt, ka, 0, 3, 0, 0)
    assert L.sunhip_update_scalars_faces(dg.h, C.byref(S0), b"s") != 0  # TVD schemes only
    dg.close()

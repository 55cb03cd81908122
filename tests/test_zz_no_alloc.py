"""The allocation-free free-surface solve (sunhip_h_solve_noalloc, through
the C ABI of libsunhip.so) against sunhip_h_solve, which it must reproduce
bit for bit with the same iteration count: two calls in a row on one device,
the second on a smaller grid (the shared static scratch keeps the first
call's contents), both preconditioners, the device grid memory unchanged by
the calls; and a grid beyond the scratch maximum is refused before any
kernel runs."""
import numpy as np
import pytest

from suntans_amd import cases, device

pytestmark = pytest.mark.gpu


def _grid(st):
    st.set("D", cases.depth_integrated_D(st))
    dg = device.DeviceGrid(st)
    for n in ("h", "htmp", "D"):
        dg.upload(n, st.get(n))
    return dg


def _solve(dg, st, fn, hprecond):
    """The solve from h, its device memory right after the call (before the
    download, whose staging is the transfer's, not the solve's) and h."""
    dg.upload("h", st.get("h"))
    r = getattr(dg, fn)(st.get_prop("grav"), st.get_prop("theta"), st.get_prop("dt"), maxiters=1000,
                        precond=hprecond, epsilon=1e-10)
    r["bytes_after"] = dg.bytes()
    return r, dg.download("h", st.Nc)


@pytest.mark.parametrize("hprecond", [1, 0])
def test_noalloc_matches_h_solve(gpu_available, hprecond):
    big = cases.ensemble_member(41, 150000, 1.5, 30)
    small = cases.tiny(11, nx=23, ny=9, Nkmax=13)
    out = []
    for st in (big, small):  # the second call is smaller
        dg = _grid(st)
        before = dg.bytes()
        rn, hn = _solve(dg, st, "h_solve_noalloc", hprecond)
        assert rn["bytes_after"] == before  # no device memory taken by the call
        rr, hr = _solve(dg, st, "h_solve", hprecond)
        dg.close()
        print(f"Nc={st.Nc} hprecond={hprecond}: noalloc {rn['niters']} iterations, resid {rn['resid']:.3e}; "
              f"h_solve {rr['niters']}")
        assert rn["converged"] == 1
        assert rn["niters"] == rr["niters"]
        assert np.array_equal(hn, hr)
        out.append(hn)
    # again on the big grid after the small one: the result repeats
    dg = _grid(big)
    rn, hn = _solve(dg, big, "h_solve_noalloc", hprecond)
    dg.close()
    assert np.array_equal(hn, out[0])


def test_noalloc_refuses_beyond_maximum(gpu_available):
    st = cases.tiny(3, nx=500, ny=480, Nkmax=6, min_layers=1)  # 480,000 triangles: 34.6 MB of scratch
    dg = _grid(st)
    with pytest.raises(RuntimeError, match="maximum"):
        dg.h_solve_noalloc(st.get_prop("grav"), st.get_prop("theta"), st.get_prop("dt"))
    dg.close()


def test_noalloc_needs_resident_inputs(gpu_available):
    st = cases.tiny(2)
    dg = device.DeviceGrid(st)
    with pytest.raises(RuntimeError, match="upload"):
        dg.h_solve_noalloc(st.get_prop("grav"), st.get_prop("theta"), st.get_prop("dt"))
    dg.close()

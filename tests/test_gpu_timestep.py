"""One nonhydrostatic time step with every field resident on the GPU, in
the reference's order (Solve, main/phys.c:1145-1306): SetFluxHeight,
UPredictor (its open-boundary hooks do nothing on these closed grids),
UpdateDZ, Continuity into wnew, UpdateScalars (salinity), WPredictor,
ComputeQSource, CGSolveQ, Corrector and Continuity, against the same
sequence of oracle calls.  The free-surface and pressure solves sum their
dot products in another order, so the fields agree to 1e-9 (relative
2-norm over the wet layers) with the solves' iteration counts within +-1,
and the discrete state (ctop, etop, wet and dry cells) is identical."""
import numpy as np
import pytest

from suntans_amd import cases, device

import _grids
import _impls
from _trajectory import device_step, upload_state

pytestmark = pytest.mark.gpu

MAKERS = [
    ("lockexchange", lambda: cases.lock_exchange(seed=7)),
    ("tiny_wide", lambda: cases.tiny(11, nx=23, ny=9, Nkmax=13)),
    ("hexquad7", lambda: cases.tiny(24, nx=14, ny=9, Nkmax=11, planform="hex", split_every=4, maxfaces=7)),
] + _grids.REAL
IDS = [m[0] for m in MAKERS]


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def oracle_step(o, a):
    o.set_flux_height(a)
    assert o.upredictor(a) == 0
    o.update_dz(a, 0)
    o.continuity(a, w="wnew")
    o.update_salinity(a)
    o.wpredictor(a)
    o.compute_qsource(a)
    o.cgsolveq(a)
    o.corrector(a, qc="qc")
    o.continuity(a, w="w")


@pytest.mark.parametrize("nonlinear", [1, 2])
@pytest.mark.parametrize("name,make", MAKERS, ids=IDS)
def test_resident_time_step(gpu_available, name, make, nonlinear):
    o = _impls.oracle()
    a = make()
    cases.timestep_inputs(a, np.random.default_rng(29), nonlinear=nonlinear)
    dg = device.DeviceGrid(a)
    upload_state(dg, a)
    sh, sq = device_step(dg, a)
    oracle_step(o, a)
    assert abs(sh["niters"] - o.stats("h")[0]) <= 1
    assert abs(sq["niters"] - o.stats("q")[0]) <= 1
    ext = dg.extents()
    for f in ("ctop", "ctopold", "etop", "etopold"):
        assert np.array_equal(ext[f], a.ints(f)), f
    assert np.array_equal(dg.download("active", a.Nc) != 0, a.active() != 0)
    m = a.wet_mask()
    assert relerr(dg.download("h", a.Nc), a.get("h")) <= 1e-9
    assert np.array_equal(dg.download("dzz", len(m)), a.get("dzz")) or \
        relerr(dg.download("dzz", len(m)), a.get("dzz")) <= 1e-9
    for dev_name, st_name, wet in (("u", "u", a.edge_wet_mask()), ("w", "w", None), ("s", "s", m),
                                   ("q", "qc", m), ("qp", "q", m), ("Cn_W", "Cn_W", None), ("wnew", "wnew", None)):
        want = a.get(st_name)
        got = dg.download(dev_name, len(want))
        if wet is not None:
            want, got = want[wet], got[wet]
        assert np.isfinite(got).all(), dev_name
        assert relerr(got, want) <= 1e-9, (dev_name, relerr(got, want))
    dg.close()


@pytest.mark.parametrize("name,make", [MAKERS[0]] + _grids.REAL[:1], ids=[IDS[0], _grids.REAL[0][0]])
def test_resident_time_steps(gpu_available, name, make):
    """Three consecutive steps (Adams-Bashforth steps 3, 4, 5) with the state
    left on the device between them, against three oracle steps: the
    solves' rounding differences stay within 1e-9 and never flip the
    discrete state."""
    o = _impls.oracle()
    a = make()
    cases.timestep_inputs(a, np.random.default_rng(31), nonlinear=2)
    source = a.get("utmp")
    dg = device.DeviceGrid(a)
    upload_state(dg, a)
    for step in (3, 4, 5):
        a.set_prop(n=step)
        a.set("utmp", source)
        dg.upload("utmp", source)
        sh, sq = device_step(dg, a)
        oracle_step(o, a)
        assert abs(sh["niters"] - o.stats("h")[0]) <= 1 and abs(sq["niters"] - o.stats("q")[0]) <= 1, step
        ext = dg.extents()
        for f in ("ctop", "etop"):
            assert np.array_equal(ext[f], a.ints(f)), (step, f)
    m = a.wet_mask()
    for dev_name, st_name, wet in (("u", "u", a.edge_wet_mask()), ("w", "w", None), ("s", "s", m),
                                   ("qp", "q", m), ("h", "h", None)):
        want = a.get(st_name)
        got = dg.download(dev_name, len(want))
        if wet is not None:
            want, got = want[wet], got[wet]
        assert relerr(got, want) <= 1e-9, (dev_name, relerr(got, want))
    dg.close()

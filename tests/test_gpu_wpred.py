"""WPredictor (main/phys.c:2025-2290) on the GPU against the oracle, bit for
bit: through the drop-in sunhip_WPredictor (the reference's signature) on
closed and open grids for schemes 0, 1, 2 and 5, the first, second and
later steps, and through the device API on a grid kept resident for two
steps.  The refused configurations (scheme 4, conserveMomentum, type-4
edges) return an error from the device API."""
import numpy as np
import pytest

from suntans_amd import cases, device

import _grids
import _impls

pytestmark = pytest.mark.gpu

OPEN = cases.open_ends()
MAKERS = [
    ("tiny0", lambda: cases.tiny(0)),
    ("tiny_wide", lambda: cases.tiny(11, nx=23, ny=9, Nkmax=13)),
    ("lockexchange", lambda: cases.lock_exchange(seed=7)),
    ("shallow40", lambda: cases.tiny(40, nx=12, ny=8, Nkmax=6, min_layers=0.0)),
    ("quadtri", lambda: cases.tiny(22, nx=12, ny=10, Nkmax=9, planform="quad", split_every=3)),
    ("hexquad7", lambda: cases.tiny(24, nx=14, ny=9, Nkmax=11, planform="hex", split_every=4, maxfaces=7)),
    ("open_tri", lambda: cases.tiny(3, nx=16, ny=9, Nkmax=10, open_marks=OPEN)),
    ("open_hex", lambda: cases.tiny(25, nx=12, ny=9, Nkmax=9, planform="hex", split_every=4, maxfaces=7,
                                    open_marks=OPEN)),
] + _grids.REAL
IDS = [m[0] for m in MAKERS]
VARIANTS = {"central": dict(nonlinear=2), "upwind": dict(nonlinear=1), "none": dict(nonlinear=0),
            "tvd": dict(nonlinear=5, tvd=4), "tvd1": dict(nonlinear=5, tvd=1),
            "stress": dict(nonlinear=2, stress=True), "first": dict(step=1), "second": dict(step=2),
            "ab2": dict(AB=2, nonlinear=1), "inviscid": dict(CdW=0.0, nu_H=0.0)}
FIELDS = ("w", "wtmp", "wc", "stmp", "Cn_W", "Cn_W2", "ut")
TVD_FIELDS = ("SfHp", "SfHm", "gradSx", "gradSy")


def same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("name,make", MAKERS, ids=IDS)
def test_wpredictor_bitwise(gpu_available, name, make, variant):
    o, d = _impls.oracle(), _impls.dropin()
    a = make()
    cases.wpredictor_inputs(a, np.random.default_rng(17), **VARIANTS[variant])
    b = a.clone()
    o.wpredictor(a)
    d.wpredictor(b)
    assert np.isfinite(a.get("w")).all()
    for f in FIELDS + (TVD_FIELDS if VARIANTS[variant].get("nonlinear") == 5 else ()):
        assert same(a.get(f), b.get(f)), f


def _device(st):
    dg = device.DeviceGrid(st)
    for f in ("w", "wtmp", "Cn_W", "Cn_W2", "dzz", "nu_tv", "nu_lax", "wc", "utmp2"):
        dg.upload(f, st.get(f))
    dg.upload("qp", st.get("q"))
    dg.upload("ue", st.get("ut"))
    dg.set_face_distances(device.face_distances(st))
    return dg


@pytest.mark.parametrize("name,make", [MAKERS[2], MAKERS[5]], ids=[IDS[2], IDS[5]])
def test_wpredictor_resident_two_steps(gpu_available, name, make):
    """Two consecutive steps with the state left on the device (Cn_W and
    Cn_W2 carried over), against two oracle steps."""
    o = _impls.oracle()
    a = make()
    cases.wpredictor_inputs(a, np.random.default_rng(23), nonlinear=2)
    dg = _device(a)
    for step in (2, 3):
        a.set_prop(n=step)
        o.wpredictor(a)
        dg.w_predict(device.WPredParams.of(a))
    for f in ("w", "wtmp", "wc", "stmp", "Cn_W", "Cn_W2"):
        assert same(dg.download(f, len(a.get(f))), a.get(f)), f
    assert same(dg.download("ue", len(a.get("ut"))), a.get("ut"))
    dg.close()


def test_wpredictor_refusals(gpu_available):
    st = cases.tiny(1)
    cases.wpredictor_inputs(st, np.random.default_rng(1))
    dg = _device(st)
    for kw in (dict(nonlinear=4), dict(conserveMomentum=1)):
        P = device.WPredParams.of(st)
        for k, v in kw.items():
            setattr(P, k, v)
        with pytest.raises(RuntimeError):
            dg.w_predict(P)
    dg.close()
    st4 = cases.lock_exchange(seed=8, nx=20, ny=5, Nkmax=10, open_marks=cases.open_ends(4, 2))
    cases.wpredictor_inputs(st4, np.random.default_rng(1))
    dg = _device(st4)
    with pytest.raises(RuntimeError, match="type-4"):
        dg.w_predict(device.WPredParams.of(st4))
    dg.close()

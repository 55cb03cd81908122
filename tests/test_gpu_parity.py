"""HIP path (through the reference-signature drop-ins in libsuntans.so) against
the oracle, on the same seeded inputs.

Tolerances:
* OperatorQ, Preconditioner, OperatorH, HorizontalFaceScalars, UpdateScalars:
  bit-identical (element-wise arithmetic keeps the reference's operation order
  and is built without FMA contraction);
* CGSolveQ / CGSolve converged: relative 2-norm difference <= 1e-10 and the
  reference's iteration counter within +-1 (the device sums the dot products
  in a different order, so iterates differ by rounding);
* a fixed number of PCG iterations: relative 2-norm difference <= 1e-12.
"""
import numpy as np
import pytest

from suntans_amd import cases

import _grids
import _impls

pytestmark = pytest.mark.gpu

SEEDS = [0, 1, 2, 3]


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _pair(make):
    a = make()
    return a, a.clone()


def _lock():
    return cases.lock_exchange(seed=7)


def _makers():
    out = [(f"tiny{s}", (lambda s=s: cases.tiny(s))) for s in SEEDS]
    out.append(("tiny_wide", lambda: cases.tiny(11, nx=23, ny=9, Nkmax=13)))
    out.append(("lockexchange", _lock))
    out += POLYGONS
    out += _grids.REAL  # the reference's own example grids, with open boundaries
    return out


# quadrilateral, hybrid (quads + triangles, hexagons + quads) and hexagonal
# planforms: the 4-, 6- and 8-face kernel instantiations
POLYGONS = [
    ("quad", lambda: cases.tiny(21, nx=13, ny=11, Nkmax=9, planform="quad")),
    ("quadtri", lambda: cases.tiny(22, nx=12, ny=10, Nkmax=9, planform="quad", split_every=3)),
    ("hex", lambda: cases.tiny(23, nx=12, ny=10, Nkmax=9, planform="hex")),
    ("hexquad7", lambda: cases.tiny(24, nx=14, ny=9, Nkmax=11, planform="hex", split_every=4, maxfaces=7)),
]


MAKERS = _makers()


@pytest.mark.parametrize("name,make", MAKERS, ids=[m[0] for m in MAKERS])
def test_operator_q_bitwise(gpu_available, name, make):
    o, d = _impls.oracle(), _impls.dropin()
    a, b = _pair(make)
    rng = np.random.default_rng(5)
    x = rng.normal(size=a.get("stmp").shape)
    for st in (a, b):
        o.qcoef(st)               # coefficients into wtmp (same bits for both)
        st.set("stmp", x)
        st.set("stmp2", 0.0)
    o.operatorq(a)
    d.operatorq(b)
    m = a.wet_mask()
    ya, yb = a.get("stmp2")[m], b.get("stmp2")[m]
    assert np.array_equal(ya, yb), f"max diff {np.abs(ya - yb).max()}"


@pytest.mark.parametrize("name,make", MAKERS, ids=[m[0] for m in MAKERS])
def test_preconditioner_bitwise(gpu_available, name, make):
    o, d = _impls.oracle(), _impls.dropin()
    a, b = _pair(make)
    rng = np.random.default_rng(6)
    x = rng.normal(size=a.get("stmp").shape)
    for st in (a, b):
        o.qcoef(st)
        st.set("stmp", x)
        st.set("uold", 0.0)
    o.preconditioner(a)
    d.preconditioner(b)
    m = a.wet_mask()
    ua, ub = a.get("uold")[m], b.get("uold")[m]
    assert np.array_equal(ua, ub), f"max diff {np.abs(ua - ub).max()}"


@pytest.mark.parametrize("name,make", MAKERS, ids=[m[0] for m in MAKERS])
def test_operator_h_bitwise(gpu_available, name, make):
    o, d = _impls.oracle(), _impls.dropin()
    a, b = _pair(make)
    rng = np.random.default_rng(8)
    x = rng.normal(size=a.Nc)
    for st in (a, b):
        st.set("D", cases.depth_integrated_D(st))
        o.hcoef(st)
        st.set("htmp2", x)
        st.set("htmp3", 0.0)
    o.operatorh(a)
    d.operatorh(b)
    assert np.array_equal(a.get("htmp3"), b.get("htmp3"))


@pytest.mark.parametrize("name,make", MAKERS, ids=[m[0] for m in MAKERS])
@pytest.mark.parametrize("qprecond", [2, 0])
def test_cgsolveq_converged(gpu_available, name, make, qprecond):
    if name == "iwaves" and qprecond == 0:
        pytest.skip("unpreconditioned CG does not converge in qmaxiters on the 1 m deep, 2 km wide tank cells")
    o, d = _impls.oracle(), _impls.dropin()
    a, b = _pair(make)
    for st in (a, b):
        st.set_prop(qprecond=qprecond, qepsilon=1e-10, qmaxiters=2000, resnorm=0)
    o.cgsolveq(a)
    d.cgsolveq(b)
    no, _ = o.stats("q")
    nd, resid = d.stats("q")
    m = a.wet_mask()
    err = relerr(b.get("qc")[m], a.get("qc")[m])
    print(f"{name} qprecond={qprecond}: iters oracle={no} device={nd} rel err={err:.3e}")
    if qprecond == 2:
        assert abs(no - nd) <= 1
    else:
        # Unpreconditioned CG needs about as many iterations as unknowns here;
        # in that regime the count is decided by rounding (loss of
        # orthogonality), so only a 5% band is meaningful.
        assert abs(no - nd) <= max(1, 0.05 * no)
    assert err <= 1e-10


FIXED = [m for m in MAKERS if m[0] in ("tiny_wide", "lockexchange", "quadtri", "hexquad7")]


@pytest.mark.parametrize("name,make", FIXED, ids=[m[0] for m in FIXED])
def test_cgsolveq_fixed_iterations(gpu_available, name, make):
    o, d = _impls.oracle(), _impls.dropin()
    a, b = _pair(make)
    for st in (a, b):
        st.set_prop(qprecond=2, qepsilon=0.0, qmaxiters=6)
    o.cgsolveq(a)
    d.cgsolveq(b)
    m = a.wet_mask()
    err = relerr(b.get("qc")[m], a.get("qc")[m])
    assert o.stats("q")[0] == d.stats("q")[0] == 6
    assert err <= 1e-12, err


@pytest.mark.parametrize("name,make", MAKERS, ids=[m[0] for m in MAKERS])
@pytest.mark.parametrize("hprecond", [1, 0])
def test_cgsolve_converged(gpu_available, name, make, hprecond):
    o, d = _impls.oracle(), _impls.dropin()
    a, b = _pair(make)
    for st in (a, b):
        st.set("D", cases.depth_integrated_D(st))
        st.set_prop(hprecond=hprecond, epsilon=1e-10, maxiters=1000)
    o.cgsolve(a)
    d.cgsolve(b)
    no, _ = o.stats("h")
    nd, _ = d.stats("h")
    err = relerr(b.get("h"), a.get("h"))
    print(f"{name} hprecond={hprecond}: iters oracle={no} device={nd} rel err={err:.3e}")
    assert abs(no - nd) <= 1
    assert err <= 1e-10


@pytest.mark.parametrize("name,make", MAKERS, ids=[m[0] for m in MAKERS])
@pytest.mark.parametrize("tvd", [1, 2, 3, 4])
def test_face_scalars_bitwise(gpu_available, name, make, tvd):
    o, d = _impls.oracle(), _impls.dropin()
    a, b = _pair(make)
    o.face_scalars(a, tvd)
    d.face_scalars(b, tvd)
    em = a.edge_wet_mask()
    for f in ("SfHp", "SfHm"):
        assert np.array_equal(a.get(f)[em], b.get(f)[em]), f
    m = a.wet_mask()
    for f in ("gradSx", "gradSy"):
        assert np.array_equal(a.get(f)[m], b.get(f)[m]), f


@pytest.mark.parametrize("name,make", MAKERS, ids=[m[0] for m in MAKERS])
@pytest.mark.parametrize("tvd", [0, 1, 3, 4])
def test_update_scalars_bitwise(gpu_available, name, make, tvd):
    o, d = _impls.oracle(), _impls.dropin()
    a, b = _pair(make)
    for st in (a, b):
        st.set_prop(TVDsalt=tvd, n=2)
    o.update_salinity(a)
    d.update_salinity(b)
    sa, sb = a.get("s"), b.get("s")
    assert np.array_equal(sa, sb), f"max diff {np.abs(sa - sb).max()}"
    assert np.array_equal(a.get("Cn_R"), b.get("Cn_R"))


SRC_GRIDS = [("tiny0", lambda: cases.tiny(0)), ("tiny_wide", lambda: cases.tiny(11, nx=23, ny=9, Nkmax=13)),
             ("lockexchange", _lock),
             ("hexquad7", lambda: cases.tiny(24, nx=14, ny=9, Nkmax=11, planform="hex", split_every=4, maxfaces=7))]


@pytest.mark.parametrize("name,make", SRC_GRIDS, ids=[g[0] for g in SRC_GRIDS])
@pytest.mark.parametrize("flags", [1, 2, 3, 12, 15])
@pytest.mark.parametrize("tvd", [0, 3, 4])
@pytest.mark.parametrize("rose", [False, True], ids=["ctop_kept", "surface_rose"])
def test_update_scalars_sources_bitwise(gpu_available, name, make, flags, tvd, rose):
    """The drop-in UpdateScalars with its source terms (src1 = uold, src2 =
    wtmp, Ftop = htmp, Fbot = htmp2, alpha_top = alpha_bot = 1) bit for bit
    against the oracle.  ctop_kept: ctopold lowered to ctop, where the
    oracle is pinned to the reference (test_oracle_vs_ref.py); surface_rose:
    the state as built, with columns whose surface rose, where both follow
    the choice stated in include/suntans_hip.h."""
    o, d = _impls.oracle(), _impls.dropin()
    a = make()
    rng = np.random.default_rng(41)
    a.set("uold", rng.normal(0, 1e-4, a.get("uold").shape))
    a.set("wtmp", rng.normal(0, 1e-3, a.get("wtmp").shape))
    a.set("htmp", rng.normal(0, 1e-2, a.Nc))
    a.set("htmp2", rng.normal(0, 1e-2, a.Nc))
    if not rose:
        a.set_ints("ctopold", np.minimum(a.ints("ctopold"), a.ints("ctop")))
    elif not (a.ints("ctopold") > a.ints("ctop")).any():
        pytest.skip("no column of this state has a risen surface")
    b = a.clone()
    for st in (a, b):
        st.set_prop(TVDsalt=tvd, n=2)
    o.update_salinity_src(a, flags, 1, 1)
    d.update_salinity_src(b, flags, 1, 1)
    sa, sb = a.get("s"), b.get("s")
    assert np.array_equal(sa, sb), f"max diff {np.abs(sa - sb).max()}"
    assert np.array_equal(a.get("Cn_R"), b.get("Cn_R"))


OPEN_SCAL = [("open_tri", lambda: cases.tiny(3, nx=8, ny=5, open_marks=cases.open_ends())),
             ("open_hex", lambda: cases.tiny(25, nx=7, ny=6, planform="hex", split_every=4, maxfaces=7,
                                             open_marks=cases.open_ends())),
             ("open_lock", lambda: cases.lock_exchange(seed=8, nx=20, ny=5, Nkmax=10, open_marks=cases.open_ends()))]


@pytest.mark.parametrize("name,make", OPEN_SCAL, ids=[m[0] for m in OPEN_SCAL])
@pytest.mark.parametrize("tvd", [0, 1, 3, 4])
def test_update_scalars_open_bitwise(gpu_available, name, make, tvd):
    """The drop-in UpdateScalars on grids with open boundaries and seeded
    boundary values, first-order upwind included, bit for bit against the
    oracle (pinned to the reference in test_oracle_vs_ref.py)."""
    o, d = _impls.oracle(), _impls.dropin()
    a = make()
    a.set("boundary_s", np.random.default_rng(43).uniform(0, 1, a.get("boundary_s").shape))
    b = a.clone()
    for st in (a, b):
        st.set_prop(TVDsalt=tvd, n=2)
    o.update_salinity(a)
    d.update_salinity(b)
    sa, sb = a.get("s"), b.get("s")
    assert np.array_equal(sa, sb), f"max diff {np.abs(sa - sb).max()}"
    assert np.array_equal(a.get("Cn_R"), b.get("Cn_R"))

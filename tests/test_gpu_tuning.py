"""Every kernel tuning variant (sunhip_set_tuning) solves the same problem.
Unroll, store-hint, register-budget, register-head, block-staging and z-level
window (zreg: coefficients computed from the layer thicknesses where they
equal the stored ones) variants are bit-identical to the default configuration (same operations in the same
order, same reduction order).  Variants that regroup a dot product only
change its rounding: tiles per block (block partial sums of 8 or 16 tiles)
rzid (r.M^-1 r as sum d'^2/b') and band (partial sums per level band); for
those the iteration count is within +-1 and q within 1e-12."""
import numpy as np
import pytest

from suntans_amd import cases, device

pytestmark = pytest.mark.gpu

BITWISE = [("xchg", 4), ("xchg", 5), ("xchg", 6), ("unroll", 1), ("unroll", 2), ("unroll", 8), ("prec_unroll", 1),
           ("prec_unroll", 4), ("ntstore", 0),
           ("minwaves", 3), ("minwaves", 4), ("blocklds", 1), ("prec_head", 0), ("prec_head", 8),
           ("prec_head", 16), ("prec_head", 24), ("prec_head", 32), ("xchg", 1), ("xchg", 2), ("xchg", 3),
           ("zreg", 0)]
ROUNDING = [("tiles", 8), ("tiles", 16), ("rzid", 0), ("band", 8), ("band", 16), ("band", 32), ("order", 0)]


def solve(st, tuning):
    dg = device.DeviceGrid(st)
    device.member_upload(dg, st)
    for k, v in tuning:
        dg.tune(k, v)
    s = dg.q_solve(epsilon=1e-10)
    q = dg.download("q", dg.sizes()["Nflat"])
    dg.close()
    return q, s["niters"]


@pytest.fixture(scope="module")
def grids():
    return [cases.tiny(11, nx=23, ny=9, Nkmax=13), cases.lock_exchange(seed=7),
            cases.ensemble_member(6, 9000, 1.2, 40),
            cases.tiny(24, nx=14, ny=9, Nkmax=40, planform="hex", split_every=4, maxfaces=7)]


@pytest.mark.parametrize("key,value", BITWISE, ids=[f"{k}={v}" for k, v in BITWISE])
def test_variant_bitwise(gpu_available, grids, key, value):
    for st in grids:
        q0, n0 = solve(st, [])
        q1, n1 = solve(st, [(key, value)])
        assert n0 == n1
        assert np.array_equal(q0, q1)


@pytest.mark.parametrize("head", [16, 24])
@pytest.mark.parametrize("rzid", [0, 1])
def test_prec_lds_bitwise(gpu_available, grids, head, rzid):
    """The preconditioner's 12 levels below its register head kept in LDS
    (columns of 13 to 50 levels: shorter than the head, ending inside the
    LDS stage and below it) give the same bits, in the FP64 solve and in the
    mixed form's FP32 inner solve."""
    for st in grids:
        base = [("prec_head", head), ("rzid", rzid)]
        q0, n0 = solve(st, base + [("prec_lds", 0)])
        q1, n1 = solve(st, base + [("prec_lds", 12)])
        assert n0 == n1
        assert np.array_equal(q0, q1)
        a = solve_mixed(st, base + [("prec_lds", 0)])
        b = solve_mixed(st, base + [("prec_lds", 12)])
        assert a[1:] == b[1:]
        assert np.array_equal(a[0], b[0])


@pytest.mark.parametrize("key,value", ROUNDING, ids=[f"{k}={v}" for k, v in ROUNDING])
def test_variant_rounding(gpu_available, grids, key, value):
    for st in grids:
        q0, n0 = solve(st, [])
        q1, n1 = solve(st, [(key, value)])
        assert abs(n0 - n1) <= 1
        assert np.linalg.norm(q1 - q0) <= 1e-12 * np.linalg.norm(q0)


def solve_mixed(st, tuning):
    dg = device.DeviceGrid(st)
    device.member_upload(dg, st)
    for k, v in tuning:
        dg.tune(k, v)
    s = dg.q_solve_mixed(epsilon=1e-10, inner_tol=1e-4)
    q = dg.download("q", dg.sizes()["Nflat"])
    dg.close()
    return q, s["niters"], s["inner_iters"]


@pytest.mark.parametrize("xchg", [0, 2])
def test_zreg_bitwise_mixed(gpu_available, grids, xchg):
    """The z-level windows in the FP32 inner kernels and the FP64 residual."""
    for st in grids:
        a = solve_mixed(st, [("xchg", xchg), ("zreg", 0)])
        b = solve_mixed(st, [("xchg", xchg), ("zreg", 1)])
        assert a[1:] == b[1:]
        assert np.array_equal(a[0], b[0])


@pytest.mark.parametrize("head", [0, 16, 24])
def test_zreg_bitwise_heads(gpu_available, grids, head):
    """The windows with every preconditioner register head, exchange kernel on."""
    for st in grids:
        q0, n0 = solve(st, [("xchg", 2), ("prec_head", head), ("zreg", 0)])
        q1, n1 = solve(st, [("xchg", 2), ("prec_head", head), ("zreg", 1)])
        assert n0 == n1
        assert np.array_equal(q0, q1)


def scalar_step(st, group, tvd, head=0):
    """One UpdateScalars on the device with tvd_group and tvd_sums_group =
    group and tvd_head = head; returns the new scalar and Cn."""
    dg = device.DeviceGrid(st)
    for f in ("s", "u", "utmp2", "dzf", "dzz", "dzzold", "kappa_tv", "wnew", "wtmp2"):
        dg.upload(f, st.get(f))
    dg.upload("active", st.active().astype(np.float64))
    dg.tune("tvd_group", group)
    dg.tune("tvd_sums_group", group)
    dg.tune("tvd_head", head)
    dg.update_scalars(st.get_prop("theta"), st.get_prop("dt"), 1e-4, tvd, step=3)
    n = len(st.get("s"))
    out = dg.download("s", n), dg.download("Cn", n)
    dg.close()
    return out


SCALAR_GRIDS = [
    ("tiny1", lambda: cases.tiny(1)),
    ("lockexchange", lambda: cases.lock_exchange(seed=7)),
    ("hexquad7", lambda: cases.tiny(24, nx=14, ny=9, Nkmax=11, planform="hex", split_every=4, maxfaces=7)),
]


@pytest.mark.parametrize("tvd", [0, 1, 3, 4])
@pytest.mark.parametrize("name,make", SCALAR_GRIDS, ids=[g[0] for g in SCALAR_GRIDS])
def test_tvd_group_bitwise(gpu_available, name, make, tvd):
    """The inflow sums and the column update walked 1, 2 or 4 levels per
    load group give the same bits (same operations in the same order)."""
    st = make()
    ref = scalar_step(st, 1, tvd)
    for group in (2, 4):
        got = scalar_step(st, group, tvd)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), group


@pytest.mark.parametrize("tvd", [1, 4])
@pytest.mark.parametrize("name,make", SCALAR_GRIDS, ids=[g[0] for g in SCALAR_GRIDS])
def test_tvd_head_bitwise(gpu_available, name, make, tvd):
    """The column update with its top 9 or 17 Thomas rows kept in registers
    (columns shallower and deeper than the head) gives the same bits as with
    every row through memory."""
    st = make()
    ref = scalar_step(st, 2, tvd, 0)
    for head in (8, 16):
        got = scalar_step(st, 2, tvd, head)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), head


def h_solve(st, tuning):
    dg = device.DeviceGrid(st)
    dg.upload("D", cases.depth_integrated_D(st))
    dg.upload("htmp", st.get("htmp"))
    dg.upload("h", np.zeros(st.Nc))
    for k, v in tuning:
        dg.tune(k, v)
    g, th, dt = st.get_prop("grav"), st.get_prop("theta"), st.get_prop("dt")
    s = dg.h_solve(g, th, dt, maxiters=2000, precond=1, epsilon=1e-10)
    h = dg.download("h", st.Nc)
    dg.close()
    return h, s["niters"]


@pytest.fixture(scope="module")
def h_grid():
    # 320,000 cells: 256 blocks of 256 threads give each thread 4 or 5 cells,
    # so the four batched cells and the loop after them both run
    return cases.large_solve(seed=3, nx=400, ny=400, Nkmax=3, target_N=6.0e5)


@pytest.mark.parametrize("blocks", [256, 512, 2048])
def test_h_blocks_rounding(gpu_available, h_grid, blocks):
    """h_blocks regroups the free-surface dot products' block sums: counts
    within +-1 and h within 1e-12 of the default."""
    h0, n0 = h_solve(h_grid, [])
    h1, n1 = h_solve(h_grid, [("h_blocks", blocks)])
    assert abs(n0 - n1) <= 1
    assert np.linalg.norm(h1 - h0) <= 1e-12 * np.linalg.norm(h0)

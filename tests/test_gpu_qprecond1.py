"""qprecond=1 on the device (ConditionQ main/phys.c:3629-3687, the
preconditioned branch of QCoefficients :3426-3454, OperatorQC :3364-3409,
CGSolveQ :2423-2433, 2539-2546) against the oracle.

Solve calls CGSolveQ with c = phys->stmp3 (main/phys.c:1280), the rows
CGSolveQ also uses as its residual r (:2419).  The iterations are then a
Jacobi-scaled CG on the scaled unknown x~ = c q, but the closing x /= c
divides by the final residual instead of the scaling.  The device does what
the reference does in both cases: with c aliasing r (qc_alias) it divides by
its own final residual, with separate c rows by c.

Checks:
* set-up, bit for bit (no iteration): the conditioning c, the initial
  residual r = src/c - L_C (c q) and q = (c q)/r;
* exactly n iterations (separate c): q within 1e-12; with the alias the
  scaled unknown x~ = q r and the residual r within 1e-12;
* converged (1e-10): separate c, q within 1e-10; with the alias, x~ and r are
  compared instead of q = x~/r (a final residual component is the rounding
  noise of two solves, so its quotient is not comparable); iteration counts
  within 5% (a Jacobi-scaled CG needs hundreds of iterations here, where the
  count is decided by rounding, as for qprecond=0);
* the drop-in (reference signature) runs both calls and gives the device
  ABI's q bit for bit;
* open grids whose type-3 boundary cells hold source and pressure values:
  the reference's iterations read a boundary neighbour's p, which aliases
  src there (a constant term in every operator application, so these solves
  do not converge); qprecond 0, 1 and 2 against the oracle after exactly 2
  and 3 iterations, within 1e-12.  (Without the self-correction of a true CG
  the rounding of the dot products grows quickly here: two summation orders
  of the oracle's own loop differ by 1.5e-10 after 8 iterations at
  qprecond 0.)
"""
import ctypes as C

import numpy as np
import pytest

from suntans_amd import cases, device

import _impls
from test_gpu_parity import MAKERS

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dev(st, c):
    dg = device.DeviceGrid(st)
    for n, f in (("q", "qc"), ("src", "stmp"), ("dzz", "dzz"), ("dzf", "dzf"), ("D", "D"), ("qcond", c)):
        dg.upload(n, st.get(f))
    return dg


def run_dev(st, c, alias, n=None, eps=1e-10):
    dg = dev(st, c)
    r = dg.q_solve(maxiters=5000, precond=1, epsilon=eps, fixed=n, qc_alias=int(alias))
    N = len(st.get("qc"))
    out = {f: dg.download(f, N) for f in ("q", "r", "qcond")}
    dg.close()
    return r, out


def run_oracle(st, c, n=None, eps=1e-10):
    o = _impls.oracle()
    a = st.clone()
    if n is None:
        a.set_prop(qprecond=1, qepsilon=eps, qmaxiters=5000, resnorm=0)
    else:
        a.set_prop(qprecond=1, qepsilon=0.0, qmaxiters=n, resnorm=0)
    o.cgsolveq(a, c=c)
    return o.stats("q")[0], a


@pytest.mark.parametrize("name,make", MAKERS, ids=[m[0] for m in MAKERS])
def test_pc1_setup_bitwise(gpu_available, name, make):
    st = make()
    st.set_prop(qprecond=1, resnorm=0)
    n_o, a = run_oracle(st, "stmp3", n=0)
    r, d = run_dev(st, "stmp3", alias=True, n=0)
    m = st.wet_mask()
    cond = st.clone()
    _impls.oracle().fn("ConditionQ")(cond.ptr("uold"), cond.g, cond.p, cond.prop)
    assert np.array_equal(d["qcond"][m], cond.get("uold")[m])
    assert np.array_equal(d["r"][m], a.get("stmp3")[m])   # r, the aliased c rows
    assert np.array_equal(d["q"][m], a.get("qc")[m])      # (c q) / r


FIXED = [m for m in MAKERS if m[0] in ("tiny_wide", "lockexchange", "quadtri", "hexquad7", "mbay")]


@pytest.mark.parametrize("n", [1, 3])
@pytest.mark.parametrize("name,make", FIXED, ids=[m[0] for m in FIXED])
def test_pc1_fixed_iterations(gpu_available, name, make, n):
    st = make()
    n_o, a = run_oracle(st, "uold", n=n)
    r, d = run_dev(st, "uold", alias=False, n=n)
    m = st.wet_mask()
    assert r["niters"] == n_o == n
    assert relerr(d["q"][m], a.get("qc")[m]) <= 1e-12
    n_o, a = run_oracle(st, "stmp3", n=n)
    r, d = run_dev(st, "stmp3", alias=True, n=n)
    ra = a.get("stmp3")[m]
    assert relerr(d["r"][m], ra) <= 1e-12
    assert relerr(d["q"][m] * d["r"][m], a.get("qc")[m] * ra) <= 1e-12


@pytest.mark.parametrize("name,make", MAKERS, ids=[m[0] for m in MAKERS])
def test_pc1_converged(gpu_available, name, make):
    st = make()
    m = st.wet_mask()
    n_o, a = run_oracle(st, "uold")
    r, d = run_dev(st, "uold", alias=False)
    err = relerr(d["q"][m], a.get("qc")[m])
    print(f"{name} separate c: iterations oracle={n_o} device={r['niters']} rel err {err:.3e}")
    assert r["converged"] == 1
    assert abs(r["niters"] - n_o) <= max(1, 0.05 * n_o)
    assert err <= 1e-10
    n_o, a = run_oracle(st, "stmp3")
    r, d = run_dev(st, "stmp3", alias=True)
    ra = a.get("stmp3")[m]
    ex = relerr(d["q"][m] * d["r"][m], a.get("qc")[m] * ra)
    print(f"{name} c = r: iterations oracle={n_o} device={r['niters']} scaled unknown rel err {ex:.3e}")
    assert abs(r["niters"] - n_o) <= max(1, 0.05 * n_o)
    assert ex <= 1e-10


DROPIN = [m for m in MAKERS if m[0] in ("tiny0", "lockexchange", "hexquad7", "mbay")]


@pytest.mark.parametrize("name,make", DROPIN, ids=[m[0] for m in DROPIN])
def test_pc1_dropin(gpu_available, name, make):
    d = _impls.dropin()
    st = make()
    st.set_prop(qprecond=1, qepsilon=1e-10, qmaxiters=5000, resnorm=0)
    m = st.wet_mask()
    for c, alias in (("stmp3", True), ("uold", False)):
        b = st.clone()
        d.cgsolveq(b, c=c)
        nd = d.stats("q")[0]
        r, dv = run_dev(st, c, alias=alias)
        assert nd == r["niters"]
        assert np.array_equal(b.get("qc")[m], dv["q"][m])


OPEN = [("open_tri", lambda: cases.tiny(3, nx=16, ny=9, Nkmax=10, open_marks=cases.open_ends())),
        ("open_lock", lambda: cases.lock_exchange(seed=8, nx=20, ny=5, Nkmax=10, open_marks=cases.open_ends(3, 3)))]


def _boundary_values(st, rng):
    """Nonzero source, pressure and conditioning rows in the type-3 cells."""
    cd, cp, nk = st.ints("celldist"), st.ints("cellp"), st.ints("Nk")
    off = np.concatenate([[0], np.cumsum(nk)])
    bnd = cp[cd[1]:cd[2]]
    assert len(bnd) > 0
    for f, lo, hi in (("stmp", -1e-3, 1e-3), ("qc", -1e-3, 1e-3), ("uold", 0.5, 2.0)):
        v = st.get(f)
        for i in bnd:
            v[off[i]:off[i + 1]] = rng.uniform(lo, hi, nk[i])
        st.set(f, v)


@pytest.mark.parametrize("n", [2, 3])
@pytest.mark.parametrize("pc", [2, 0, 1])
@pytest.mark.parametrize("name,make", OPEN, ids=[m[0] for m in OPEN])
def test_boundary_cells(gpu_available, name, make, pc, n):
    o, d = _impls.oracle(), _impls.dropin()
    st = make()
    _boundary_values(st, np.random.default_rng(4))
    st.set_prop(qprecond=pc, qepsilon=0.0, qmaxiters=n, resnorm=0)
    a, b = st.clone(), st.clone()
    o.cgsolveq(a, c="uold")
    d.cgsolveq(b, c="uold")
    no, nd = o.stats("q")[0], d.stats("q")[0]
    m = st.wet_mask()
    err = relerr(b.get("qc")[m], a.get("qc")[m])
    print(f"{name} qprecond={pc} n={n}: rel err {err:.3e}")
    assert no == nd == n
    assert err <= 1e-12

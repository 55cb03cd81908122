"""Checks of the oracle that do not need the reference: dense brute force on
small grids, closed forms and library routines (so that a dropped term, a
wrong sign or index, or a transposed operand fails here)."""
import numpy as np
import pytest
import scipy.linalg

from suntans_amd import cases

import _impls


def dense_operator(st, op, n, xname, yname):
    """Matrix of y = op(x) over the wet cell-layers, column by column."""
    m = st.wet_mask()
    idx = np.flatnonzero(m)
    A = np.zeros((len(idx), len(idx)))
    base = np.zeros(n)
    for j, col in enumerate(idx):
        x = base.copy()
        x[col] = 1.0
        st.set(xname, x)
        st.set(yname, 0.0)
        op(st)
        A[:, j] = st.get(yname)[idx]
    return A, idx


@pytest.fixture(scope="module")
def tiny():
    return cases.tiny(3)


def test_operator_q_symmetric_negative_definite(tiny):
    o = _impls.oracle()
    st = tiny.clone()
    o.qcoef(st)
    A, _ = dense_operator(st, lambda s: o.operatorq(s), len(st.get("stmp")), "stmp", "stmp2")
    assert np.allclose(A, A.T, rtol=1e-13, atol=1e-12 * np.abs(A).max())
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    assert ev.max() < 0


def test_operator_q_of_constant_is_top_dirichlet_term(tiny):
    """L(1): horizontal differences vanish, interior and bottom rows vanish,
    the top row keeps -2 coef[ctop] = -2 Ac/dzz[ctop]."""
    o = _impls.oracle()
    st = tiny.clone()
    o.qcoef(st)
    st.set("stmp", 1.0)
    o.operatorq(st)
    cell, k = st.cell_layers()
    ctop = st.ints("ctop")
    y = st.get("stmp2")
    Ac, dzz = st.get("Ac"), st.get("dzz")
    top = k == ctop[cell]
    expect = -2 * Ac[cell[top]] / dzz[top]
    assert np.allclose(y[top], expect, rtol=1e-14)
    interior = (k > ctop[cell])
    assert np.all(np.abs(y[interior]) <= 1e-12 * np.abs(expect).max())


def test_preconditioner_inverts_vertical_operator(tiny):
    """With D = 0 the operator is the vertical part the preconditioner
    inverts: L_v(M^-1 x) = x."""
    o = _impls.oracle()
    st = tiny.clone()
    o.qcoef(st)
    rng = np.random.default_rng(0)
    x = rng.normal(size=st.get("stmp").shape)
    st.set("stmp", x)
    o.preconditioner(st, x="stmp", xc="uold")
    st.set("D", 0.0)
    o.operatorq(st, x="uold", y="stmp2")
    m = st.wet_mask()
    assert np.allclose(st.get("stmp2")[m], x[m], rtol=1e-10, atol=1e-10 * np.abs(x).max())


def test_trisolve_matches_banded_solver():
    o = _impls.oracle()
    import ctypes as C
    rng = np.random.default_rng(1)
    for n in (1, 2, 3, 7, 50):
        a = rng.uniform(0.1, 1, n)
        c = rng.uniform(0.1, 1, n)
        b = -(a + c) - rng.uniform(0.1, 1, n)
        d = rng.normal(size=n)
        ab = np.zeros((3, n))
        ab[0, 1:] = c[:-1]
        ab[1] = b
        ab[2, :-1] = a[1:]
        want = scipy.linalg.solve_banded((1, 1), ab, d)
        u = np.zeros(n)
        bb, dd = b.copy(), d.copy()
        o.fn("TriSolve")(a.ctypes.data_as(C.c_void_p), bb.ctypes.data_as(C.c_void_p),
                         c.ctypes.data_as(C.c_void_p), dd.ctypes.data_as(C.c_void_p),
                         u.ctypes.data_as(C.c_void_p), n)
        assert np.allclose(u, want, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("qprecond", [2, 0])
def test_cgsolveq_matches_dense_solve(tiny, qprecond):
    o = _impls.oracle()
    st = tiny.clone()
    o.qcoef(st)
    A, idx = dense_operator(st, lambda s: o.operatorq(s), len(st.get("stmp")), "stmp", "stmp2")
    st = tiny.clone()
    src = st.get("stmp")
    q0 = st.get("qc")
    want = np.linalg.solve(A, src[idx])
    st.set_prop(qprecond=qprecond, qepsilon=1e-14, qmaxiters=5000)
    o.cgsolveq(st)
    got = st.get("qc")[idx]
    assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want)
    # the warm start only changes the path, not the answer
    assert not np.allclose(q0[idx], want)


def test_operator_h_of_constant_is_area(tiny):
    o = _impls.oracle()
    st = tiny.clone()
    st.set("D", cases.depth_integrated_D(st))
    o.hcoef(st)
    st.set("htmp2", 2.5)
    o.operatorh(st)
    assert np.allclose(st.get("htmp3"), 2.5 * st.get("Ac"), rtol=1e-13)


@pytest.mark.parametrize("hprecond", [1, 0])
def test_cgsolve_matches_dense_solve(tiny, hprecond):
    o = _impls.oracle()
    st = tiny.clone()
    st.set("D", cases.depth_integrated_D(st))
    o.hcoef(st)
    Nc = st.Nc
    A = np.zeros((Nc, Nc))
    for j in range(Nc):
        e = np.zeros(Nc)
        e[j] = 1
        st.set("htmp2", e)
        o.operatorh(st)
        A[:, j] = st.get("htmp3")
    assert np.allclose(A, A.T, rtol=1e-13)
    assert np.linalg.eigvalsh(A).min() > 0
    st = tiny.clone()
    st.set("D", cases.depth_integrated_D(st))
    b = st.get("htmp")
    st.set_prop(hprecond=hprecond, epsilon=1e-14, maxiters=5000)
    o.cgsolve(st)
    want = np.linalg.solve(A, b)
    assert np.linalg.norm(st.get("h") - want) <= 1e-11 * np.linalg.norm(want)


@pytest.mark.parametrize("r,superbee,minmod", [(-1, 0, 0), (0, 0, 0), (0.5, 1, 0.5), (1, 1, 1), (2, 2, 1),
                                               (3, 2, 1), (1.5, 1.5, 1)])
def test_psi_closed_forms(r, superbee, minmod):
    o = _impls.oracle()
    import ctypes as C
    f = o.fn("Psi")
    f.argtypes = [C.c_double, C.c_int]
    f.restype = C.c_double
    assert f(r, 1) == 0
    assert f(r, 2) == 1
    assert f(r, 3) == superbee
    assert f(r, 4) == minmod


@pytest.mark.parametrize("tvd", [0, 1, 3, 4])
def test_still_water_keeps_any_scalar(tvd):
    """No flow, no diffusion, fixed layers: UpdateScalars returns s."""
    o = _impls.oracle()
    st = cases.tiny(5, move_surface=False)
    for f in ("u", "utmp2", "w", "wtmp2", "wnew"):
        st.set(f, 0.0)
    st.set("dzzold", st.get("dzz"))
    st.set_prop(TVDsalt=tvd, kappa_s=0.0, n=2)
    s0 = st.get("s")
    o.update_salinity(st)
    assert np.allclose(st.get("s"), s0, rtol=1e-15, atol=0)


STEP_CASES = [
    ("tiny3", lambda **kw: cases.tiny(3, **kw)),
    ("quadtri", lambda **kw: cases.tiny(22, nx=7, ny=5, planform="quad", split_every=3, **kw)),
    ("hexquad7", lambda **kw: cases.tiny(24, nx=7, ny=6, planform="hex", split_every=4, maxfaces=7, **kw)),
]


@pytest.mark.parametrize("name,make", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_continuity_zeroes_the_source(name, make):
    """w from Continuity makes the discrete divergence that ComputeQSource
    measures vanish: the two are the same theta-weighted flux balance."""
    o = _impls.oracle()
    st = make()
    m = st.wet_mask()
    o.compute_qsource(st)
    before = np.abs(st.get("stmp")[m]).max()
    o.continuity(st)
    o.compute_qsource(st)
    after = np.abs(st.get("stmp")[m]).max()
    assert before > 1.0
    assert after <= 1e-14 * before, (before, after)


@pytest.mark.parametrize("name,make", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_projection_zeroes_the_source(name, make):
    """Corrector with the converged solution of L qc = src removes the
    divergence: ComputeQSource(u', w') = src - L(qc).  On a flat surface
    (etop = max ctop) the corrected edge layers are exactly the operator's,
    so the identity holds to rounding; a wrong sign, factor or surface ghost
    term in Corrector leaves an O(1) source."""
    o = _impls.oracle()
    st = make(move_surface=False)
    m = st.wet_mask()
    o.compute_qsource(st)
    before = np.abs(st.get("stmp")[m]).max()
    st.set("qc", 0.0)
    st.set_prop(qepsilon=1e-14, qmaxiters=20000, qprecond=2)
    o.cgsolveq(st)
    q0 = st.get("q")
    o.compute_qsource(st)  # CGSolveQ clobbers its source
    o.corrector(st)
    o.compute_qsource(st)
    after = np.abs(st.get("stmp")[m]).max()
    assert after <= 1e-12 * before, (before, after)
    # the pressure accumulates the correction on multi-layer columns
    cell, k = st.cell_layers()
    multi = (st.ints("Nk") - st.ints("ctop") > 1)[cell]
    assert np.array_equal(st.get("q")[m & multi], (q0 + st.get("qc"))[m & multi])


def test_upredictor_inviscid_closed_form():
    """No viscosity, drag, wind or momentum advection: every vertical system
    is the identity, so E = 1, D = sum of dzf over the edge's wet layers
    (the depth-integrated coefficient, main/phys.c:2936-2939), U*** = U**
    (the source minus the explicit free-surface gradient), and after the
    free-surface solve u = U*** - g theta dt (h1 - h2)/dg."""
    o = _impls.oracle()
    st = cases.tiny(11, nx=23, ny=9, Nkmax=13)
    cases.upredictor_inputs(st, np.random.default_rng(3), noslip=False, nonlinear=0)
    st.set("nu_tv", 0.0)
    st.set("nu_lax", 0.0)
    st.set("CdB", 0.0)
    st.set("CdT", 0.0)
    st.set("tau_T", 0.0)
    st.set_prop(nu=0.0)
    g, th, dt = st.get_prop("grav"), st.get_prop("theta"), st.get_prop("dt")
    edge, k = st.edge_layers()
    grad = st.ints("grad").reshape(-1, 2)
    etop, nke, mark = st.ints("etop"), st.ints("Nke"), st.ints("mark")
    comp = (mark[edge] == 0) & (k >= etop[edge]) & (k < nke[edge])
    h0, dg = st.get("h"), st.get("dg")
    grad_h = (h0[grad[edge, 0]] - h0[grad[edge, 1]]) / dg[edge]
    expect = st.get("utmp") - g * (1 - th) * dt * grad_h
    assert o.upredictor_begin(st) == 0
    assert np.array_equal(st.get("ut")[comp], np.ones(comp.sum()))
    assert np.allclose(st.get("utmp")[comp], expect[comp], rtol=1e-15, atol=1e-17)
    D = st.get("D")
    wet = comp & (st.get("dzf") > 0)
    depth = np.bincount(edge[wet], weights=st.get("dzf")[wet], minlength=st.Ne)
    assert np.allclose(D[mark == 0], depth[mark == 0], rtol=1e-14)
    ustar = st.get("utmp").copy()
    assert o.upredictor_solve(st) == 0
    h1 = st.get("h")
    grad_h1 = (h1[grad[edge, 0]] - h1[grad[edge, 1]]) / dg[edge]
    assert np.allclose(st.get("u")[comp], (ustar - g * th * dt * grad_h1)[comp], rtol=1e-14, atol=1e-16)


def test_update_dz_column_heights():
    """After UpdateDZ each wet column's layer heights add up to h + dv, the
    cells above ctop are empty and the ctop cell is not, and etop is the
    shallower ctop of an edge's two cells."""
    o = _impls.oracle()
    st = cases.tiny(7, nx=12, ny=8, Nkmax=10)
    rng = np.random.default_rng(4)
    st.set("h", st.get("h") + rng.uniform(-2.5, 1.5, st.Nc))
    o.update_dz(st, 0)
    cell, k = st.cell_layers()
    dzz, h, dv = st.get("dzz"), st.get("h"), st.get("dv")
    ctop, nk = st.ints("ctop"), st.ints("Nk")
    total = np.bincount(cell, weights=dzz, minlength=st.Nc)
    wet = h > -dv
    assert np.allclose(total[wet], (h + dv)[wet], rtol=0, atol=1e-12 * dv.max())
    assert np.all(dzz[k < ctop[cell]] == 0)
    top = (k == ctop[cell]) & wet[cell] & (ctop[cell] < nk[cell])
    assert np.all(dzz[top] >= 0)
    grad = st.ints("grad").reshape(-1, 2)
    etop = st.ints("etop")
    both = (grad[:, 0] >= 0) & (grad[:, 1] >= 0)
    assert np.array_equal(etop[both], np.minimum(ctop[grad[both, 0]], ctop[grad[both, 1]]))


def _edge_cells(st):
    grad = st.ints("grad").reshape(-1, 2)
    return grad[:, 0], grad[:, 1]


def test_central_face_values_of_a_linear_field():
    """On rectangles (each edge midpoint on the segment joining the two cell
    centres) the distance-weighted face value of a field linear in x, y and
    k is the field at the edge midpoint."""
    o = _impls.oracle()
    st = cases.tiny(21, nx=6, ny=5, planform="quad")
    cases.wpredictor_inputs(st, np.random.default_rng(2), nonlinear=2, stress=True)
    cell, k = st.cell_layers()
    xv, yv = st.get("xv"), st.get("yv")
    st.set("wc", 2e-3 * xv[cell] - 3e-3 * yv[cell] + 0.25 * k)
    o.momentum_faces(st)
    edge, ke = st.edge_layers()
    xe, ye = st.get("xe"), st.get("ye")
    nc1, nc2 = _edge_cells(st)
    ctop = st.ints("ctop")
    comp = (st.ints("mark") == 0)[edge]
    kmin = np.maximum(ctop[nc1[edge].clip(0)], ctop[nc2[edge].clip(0)])
    m = comp & (ke >= kmin) & (ke < st.ints("Nke")[edge])
    want = 2e-3 * xe[edge] - 3e-3 * ye[edge] + 0.25 * ke
    assert m.sum() > 100
    assert np.allclose(st.get("ut")[m], want[m], rtol=1e-12, atol=1e-12)
    assert np.all(st.get("ut")[comp & (ke < kmin)] == 0)


def test_upwind_face_value_comes_from_the_outflow_cell():
    """Scheme 1: the face value is wc of the cell the flow leaves, the one
    whose outward normal velocity normal * U is positive (when U = 0, the
    edge's first cell, grad[2j])."""
    o = _impls.oracle()
    st = cases.tiny(4, nx=9, ny=6)
    cases.wpredictor_inputs(st, np.random.default_rng(3), nonlinear=1, stress=True)
    o.momentum_faces(st)
    edge, ke = st.edge_layers()
    nc1, nc2 = _edge_cells(st)
    nke, ctop, mf = st.ints("Nke"), st.ints("ctop"), st.scalar("maxfaces")
    face, normal = st.ints("face").reshape(-1, mf), st.ints("normal").reshape(-1, mf)
    # outward normal sign of each edge seen from its first cell
    n1 = np.zeros(st.Ne, int)
    for i in range(st.Nc):
        for f in range(st.ints("nfaces")[i]):
            if nc1[face[i, f]] == i:
                n1[face[i, f]] = normal[i, f]
    wc, ut, U = st.get("wc"), st.get("ut"), st.get("utmp2")
    cstart = np.concatenate([[0], np.cumsum(st.ints("Nk"))[:-1]])
    checked = 0
    for idx in range(len(edge)):
        j, k = edge[idx], ke[idx]
        if st.ints("mark")[j] != 0 or k < max(ctop[nc1[j]], ctop[nc2[j]]) or k >= nke[j]:
            continue
        src = nc1[j] if n1[j] * U[idx] > 0 or U[idx] == 0 else nc2[j]
        assert ut[idx] == wc[cstart[src] + k]
        checked += 1
    assert checked > 100


def test_horizontal_diffusion_conserves_momentum():
    """Without advection and sidewall drag the horizontal diffusion only
    moves w between neighbours: per level, sum of Ac * stmp over the cells
    vanishes (the reference's own conservation check, main/phys.c:2147-2156)."""
    o = _impls.oracle()
    st = cases.tiny(6, nx=10, ny=7)
    cases.wpredictor_inputs(st, np.random.default_rng(4), nonlinear=0, CdW=0.0, nu_H=1.0)
    o.wpredictor(st)
    cell, k = st.cell_layers()
    flux = st.get("Ac")[cell] * st.get("stmp")
    scale = np.abs(flux).max()
    assert scale > 0
    for lev in range(k.max() + 1):
        assert abs(flux[k == lev].sum()) <= 1e-13 * scale


def test_wpredictor_without_sources_or_viscosity():
    """No advection, diffusion or drag and zero viscosity: w becomes
    w + f2 Cn_W + f3 Cn_W2 (Adams-Bashforth 3 weights -4/3 and 5/12) minus
    dt times the vertical gradient of q between layer centres (q = 0 at the
    free surface, half a layer above the top centre); Cn_W2 takes the old
    Cn_W and Cn_W becomes 0."""
    o = _impls.oracle()
    st = cases.tiny(7, nx=8, ny=6)
    cases.wpredictor_inputs(st, np.random.default_rng(5), nonlinear=0, CdW=0.0, nu_H=0.0)
    st.set("nu_tv", 0.0)
    st.set("nu_lax", 0.0)
    st.set_prop(nu=0.0)
    w0, cn, cn2, q, dzz = (st.get(f) for f in ("w", "Cn_W", "Cn_W2", "q", "dzz"))
    o.wpredictor(st)
    dt = st.get_prop("dt")
    nk, ctop = st.ints("Nk"), st.ints("ctop")
    wstart = np.concatenate([[0], np.cumsum(nk + 1)[:-1]])
    cstart = np.concatenate([[0], np.cumsum(nk)[:-1]])
    w = st.get("w")
    for i in range(st.Nc):
        for k in range(ctop[i], nk[i]):
            iw, ic = wstart[i] + k, cstart[i] + k
            if k == ctop[i]:
                grad = -q[ic] / (0.5 * dzz[ic])
            else:
                grad = (q[ic - 1] - q[ic]) / (0.5 * (dzz[ic - 1] + dzz[ic]))
            want = w0[iw] - 4.0 / 3.0 * cn[iw] + 5.0 / 12.0 * cn2[iw] - dt * grad
            assert abs(w[iw] - want) <= 1e-12 * max(1.0, abs(want)), (i, k)
            assert st.get("Cn_W2")[iw] == cn[iw] and st.get("Cn_W")[iw] == 0


@pytest.mark.parametrize("planform", ["tri", "quad", "hex"])
def test_face_distances_match_the_oracle(planform):
    """The host-side face distances handed to the device (numpy) equal the
    oracle's (the reference's expression, main/phys.c:4272-4273) bit for bit."""
    import ctypes as C
    from suntans_amd import device
    o = _impls.oracle()
    st = cases.tiny(8, nx=7, ny=5, planform=planform, open_marks=cases.open_ends())
    o.L.oracle_face_def1.restype = C.c_double
    o.L.oracle_face_def1.argtypes = [C.c_void_p, C.c_int]
    want = np.array([o.L.oracle_face_def1(st.g, j) for j in range(st.Ne)])
    assert np.array_equal(device.face_distances(st), want)

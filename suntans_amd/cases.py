"""Seeded synthetic workloads for the elliptic solves and scalar transport.

* :func:`tiny`          small ragged grids with moving free surface (parity)
* :func:`lock_exchange` 10 km x 1 km channel, 100x10 quads as 2000 triangles,
                        50 uniform levels over 20 m, salinity step at x=5 km
* :func:`large_solve`   100 km x 50 km, 1000x500 quads as 1.0M triangles,
                        depth 10-200 m from 8 Gaussian bumps, 100 levels of
                        2 m (~50M pressure unknowns)
* :func:`ensemble_member` one member of the 64-member ensemble

Every random number comes from ``numpy.random.default_rng(seed)``.  Velocity
fields are sums of a few low-wavenumber modes, zero on closed boundary edges
and outside each edge's wet layers; vertical velocity follows continuity
where a step needs a divergence-free field.
"""
from __future__ import annotations

import os

import numpy as np

from .state import State


def _modes(rng, n, amp):
    return [(rng.normal(0, amp), rng.uniform(0.5, 3.0), rng.uniform(0.5, 3.0), rng.uniform(0.2, 1.5),
             rng.uniform(0, 2 * np.pi)) for _ in range(n)]


def smooth(rng, x, y, z, Lx, Ly, H, amp=1.0, n=5):
    """Sum of n random low-wavenumber modes at (x, y, z)."""
    out = np.zeros(np.broadcast(x, y, z).shape)
    for a, kx, ky, kz, ph in _modes(rng, n, amp):
        out += a * np.sin(kx * np.pi * x / Lx + ph) * np.cos(ky * np.pi * y / Ly + 0.5 * ph) * \
            np.cos(kz * np.pi * z / H + 0.3 * ph)
    return out


def smooth_columns(rng, xc, yc, ic, zl, il, Lx, Ly, H, amp=1.0, n=5):
    """smooth() at (xc[ic], yc[ic], zl[il]) with the trigonometry evaluated
    once per column and once per level; the products are formed in the same
    order, so the values are identical to smooth() on the expanded points."""
    out = np.zeros(len(ic))
    for a, kx, ky, kz, ph in _modes(rng, n, amp):
        f = a * np.sin(kx * np.pi * xc / Lx + ph) * np.cos(ky * np.pi * yc / Ly + 0.5 * ph)
        g = np.cos(kz * np.pi * zl / H + 0.3 * ph)
        out += f[ic] * g[il]
    return out


def bumps_depth(rng, Lx, Ly, dmin, dmax, nb=8, mean=None):
    """Depth function: clip(base + sum of nb Gaussian bumps, dmin, dmax)."""
    cx = rng.uniform(0, Lx, nb)
    cy = rng.uniform(0, Ly, nb)
    sx = rng.uniform(0.08, 0.25, nb) * Lx
    sy = rng.uniform(0.08, 0.25, nb) * Ly
    amp = rng.uniform(-1.0, 1.0, nb)

    def raw(x, y):
        v = np.zeros_like(x)
        for i in range(nb):
            v += amp[i] * np.exp(-((x - cx[i]) / sx[i]) ** 2 - ((y - cy[i]) / sy[i]) ** 2)
        return v

    def depth(x, y):
        r = raw(x, y)
        r = (r - r.min()) / max(r.max() - r.min(), 1e-30)
        target = 0.5 if mean is None else (mean - dmin) / (dmax - dmin)
        # power law that puts the mean of the normalised field at target
        lo, hi = 0.05, 20.0
        for _ in range(60):
            g = np.sqrt(lo * hi)
            if np.mean(r ** g) > target:
                lo = g
            else:
                hi = g
        d = dmin + (dmax - dmin) * r ** np.sqrt(lo * hi)
        d = np.clip(d, dmin, dmax)
        d[np.argmax(d)] = dmax  # the deepest column uses every level
        return d

    return depth


def fill_velocities(st: State, rng, amp_u=0.1, amp_w=1e-3, zero_w_bottom=True,
                    edge_names=("u", "utmp2"), cell_names=("w", "wtmp2", "wnew")):
    """u, utmp2 on wet edge layers (zero on closed boundaries), w, wtmp2, wnew."""
    Lx, Ly = st.Lx, st.Ly
    H = float(np.max(st.get("dv"))) + 1e-9
    edge, k = st.edge_layers()
    xe, ye = st.get("xe"), st.get("ye")
    dz = st.get("dz")
    zc = np.concatenate([[0.0], np.cumsum(dz)])
    zmid = 0.5 * (zc[:-1] + zc[1:])
    mark = st.ints("mark")
    wet = st.edge_wet_mask() & (mark[edge] == 0)
    kz = np.minimum(k, len(zmid) - 1)
    for name in edge_names:
        v = smooth_columns(rng, xe, ye, edge, zmid, kz, Lx, Ly, H, amp_u)
        st.set(name, np.where(wet, v, 0.0))
    cell, kw = st.cell_layers(extra=1)
    xv, yv = st.get("xv"), st.get("yv")
    nk = st.ints("Nk")
    kwz = np.minimum(kw, len(zc) - 1)
    for name in cell_names:
        v = smooth_columns(rng, xv, yv, cell, zc, kwz, Lx, Ly, H, amp_w)
        if zero_w_bottom:
            v = np.where(kw == nk[cell], 0.0, v)
        st.set(name, v)


def fill_scalar(st: State, rng, kind="random", noise=1e-3):
    cell, k = st.cell_layers()
    xv = st.get("xv")
    if kind == "step":
        s = (xv[cell] >= 0.5 * st.Lx).astype(float) + rng.normal(0, noise, len(cell))
    else:
        s = rng.uniform(0, 1, len(cell))
    st.set("s", s)
    return s


def _finish(st: State, rng, dt, warm_q=True, h_amp=0.0):
    st.set_prop(dt=dt)
    st.set_flux_height()
    # pressure source from the predicted velocities (ComputeQSource)
    st.compute_qsource()
    if warm_q:
        cell, k = st.cell_layers()
        st.set("qc", rng.normal(0, 1e-3, len(cell)) * st.wet_mask())
    # free-surface right-hand side and coefficient D = sum_k dzf (inviscid)
    Ac, xv, yv = st.get("Ac"), st.get("xv"), st.get("yv")
    st.set("htmp", Ac * (st.get("h") + smooth(rng, xv, yv, 0.0, st.Lx, st.Ly, 1.0, 1e-2)))
    st.set("h", rng.normal(0, 1e-3, st.Nc))
    st.set("kappa_tv", 0.0)


def depth_integrated_D(st: State) -> np.ndarray:
    edge, k = st.edge_layers()
    dzf = st.get("dzf")
    return np.bincount(edge, weights=dzf * st.edge_wet_mask(), minlength=st.Ne)


def open_ends(west: int = 2, east: int = 3, frac: float = 0.1):
    """open_marks for State: closed boundary edges whose midpoints lie in the
    westmost (eastmost) frac of the boundary's x-range take mark west (east)."""
    def marks(xe, ye, mark):
        out = mark.copy()
        b = mark == 1
        lo, hi = xe[b].min(), xe[b].max()
        span = hi - lo
        out[b & (xe <= lo + frac * span)] = west
        out[b & (xe >= hi - frac * span)] = east
        return out
    return marks


def tiny(seed=0, nx=4, ny=3, Nkmax=8, move_surface=True, dt=5.0, min_layers=3, planform="tri",
         split_every=0, maxfaces=None, open_marks=None, grid_params=None) -> State:
    """Small grid with ragged columns, dry top layers and a moved surface.
    planform/split_every/maxfaces choose the cell shapes, open_marks opens
    boundary edges (State); grid_params (smoothbot, dzsmall, fixdzz) are set
    before the first UpdateDZ."""
    rng = np.random.default_rng(seed)
    dz = np.full(Nkmax, 1.0)
    H = Nkmax * 1.0

    def depth(x, y):
        d = rng.uniform(min_layers + 2.1, H, x.shape)
        d[rng.integers(len(d))] = H
        return d

    st = State(nx, ny, 400.0 * nx, 300.0 * ny, Nkmax, dz, depth, planform=planform, split_every=split_every,
               maxfaces=maxfaces, open_marks=open_marks)
    if grid_params:
        st.set_grid_param(**grid_params)
    st.update_dz(-1)
    h1 = rng.uniform(-1.6, 0.3, st.Nc) if move_surface else np.zeros(st.Nc)
    st.set("h", h1)
    st.update_dz(1)
    if move_surface:
        st.set("h", h1 + rng.uniform(-1.0, 1.0, st.Nc))
        st.update_dz(0)
    fill_velocities(st, rng)
    fill_scalar(st, rng)
    _finish(st, rng, dt)
    return st


def lock_exchange(seed=0, nx=100, ny=10, Nkmax=50, depth=20.0, dt=5.0, open_marks=None) -> State:
    rng = np.random.default_rng(seed)
    st = State(nx, ny, 10e3, 1e3, Nkmax, np.full(Nkmax, depth / Nkmax), depth, open_marks=open_marks)
    st.update_dz(-1)
    st.update_dz(1)
    fill_velocities(st, rng, amp_u=0.05, amp_w=5e-4)
    fill_scalar(st, rng, kind="step")
    _finish(st, rng, dt)
    st.set_prop(beta=1.0, TVDsalt=4)
    return st


def large_solve(seed=0, nx=1000, ny=500, Nkmax=100, dz=2.0, mean_depth=None, dt=10.0,
                target_N=5.0e7) -> State:
    """The ~50M-unknown solve: depth tuned so that sum(Nk) is close to target_N."""
    rng = np.random.default_rng(seed)
    Nc = 2 * nx * ny
    mean = mean_depth if mean_depth is not None else target_N / Nc * dz - 0.5 * dz
    depth = bumps_depth(rng, 100e3 * nx / 1000, 50e3 * ny / 500, 10.0, Nkmax * dz, mean=mean)
    st = State(nx, ny, 100e3 * nx / 1000, 50e3 * ny / 500, Nkmax, np.full(Nkmax, dz), depth)
    st.update_dz(-1)
    st.update_dz(1)
    fill_velocities(st, rng, amp_u=0.1, amp_w=1e-3, edge_names=("u",), cell_names=("w",))
    fill_scalar(st, rng)
    st.set_prop(dt=dt)
    st.set_flux_height()
    st.compute_qsource()
    st.set("qc", 0.0)
    Ac, xv, yv = st.get("Ac"), st.get("xv"), st.get("yv")
    st.set("htmp", Ac * smooth(rng, xv, yv, 0.0, st.Lx, st.Ly, 1.0, 1e-2))
    st.set("h", 0.0)
    return st


def ensemble_member(seed: int, Nc_target: int, aspect: float, Nkmax: int, dt=10.0) -> State:
    """One member: a triangulated rectangle of ~Nc_target cells, Gaussian-bump
    bathymetry with Nkmax levels of 2 m."""
    rng = np.random.default_rng(seed)
    nquad = max(Nc_target // 2, 4)
    ny = max(int(round(np.sqrt(nquad / aspect))), 2)
    nx = max(int(round(nquad / ny)), 2)
    dz = 2.0
    Lx, Ly = 100.0 * nx, 100.0 * ny
    depth = bumps_depth(rng, Lx, Ly, 3.2 * dz, Nkmax * dz, mean=0.55 * Nkmax * dz)
    st = State(nx, ny, Lx, Ly, Nkmax, np.full(Nkmax, dz), depth)
    st.update_dz(-1)
    st.update_dz(1)
    fill_velocities(st, rng, amp_u=0.1, amp_w=1e-3)
    fill_scalar(st, rng)
    _finish(st, rng, dt, warm_q=False)
    return st


def upredictor_inputs(st: State, rng, noslip: bool = True, nonlinear: int = 1, thetaM: float = 0.5,
                      lax: float = 1.0) -> State:
    """Inputs of UPredictor on a closed grid: the horizontal momentum source in
    utmp, wind stress, drag coefficients (with no-slip, -1, on a third of the
    edges of more than one wet layer when noslip), eddy viscosities, and the
    properties it reads.  Single-layer edges never get no-slip (there the
    reference reads a coefficient left from the previous edge)."""
    edge, k = st.edge_layers()
    st.set("utmp", rng.normal(0, 0.05, len(edge)) * st.edge_wet_mask())
    ne = st.Ne
    etop, nke = st.ints("etop"), st.ints("Nke")
    multi = (nke - etop) > 1
    cdb = rng.uniform(1e-3, 5e-3, ne)
    cdt = rng.uniform(0.0, 2e-3, ne)
    if noslip:
        cdb[multi & (rng.random(ne) < 0.33)] = -1.0
        cdt[multi & (rng.random(ne) < 0.33)] = -1.0
    st.set("CdB", cdb)
    st.set("CdT", cdt)
    st.set("tau_T", rng.normal(0, 1e-4, ne))
    n = len(st.get("nu_tv"))
    st.set("nu_tv", rng.uniform(0, 1e-3, n))
    st.set("nu_lax", rng.uniform(0, 1e-4, n))
    st.set_prop(nu=1e-4, laxWendroff_Vertical=lax, nonlinear=nonlinear, thetaM=thetaM, cgsolver=1,
                epsilon=1e-10, maxiters=1000, hprecond=1, grav=9.81)
    return st


def wpredictor_inputs(st: State, rng, nonlinear: int = 2, step: int = 3, AB: int = 3, CdW: float = 5e-3,
                      nu_H: float = 0.1, conserve: int = 0, tvd: int = 4, stress: bool = False) -> State:
    """Inputs of WPredictor: the Adams-Bashforth histories Cn_W/Cn_W2, eddy
    and Lax-Wendroff viscosities, the old edge velocity utmp2 (= u, as
    UPredictor leaves it), stale face values ut and cell-centred wc (what
    the reference reads where it does not write them this step), open
    boundary rows of w, and the properties it reads.  stress: utmp2 random on
    every edge-layer, dry and closed ones included, so that every product of
    a face the reference reads counts."""
    nw = len(st.get("Cn_W"))
    st.set("Cn_W", rng.normal(0, 1e-5, nw))
    st.set("Cn_W2", rng.normal(0, 1e-5, nw))
    n = len(st.get("nu_tv"))
    st.set("nu_tv", rng.uniform(0, 1e-3, n))
    st.set("nu_lax", rng.uniform(0, 1e-4, n))
    st.set("wc", rng.normal(0, 1e-3, n))
    ne = len(st.get("ut"))
    st.set("ut", rng.normal(0, 1e-2, ne))
    st.set("utmp2", rng.normal(0, 0.05, ne) if stress else st.get("u"))
    nb = len(st.get("boundary_w"))
    if nb:
        st.set("boundary_w", rng.normal(0, 1e-3, nb))
    st.set_prop(nu=1e-4, nu_H=nu_H, CdW=CdW, laxWendroff_Vertical=1.0, nonlinear=nonlinear, n=step, AB=AB,
                conserveMomentum=conserve, TVDmomentum=tvd)
    return st


def timestep_inputs(st: State, rng, nonlinear: int = 1, tvd: int = 4) -> State:
    """Inputs of one nonhydrostatic time step after HorizontalSource: those of
    UPredictor (the horizontal source in utmp, drag, viscosities), the
    vertical-momentum histories Cn_W/Cn_W2, the nonhydrostatic pressure q,
    stale face values and the properties of WPredictor, UpdateScalars
    (salinity, TVD limiter tvd) and CGSolveQ.  u, utmp2 and w stay as built."""
    upredictor_inputs(st, rng, nonlinear=nonlinear)
    nw = len(st.get("Cn_W"))
    st.set("Cn_W", rng.normal(0, 1e-5, nw))
    st.set("Cn_W2", rng.normal(0, 1e-5, nw))
    n = len(st.get("wc"))
    st.set("wc", rng.normal(0, 1e-3, n))
    st.set("q", rng.normal(0, 1e-3, n) * st.wet_mask())
    st.set("ut", rng.normal(0, 1e-2, len(st.get("ut"))))
    st.set_prop(nu_H=0.1, CdW=5e-3, n=3, AB=3, TVDmomentum=tvd, conserveMomentum=0, beta=1.0, TVDsalt=tvd,
                kappa_s=1e-5, qprecond=2, qepsilon=1e-10, qmaxiters=2000, resnorm=0)
    return st


def from_files(datadir: str, seed: int = 0, dt: float = 5.0, h_amp: float = 0.5, amp_u: float = 0.1) -> State:
    """A grid from run-phase files (State.load) with a moved free surface,
    seeded velocities and scalar, and the solver inputs of tiny()."""
    return _seeded(State.load(datadir), seed, dt, h_amp, amp_u)


def from_planform(planform_dir: str, seed: int = 0, depth=None, dt: float = 5.0, h_amp: float = 0.5,
                  amp_u: float = 0.1) -> State:
    """A grid generated from a planform (State.generate on a scratch copy of
    planform_dir, so the run-phase files it writes land there), seeded as
    from_files seeds it."""
    import shutil
    import tempfile
    work = tempfile.mkdtemp(prefix="sun_planform_")
    try:
        for f in os.listdir(planform_dir):
            src = os.path.join(planform_dir, f)
            if os.path.isfile(src):
                shutil.copyfile(src, os.path.join(work, f))
        st = State.generate(work, depth=depth)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return _seeded(st, seed, dt, h_amp, amp_u)


def _seeded(st: State, seed: int, dt: float, h_amp: float, amp_u: float) -> State:
    rng = np.random.default_rng(seed)
    st.update_dz(-1)
    h1 = rng.uniform(-h_amp, 0.3 * h_amp, st.Nc)
    st.set("h", h1)
    st.update_dz(1)
    st.set("h", h1 + rng.uniform(-0.5 * h_amp, 0.5 * h_amp, st.Nc))
    st.update_dz(0)
    fill_velocities(st, rng, amp_u=amp_u)
    fill_scalar(st, rng)
    _finish(st, rng, dt)
    return st


def warm_start(datadir: str, restart: str = None, dt: float = 5.0, step: int = None) -> State:
    """A simulation restarted from the reference's files: the run-phase grid
    (State.load) and, when given, a restart record (State.load_restart), with
    the per-step inputs of the path derived from the restored state as a
    step would find them: flux heights (SetFluxHeight), the pressure source
    of the restored velocities taken as both old and new (ComputeQSource,
    which also sets D = df/dg), the free-surface source Ac h, the previous
    pressure correction as the warm start."""
    st = State.load(datadir)
    st.set_prop(dt=dt)
    n = st.load_restart(restart) if restart else None
    if not restart:
        st.update_dz(-1)
    st.set_prop(n=(step if step is not None else (n if n is not None else 1)))
    st.set("utmp2", st.get("u"))
    st.set("wtmp2", st.get("w"))
    st.set("wnew", st.get("w"))
    st.set_flux_height()
    st.compute_qsource()
    st.set("htmp", st.get("Ac") * st.get("h"))
    return st

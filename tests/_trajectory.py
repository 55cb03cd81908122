"""The lock-exchange trajectory: 20 nonhydrostatic time steps of the
lock-exchange channel (10 km x 1 km, 100x10 quads as 2000 triangles, 50
levels over 20 m, salinity step at x = 5 km with seeded noise, dt = 5 s,
epsilon = qepsilon = 1e-10), each step in the reference's order (Solve,
main/phys.c:1145-1306): SetFluxHeight, UPredictor (its free-surface CGSolve
inside), UpdateDZ, Continuity into wnew, UpdateScalars (TVD superbee
salinity), WPredictor, ComputeQSource, CGSolveQ (qprecond 2), Corrector,
Continuity.  The explicit horizontal momentum source (HorizontalSource) is a
fixed seeded field restored before every step; the Adams-Bashforth step
counter n runs 1..20.

Two implementations take part in a trajectory: ``hot`` runs the three
functions of the path (CGSolve, UpdateScalars, CGSolveQ) and ``host`` the
rest of the step, as the reference's own host code would around a drop-in.
UPredictor is one static function in the reference, so its parts before and
after CGSolve always come from the oracle's restatement
(oracle_UPredictor_begin, _rhs, _finish), which matches the reference bit
for bit on these grids."""
from __future__ import annotations

import time

import numpy as np

from suntans_amd import cases, device

import _impls

STEPS = 20


def initial():
    """The trajectory's starting state (deterministic)."""
    st = cases.lock_exchange(seed=7)
    cases.timestep_inputs(st, np.random.default_rng(31), nonlinear=2)
    return st


def run(st, hot, host, steps=STEPS, on_step=None):
    """Advance st by `steps` steps.  Returns one dict per step: the iteration
    counts and final residuals of the two solves (from hot.stats) and the
    wall time of each hot call in seconds.  on_step(n, st) runs after each
    step."""
    o = _impls.oracle()
    source = st.get("utmp")
    rows = []
    for n in range(1, steps + 1):
        st.set_prop(n=n)
        st.set("utmp", source)
        row = {"n": n}
        host.set_flux_height(st)
        assert o.upredictor_begin(st) == 0
        o.upredictor_rhs(st)
        t = time.perf_counter()
        hot.cgsolve(st)
        row["cgsolve_s"] = time.perf_counter() - t
        row["h_iters"], row["h_resid"] = hot.stats("h")
        o.upredictor_finish(st)
        host.update_dz(st, 0)
        host.continuity(st, w="wnew")
        t = time.perf_counter()
        hot.update_salinity(st)
        row["update_scalars_s"] = time.perf_counter() - t
        host.wpredictor(st)
        host.compute_qsource(st)
        t = time.perf_counter()
        hot.cgsolveq(st)
        row["cgsolveq_s"] = time.perf_counter() - t
        row["q_iters"], row["q_resid"] = hot.stats("q")
        host.corrector(st, qc="qc")
        host.continuity(st, w="w")
        if isinstance(hot, _impls.RefImpl) and hot.history:
            row["h_hist"] = hot.residuals("h")
            row["q_hist"] = hot.residuals("q")
        rows.append(row)
        if on_step is not None:
            on_step(n, st)
    return rows


# fields compared at the end of the trajectory: (state name, wet mask)
FINAL_FIELDS = ("h", "qc", "q", "s", "u", "w")


def final_mask(st, name):
    if name in ("qc", "q", "s"):
        return st.wet_mask()
    if name == "u":
        return st.edge_wet_mask()
    return None


def device_step(dg, a):
    """One step of the same sequence with every field resident on the GPU
    (a supplies the properties); returns the two solves' statistics."""
    g = a.get_prop
    th, dt = g("theta"), g("dt")
    P = device.UPredParams(dt, th, g("grav"), g("nu"), g("laxWendroff_Vertical"), g("thetaM"), int(g("nonlinear")))
    dg.set_flux_height(a)
    dg.u_predict(P)
    sh = dg.u_solve(P, maxiters=int(g("maxiters")), precond=int(g("hprecond")), epsilon=g("epsilon"))
    dg.update_dz(a, 0)
    dg.continuity(th, "wnew")
    dg.update_scalars(th, dt, g("kappa_s"), int(g("TVDsalt")), step=int(g("n")))
    dg.w_predict(device.WPredParams.of(a))
    dg.compute_qsource(th, dt, "src")
    sq = dg.q_solve(maxiters=int(g("qmaxiters")), precond=2, epsilon=g("qepsilon"))
    dg.corrector(dt, "q", "qp")
    dg.continuity(th, "w")
    return sh, sq


def upload_state(dg, a):
    """Every field a resident step reads, from the host state a."""
    for f in ("u", "utmp", "utmp2", "dzf", "dzz", "dzzold", "s", "kappa_tv", "nu_tv", "nu_lax", "wc", "w", "wnew",
              "wtmp2", "wtmp", "Cn_W", "Cn_W2", "h", "dv", "hcorr", "CdB", "CdT", "tau_T", "D"):
        dg.upload(f, a.get(f))
    dg.upload("ue", a.get("ut"))
    dg.upload("q", a.get("qc"))
    dg.upload("qp", a.get("q"))
    dg.upload("active", a.active().astype(np.float64))
    dg.set_face_distances(device.face_distances(a))

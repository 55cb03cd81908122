#!/usr/bin/env python3
"""Benchmark of the nonhydrostatic pressure Poisson PCG on one MI355X.

Workload (the large single-GPU solve): a 100 km x 50 km basin of 1000x500
quads split into 1.0M triangles, depth 10-200 m from 8 seeded Gaussian bumps
on 100 levels of 2 m (ragged columns, ~5.0e7 pressure unknowns), source
= ComputeQSource of seeded smooth random u and w.  One step = one FP64
CGSolveQ with the vertical-line preconditioner (qprecond=2) from q=0 to
sqrt(r.r/r0.r0) < 1e-10, all fields resident in HBM.

Beside the headline FP64 figure the JSON line carries the mixed-precision
form on the same grid and source (FP32 inner PCG in FP64 refinement: speed-up
and relative error against the FP64 solve and against a tight 1e-14 FP64
solve), the free-surface solve (FP64 and mixed) on the same grid, per-solve
latency, the roofline of the dominant kernel and the CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
PMC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_large_solve.txt")

# the FP64 kernels of the workload's iteration (triangles, z-level windows)
KERNEL_NAMES = {"qapply": "k_qapply_x<3, 4, double, true, 4, true>", "qprec": "k_qprec<true, 4, double, true, 24, true>"}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_case(args):
    from suntans_amd import cases
    t = time.time()
    st = cases.large_solve(seed=args.seed, nx=args.nx, ny=args.ny)
    log(f"built grid: Nc={st.Nc} Ne={st.Ne} N={st.N} E_k={st.scalar('Ek')} in {time.time() - t:.1f}s")
    return st


def q_bytes(N: int, Ek: int) -> dict:
    """Algorithmic HBM bytes (FP64): each distinct array read or written once
    per kernel, neighbour gathers counted once."""
    return {
        # CGSolveQ iteration as the reference structures it: 8(14N + E_k)
        "iteration_ref": 8 * (14 * N + Ek),
        # apply kernel: read rtmp, p_old, x, coef, dzf(E_k); write p_new, z, x
        "qapply": 8 * (7 * N + Ek),
        # precondition kernel: read r, z, coef; write r, rtmp (the stored
        # Thomas pivots b' are this build's choice and are not counted)
        "qprec": 8 * (5 * N),
    }


def relerr(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _checkers():
    """The CPU checker: the reference compiled from its own sources
    (oracle/_ref) where present, else the oracle's restatement."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _impls
    impl = _impls.reference()
    if impl is not None:
        return impl, "reference"
    return _impls.oracle(), "port"


def cpu_model() -> str:
    """The host CPU as lscpu names it (model, family/model/stepping, sockets
    x cores, maximum clock), else /proc/cpuinfo's model name."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        f = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        f = {k.strip(): v.strip() for k, v in f.items()}
        parts = [f.get("Model name", "unknown"),
                 f"family {f.get('CPU family', '?')} model {f.get('Model', '?')} stepping {f.get('Stepping', '?')}",
                 f"{f.get('Socket(s)', '?')} socket(s) x {f.get('Core(s) per socket', '?')} cores x "
                 f"{f.get('Thread(s) per core', '?')} threads"]
        if "CPU max MHz" in f:
            parts.append(f"max {f['CPU max MHz']} MHz")
        return ", ".join(parts)
    except (OSError, subprocess.SubprocessError, ValueError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    """Cores this process may run on, capped at 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _cpu_q_solve(impl, st, src0, maxiters, eps):
    """One CGSolveQ (qprecond 2) from q = 0 by the CPU checker; returns
    (seconds, iterations, q)."""
    st.set("qc", 0.0)
    st.set("stmp", src0)
    st.set_prop(qprecond=2, qepsilon=eps, qmaxiters=maxiters, resnorm=0)
    t = time.perf_counter()
    impl.cgsolveq(st)
    sec = time.perf_counter() - t
    return sec, impl.stats("q")[0], st.get("qc")


def cpu_baseline(st, x64, x_fixed, gpu_iters: int, x_h=None):
    """The reference's CGSolveQ on the full workload on one host core (from
    q = 0 to 1e-10), and the device's result against it: the converged q
    (x64) and the q after exactly 1 and 3 PCG iterations (x_fixed).  Also
    the free-surface CGSolve of the same grid (hprecond 1, D depth-
    integrated, h = 0, 1e-10), against the device's h (x_h) when given."""
    impl, kind = _checkers()
    src0 = st.get("stmp")
    m = st.wet_mask()
    fixed = {}
    for n in (1, 3):
        sec, _, qr = _cpu_q_solve(impl, st, src0, n, 0.0)
        fixed[n] = {"cpu_s": sec, "rel_err_q": relerr(x_fixed[n][m], qr[m])}
        log(f"cpu {kind}: {n} iteration(s) in {sec:.2f}s, device rel err {fixed[n]['rel_err_q']:.2e}")
    sec, its, qr = _cpu_q_solve(impl, st, src0, 2000, 1e-10)
    err = relerr(x64[m], qr[m])
    log(f"cpu {kind}: full solve {its} iterations in {sec:.1f}s; device q rel err {err:.2e}")
    st.set("stmp", src0)
    st.set("qc", 0.0)
    from suntans_amd import cases
    D0, h0 = st.get("D"), st.get("h")
    st.set("D", cases.depth_integrated_D(st))
    st.set("h", 0.0)
    st.set_prop(hprecond=1, epsilon=1e-10, maxiters=1000, resnorm=0)
    t = time.perf_counter()
    impl.cgsolve(st)
    hsec = time.perf_counter() - t
    hits = impl.stats("h")[0]
    herr = relerr(x_h, st.get("h")) if x_h is not None else None
    log(f"cpu {kind}: CGSolve {hits} iterations in {hsec:.2f}s; device h rel err {herr}")
    st.set("D", D0)
    st.set("h", h0)
    return {
        "value": 1.0 / sec, "unit": "solves/s", "cores": 1, "kind": kind, "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
        "cgsolve_s": hsec, "cgsolve_iterations": hits, "rel_err_h_vs_reference": herr,
        "sample": f"one full CGSolveQ of the workload from q=0 to 1e-10 on one core ({its} iterations, {sec:.1f} s)",
        "s_per_solve": sec, "iterations": its, "device_iterations": gpu_iters,
        "rel_err_q_vs_reference": err,
        "fixed_iterations_rel_err_q": {str(n): v["rel_err_q"] for n, v in fixed.items()},
    }


def _core_worker(seed, nx, ny, target_N, barrier, out):
    """One process of the all-core leg: build a reduced large-solve case,
    wait for the others, solve it to 1e-10 by the CPU checker."""
    from suntans_amd import cases
    impl, _ = _checkers()
    st = cases.large_solve(seed=seed, nx=nx, ny=ny, target_N=target_N)
    st.set_prop(qprecond=2, qepsilon=1e-10, qmaxiters=2000, resnorm=0)
    barrier.wait()
    t0 = time.time()
    impl.cgsolveq(st)
    out.put((st.N, impl.stats("q")[0], t0, time.time()))


def cpu_all_cores(full_N: int, full_iters: int, one_core_s: float, nx=250, ny=125):
    """All-core throughput of the CPU checker: one process per core, each
    solving its own reduced case of the workload (nx x ny quads, the same
    bathymetry statistics) to 1e-10 at the same time.  Unknown-iterations per
    second over the whole set, stated also as full-workload solves/s."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    P = host_cores()
    tN = 5.0e7 * (nx * ny) / (1000 * 500)
    barrier, out = ctx.Barrier(P), ctx.Queue()
    procs = [ctx.Process(target=_core_worker, args=(1000 + i, nx, ny, tN, barrier, out)) for i in range(P)]
    for p in procs:
        p.start()
    res = [out.get(timeout=900) for _ in procs]
    for p in procs:
        p.join()
    start, end = min(r[2] for r in res), max(r[3] for r in res)
    work = sum(r[0] * r[1] for r in res)
    rate = work / (end - start)
    one = full_N * full_iters / one_core_s
    log(f"cpu all cores: {P} processes, {work:.3e} unknown-iterations in {end - start:.1f}s")
    return {
        "cores": P, "unknown_iterations_per_s": rate, "one_core_unknown_iterations_per_s": one,
        "full_workload_solves_per_s": rate / (full_N * full_iters),
        "sample": (f"{P} processes, each one CGSolveQ of its own {nx}x{ny}-quad case (N {min(r[0] for r in res)}"
                   f"-{max(r[0] for r in res)}, seeded bathymetry) to 1e-10, run together; wall {end - start:.1f} s"),
    }


def lockexchange_leg():
    """The lock-exchange trajectory (2000 cells x 50 levels, 20 steps, dt 5 s)
    twice: once with every function from the CPU checker, once with CGSolve,
    UpdateScalars and CGSolveQ from the drop-ins (libsuntans.so; host pack +
    PCIe + device per call) and the rest of the step from the checker.  Per
    call: the checker's time on one core and the drop-in's; the iteration
    counts and final fields of the two trajectories compared."""
    import _impls
    import _trajectory
    impl, kind = _checkers()
    d = _impls.dropin()
    a = _trajectory.initial()
    ra = _trajectory.run(a, impl, impl)
    b = _trajectory.initial()
    rb = _trajectory.run(b, d, impl)
    out = {"kind": kind, "steps": len(ra), "cells": a.Nc, "unknowns": a.N}
    for key in ("cgsolve_s", "update_scalars_s", "cgsolveq_s"):
        name = key[:-2]
        out[name] = {"cpu_ms_p50": float(np.median([r[key] for r in ra])) * 1e3,
                     "dropin_ms_p50": float(np.median([r[key] for r in rb[1:]])) * 1e3,
                     "dropin_first_call_ms": rb[0][key] * 1e3}
        out[name]["speedup"] = out[name]["cpu_ms_p50"] / out[name]["dropin_ms_p50"]
    out["h_iterations_max_diff"] = int(max(abs(x["h_iters"] - y["h_iters"]) for x, y in zip(ra, rb)))
    out["q_iterations_max_diff"] = int(max(abs(x["q_iters"] - y["q_iters"]) for x, y in zip(ra, rb)))
    errs = {}
    for name in _trajectory.FINAL_FIELDS:
        ma = _trajectory.final_mask(a, name)
        va, vb = a.get(name), b.get(name)
        errs[name] = relerr(vb[ma], va[ma]) if ma is not None else relerr(vb, va)
    out["final_rel_err"] = errs
    log(f"lock exchange: {out}")
    return out


def traffic_for(kernel: str):
    """HBM bytes per launch of the kernel from the committed rocprofv3 PMC
    summary (copy-calibrated FETCH_SIZE + WRITE_SIZE), None if absent."""
    from suntans_amd import profiling
    if not os.path.exists(PMC_SUMMARY):
        return None
    t = profiling.load_traffic(PMC_SUMMARY)
    key = kernel.replace(" ", "")
    for name, (f, w) in t.items():
        if name.replace(" ", "") == key:
            return f + w
    return None


def tvd_leg(st, dg, N, Ek, steps=5):
    """One TVD superbee scalar step (HorizontalFaceScalars + UpdateScalars)
    on the large grid, all fields resident; per-kernel times against their
    algorithmic bytes (8 B each): face sums 3E_k + 3N, face values and the
    column's flux terms 7E_k + 3N, column update 9N + 2E_k."""
    u, w = st.get("u"), st.get("w")
    dg.upload("s", st.get("s"))
    dg.upload("u", u)
    dg.upload("utmp2", u)
    dg.upload("wnew", w)
    dg.upload("wtmp2", w)
    dg.upload("dzzold", st.get("dzz"))
    dg.upload("kappa_tv", st.get("kappa_tv"))
    dg.upload("active", np.ones(st.Nc))
    dg.copy("s0", "s")
    th, dt, ka = st.get_prop("theta"), st.get_prop("dt"), st.get_prop("kappa_s")
    dg.update_scalars(th, dt, ka, 3)
    t = time.perf_counter()
    for i in range(steps):
        if i == steps - 1:
            dg.profile(True)
        dg.copy("s", "s0")
        dg.update_scalars(th, dt, ka, 3)
    wall = (time.perf_counter() - t) / steps
    kt = dg.kernel_times()
    dg.profile(False)
    # faces: u, utmp2, dzf, the two cells' values and inflow sums in, SfHp, SfHm and the column's
    # flux terms (one per edge-level and side) out; column: its nine cell streams and those terms
    nb = {"tvd_sums": 8 * (3 * Ek + 3 * N), "tvd_faces": 8 * (7 * Ek + 3 * N), "tvd_column": 8 * (9 * N + 2 * Ek)}
    res = {"step_ms": wall * 1e3, "steps_per_s": 1 / wall}
    for k, b in nb.items():
        if k in kt:
            ms = kt[k][0] / kt[k][1]
            res[k] = {"ms": ms, "bytes": b, "GBps": b / (ms * 1e-3) / 1e9, "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    log(f"tvd: {res}")
    return res


def upred_leg(st, dg, N, Ek, steps=3):
    """UPredictor on the large grid, all fields resident: the per-edge
    vertical systems (viscosity, drag, vertical advection at thetaM = 0.5)
    solved for U*** and E, then the free-surface source, PCG, drying and
    velocity update.  Seeded inputs as cases.upredictor_inputs draws them."""
    from suntans_amd import device
    rng = np.random.default_rng(11)
    u = st.get("u")
    ne, nflat = st.Ne, len(st.get("dzz"))
    dg.upload("u", u)
    dg.upload("u0", u)
    dg.upload("w", st.get("w"))
    dg.upload("utmp", rng.normal(0, 0.05, len(u)) * st.edge_wet_mask())
    dg.copy("up_src", "utmp")
    dg.upload("nu_tv", rng.uniform(0, 1e-3, nflat))
    dg.upload("nu_lax", rng.uniform(0, 1e-4, nflat))
    dg.upload("CdB", rng.uniform(1e-3, 5e-3, ne))
    dg.upload("CdT", rng.uniform(0.0, 2e-3, ne))
    dg.upload("tau_T", rng.normal(0, 1e-4, ne))
    dg.upload("dv", st.get("dv"))
    dg.upload("h", np.zeros(st.Nc))
    P = device.UPredParams(st.get_prop("dt"), st.get_prop("theta"), st.get_prop("grav"), 1e-4, 1.0, 0.5, 1)
    tp, ts, it = [], [], []
    for _ in range(steps + 1):
        dg.copy("utmp", "up_src")
        dg.copy("u", "u0")
        dg.scale("h", "dv", 0.0)
        dg.sync()
        t = time.perf_counter()
        dg.u_predict(P)
        t1 = time.perf_counter()
        st_h = dg.u_solve(P, maxiters=1000, precond=1, epsilon=1e-10)
        t2 = time.perf_counter()
        tp.append(t1 - t)
        ts.append(t2 - t1)
        it.append(st_h["niters"])
    res = {"predict_ms": float(np.median(tp[1:])) * 1e3, "solve_ms": float(np.median(ts[1:])) * 1e3,
           "h_iterations": it[-1], "edges": ne, "edge_layers": int(len(u))}
    log(f"upredictor: {res}")
    return res


def wpred_leg(st, dg, N, Ek, steps=3):
    """WPredictor on the large grid, all fields resident: central
    differencing (nonlinear=2) at an Adams-Bashforth-3 step, horizontal
    diffusion and sidewall drag, implicit vertical diffusion.  Per-kernel
    times against algorithmic bytes (8 B each): cell source 2E_k + 4N (face
    values and velocities once per edge-layer; w, dzz, nu_lax, stmp), column
    12N (w, Cn_W, Cn_W2, q, dzz, stmp, nu_lax, nu_tv in; w, wtmp, Cn_W, Cn_W2
    out)."""
    from suntans_amd import device
    rng = np.random.default_rng(13)
    nw, nflat = N + st.Nc, len(st.get("dzz"))
    u = st.get("u")
    dg.upload("w", st.get("w"))
    dg.copy("w0", "w")
    dg.upload("Cn_W", rng.normal(0, 1e-5, nw))
    dg.copy("wcn0", "Cn_W")
    dg.upload("Cn_W2", rng.normal(0, 1e-5, nw))
    dg.copy("wcn20", "Cn_W2")
    dg.upload("qp", rng.normal(0, 1e-3, nflat))
    dg.upload("nu_tv", rng.uniform(0, 1e-3, nflat))
    dg.upload("nu_lax", rng.uniform(0, 1e-4, nflat))
    dg.upload("utmp2", u)
    dg.set_face_distances(device.face_distances(st))
    P = device.WPredParams(st.get_prop("dt"), st.get_prop("theta"), 1e-4, 0.1, 5e-3, 1.0, 2, 3, 3, 4, 0)
    walls = []
    for i in range(steps + 1):
        dg.copy("w", "w0")
        dg.copy("Cn_W", "wcn0")
        dg.copy("Cn_W2", "wcn20")
        dg.sync()
        if i == steps:
            dg.profile(True)
        t = time.perf_counter()
        dg.w_predict(P)
        walls.append(time.perf_counter() - t)
    kt = dg.kernel_times()
    dg.profile(False)
    res = {"ms": float(np.median(walls[1:])) * 1e3}
    # "other" holds the centres and face-value launches, two per call: time per call
    per_call = {"wp_source": 1, "wp_column": 1, "other": 2}
    for k, b in {"wp_source": 8 * (2 * Ek + 4 * N), "wp_column": 8 * 12 * N, "other": 8 * (2 * Ek + 5 * N)}.items():
        if k in kt:
            ms = kt[k][0] / kt[k][1] * per_call[k]
            res[k] = {"ms": ms, "bytes": b, "GBps": b / (ms * 1e-3) / 1e9, "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    log(f"wpredictor: {res}")
    return res


def step_leg(st, dg, steps=3, mixed=False, inner_tol=1e-4):
    """Consecutive nonhydrostatic time steps on the large grid with every
    field resident, in the reference's order (Solve, main/phys.c:1145-1306):
    SetFluxHeight, UPredictor (free-surface PCG inside), UpdateDZ, Continuity
    into wnew, UpdateScalars (TVD superbee salinity), WPredictor,
    ComputeQSource, CGSolveQ (qprecond=2), Corrector, Continuity.  The
    explicit horizontal momentum source (HorizontalSource, host code) is a
    fixed seeded field, and the open-boundary hooks do nothing on the closed
    grid.  Stage times are host wall clock around synchronised calls.  mixed:
    the pressure solve in the mixed-precision form (same 1e-10 FP64
    residual)."""
    from suntans_amd import device
    rng = np.random.default_rng(17)
    nw, nflat, ne = st.N + st.Nc, len(st.get("dzz")), st.Ne
    u = st.get("u")
    dg.set_extents(st)
    for f in ("u", "utmp2", "dzf", "dzz", "dzzold", "s", "h", "dv"):
        dg.upload(f, st.get(f))
    w = st.get("w")
    for f in ("w", "wnew", "wtmp2"):
        dg.upload(f, w)
    dg.upload("utmp", rng.normal(0, 0.05, len(u)) * st.edge_wet_mask())
    dg.copy("up_src", "utmp")
    dg.upload("kappa_tv", np.zeros(nflat))
    dg.upload("nu_tv", rng.uniform(0, 1e-3, nflat))
    dg.upload("nu_lax", rng.uniform(0, 1e-4, nflat))
    dg.upload("CdB", rng.uniform(1e-3, 5e-3, ne))
    dg.upload("CdT", rng.uniform(0.0, 2e-3, ne))
    dg.upload("tau_T", rng.normal(0, 1e-4, ne))
    dg.upload("hcorr", np.zeros(st.Nc))
    dg.upload("active", np.ones(st.Nc))
    dg.upload("Cn_W", np.zeros(nw))
    dg.upload("Cn_W2", np.zeros(nw))
    dg.upload("qp", np.zeros(nflat))
    dg.upload("q", np.zeros(nflat))
    dg.set_face_distances(device.face_distances(st))
    g = st.get_prop
    th, dt = g("theta"), g("dt")
    PU = device.UPredParams(dt, th, g("grav"), 1e-4, 1.0, 0.5, 2)
    stages = ("flux_height", "u_predict", "h_solve", "update_dz", "continuity_wnew", "scalars", "w_predict",
              "q_source", "q_solve", "corrector", "continuity")
    rows, its = [], []
    for i in range(steps + 1):
        PW = device.WPredParams(dt, th, 1e-4, 0.1, 5e-3, 1.0, 2, min(i + 1, 3), 3, 4, 0)
        tm = {}

        def run(name, fn):
            t = time.perf_counter()
            r = fn()
            dg.sync()
            tm[name] = (time.perf_counter() - t) * 1e3
            return r

        dg.copy("utmp", "up_src")
        dg.sync()
        run("flux_height", lambda: dg.set_flux_height(st))
        run("u_predict", lambda: dg.u_predict(PU))
        sh = run("h_solve", lambda: dg.u_solve(PU, maxiters=1000, precond=1, epsilon=1e-10))
        run("update_dz", lambda: dg.update_dz(st, 0))
        run("continuity_wnew", lambda: dg.continuity(th, "wnew"))
        run("scalars", lambda: dg.update_scalars(th, dt, 0.0, 4, step=i + 1))
        run("w_predict", lambda: dg.w_predict(PW))
        run("q_source", lambda: dg.compute_qsource(th, dt, "src"))
        if mixed:
            sq = run("q_solve", lambda: dg.q_solve_mixed(epsilon=1e-10, inner_tol=inner_tol))
        else:
            sq = run("q_solve", lambda: dg.q_solve(maxiters=2000, precond=2, epsilon=1e-10))
        run("corrector", lambda: dg.corrector(dt, "q", "qp"))
        run("continuity", lambda: dg.continuity(th, "w"))
        rows.append(tm)
        its.append((sh["niters"], sq["niters"]))
    med = {k: float(np.median([r[k] for r in rows[1:]])) for k in stages}
    total = [sum(r.values()) for r in rows[1:]]
    res = {"pressure_solve": "mixed" if mixed else "fp64",
           "step_ms": float(np.median(total)), "steps_per_s": 1e3 / float(np.median(total)),
           "stages_ms": med, "h_iterations": [x[0] for x in its], "q_iterations": [x[1] for x in its],
           "w_finite": bool(np.isfinite(dg.download("w", nw)).all())}
    log(f"time step: {res}")
    return res


def pcie_bandwidth(dg, nbytes=1 << 30, reps=5):
    """Host-to-device and device-to-host copy rates (GB/s) between pinned host
    memory and HBM."""
    h2d, d2h = dg.pcie_probe(nbytes, reps)
    return {"h2d": nbytes / h2d / 1e6, "d2h": nbytes / d2h / 1e6}


def dropin_large_leg(st, dg, resident_ms):
    """The reference signatures (libsuntans.so) on the full workload's host
    state: CGSolveQ, CGSolve and UpdateScalars per call, each = pack the
    reference's ragged arrays, copy to HBM, solve, copy back and unpack.
    Beside each: the resident device time of the same call and the PCIe
    floor (the bytes the call moves / the measured pinned copy rates, one
    direction after the other: a call whose downloads overlap its uploads,
    as UpdateScalars' face values do, can come in under it).
    Bytes per call (8 B each): CGSolveQ in q, src, dzz (N each), dzf (E_k),
    D (Ne), out q (N); CGSolve in h, htmp, D (Nc, Nc, Ne), out h (Nc);
    UpdateScalars in s, dzz, dzzold, kappa_tv (N), wnew, wtmp2 (N + Nc), u,
    utmp2, dzf (E_k), active (Nc), out s, Cn (N), SfHp, SfHm (E_k)."""
    import ctypes as C
    from suntans_amd import cases, native
    L = native.dropin_lib()
    bw = pcie_bandwidth(dg)
    N, Ek, Nc, Ne = st.N, st.scalar("Ek"), st.Nc, st.Ne
    moved = {"CGSolveQ": (8 * (3 * N + Ek + Ne), 8 * N), "CGSolve": (8 * (2 * Nc + Ne), 8 * Nc),
             "UpdateScalars": (8 * (6 * N + 2 * Nc + 3 * Ek + Nc), 8 * (2 * N + 2 * Ek))}
    src0, D0, s0, h0 = st.get("stmp"), st.get("D"), st.get("s"), st.get("h")
    Dh = cases.depth_integrated_D(st)
    n_ = C.c_int()
    r_ = C.c_double()
    ms_ = C.c_double()

    def q_call():
        st.set("qc", 0.0)
        st.set("stmp", src0)
        st.set("D", D0)
        t = time.perf_counter()
        L.CGSolveQ(st.ptr("qc"), st.ptr("stmp"), st.ptr("stmp3"), st.g, st.p, st.prop, 0, 1, 0)
        return time.perf_counter() - t

    def h_call():
        st.set("h", h0)
        st.set("D", Dh)
        t = time.perf_counter()
        L.CGSolve(st.g, st.p, st.prop, 0, 1, 0)
        return time.perf_counter() - t

    def s_call():
        st.set("s", s0)
        g = st.get_prop
        t = time.perf_counter()
        L.UpdateScalars(st.g, st.p, st.prop, st.ptr("wnew"), st.ptr("s"), st.ptr("boundary_s"), st.ptr("Cn_R"),
                        g("kappa_s"), g("kappa_sH"), st.ptr("kappa_tv"), g("theta"), None, None, None, None,
                        0, 0, 0, 0, 1, 3)
        return time.perf_counter() - t

    st.set_prop(qprecond=2, qepsilon=1e-10, qmaxiters=2000, resnorm=0, hprecond=1, epsilon=1e-10, maxiters=1000,
                n=3)
    out = {"pcie_GBps": bw}
    for name, fn, which in (("CGSolveQ", q_call, 0), ("CGSolve", h_call, 1), ("UpdateScalars", s_call, None)):
        ts = [fn() for _ in range(4)]
        dev = None
        if which is not None:
            L.sunhip_last_solve(which, C.byref(n_), C.byref(r_), C.byref(ms_))
            dev = ms_.value
        up, down = moved[name]
        floor = (up / bw["h2d"] + down / bw["d2h"]) / 1e6
        call = float(np.median(ts[1:])) * 1e3
        res = resident_ms.get(name)
        out[name] = {"call_ms_p50": call, "first_call_ms": ts[0] * 1e3, "resident_ms": res,
                     "device_ms_last_call": dev, "bytes_h2d": up, "bytes_d2h": down, "pcie_floor_ms": floor,
                     "overhead_ms": call - res if res is not None else None,
                     "overhead_minus_floor_ms": call - res - floor if res is not None else None}
        log(f"drop-in {name}: {out[name]}")
    st.set("stmp", src0)
    st.set("D", D0)
    st.set("s", s0)
    st.set("h", h0)
    st.set("qc", 0.0)
    return out


def ensemble_specs(n: int, seed: int):
    """Member sizes of the ensemble: cells U[5e3, 2e5], aspect U[1, 4],
    levels U{20..100} (seeded)."""
    rng = np.random.default_rng(seed)
    return [(seed * 1000 + i, int(rng.uniform(5e3, 2e5)), float(rng.uniform(1, 4)), int(rng.integers(20, 101)))
            for i in range(n)]


def build_members(specs, threads: int):
    """Build the members' host states on a thread pool and their device grids
    (everything a time step reads) on the calling thread, freeing each host
    state once uploaded.  Returns serve.Sim objects."""
    from concurrent.futures import ThreadPoolExecutor
    from suntans_amd import cases, serve
    sims = []
    with ThreadPoolExecutor(threads) as ex:
        for i, st in enumerate(ex.map(lambda sp: cases.ensemble_member(*sp), specs)):
            sims.append(serve.sim_from_state(i, st, keep_host=False))
            del st
    return sims


def ensemble_leg(args, sims):
    """64 independent members, each taking the nonhydrostatic part of a time
    step on the device: the Poisson source of its predicted velocities
    (ComputeQSource), a mixed-precision pressure solve (FP32 inner PCG in
    FP64 refinement), Corrector + Continuity (divergence-free u and w), then
    one TVD superbee scalar step transported by them.  Every stage runs all
    members in shared launches (one persistent-kernel batch).  Also timed:
    the solves alone (mixed and FP64 batches, mixed one grid at a time)."""
    from suntans_amd import device
    dgs = [s.dg for s in sims]
    N = sum(s.N for s in sims)
    b = device.Batch()
    theta = [s.phys["theta"] for s in sims]
    dt = [s.phys["dt"] for s in sims]
    for dg in dgs:  # base velocities, restored before every timed run
        dg.copy("u0", "u")
        dg.copy("w0", "w")

    def reset():
        for dg in dgs:
            dg.copy("src", "src0")
            dg.copy("D", "Dq")
            dg.copy("u", "u0")
            dg.copy("w", "w0")
            dg.scale("q", "src0", 0.0)
            dg.scale("qp", "src0", 0.0)

    def timed(fn):
        reset()
        for dg in dgs:
            dg.sync()
        b.sync()
        t = time.perf_counter()
        r = fn()
        b.sync()
        for dg in dgs:
            dg.sync()
        return r, time.perf_counter() - t

    tvd = [(s.phys["theta"], s.phys["dt"], s.phys["kappa_s"], 3, 2) for s in sims]

    def scalar_step():  # transported by the corrected velocities
        for dg in dgs:
            dg.copy("wnew", "w")
        b.update_scalars(dgs, tvd)

    def member_step():
        b.compute_qsource(dgs, theta, dt)
        st = b.q_solve_mixed(dgs, epsilon=1e-10, inner_tol=args.inner_tol)
        b.correct(dgs, theta, dt)
        scalar_step()
        return st

    timed(lambda: b.q_solve_mixed(dgs, epsilon=1e-10, inner_tol=args.inner_tol))  # warm-up, autotuning
    mixed, t_mixed = timed(lambda: b.q_solve_mixed(dgs, epsilon=1e-10, inner_tol=args.inner_tol))
    tuning = b.tuning()
    fp64, t_fp64 = timed(lambda: b.q_solve(dgs, epsilon=1e-10))
    timed(member_step)  # warm-up of the stencil and scalar launches
    step, t_step = timed(member_step)
    # stage times of one member step (each stage synchronised on its own)
    reset()
    b.sync()
    stages = {}
    for name, fn in (("source", lambda: b.compute_qsource(dgs, theta, dt)),
                     ("mixed_solve", lambda: b.q_solve_mixed(dgs, epsilon=1e-10, inner_tol=args.inner_tol)),
                     ("corrector_continuity", lambda: b.correct(dgs, theta, dt)),
                     ("tvd_superbee", scalar_step)):
        t = time.perf_counter()
        fn()
        b.sync()
        stages[name + "_ms"] = (time.perf_counter() - t) * 1e3
    for dg in dgs:
        b.detach(dg)
    # the same mixed solves one grid at a time (each on its own stream)
    reset()
    for dg in dgs:
        dg.sync()
    t = time.perf_counter()
    [dg.q_solve_mixed(epsilon=1e-10, inner_tol=args.inner_tol) for dg in dgs]
    t_seq = time.perf_counter() - t
    b.close()
    return {
        "workload": (f"{len(sims)} members, cells U[5e3,2e5] (aspect U[1,4]), levels U{{20..100}}, "
                     "Gaussian-bump depth; N total = %d; member step = ComputeQSource + mixed CGSolveQ + "
                     "Corrector + Continuity + TVD superbee UpdateScalars" % N),
        "members": len(sims), "N_total": N,
        "Nc_range": [min(s.Nc for s in sims), max(s.Nc for s in sims)],
        "member_step_batched_ms": t_step * 1e3, "member_steps_per_s": len(sims) / t_step,
        "member_step_stages": stages,
        "member_step_all_converged": all(m["converged"] for m in step),
        "mixed_batched_ms": t_mixed * 1e3, "mixed_solves_per_s": len(sims) / t_mixed,
        "mixed_all_converged": all(m["converged"] for m in mixed),
        "mixed_max_true_resid": max(m["resid"] for m in mixed),
        "mixed_inner_iterations_total": int(sum(m["inner_iters"] for m in mixed)),
        "fp64_batched_ms": t_fp64 * 1e3, "fp64_solves_per_s": len(sims) / t_fp64,
        "fp64_iterations_max": max(f["niters"] for f in fp64),
        "mixed_one_at_a_time_ms": t_seq * 1e3,
        "batching_speedup": t_seq / t_mixed,
        "autotuned": tuning,
        "unknowns_per_s_mixed": N / t_mixed,
    }


def serving_leg(args, sims, large_sim):
    """Poisson arrivals of time-step requests (free-surface solve +
    warm-started pressure solve + TVD step) over the ensemble members, 10% of
    them for the large grid; admission by a 250 GB HBM budget.  Swept once
    with FP64 pressure solves and once with the mixed-precision form (same
    1e-10 FP64 residual)."""
    from suntans_amd import serve
    allsims = sims + [large_sim]
    sids = [s.sid for s in allsims]
    weights = [0.9 / len(sims)] * len(sims) + [0.1]
    out = {
        "workload": (f"{len(sims)} ensemble members + 1 grid of the large solve (10% of requests); each request = "
                     "CGSolve + CGSolveQ (q warm start on device) + TVD superbee step, 1e-10"),
        "hbm_budget_GB": 250, "max_slots": 128, "requests_per_rate": args.serve_requests,
    }
    for name, mixed in (("fp64", False), ("mixed", True)):
        res = []
        for rate in args.serve_rates:
            arr = serve.poisson_arrivals(rate, args.serve_requests, sids, weights, seed=int(rate))
            srv = serve.Server(allsims, hbm_budget=250e9, max_slots=128, mixed=mixed)
            done = srv.run(arr)
            srv.close()
            row = {"rate_per_s": rate, **serve.summarize(done)}
            log(f"serving {name} {row}")
            res.append(row)
        out["rates_" + name] = res
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--nx", type=int, default=1000)
    ap.add_argument("--ny", type=int, default=500)
    ap.add_argument("--mixed-steps", type=int, default=3)
    ap.add_argument("--inner-tol", type=float, default=1e-4)
    ap.add_argument("--skip-cpu", action="store_true", help="leave out the CPU baseline leg")
    ap.add_argument("--skip-extra", action="store_true",
                    help="only the FP64 pressure solves (for short profiled runs)")
    ap.add_argument("--skip-allcore", action="store_true", help="leave out the all-core CPU leg")
    ap.add_argument("--skip-steps", action="store_true", help="leave out the whole-time-step legs")
    ap.add_argument("--skip-dropin", action="store_true", help="leave out the drop-in per-call leg")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--skip-ensemble", action="store_true", help="leave out the ensemble and serving legs")
    ap.add_argument("--skip-serve", action="store_true", help="leave out the serving leg")
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[10, 50, 200, 1000, 2000])
    ap.add_argument("--serve-requests", type=int, default=500)
    ap.add_argument("--build-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run the calibration copy kernel (for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by whole columns across GPUs, which is not built here)", file=sys.stderr)
        sys.exit(2)

    from suntans_amd import cases, device

    st = build_case(args)
    N, Ek = st.N, st.scalar("Ek")
    if device.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(3)
    t = time.time()
    dg = device.DeviceGrid(st)
    for name, f in (("q", "qc"), ("src", "stmp"), ("dzz", "dzz"), ("dzf", "dzf"), ("D", "D")):
        dg.upload(name, st.get(f))
    dg.copy("q0", "q")
    dg.sync()
    nflat = dg.sizes()["Nflat"]
    log(f"uploaded in {time.time() - t:.1f}s, device bytes {dg.bytes() / 2**30:.2f} GiB")

    prm = dict(maxiters=2000, precond=2, resnorm=0, epsilon=1e-10)
    for w in range(args.warmup):
        dg.copy("q", "q0")
        s = dg.q_solve(**prm)
        log(f"warmup {w}: {s['niters']} iterations, resid {s['resid']:.2e}, {s['device_ms']:.1f} ms")

    # ---- timed FP64 solves ------------------------------------------------
    # the last timed solve records an event pair around each kernel launch
    # (per-kernel durations); each record costs a few us between launches
    dg.sync()
    stats = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        if k == args.steps - 1:
            dg.profile(True)
        dg.copy("q", "q0")
        stats.append(dg.q_solve(**prm))
    dg.sync()
    wall = time.perf_counter() - t0
    kt = dg.kernel_times()
    dg.profile(False)
    iters = [s["niters"] for s in stats]
    lat = np.array([s["setup_ms"] + s["device_ms"] for s in stats])
    log(f"timed {args.steps} solves in {wall:.3f}s; iterations {iters}; kernels {kt}")
    ms_per_step = wall / args.steps * 1e3
    its = float(np.mean(iters)) + 1  # the loop ran n+1 times when it broke at n
    B = q_bytes(N, Ek)
    iter_ms = float(np.mean([s["device_ms"] for s in stats])) / its
    dom = max(("qapply", "qprec"), key=lambda k: kt.get(k, (0, 0))[0])
    dom_ms = kt[dom][0] / max(kt[dom][1], 1)
    achieved = B[dom] / (dom_ms * 1e-3) / 1e9
    x64 = dg.download("q", nflat)
    traffic = traffic_for(KERNEL_NAMES[dom])
    x_h = None
    # the device's q after exactly 1 and 3 iterations, for the comparison with the reference
    x_fixed = {}
    for n in (1, 3):
        dg.copy("q", "q0")
        dg.q_solve(maxiters=n, precond=2, epsilon=0.0, fixed=n)
        x_fixed[n] = dg.download("q", nflat)

    out = {
        "metric": "Poisson-CG solves/s at 1e-10 rel resid (FP64 CGSolveQ, vertical-line PCG)",
        "value": args.steps / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded bathymetry and velocities)",
        "config": {
            "workload": (f"nonhydrostatic pressure Poisson PCG, {args.nx}x{args.ny} quads as "
                         f"{st.Nc} triangles, 100 levels of 2 m, depth 10-200 m, N={N} unknowns, "
                         "qprecond=2, qepsilon=1e-10, q0=0"),
            "N": N, "E_k": Ek, "Nc": st.Nc,
            "tile_padding_ratio": dg.sizes()["NT"] / N,
            "parallelism": "single GPU",
        },
        "iterations": iters,
        "cg_iterations_per_s": its * args.steps / wall,
        "latency_ms_p50": float(np.percentile(lat, 50)),
        "latency_ms_p99": float(np.percentile(lat, 99)),
        "iteration_ms": iter_ms,
        "iteration_GBps_ref_bytes": B["iteration_ref"] / (iter_ms * 1e-3) / 1e9,
        "iteration_roofline_frac_ref_bytes": B["iteration_ref"] / (iter_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "kernels_ms_per_launch": {k: v[0] / max(v[1], 1) for k, v in kt.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": KERNEL_NAMES[dom],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": B[dom],
            "launch_ms": dom_ms,
        },
    }

    if not args.skip_extra:
        # ---- tight FP64 reference solution --------------------------------
        dg.copy("q", "q0")
        tight = dg.q_solve(maxiters=5000, precond=2, resnorm=0, epsilon=1e-14)
        xt = dg.download("q", nflat)
        # ---- mixed precision on the same grid and source ------------------
        dg.copy("q", "q0")
        dg.q_solve_mixed(epsilon=1e-10, inner_tol=args.inner_tol)  # warm-up
        mix = []
        t0 = time.perf_counter()
        for k in range(args.mixed_steps):
            if k == args.mixed_steps - 1:
                dg.profile(True)
            dg.copy("q", "q0")
            mix.append(dg.q_solve_mixed(epsilon=1e-10, inner_tol=args.inner_tol))
        dg.sync()
        mwall = time.perf_counter() - t0
        mkt = dg.kernel_times()
        dg.profile(False)
        xm = dg.download("q", nflat)
        mlat = np.array([s["setup_ms"] + s["device_ms"] for s in mix])
        out["mixed"] = {
            "value": args.mixed_steps / mwall, "unit": "solves/s",
            "speedup_vs_fp64": (args.mixed_steps / mwall) / out["value"],
            "refinement_steps": mix[0]["niters"], "inner_fp32_iterations": mix[0]["inner_iters"],
            "true_resid": mix[0]["resid"], "inner_tol": args.inner_tol,
            "rel_err_q_vs_fp64": relerr(xm, x64), "rel_err_q_vs_tight": relerr(xm, xt),
            "rel_err_fp64_vs_tight": relerr(x64, xt), "tight_iterations": tight["niters"],
            "latency_ms_p50": float(np.percentile(mlat, 50)), "latency_ms_p99": float(np.percentile(mlat, 99)),
            "kernels_ms_per_launch": {k: v[0] / max(v[1], 1) for k, v in mkt.items()},
        }
        log(f"mixed: {out['mixed']}")
        dg.copy("q", "q0")

        # ---- free surface on the same grid (D = depth-integrated) ---------
        dg.upload("D", cases.depth_integrated_D(st))
        dg.upload("htmp", st.get("htmp"))
        g, th, dt = st.get_prop("grav"), st.get_prop("theta"), st.get_prop("dt")
        zero = np.zeros(st.Nc)

        def hrun(fn, n=3, **kw):
            res = []
            for _ in range(n):
                dg.upload("h", zero)
                res.append(fn(g, th, dt, **kw))
            return res, dg.download("h", st.Nc)

        h64, xh = hrun(dg.h_solve, maxiters=1000, precond=1, epsilon=1e-10)
        x_h = xh
        ht, xht = hrun(dg.h_solve, n=1, maxiters=10000, precond=1, epsilon=1e-14)
        hm, xhm = hrun(dg.h_solve_mixed, epsilon=1e-10, inner_tol=args.inner_tol, precond=1)
        ms = lambda rs: float(np.median([s["setup_ms"] + s["device_ms"] for s in rs]))
        out["free_surface"] = {
            "Nc": st.Nc, "iterations": h64[0]["niters"], "fp64_ms_p50": ms(h64),
            "mixed_ms_p50": ms(hm), "mixed_refinement_steps": hm[0]["niters"],
            "mixed_inner_iterations": hm[0]["inner_iters"],
            "rel_err_h_mixed_vs_fp64": relerr(xhm, xh), "rel_err_h_mixed_vs_tight": relerr(xhm, xht),
            "rel_err_h_fp64_vs_tight": relerr(xh, xht),
        }
        dg.upload("D", st.get("D"))
        out["tvd"] = tvd_leg(st, dg, N, Ek)
        out["upredictor"] = upred_leg(st, dg, N, Ek)
        out["wpredictor"] = wpred_leg(st, dg, N, Ek)
        if not args.skip_steps:
            out["time_step"] = step_leg(st, dg)
            out["time_step_mixed"] = step_leg(st, dg, mixed=True, inner_tol=args.inner_tol)
        if not args.skip_dropin:
            out["dropin"] = dropin_large_leg(st, dg, {"CGSolveQ": ms_per_step,
                                                      "CGSolve": out["free_surface"]["fp64_ms_p50"],
                                                      "UpdateScalars": out["tvd"]["step_ms"]})

    if args.copy_probe:
        cp_ms = dg.copy_probe(1 << 30, 20)
        log(f"copy probe: 1 GiB in {cp_ms:.3f} ms = {2 * (1 << 30) / cp_ms / 1e6:.0f} GB/s")
        out["copy_probe_GBps"] = 2 * (1 << 30) / cp_ms / 1e6

    if not args.skip_ensemble:
        from suntans_amd import serve
        t = time.time()
        sims = build_members(ensemble_specs(args.members, 7), args.build_threads)
        log(f"built {len(sims)} members (N total {sum(x.N for x in sims)}) in {time.time() - t:.1f}s")
        out["ensemble"] = ensemble_leg(args, sims)
        log(f"ensemble: {out['ensemble']}")
        if not args.skip_serve:
            large = serve.sim_from_state(len(sims), st, keep_host=False)
            out["serving"] = serving_leg(args, sims, large)
            large.dg.close()
        for x in sims:
            x.dg.close()

    if not args.skip_cpu:
        cb = cpu_baseline(st, x64, x_fixed, iters[-1], x_h)
        out["rel_err_vs_reference"] = {"converged_q": cb["rel_err_q_vs_reference"],
                                       **{f"q_after_{k}_iterations": v
                                          for k, v in cb["fixed_iterations_rel_err_q"].items()},
                                       "iterations_reference": cb["iterations"], "iterations_device": iters[-1]}
        if not args.skip_allcore:
            cb["all_cores"] = cpu_all_cores(N, cb["iterations"], cb["s_per_solve"])
        cb["lockexchange"] = lockexchange_leg()
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

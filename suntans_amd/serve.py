"""Continuous-batching serving of time-step requests on one GPU.

Each request asks one simulation (an ensemble member, or a large grid) for
one time step of the path this package accelerates:

    free-surface PCG (CGSolve)  ->  nonhydrostatic pressure PCG (CGSolveQ,
    warm-started from the simulation's previous q, kept on the device; FP64,
    or the mixed-precision form to the same 1e-10 FP64 residual)
    ->  one TVD scalar step (UpdateScalars)

A step's forcing is built on the device from the simulation's resident base
fields (``htmp = f_k * htmp0``, ``src = f_k * src0``), so a request moves no
field data over PCIe.

The loop (:class:`Server`) is iteration-level continuous batching on top of
the batched engine (``sunhip_batch_*``):

* requests that arrive while others are on the GPU are admitted as soon as a
  slot is free and their simulation is resident, between two chunks of CG
  iterations; their solves join the same launches as the ones in flight;
* a simulation is made resident on admission when its HBM footprint fits the
  budget, evicting idle simulations (least recently used first, their q, h and
  scalar saved to host memory and restored on readmission) when it does not;
  a request that cannot be placed waits in the queue;
* requests of one simulation run in arrival order, one at a time.

Per request the latency is completion minus arrival (queueing included);
completion is when its results are resident on the device.  Because every
member's arithmetic is independent of the batch it runs in, a simulation's
results are bit-identical whether its requests run alone or in the mix.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import cases, device


def forcing(sid: int, step: int) -> float:
    """Deterministic forcing factor of a simulation's step."""
    return 1.0 + 0.1 * np.sin(0.7 * step + 0.37 * sid)


def estimate_bytes(st) -> int:
    """HBM estimate of a resident simulation before it is built (about 34
    FP64 values per cell-layer for the solver, scalar, velocity and face
    fields)."""
    return int(8 * 34 * st.N * 1.1)


def make_device_grid(st, device_id: int = 0) -> device.DeviceGrid:
    """A DeviceGrid holding everything a time step reads: solver inputs, the
    two coefficient fields D (free surface: depth-integrated; pressure:
    df/dg), base forcings, the scalar-transport inputs and the predicted
    vertical velocity w (the velocities the pressure source and Corrector
    read)."""
    dg = device.DeviceGrid(st, device_id)
    dg.upload("w", st.get("w"))
    dg.upload("q", st.get("qc"))
    dg.upload("src0", st.get("stmp"))
    dg.upload("dzf", st.get("dzf"))
    dg.upload("Dq", st.get("D"))
    dg.upload("Dh", cases.depth_integrated_D(st))
    dg.upload("htmp0", st.get("htmp"))
    device.member_upload_tvd(dg, st)
    return dg


@dataclass
class Sim:
    sid: int
    N: int
    Nc: int
    phys: dict                       # grav, theta, dt, kappa, TVD
    make_grid: object = None         # () -> DeviceGrid, to (re)admit
    dg: object = None                # resident DeviceGrid or None
    est_bytes: int = 0
    step: int = 0
    busy: bool = False
    last_used: float = 0.0
    saved: dict = field(default_factory=dict)   # q, s, h kept while evicted

    def footprint(self) -> int:
        return max(self.dg.bytes(), self.est_bytes) if self.dg is not None else self.est_bytes


@dataclass
class Request:
    rid: int
    sid: int
    t_arrival: float
    step: int = -1
    t_admit: float = 0.0
    t_done: float = 0.0
    h_iters: int = 0
    q_iters: int = 0


def sim_from_state(sid: int, st, keep_host: bool, device_id: int = 0, resident: bool = True) -> Sim:
    phys = {k: st.get_prop(k) for k in ("grav", "theta", "dt", "kappa_s")}
    s = Sim(sid=sid, N=st.N, Nc=st.Nc, phys=phys, est_bytes=estimate_bytes(st))
    if keep_host:
        s.make_grid = lambda st=st: make_device_grid(st, device_id)
    if resident:
        s.dg = make_device_grid(st, device_id)
        if not keep_host:
            s.dg.st = None
    return s


def sim_from_files(sid: int, datadir: str, restart: str = None, keep_host: bool = False, device_id: int = 0,
                   dt: float = 5.0) -> Sim:
    """A simulation warm-started from the reference's own files: the
    run-phase grid its GetGrid wrote in datadir and a restart record (its
    store file) holding h, u, w, q, qc, s and the Adams-Bashforth histories
    (cases.warm_start)."""
    return sim_from_state(sid, cases.warm_start(datadir, restart, dt=dt), keep_host, device_id)


class Server:
    """Continuous-batching loop over a set of simulations on one GPU."""

    def __init__(self, sims, hbm_budget: float = 250e9, max_slots: int = 128, device_id: int = 0,
                 epsilon: float = 1e-10, tvd: int = 3, mixed: bool = False, inner_tol: float = 1e-4):
        self.mixed = mixed
        self.inner_tol = inner_tol
        self.sims = {s.sid: s for s in sims}
        self.budget = hbm_budget
        self.max_slots = max_slots
        self.eps = epsilon
        self.tvd = tvd
        self.batch = device.Batch(device_id)
        self.evictions = 0
        self.admissions = 0

    def close(self):
        self.batch.close()

    # -- residency ----------------------------------------------------------
    def _resident_bytes(self) -> int:
        return sum(s.footprint() for s in self.sims.values() if s.dg is not None)

    def _evict(self, s: Sim):
        nflat = s.dg.sizes()["Nflat"]
        s.saved = {"q": s.dg.download("q", nflat), "s": s.dg.download("s", nflat), "h": s.dg.download("h", s.Nc)}
        s.dg.close()
        s.dg = None
        self.evictions += 1

    def _make_resident(self, s: Sim) -> bool:
        """Make s resident, evicting idle rebuildable simulations (least
        recently used first) while the budget is short.  False when it does
        not fit now; an error when it never can (it cannot be rebuilt)."""
        if s.dg is not None:
            return True
        if s.make_grid is None:
            raise RuntimeError(f"simulation {s.sid} is not resident and has no host state to rebuild it from")
        need = s.est_bytes
        if need > self.budget:
            raise RuntimeError(f"simulation {s.sid} needs {need / 1e9:.1f} GB, over the {self.budget / 1e9:.1f} GB "
                               "budget")
        while self._resident_bytes() + need > self.budget:
            # only simulations that can be rebuilt may leave the device
            idle = [x for x in self.sims.values() if x.dg is not None and not x.busy and x.make_grid is not None]
            if not idle:
                return False
            self._evict(min(idle, key=lambda x: x.last_used))
        s.dg = s.make_grid()
        for name, arr in s.saved.items():
            s.dg.upload(name, arr)
        s.saved = {}
        self.admissions += 1
        return True

    # -- stages ---------------------------------------------------------------
    def _start_h(self, r: Request, tag: int):
        s = self.sims[r.sid]
        dg = s.dg
        f = forcing(s.sid, r.step)
        dg.copy("D", "Dh")
        dg.scale("htmp", "htmp0", f)
        dg.scale("h", "htmp0", 0.0)
        p = s.phys
        self.batch.begin_h(dg, tag, p["grav"], p["theta"], p["dt"], epsilon=self.eps)

    def _start_q(self, r: Request, tag: int):
        s = self.sims[r.sid]
        dg = s.dg
        dg.copy("D", "Dq")
        dg.scale("src", "src0", forcing(s.sid, r.step))
        if self.mixed:
            self.batch.begin_q_mixed(dg, tag, epsilon=self.eps, inner_tol=self.inner_tol)
        else:
            self.batch.begin_q(dg, tag, epsilon=self.eps)

    def _scalar_step(self, r: Request):
        s = self.sims[r.sid]
        p = s.phys
        s.dg.update_scalars(p["theta"], p["dt"], p["kappa_s"], self.tvd, step=2)

    # -- loop -------------------------------------------------------------------
    def run(self, arrivals, poll_every: int = 1):
        """arrivals: list of (t_arrival seconds from start, sid), sorted.
        Returns the completed requests."""
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        queue = deque()
        inflight = {}
        done = []
        i = 0
        rid = 0
        while i < len(arrivals) or queue or inflight:
            now = clock()
            while i < len(arrivals) and arrivals[i][0] <= now:
                queue.append(Request(rid=rid, sid=arrivals[i][1], t_arrival=arrivals[i][0]))
                rid += 1
                i += 1
            # admission between chunks: FIFO, one request per simulation at a time
            admitted = 0
            for r in list(queue):
                if len(inflight) >= self.max_slots:
                    break
                s = self.sims[r.sid]
                if s.busy or not self._make_resident(s):
                    continue
                queue.remove(r)
                s.busy = True
                r.step = s.step
                s.step += 1
                r.t_admit = clock()
                inflight[2 * r.rid] = r
                self._start_h(r, 2 * r.rid)
                admitted += 1
            if queue and not inflight and not admitted:
                # every simulation is idle, so nothing can free memory for the
                # queued requests: they would wait forever
                raise RuntimeError(f"request for simulation {queue[0].sid} cannot be placed within the "
                                   f"{self.budget / 1e9:.1f} GB budget (resident simulations that cannot be "
                                   "rebuilt hold it)")
            if inflight:
                self.batch.advance(poll_every)
                for tag, kind, st in self.batch.poll():
                    r = inflight.pop(tag)
                    if kind == "h":
                        r.h_iters = st["niters"]
                        inflight[tag + 1] = r
                        self._start_q(r, tag + 1)
                    else:
                        r.q_iters = st["niters"]
                        self._scalar_step(r)
                        r.t_done = clock()
                        s = self.sims[r.sid]
                        s.busy = False
                        s.last_used = r.t_done
                        done.append(r)
            elif i < len(arrivals):
                wait = arrivals[i][0] - clock()
                if wait > 0:
                    time.sleep(min(wait, 0.05))
        return done


def poisson_arrivals(rate: float, n: int, sids, weights, seed: int = 0):
    """n arrivals of a Poisson process of the given rate (seeded exponential
    gaps), each for a simulation drawn with the given weights."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n)
    t = np.cumsum(gaps) - gaps[0]
    who = rng.choice(np.asarray(sids), size=n, p=np.asarray(weights) / np.sum(weights))
    return [(float(a), int(b)) for a, b in zip(t, who)]


def summarize(done, wall: float = None) -> dict:
    lat = np.array([(r.t_done - r.t_arrival) * 1e3 for r in done])
    first = min(r.t_arrival for r in done)
    last = max(r.t_done for r in done)
    span = wall if wall is not None else (last - first)
    return {
        "requests": len(done),
        "solves_per_s": 2 * len(done) / span,        # free-surface + pressure solve per request
        "requests_per_s": len(done) / span,
        "latency_ms_p50": float(np.percentile(lat, 50)),
        "latency_ms_p99": float(np.percentile(lat, 99)),
        "latency_ms_max": float(lat.max()),
        "mean_q_iters": float(np.mean([r.q_iters for r in done])),
        "mean_h_iters": float(np.mean([r.h_iters for r in done])),
    }

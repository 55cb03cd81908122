"""Batched solves of many grids (sunhip_batch_*, through the C ABI).

* Each member of a batch matches the oracle to the same bar as the
  single-grid drop-in: q and h within 1e-10 relative 2-norm, iteration counts
  within +-1 (the per-member reduction order differs from the oracle's);
  scalar steps bit-identical.
* A member's result is bit-identical alone and in any mix, for every launch
  configuration the autotuner may pick, and when it joins a batch that is
  already iterating (continuous batching).
* Mixed precision over a batch: every member's FP64 true residual (recomputed
  by the oracle's operator) < 1e-10 and error vs a tight oracle solve <= 1e-8.
"""
import numpy as np
import pytest

from suntans_amd import cases, device

import _impls

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def members():
    out = [cases.tiny(s) for s in (0, 1, 2)]
    out.append(cases.tiny(11, nx=23, ny=9, Nkmax=13))
    out.append(cases.lock_exchange(seed=7))
    out.append(cases.ensemble_member(5, 6000, 2.5, 24))
    out.append(cases.ensemble_member(6, 9000, 1.2, 40))
    # other planforms: a batch runs the face-count class of its largest
    # maxfaces, so triangles also run the 8-face kernels when mixed with these
    out.append(cases.tiny(22, nx=12, ny=10, Nkmax=9, planform="quad", split_every=3))
    out.append(cases.tiny(23, nx=12, ny=10, Nkmax=9, planform="hex"))
    out.append(cases.tiny(24, nx=14, ny=9, Nkmax=11, planform="hex", split_every=4, maxfaces=7))
    return out


@pytest.fixture(scope="module")
def states():
    return members()


def grids_for(states, which=("q", "src", "dzz", "dzf", "D")):
    dgs = []
    for st in states:
        dg = device.DeviceGrid(st)
        device.member_upload(dg, st, which)
        dgs.append(dg)
    return dgs


def test_batch_q_matches_oracle(gpu_available, states):
    o = _impls.oracle()
    b = device.Batch()
    dgs = grids_for(states)
    stats = b.q_solve(dgs)
    for st, dg, s in zip(states, dgs, stats):
        a = st.clone()
        a.set_prop(qprecond=2, qepsilon=1e-10, qmaxiters=2000, resnorm=0)
        o.cgsolveq(a)
        m = a.wet_mask()
        q = dg.download("q", len(m))
        err = relerr(q[m], a.get("qc")[m])
        print(f"Nc={st.Nc}: iters oracle={o.stats('q')[0]} batch={s['niters']} err={err:.2e}")
        assert s["converged"] == 1
        assert abs(o.stats("q")[0] - s["niters"]) <= 1
        assert err <= 1e-10
    for dg in dgs:
        dg.close()
    b.close()


def _solve_q(states, idx, tuning=None):
    b = device.Batch()
    if tuning:
        b.set_tuning(*tuning)
    dgs = grids_for([states[i] for i in idx])
    stats = b.q_solve(dgs)
    out = [(dg.download("q", dg.sizes()["Nflat"]), s["niters"]) for dg, s in zip(dgs, stats)]
    for dg in dgs:
        dg.close()
    b.close()
    return out


def test_batch_alone_equals_mix(gpu_available, states):
    alone = [_solve_q(states, [i])[0] for i in range(len(states))]
    mix = _solve_q(states, list(range(len(states))))
    rev = _solve_q(states, list(range(len(states)))[::-1])[::-1]
    for a, m, r in zip(alone, mix, rev):
        assert a[1] == m[1] == r[1]
        assert np.array_equal(a[0], m[0]) and np.array_equal(a[0], r[0])


@pytest.mark.parametrize("tuning", [(1, 0, 2), (2, 4, 4), (4, 8, 8), (4, 0, 16)])
@pytest.mark.parametrize("idx", [(0, 3, 5), (0, 7, 9)], ids=["tri", "polygons"])
def test_batch_independent_of_launch_configuration(gpu_available, states, tuning, idx):
    ref = _solve_q(states, list(idx), (4, 0, 0))
    got = _solve_q(states, list(idx), tuning)
    for a, g in zip(ref, got):
        assert a[1] == g[1]
        assert np.array_equal(a[0], g[0])


def test_batch_join_while_running(gpu_available, states):
    """Continuous batching: a solve admitted after others have iterated for a
    while ends bit-identical to the same solve alone."""
    alone = [_solve_q(states, [i])[0] for i in (3, 4, 6)]
    b = device.Batch()
    dgs = grids_for([states[i] for i in (3, 4, 6)])
    done = {}
    b.begin_q(dgs[0], 0)
    b.advance(6)
    for tag, kind, s in b.poll():
        done[tag] = s
    b.begin_q(dgs[1], 1)
    b.advance(4)
    b.begin_q(dgs[2], 2)
    while b.active():
        b.advance(4)
        for tag, kind, s in b.poll():
            done[tag] = s
    for i, (dg, a) in enumerate(zip(dgs, alone)):
        q = dg.download("q", dg.sizes()["Nflat"])
        assert done[i]["niters"] == a[1]
        assert np.array_equal(q, a[0])
        dg.close()
    b.close()


def test_batch_h_matches_oracle(gpu_available, states):
    o = _impls.oracle()
    sts = [st.clone() for st in states]
    for st in sts:
        st.set("D", cases.depth_integrated_D(st))
    dgs = grids_for(sts, ("h", "htmp", "D"))
    b = device.Batch()
    phys = [(st.get_prop("grav"), st.get_prop("theta"), st.get_prop("dt")) for st in sts]
    stats = b.h_solve(dgs, phys)
    for st, dg, s in zip(sts, dgs, stats):
        a = st.clone()
        a.set_prop(hprecond=1, epsilon=1e-10, maxiters=1000)
        o.cgsolve(a)
        h = dg.download("h", st.Nc)
        err = relerr(h, a.get("h"))
        assert abs(o.stats("h")[0] - s["niters"]) <= 1
        assert err <= 1e-10, err
        dg.close()
    b.close()


def test_batch_mixed(gpu_available, states):
    o = _impls.oracle()
    dgs = grids_for(states)
    b = device.Batch()
    stats = b.q_solve_mixed(dgs, epsilon=1e-10, inner_tol=1e-4)
    for st, dg, s in zip(states, dgs, stats):
        m = st.wet_mask()
        q = dg.download("q", len(m))
        s1 = st.clone()
        o.qcoef(s1)
        s1.set("uold", q)
        o.operatorq(s1, x="uold", y="stmp2")
        s0 = st.clone()
        o.qcoef(s0)
        s0.set("uold", st.get("qc"))
        o.operatorq(s0, x="uold", y="stmp2")
        res = np.linalg.norm((st.get("stmp") - s1.get("stmp2"))[m]) / \
            np.linalg.norm((st.get("stmp") - s0.get("stmp2"))[m])
        tight = st.clone()
        tight.set_prop(qepsilon=1e-14, qmaxiters=20000)
        o.cgsolveq(tight)
        err = relerr(q[m], tight.get("qc")[m])
        print(f"Nc={st.Nc}: outer={s['niters']} inner={s['inner_iters']} resid={res:.2e} err={err:.2e}")
        assert s["converged"] == 1
        assert res < 1e-10 * (1 + 1e-3)
        assert err <= 1e-8
        dg.close()
    b.close()


def test_batch_update_scalars_bitwise(gpu_available, states):
    """One batched scalar step over members with different TVD schemes
    (batched face sums, face values and column kernels) equals the oracle's
    UpdateScalars of each member bit for bit."""
    o = _impls.oracle()
    sts = [st.clone() for st in states]
    schemes = [3, 0, 4, 1, 3, 4, 3, 4, 1, 3][:len(sts)]
    for st, tvd in zip(sts, schemes):
        st.set_prop(TVDsalt=tvd, n=2)
    dgs = []
    for st in sts:
        dg = device.DeviceGrid(st)
        device.member_upload_tvd(dg, st)
        dgs.append(dg)
    b = device.Batch()
    b.update_scalars(dgs, [(st.get_prop("theta"), st.get_prop("dt"), st.get_prop("kappa_s"), tvd, 2)
                           for st, tvd in zip(sts, schemes)])
    for st, dg in zip(sts, dgs):
        a = st.clone()
        o.update_salinity(a)
        s = dg.download("s", len(a.get("s")))
        m = a.wet_mask()
        assert np.array_equal(s[m], a.get("s")[m])
        dg.close()
    b.close()


def _solve_mixed(states, idx):
    b = device.Batch()
    dgs = grids_for([states[i] for i in idx])
    stats = b.q_solve_mixed(dgs, epsilon=1e-10, inner_tol=1e-4)
    out = [(dg.download("q", dg.sizes()["Nflat"]), s["inner_iters"], s["niters"]) for dg, s in zip(dgs, stats)]
    for dg in dgs:
        dg.close()
    b.close()
    return out


def test_batch_mixed_alone_equals_mix(gpu_available, states):
    """Each member refines on its own schedule, so its result does not depend
    on the others: bit-identical alone, in a mix and in reverse order."""
    alone = [_solve_mixed(states, [i])[0] for i in range(len(states))]
    mix = _solve_mixed(states, list(range(len(states))))
    rev = _solve_mixed(states, list(range(len(states)))[::-1])[::-1]
    for a, m, r in zip(alone, mix, rev):
        assert a[1:] == m[1:] == r[1:]
        assert np.array_equal(a[0], m[0]) and np.array_equal(a[0], r[0])


def test_batch_mixed_joins_fp64_batch(gpu_available, states):
    """A mixed-precision solve admitted into a batch that is running FP64 and
    free-surface solves ends bit-identical to the same mixed solve alone."""
    alone = _solve_mixed(states, [6])[0]
    b = device.Batch()
    dg_f64 = grids_for([states[5]])[0]
    sth = states[4].clone()
    sth.set("D", cases.depth_integrated_D(sth))
    dg_h = grids_for([sth], ("h", "htmp", "D"))[0]
    dg_m = grids_for([states[6]])[0]
    b.begin_q(dg_f64, 0)
    b.begin_h(dg_h, 1, sth.get_prop("grav"), sth.get_prop("theta"), sth.get_prop("dt"))
    b.advance(4)
    done = {t: s for t, k, s in b.poll()}
    b.begin_q_mixed(dg_m, 2)
    while b.active():
        b.advance(4)
        for t, k, s in b.poll():
            done[t] = s
    q = dg_m.download("q", dg_m.sizes()["Nflat"])
    assert done[2]["inner"] == alone[1] and done[2]["outer"] == alone[2]
    assert np.array_equal(q, alone[0])
    for dg in (dg_f64, dg_h, dg_m):
        dg.close()
    b.close()

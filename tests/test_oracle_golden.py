"""The oracle (and, on a GPU, the drop-ins) against golden fixtures written by
the real reference (tests/golden/make_golden.py).  These pin the oracle where
the reference tree is absent.

CPU: the oracle must reproduce every fixture bit for bit, iteration counts
included.  GPU: operators and scalar transport bit for bit; converged solves
within 1e-10 relative 2-norm and the iteration counter within +-1 (qprecond=2,
hprecond 0/1) or a 5% band (qprecond=0, a rounding-dominated regime).
"""
import glob
import hashlib
import os

import numpy as np
import pytest

from suntans_amd import cases

import _impls
from golden_io import inputs_digest, restore

HERE = os.path.dirname(os.path.abspath(__file__))
TINY = sorted(glob.glob(os.path.join(HERE, "golden", "tiny*.npz")))
LOCK = os.path.join(HERE, "golden", "lockexchange.npz")


def _run_all(impl, make):
    """Same sequence of calls as tests/golden/make_golden.py."""
    out = {}
    for pc in (2, 0):
        st = make()
        st.set_prop(qprecond=pc)
        impl.cgsolveq(st)
        out[f"q_pc{pc}"] = st.get("qc")
        out[f"q_pc{pc}_iters"] = impl.stats("q")[0]
    for pc in (1, 0):
        st = make()
        st.set("D", cases.depth_integrated_D(st))
        st.set_prop(hprecond=pc)
        impl.cgsolve(st)
        out[f"h_pc{pc}"] = st.get("h")
        out[f"h_pc{pc}_iters"] = impl.stats("h")[0]
    st = make()
    rng = np.random.default_rng(123)
    st.set("stmp", rng.normal(size=st.get("stmp").shape))
    st.set("htmp2", rng.normal(size=st.Nc))
    st.set("D", cases.depth_integrated_D(st))
    o = _impls.oracle()
    o.qcoef(st)        # coefficients by the checker for every implementation
    o.hcoef(st)
    impl.operatorq(st)
    impl.preconditioner(st)
    impl.operatorh(st)
    out["opq_y"] = st.get("stmp2")
    out["prec_y"] = st.get("uold")
    out["oph_y"] = st.get("htmp3")
    out["_wet"] = st.wet_mask()
    for tvd in (0, 1, 2, 3, 4):
        st = make()
        st.set_prop(TVDsalt=tvd, n=2)
        impl.update_salinity(st)
        out[f"s_tvd{tvd}"] = st.get("s")
        out[f"Cn_tvd{tvd}"] = st.get("Cn_R")
        out[f"SfHp_tvd{tvd}"] = st.get("SfHp") * st.edge_wet_mask()
        out[f"SfHm_tvd{tvd}"] = st.get("SfHm") * st.edge_wet_mask()
    return out


BITWISE = ["opq_y", "prec_y", "oph_y"] + [f"{f}_tvd{t}" for f in ("s", "Cn", "SfHp", "SfHm") for t in range(5)]
SOLVES = ["q_pc2", "q_pc0", "h_pc1", "h_pc0"]


@pytest.mark.parametrize("path", TINY, ids=[os.path.basename(p)[:-4] for p in TINY])
def test_oracle_matches_tiny_fixture(path):
    data = np.load(path)
    got = _run_all(_impls.oracle(), lambda: restore(data))
    m = got["_wet"]
    for k in BITWISE:
        a, b = got[k], data[k]
        if k in ("opq_y", "prec_y"):
            a, b = a[m], b[m]
        assert np.array_equal(a, b), k
    for k in SOLVES:
        assert np.array_equal(got[k], data[k]), k
        assert got[k + "_iters"] == int(data[k + "_iters"][0]), k


def test_oracle_matches_lockexchange_fixture():
    data = np.load(LOCK)
    make = lambda: cases.lock_exchange(seed=7)  # noqa: E731
    assert inputs_digest(make()) == data["inputs_sha256"].tobytes(), \
        "the lock-exchange generator changed: regenerate tests/golden with make_golden.py"
    got = _run_all(_impls.oracle(), make)
    for k in SOLVES + ["opq_y", "prec_y", "oph_y", "s_tvd3", "s_tvd4", "Cn_tvd4"]:
        v = np.ascontiguousarray(got[k])
        if k in ("opq_y", "prec_y"):
            # the fixture holds the full arrays; dry layers are zero in both
            pass
        assert hashlib.sha256(v.tobytes()).digest() == data[k + "_sha256"].tobytes(), k
    for k in SOLVES:
        assert got[k + "_iters"] == int(data[k + "_iters"][0]), k


@pytest.mark.gpu
@pytest.mark.parametrize("path", TINY + [LOCK], ids=[os.path.basename(p)[:-4] for p in TINY + [LOCK]])
def test_dropin_matches_fixture(gpu_available, path):
    data = np.load(path)
    if "grid" in data:
        make = lambda: restore(data)  # noqa: E731
    else:
        make = lambda: cases.lock_exchange(seed=7)  # noqa: E731
    got = _run_all(_impls.dropin(), make)
    m = got["_wet"]

    def ref(k):
        if k in data:
            return data[k], None
        return None, (data[k + "_idx"], data[k + "_val"], float(data[k + "_norm"][0]))

    for k in BITWISE:
        full, samp = ref(k)
        a = got[k]
        if full is not None:
            b = full
            if k in ("opq_y", "prec_y"):
                a, b = a[m], b[m]
            assert np.array_equal(a, b), k
        elif samp is not None and k + "_idx" in data:
            idx, val, _ = samp
            assert np.array_equal(a[idx], val), k
    for k in SOLVES:
        full, samp = ref(k)
        a = got[k]
        if full is not None:
            err = np.linalg.norm(a - full) / max(np.linalg.norm(full), 1e-300)
        else:
            idx, val, nrm = samp
            err = max(np.linalg.norm(a[idx] - val) / max(np.linalg.norm(val), 1e-300),
                      abs(np.linalg.norm(a) - nrm) / nrm)
        assert err <= 1e-10, (k, err)
        n_ref = int(data[k + "_iters"][0])
        band = max(1, int(0.05 * n_ref)) if k == "q_pc0" else 1
        assert abs(got[k + "_iters"] - n_ref) <= band, (k, got[k + "_iters"], n_ref)

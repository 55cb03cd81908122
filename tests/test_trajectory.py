"""The oracle over the 20-step lock-exchange trajectory (tests/_trajectory.py)
against the reference's own run of it (tests/golden/lockexchange_steps.npz,
written by tests/golden/make_trajectory.py from oracle/_ref): every step's
iteration counts of CGSolve and CGSolveQ equal, the printed residuals and
residual histories equal to the 7 digits printed, and the final fields bit
for bit (h in full; qc, q, s, u and w through 2048 seeded samples, their
2-norm and their SHA-256)."""
import hashlib
import os

import numpy as np
import pytest

from golden_io import inputs_digest

import _impls
import _trajectory

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "lockexchange_steps.npz")


def fixture():
    return np.load(FIX)


def check_samples(data, key, v):
    idx = data[key + "_idx"]
    assert len(v) == int(data[key + "_len"][0]), key
    assert np.array_equal(v[idx], data[key + "_val"]), key
    assert hashlib.sha256(np.ascontiguousarray(v).tobytes()).digest() == data[key + "_sha256"].tobytes(), key


def test_inputs_match_fixture():
    st = _trajectory.initial()
    assert inputs_digest(st) == fixture()["inputs_sha256"].tobytes()


def test_oracle_trajectory_matches_reference():
    data = fixture()
    o = _impls.oracle()
    st = _trajectory.initial()
    rows = _trajectory.run(st, o, o)
    for w in ("h", "q"):
        assert [r[w + "_iters"] for r in rows] == list(data[w + "_iters"]), w
        got = np.array([r[w + "_resid"] for r in rows])
        # the reference prints "%e": 7 significant digits
        assert np.allclose(got, data[w + "_resid"], rtol=6e-7, atol=0), w
    assert np.array_equal(st.get("h"), data["h"])
    for name in _trajectory.FINAL_FIELDS:
        v = st.get(name)
        m = _trajectory.final_mask(st, name)
        check_samples(data, "final_" + name, v[m] if m is not None else v)


def test_histories_consistent():
    """Each solve's residual history ends at its reported residual, and has
    one entry per pass of the iteration loop (iterations + 1)."""
    data = fixture()
    for w in ("h", "q"):
        lens, hist = data[w + "_hist_len"], data[w + "_hist"]
        ends = np.cumsum(lens)
        assert np.array_equal(lens, data[w + "_iters"] + 1), w
        assert np.allclose(hist[ends - 1], data[w + "_resid"], rtol=6e-7), w
        assert (hist[ends - 1] < 1e-10).all(), w


@pytest.mark.skipif(_impls.reference() is None, reason="oracle/_ref is not built here")
def test_reference_still_matches_fixture():
    """Where the reference is built: its run of the trajectory is the one the
    fixture holds (the generator and the fixture agree)."""
    data = fixture()
    r = _impls.reference()
    st = _trajectory.initial()
    rows = _trajectory.run(st, r, r)
    assert [x["q_iters"] for x in rows] == list(data["q_iters"])
    assert np.array_equal(st.get("h"), data["h"])

"""The 20-step lock-exchange trajectory (tests/_trajectory.py) on the GPU
against the reference's run of it (tests/golden/lockexchange_steps.npz):

* through the drop-ins: CGSolve, UpdateScalars and CGSolveQ from
  libsuntans.so with the rest of each step on the host (the oracle's
  restatement of the reference's host code, bit-identical to it), as a
  SUNTANS build linked against the drop-ins runs;
* resident: every function of the step on the device, the fields staying in
  HBM between steps.

Per step the iteration counts of both solves within +-1 of the reference's.
At the end h, q, qc, u and w within 1e-8 and s within 1e-10 relative 2-norm
(over the wet entries, h in full, the rest over the fixture's 2048 seeded
samples): the solves stop at 1e-10 relative residual with dot products
summed in another order, and those differences carry through 20 coupled
steps.  The measured errors are printed."""
import os

import numpy as np
import pytest

from suntans_amd import device

import _impls
import _trajectory

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lockexchange_steps.npz")
TOL = {"h": 1e-8, "qc": 1e-8, "q": 1e-8, "s": 1e-10, "u": 1e-8, "w": 1e-8}


def relerr(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def check_final(data, fields, label):
    errs = {}
    errs["h"] = relerr(fields["h"], data["h"])
    for name in _trajectory.FINAL_FIELDS:
        if name == "h":
            continue
        idx = data["final_" + name + "_idx"]
        assert len(fields[name]) == int(data["final_" + name + "_len"][0]), name
        errs[name] = relerr(fields[name][idx], data["final_" + name + "_val"])
    print(label, " ".join(f"{k}={v:.2e}" for k, v in errs.items()))
    for k, v in errs.items():
        assert v <= TOL[k], (k, v)


def check_iters(data, h_iters, q_iters):
    assert np.abs(np.array(h_iters) - data["h_iters"]).max() <= 1, (h_iters, list(data["h_iters"]))
    assert np.abs(np.array(q_iters) - data["q_iters"]).max() <= 1, (q_iters, list(data["q_iters"]))


def test_dropin_trajectory(gpu_available):
    data = np.load(FIX)
    o, d = _impls.oracle(), _impls.dropin()
    st = _trajectory.initial()
    rows = _trajectory.run(st, d, o)
    check_iters(data, [r["h_iters"] for r in rows], [r["q_iters"] for r in rows])
    fields = {}
    for name in _trajectory.FINAL_FIELDS:
        v = st.get(name)
        m = _trajectory.final_mask(st, name)
        fields[name] = v[m] if m is not None else v
    check_final(data, fields, "drop-in:")


def test_resident_trajectory(gpu_available):
    data = np.load(FIX)
    a = _trajectory.initial()
    source = a.get("utmp")
    dg = device.DeviceGrid(a)
    _trajectory.upload_state(dg, a)
    hi, qi = [], []
    for n in range(1, _trajectory.STEPS + 1):
        a.set_prop(n=n)
        dg.upload("utmp", source)
        sh, sq = _trajectory.device_step(dg, a)
        hi.append(sh["niters"])
        qi.append(sq["niters"])
    check_iters(data, hi, qi)
    names = {"h": "h", "qc": "q", "q": "qp", "s": "s", "u": "u", "w": "w"}
    ext = dg.extents()
    a.set_ints("ctop", ext["ctop"])  # the host state was not stepped: masks from the device's extents
    a.set_ints("etop", ext["etop"])
    fields = {}
    for name in _trajectory.FINAL_FIELDS:
        v = dg.download(names[name], len(a.get(name)))
        m = _trajectory.final_mask(a, name)
        fields[name] = v[m] if m is not None else v
    dg.close()
    check_final(data, fields, "resident:")

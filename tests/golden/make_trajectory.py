"""Write the trajectory and full-size fixtures from the real reference.

Run where the reference was compiled from its own sources (oracle/_ref):

    python tests/golden/make_trajectory.py              # both fixtures
    python tests/golden/make_trajectory.py lockexchange_steps

* lockexchange_steps.npz -- the 20-step lock-exchange trajectory of
  tests/_trajectory.py with every function from the reference: per step the
  iteration counts, final residuals and per-iteration residual histories of
  CGSolve and CGSolveQ; at the end h in full and, for qc, q, s, u and w,
  2048 seeded samples over the wet entries, their 2-norm and a SHA-256.
* large_solve.npz -- the large single solve (1000x500 quads as 1.0M
  triangles, ~5.0e7 unknowns; cases.large_solve(seed=0)): CGSolveQ
  (qprecond 2, 1e-10) and CGSolve (hprecond 1, 1e-10, D depth-integrated)
  from zero, each with its iteration count, final residual, residual
  history, 2-norm, 4096 seeded samples and SHA-256 of the solution.  About
  a minute of one core and ~25 GB of memory.

Inputs are regenerated from their seeds; the SHA-256 of every input field
is stored so a changed generator is caught.  Data only: no code of the
reference is kept.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from suntans_amd import cases  # noqa: E402
from golden_io import inputs_digest  # noqa: E402

import _impls  # noqa: E402
import _trajectory  # noqa: E402


def sha(v):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(v).tobytes()).digest(), np.uint8)


def samples(data, key, v, n, seed):
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(len(v), size=min(n, len(v)), replace=False))
    data[key + "_idx"] = idx
    data[key + "_val"] = v[idx]
    data[key + "_norm"] = np.array([np.linalg.norm(v)])
    data[key + "_sha256"] = sha(v)
    data[key + "_len"] = np.array([len(v)])


def lockexchange_steps(ref):
    st = _trajectory.initial()
    data = {"inputs_sha256": np.frombuffer(inputs_digest(st), np.uint8)}
    ref.history = True
    try:
        rows = _trajectory.run(st, ref, ref)
    finally:
        ref.history = False
    for w in ("h", "q"):
        data[w + "_iters"] = np.array([r[w + "_iters"] for r in rows], np.int32)
        data[w + "_resid"] = np.array([r[w + "_resid"] for r in rows])
        hist = [r[w + "_hist"] for r in rows]
        data[w + "_hist_len"] = np.array([len(x) for x in hist], np.int32)
        data[w + "_hist"] = np.concatenate([np.asarray(x, float) for x in hist])
    data["h"] = st.get("h")
    for k, name in enumerate(_trajectory.FINAL_FIELDS):
        v = st.get(name)
        m = _trajectory.final_mask(st, name)
        samples(data, "final_" + name, v[m] if m is not None else v, 2048, 100 + k)
    return data


def large_solve(ref):
    st = cases.large_solve(seed=0)
    data = {"inputs_sha256": np.frombuffer(inputs_digest(st), np.uint8), "N": np.array([st.N])}
    ref.history = True
    try:
        q = st.clone()
        q.set_prop(qprecond=2, qepsilon=1e-10, qmaxiters=2000, resnorm=0)
        ref.cgsolveq(q)
        data["q_iters"] = np.array([ref.stats("q")[0]])
        data["q_resid"] = np.array([ref.stats("q")[1]])
        data["q_hist"] = np.array(ref.residuals("q"))
        v = q.get("qc")
        samples(data, "q", v[st.wet_mask()], 4096, 201)
        del q
        h = st.clone()
        h.set("D", cases.depth_integrated_D(h))
        h.set_prop(hprecond=1, epsilon=1e-10, maxiters=1000, resnorm=0)
        ref.cgsolve(h)
        data["h_iters"] = np.array([ref.stats("h")[0]])
        data["h_resid"] = np.array([ref.stats("h")[1]])
        data["h_hist"] = np.array(ref.residuals("h"))
        samples(data, "h", h.get("h"), 4096, 202)
    finally:
        ref.history = False
    return data


def main():
    ref = _impls.reference()
    if ref is None:
        sys.exit("make_trajectory.py needs oracle/_ref/libsunref.so (build it where the reference tree exists)")
    only = set(sys.argv[1:])
    for name, fn in (("lockexchange_steps", lockexchange_steps), ("large_solve", large_solve)):
        if only and name not in only:
            continue
        data = fn(ref)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **data)
        print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)", flush=True)


if __name__ == "__main__":
    main()

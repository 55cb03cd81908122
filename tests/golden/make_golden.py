"""Write the golden fixtures of the elliptic solves and scalar transport.

Run where the reference was compiled from its own sources (oracle/_ref):

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py hex_tiny   # the named ones

Every fixture is data only: the grid parameters and every input field after
set-up, and the outputs the real reference computed from them (CGSolveQ with
qprecond 2 and 0, CGSolve with hprecond 1 and 0, OperatorQ, Preconditioner,
OperatorH, HorizontalFaceScalars and UpdateScalars for TVD 0..4), with the
iteration counts the reference printed.  tests/test_oracle_golden.py rebuilds
the state from the fixture and checks the oracle (and, on a GPU, the
drop-ins) against it.
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
from golden_io import INPUT_FIELDS, INT_FIELDS, capture, inputs_digest  # noqa: E402

import _impls  # noqa: E402


def outputs(make, ref):
    out = {}
    # pressure solves
    for pc in (2, 0):
        st = make()
        st.set_prop(qprecond=pc)
        ref.cgsolveq(st)
        out[f"q_pc{pc}"] = st.get("qc")
        out[f"q_pc{pc}_iters"] = np.array([ref.stats("q")[0]])
    # free surface
    for pc in (1, 0):
        st = make()
        st.set("D", cases.depth_integrated_D(st))
        st.set_prop(hprecond=pc)
        ref.cgsolve(st)
        out[f"h_pc{pc}"] = st.get("h")
        out[f"h_pc{pc}_iters"] = np.array([ref.stats("h")[0]])
    # operators on a seeded vector
    st = make()
    rng = np.random.default_rng(123)
    st.set("stmp", rng.normal(size=st.get("stmp").shape))
    st.set("htmp2", rng.normal(size=st.Nc))
    st.set("D", cases.depth_integrated_D(st))
    ref.qcoef(st)
    ref.operatorq(st)
    ref.preconditioner(st)
    ref.hcoef(st)
    ref.operatorh(st)
    out["opq_x"] = st.get("stmp")
    out["opq_y"] = st.get("stmp2")
    out["prec_y"] = st.get("uold")
    out["oph_x"] = st.get("htmp2")
    out["oph_y"] = st.get("htmp3")
    # scalar transport
    for tvd in (0, 1, 2, 3, 4):
        st = make()
        st.set_prop(TVDsalt=tvd, n=2)
        ref.update_salinity(st)
        out[f"s_tvd{tvd}"] = st.get("s")
        out[f"Cn_tvd{tvd}"] = st.get("Cn_R")
        out[f"SfHp_tvd{tvd}"] = st.get("SfHp") * st.edge_wet_mask()
        out[f"SfHm_tvd{tvd}"] = st.get("SfHm") * st.edge_wet_mask()
    return out


def main():
    ref = _impls.reference()
    if ref is None:
        sys.exit("make_golden.py needs oracle/_ref/libsunref.so (build it where the reference tree exists)")
    specs = [(f"tiny{s}", dict(kind="tiny", seed=s)) for s in range(4)]
    specs.append(("tiny_wide", dict(kind="tiny", seed=11, nx=23, ny=9, Nkmax=13)))
    specs.append(("lockexchange", dict(kind="lock_exchange", seed=7)))
    # other planforms (sun_grid_polygons)
    specs.append(("tiny_quadtri", dict(kind="tiny", seed=22, nx=7, ny=5, planform="quad", split_every=3)))
    specs.append(("tiny_hexquad7", dict(kind="tiny", seed=24, nx=7, ny=6, planform="hex", split_every=4,
                                        maxfaces=7)))
    only = set(sys.argv[1:])
    for name, spec in specs:
        if only and name not in only:
            continue
        kw = {k: v for k, v in spec.items() if k != "kind"}
        make = (lambda kw=kw, spec=spec: getattr(cases, spec["kind"])(**kw))
        st = make()
        outs = outputs(make, ref)
        if spec["kind"] == "tiny":
            data = capture(st, kw.get("split_every", 0), kw.get("maxfaces") or 0)
            data.update(outs)
        else:
            # large case: inputs are regenerated from the seed (their digest is
            # stored); outputs as digests, norms and 2048 seeded samples
            data = {"inputs_sha256": np.frombuffer(inputs_digest(st), np.uint8)}
            rng = np.random.default_rng(99)
            for k, v in outs.items():
                if k.endswith("_iters"):
                    data[k] = v
                    continue
                idx = np.sort(rng.choice(len(v), size=min(2048, len(v)), replace=False))
                data[k + "_sha256"] = np.frombuffer(hashlib.sha256(np.ascontiguousarray(v).tobytes()).digest(), np.uint8)
                data[k + "_norm"] = np.array([np.linalg.norm(v)])
                data[k + "_idx"] = idx
                data[k + "_val"] = v[idx]
                data[k + "_len"] = np.array([len(v)])
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **data)
        print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


if __name__ == "__main__":
    main()

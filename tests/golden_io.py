"""Save and restore the complete input state of a golden fixture."""
import numpy as np

from suntans_amd.state import State

# physT / gridT real fields that the path reads
INPUT_FIELDS = ["h", "dzz", "dzzold", "dzf", "u", "utmp2", "w", "wtmp2", "wnew", "s", "qc", "stmp", "D", "htmp",
                "kappa_tv"]
INT_FIELDS = ["ctop", "ctopold", "etop", "etopold", "Nke"]
PROPS = ["dt", "theta", "epsilon", "qepsilon", "resnorm", "grav", "kappa_s", "qmaxiters", "maxiters",
         "qprecond", "hprecond", "TVDsalt", "n"]


PLANFORMS = ["tri", "quad", "hex"]


def capture(st: State, split_every: int = 0, maxfaces: int = 0) -> dict:
    """split_every/maxfaces: the planform arguments the state was built with."""
    d = {"grid": np.array([st.nx, st.ny, st.Lx, st.Ly, st.scalar("Nkmax")], np.float64),
         "dz": st.get("dz"), "dv": st.get("dv")}
    if st.planform != "tri":
        d["planform"] = np.array([PLANFORMS.index(st.planform), split_every, maxfaces], np.int32)
    for f in INPUT_FIELDS:
        d["in_" + f] = st.get(f)
    for f in INT_FIELDS:
        d["int_" + f] = st.ints(f)
    d["props"] = np.array([st.get_prop(p) for p in PROPS])
    return d


def inputs_digest(st: State) -> bytes:
    """SHA-256 over every input field and integer extent, in a fixed order."""
    import hashlib
    h = hashlib.sha256()
    for f in INPUT_FIELDS:
        h.update(np.ascontiguousarray(st.get(f)).tobytes())
    for f in INT_FIELDS:
        h.update(np.ascontiguousarray(st.ints(f)).tobytes())
    return h.digest()


def restore(data) -> State:
    nx, ny, Lx, Ly, Nkmax = data["grid"]
    kw = {}
    if "planform" in data:
        kind, split, mf = (int(v) for v in data["planform"])
        kw = dict(planform=PLANFORMS[kind], split_every=split, maxfaces=mf or None)
    st = State(int(nx), int(ny), float(Lx), float(Ly), int(Nkmax), data["dz"], data["dv"], **kw)
    for f in INT_FIELDS:
        st.set_ints(f, data["int_" + f])
    for f in INPUT_FIELDS:
        st.set(f, data["in_" + f])
    st.set_prop(**dict(zip(PROPS, data["props"])))
    return st

"""Restart records in the reference's binary format (main/physio.c:337-394
writes, :407-495 reads).  The record written here is read by the
reference's own ReadPhysicalVariables into the same fields, bit for bit;
the reader here, with the vertical grid rebuilt as the reference's reader
rebuilds it, gives the same state as the reference's reader.  Without the
reference: round trip, the record's length from the format, and refusal of
short and long files."""
import ctypes as C

import numpy as np
import pytest

from suntans_amd import cases

import _grids
import _impls

ref = _impls.reference()
RECORD = ("h", "Cn_U", "Cn_U2", "Cn_W", "Cn_W2", "Cn_R", "Cn_T", "nu_tv", "kappa_tv", "u", "w", "q", "qc",
          "s", "T", "s0")
MAKERS = [("tiny2", lambda: cases.tiny(2)), ("hexquad7", lambda: cases.tiny(24, nx=7, ny=6, planform="hex",
                                                                          split_every=4, maxfaces=7))] + _grids.REAL


def seeded(make, seed=3):
    """A state whose record fields all hold distinct seeded values."""
    st = make()
    rng = np.random.default_rng(seed)
    for f in RECORD:
        if f == "h":
            continue
        st.set(f, rng.normal(0, 1e-2, len(st.get(f))))
    st.set("h", st.get("h") + rng.uniform(-0.3, 0.3, st.Nc))
    return st


def zeroed(st):
    z = st.clone()
    for f in RECORD:
        z.set(f, 0.0)
    return z


def record_values(st):
    """Each record field restricted to what the record holds (Nke levels per
    edge row, Nk or Nk+1 per cell row)."""
    out = {"h": st.get("h")}
    cell, kw = st.cell_layers(extra=1)
    above_bed = kw < st.ints("Nk")[cell]  # Cn_W, Cn_W2: Nk of their Nk + 1 values
    edge, ke = st.edge_layers()
    ewet = ke < st.ints("Nke")[edge]
    for f in RECORD[1:]:
        v = st.get(f)
        if f in ("Cn_U", "Cn_U2", "u"):
            v = v[ewet]
        elif f in ("Cn_W", "Cn_W2"):
            v = v[above_bed]
        out[f] = v
    return out


def record_bytes(st) -> int:
    nk, nke = st.ints("Nk").astype(int), st.ints("Nke").astype(int)
    cells, edges = nk.sum(), nke.sum()
    return 4 + 8 * (st.Nc + 3 * edges + 12 * cells + st.Nc)  # w holds Nk + 1 per cell


@pytest.mark.parametrize("name,make", MAKERS, ids=[m[0] for m in MAKERS])
def test_round_trip(tmp_path, name, make):
    a = seeded(make)
    path = tmp_path / "store.dat.0"
    a.save_restart(path, 1234)
    assert path.stat().st_size == record_bytes(a)
    b = zeroed(a)
    b.update_dz(-1)
    assert b._L.sun_restart_read(str(path).encode(), b.g, b.p, None, 0) == 0
    ra, rb = record_values(a), record_values(b)
    for f in RECORD:
        assert np.array_equal(ra[f], rb[f]), f
    n = np.zeros(1, np.int32)
    c = zeroed(a)
    from suntans_amd import native
    assert c._L.sun_restart_read(str(path).encode(), c.g, c.p, native.iptr(n), 0) == 0 and n[0] == 1234
    # a short record and a longer one are refused
    data = path.read_bytes()
    (tmp_path / "short").write_bytes(data[:-8])
    (tmp_path / "long").write_bytes(data + b"\0" * 8)
    for bad in ("short", "long"):
        assert c._L.sun_restart_read(str(tmp_path / bad).encode(), c.g, c.p, None, 0) == -1


@pytest.mark.skipif(ref is None, reason="reference harness not built (no reference tree)")
@pytest.mark.parametrize("name,make", MAKERS, ids=[m[0] for m in MAKERS])
def test_reference_reads_our_record(tmp_path, name, make):
    a = seeded(make)
    path = tmp_path / "store.dat.0"
    a.save_restart(path, 77)
    theirs, ours = zeroed(a), zeroed(a)
    ref.L.ref_read_restart.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p]
    assert ref.L.ref_read_restart(theirs.g, theirs.p, theirs.prop, str(path).encode()) == 0
    assert ours.load_restart(path) == 77
    rt, ro, ra = record_values(theirs), record_values(ours), record_values(a)
    for f in RECORD:
        assert np.array_equal(rt[f], ra[f]), f
        assert np.array_equal(ro[f], rt[f]), f
    for f in ("dzz", "dzzold", "dv"):
        assert np.array_equal(ours.get(f), theirs.get(f)), f
    for f in ("ctop", "ctopold", "etop", "etopold"):
        assert np.array_equal(ours.ints(f), theirs.ints(f)), f


def test_warm_start_restores_the_record(tmp_path):
    """cases.warm_start from the Monterey Bay run-phase grid and a record
    restores every field the record holds."""
    a = seeded(_grids.REAL[0][1])
    path = tmp_path / "store.dat.0"
    a.save_restart(path, 5)
    b = cases.warm_start(_grids.path("tides"), str(path))
    ra, rb = record_values(a), record_values(b)
    for f in RECORD:
        assert np.array_equal(ra[f], rb[f]), f
    assert b.get_prop("n") == 5

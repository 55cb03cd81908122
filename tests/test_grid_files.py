"""The run-phase grid reader (sun_grid_load, the reference's ReadGrid for
one process) against the reference itself, and the grid fixtures made from
the reference's examples.

* With the reference compiled (oracle/_ref): for five examples, the
  reference's GetGrid writes its run-phase files into a scratch copy of the
  example's data directory (in a child process: it exits on a fatal grid
  error), then every array sun_grid_load reads equals the reference
  ReadGrid's bit for bit; the committed fixtures are byte for byte what
  GetGrid writes today.
* Without it: the fixtures' geometry is consistent (faces, neighbours,
  gradients, normals, closed cell boundaries, vertical extents)."""
import ctypes as C
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from suntans_amd.state import State

import _grids
import _impls

EXAMPLES = "/root/reference/examples"
ref = _impls.reference()
needs_ref = pytest.mark.skipif(ref is None or not os.path.isdir(EXAMPLES),
                               reason="reference harness or examples not available")

INTS = ("Nk", "nfaces", "cellp", "celldist", "edgedist", "edgep", "neigh", "face", "normal", "cells", "mnptr",
        "Nke", "Nkc", "grad", "gradf", "mark", "edges", "edge_id", "eptr")
REALS = ("Ac", "dv", "xv", "yv", "df", "dg", "n1", "n2", "xe", "ye", "dz", "xp", "yp", "def")


def _generate(example: str, dst: str) -> None:
    src = os.path.join(EXAMPLES, example)
    for f in os.listdir(src):
        if os.path.isfile(os.path.join(src, f)):
            shutil.copy(os.path.join(src, f), dst)
    code = ("import ctypes as C, sys; L = C.CDLL(sys.argv[1]); L.ref_set_verbose(0); "
            "L.ref_grid_generate.argtypes = [C.c_char_p]; L.ref_grid_generate(sys.argv[2].encode())")
    subprocess.run([sys.executable, "-c", code, _impls.REF_SO, dst], check=True, cwd=dst, capture_output=True)


def _used(st, name, a):
    """Only the entries the reference defines: face slots below nfaces, the
    two end points of each edge."""
    if name in ("neigh", "face", "normal", "cells", "def"):
        mf = st.scalar("maxfaces")
        m = (np.arange(mf)[None, :] < st.ints("nfaces")[:, None]).ravel()
        return a[m]
    if name == "edges":
        return a.reshape(-1, 3)[:, :2]
    return a


@needs_ref
@pytest.mark.parametrize("example", ["iwaves/rundata", "cylinder/rundata", "cylinderquad/rundata",
                                     "cavity/rundata", "tides/rundata"])
def test_loader_matches_readgrid(tmp_path, example):
    _generate(example, str(tmp_path))
    L = ref.L
    L.ref_grid_read.restype = C.c_void_p
    L.ref_grid_read.argtypes = [C.c_char_p]
    theirs = State.view(L.ref_grid_read(str(tmp_path).encode()))
    mine = State.load(str(tmp_path))
    for n in ("Nc", "Ne", "Np", "Nkmax", "maxfaces"):
        assert mine.scalar(n) == theirs.scalar(n), n
    for n in INTS:
        assert np.array_equal(_used(mine, n, mine.ints(n)), _used(theirs, n, theirs.ints(n))), n
    for n in REALS:
        a, b = _used(mine, n, mine.get(n)), _used(theirs, n, theirs.get(n))
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), n


@needs_ref
@pytest.mark.parametrize("name,example", [("tides", "tides/rundata"), ("iwaves", "iwaves/rundata")])
def test_fixture_is_the_reference_output(tmp_path, name, example):
    _generate(example, str(tmp_path))
    for f in ("celldata.dat.0", "edgedata.dat.0", "topology.dat.0", "vertspace.dat"):
        with open(tmp_path / f, "rb") as x, open(os.path.join(_grids.path(name), f), "rb") as y:
            assert x.read() == y.read(), f
    want = np.loadtxt(tmp_path / "nodes.dat.0", usecols=(1, 2))
    got = np.loadtxt(os.path.join(_grids.path(name), "nodes.dat.0"), usecols=(1, 2))
    assert np.array_equal(want, got)


@pytest.mark.parametrize("name", ["tides", "iwaves"])
def test_fixture_geometry(name):
    st = State.load(_grids.path(name))
    Nc, Ne, mf = st.Nc, st.Ne, st.scalar("maxfaces")
    nfaces = st.ints("nfaces")
    face = st.ints("face").reshape(Nc, mf)
    neigh = st.ints("neigh").reshape(Nc, mf)
    normal = st.ints("normal").reshape(Nc, mf)
    grad = st.ints("grad").reshape(Ne, 2)
    df, n1, n2 = st.get("df"), st.get("n1"), st.get("n2")
    assert (st.get("Ac") > 0).all() and (df > 0).all() and (st.get("dg") > 0).all()
    for i in range(Nc):
        for f in range(nfaces[i]):
            e = face[i, f]
            assert i in grad[e]
            other = grad[e][1] if grad[e][0] == i else grad[e][0]
            assert neigh[i, f] == other
            assert normal[i, f] == (-1 if grad[e][0] == i else 1)
    # unit edge normals (metrics rounded to 7 digits in the text files)
    assert np.allclose(n1 * n1 + n2 * n2, 1.0, atol=1e-6)
    nk, nke, nkc = st.ints("Nk"), st.ints("Nke"), st.ints("Nkc")
    inner = (grad >= 0).all(axis=1)
    assert np.array_equal(nke[inner], np.minimum(nk[grad[inner, 0]], nk[grad[inner, 1]]))
    assert np.array_equal(nkc[inner], np.maximum(nk[grad[inner, 0]], nk[grad[inner, 1]]))
    assert sorted(st.ints("cellp")) == list(range(Nc)) and sorted(st.ints("edgep")) == list(range(Ne))
    mark, edgep, ed = st.ints("mark"), st.ints("edgep"), st.ints("edgedist")
    assert set(mark[edgep[ed[0]:ed[1]]]) <= {0, 5}
    for m in range(1, 5):
        assert (mark[edgep[ed[m]:ed[m + 1]]] == m).all()

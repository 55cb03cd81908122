"""Grid generation (sun_grid_generate, the reference's one-process GetGrid,
main/grid.c:110-212) against the reference's own output.

The committed fixtures hold two examples' input planforms
(tests/golden/grids/<name>/planform: points, edges, cells and depth files
with the generation keys of their suntans.dat) beside the run-phase files the
reference's GetGrid wrote from them; the generator must write the same bytes.
Where the reference was compiled from its sources (oracle/_ref), every
example with a suntans-format planform is generated both ways here.  The
in-memory grid is checked against the written files read back, and the
configurations the generator refuses fail cleanly."""
import ctypes as C
import filecmp
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from suntans_amd.state import State

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "grids")
ROOT = os.path.dirname(HERE)
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsunref.so")
EXAMPLES = "/root/reference/examples"
RUN_FILES = ("cells.dat.0", "edges.dat.0", "celldata.dat.0", "edgedata.dat.0", "nodes.dat.0", "topology.dat.0",
             "vertspace.dat")
# examples with a suntans-format planform that the reference's GetGrid
# completes for one process: name -> data directory
EXAMPLE_GRIDS = {"tides": "tides/rundata", "iwaves": "iwaves/rundata", "cavity": "cavity/rundata",
                 "cylinder": "cylinder/rundata", "cylinderquad": "cylinderquad/rundata",
                 "heatflux": "heatflux/rundata"}


def return_depth(x, y):
    """ReturnDepth of main/initialization.c:78-92 (the depth for IntDepth 0)."""
    return 1.0


def copy_dir(src, dst):
    for f in os.listdir(src):
        p = os.path.join(src, f)
        if os.path.isfile(p):
            shutil.copyfile(p, os.path.join(dst, f))


def generate(d):
    return State.generate(d, depth=return_depth)


@pytest.mark.parametrize("name", ["tides", "iwaves"])
def test_generate_matches_fixture(name, tmp_path):
    copy_dir(os.path.join(GOLD, name, "planform"), tmp_path)
    st = generate(str(tmp_path))
    assert st.Nc > 0
    for f in RUN_FILES:
        assert filecmp.cmp(os.path.join(tmp_path, f), os.path.join(GOLD, name, f), shallow=False), f


@pytest.mark.parametrize("name", ["tides", "iwaves"])
def test_generated_grid_matches_files(name, tmp_path):
    """The grid in memory is the written files' grid: integers equal, reals
    equal to the 7 digits the files keep."""
    copy_dir(os.path.join(GOLD, name, "planform"), tmp_path)
    gen = generate(str(tmp_path))
    # the run phase's copy, from the files just written
    run = tmp_path / "run"
    run.mkdir()
    for f in RUN_FILES:
        shutil.copyfile(os.path.join(tmp_path, f), run / f)
    shutil.copyfile(os.path.join(GOLD, name, "suntans.dat"), run / "suntans.dat")
    ld = State.load(str(run))
    for k in ("Nc", "Ne", "Np", "Nkmax", "maxfaces"):
        assert gen.scalar(k) == ld.scalar(k), k
    for k in ("nfaces", "Nk", "face", "neigh", "normal", "cells", "mnptr", "Nke", "Nkc", "grad", "gradf", "mark",
              "eptr", "edge_id", "cellp", "edgep", "celldist", "edgedist", "Nkp"):
        np.testing.assert_array_equal(gen.ints(k), ld.ints(k), err_msg=k)
    for k in ("xv", "yv", "Ac", "dv", "def", "df", "dg", "n1", "n2", "xe", "ye", "xp", "yp", "dz"):
        a, b = gen.get(k), ld.get(k)
        np.testing.assert_allclose(a, b, rtol=6e-7, atol=1e-12 * max(1.0, float(np.abs(b).max())), err_msg=k)


def test_generated_momentum_neighbours(tmp_path):
    """eneigh (CreateMomentumCV, main/grid.c:1071-1092): for each edge the
    other faces of its first cell in face order, then of its second, then
    -1."""
    copy_dir(os.path.join(GOLD, "tides", "planform"), tmp_path)
    g = generate(str(tmp_path))
    Ne, mf = g.Ne, g.scalar("maxfaces")
    en = g.ints("eneigh").reshape(Ne, 2 * (mf - 1))
    grad = g.ints("grad").reshape(Ne, 2)
    face = g.ints("face").reshape(-1, mf)
    nfaces = g.ints("nfaces")
    for n in range(Ne):
        want = [e for c in grad[n] if c != -1 for e in face[c, :nfaces[c]] if e != n]
        want += [-1] * (2 * (mf - 1) - len(want))
        assert list(en[n]) == want, n


def test_generate_refusals(tmp_path):
    """IntDepth 1 needs the scaledepth keys; IntDepth 0 needs the depth
    function; a missing planform file is an error."""
    src = os.path.join(GOLD, "iwaves", "planform")

    def case(sub, edit=None, depth=return_depth):
        d = tmp_path / sub
        d.mkdir()
        copy_dir(src, d)
        if edit:
            text = (d / "suntans.dat").read_text().splitlines()
            text = [edit(line) for line in text]
            (d / "suntans.dat").write_text("\n".join(text) + "\n")
        with pytest.raises(RuntimeError):
            State.generate(str(d), depth=depth)
        assert not (d / "celldata.dat.0").exists()

    case("interp", lambda l: "IntDepth\t1\t#" if l.startswith("IntDepth") else l)
    case("nodepth", depth=None)
    case("nopoints", lambda l: "points\tabsent.dat\t#" if l.startswith("points") else l)


def test_generate_minimum_depth_and_stretch(tmp_path):
    """minimum_depth > 0 raises shallow cells, rstretch > 1 grows the layers
    geometrically from dz[0] = D (r - 1) / (r^N - 1), and stair-stepping sets
    each depth to its layers' sum (main/grid.c:476-496, 1423-1480;
    main/initialization.c:33-36)."""
    copy_dir(os.path.join(GOLD, "iwaves", "planform"), tmp_path)
    lines = (tmp_path / "suntans.dat").read_text().splitlines()
    out = []
    for l in lines:
        if l.startswith("stairstep"):
            l = "stairstep\t1\t# a value must not end its line"
        if l.startswith("Nkmax"):
            l = "Nkmax\t8\t# layers"
        if l.startswith("rstretch"):
            l = "rstretch\t1.05\t# growth"
        out.append(l)
    out.insert(1, "minimum_depth\t40\t# m")
    (tmp_path / "suntans.dat").write_text("\n".join(out) + "\n")
    st = State.generate(str(tmp_path), depth=lambda x, y: 30.0 + 70.0 * x / 1e5)
    xv, dv, Nk, dz = st.get("xv"), st.get("dv"), st.ints("Nk"), st.get("dz")
    D = max(40.0, 30.0 + 70.0 * xv.max() / 1e5)
    r, N = 1.05, 8
    want = np.array([D * (r - 1) / (r ** N - 1) * r ** k for k in range(N)])
    np.testing.assert_allclose(dz, want, rtol=1e-13)
    assert abs(dz.sum() - D) < 1e-9 * D
    for i in range(st.Nc):
        d = max(40.0, 30.0 + 70.0 * xv[i] / 1e5)
        # the fewest layers reaching the depth, and the depth their sum
        n = int(np.searchsorted(np.cumsum(dz) - 1e-9 * D, d)) + 1 if d < D else N
        assert Nk[i] == min(n, N), i
        assert abs(dv[i] - dz[:Nk[i]].sum()) < 1e-9 * D


NEEDS_REF = pytest.mark.skipif(not (os.path.exists(REF_SO) and os.path.isdir(EXAMPLES)),
                               reason="needs the reference compiled from its sources (oracle/_ref) and its examples")


@NEEDS_REF
@pytest.mark.parametrize("name", sorted(EXAMPLE_GRIDS))
def test_generate_matches_reference(name, tmp_path):
    both_ways(name, tmp_path)


@NEEDS_REF
@pytest.mark.parametrize("name,ratio", [("tides", "0.9"), ("cavity", "1.0"), ("cylinderquad", "0.95")])
def test_generate_correct_voronoi_matches_reference(name, ratio, tmp_path):
    """CorrectVoronoi > 0 (main/grid.c:3735-3821) on triangles, and on
    quadrilaterals, whose edges take the midpoint branch."""
    def edit(line):
        if line.startswith("CorrectVoronoi"):
            return "CorrectVoronoi\t1\t# edited"
        if line.startswith("VoronoiRatio"):
            return f"VoronoiRatio\t{ratio}\t# edited"
        return line
    both_ways(name, tmp_path, edit)


@NEEDS_REF
@pytest.mark.parametrize("name", ["tides", "cavity", "cylinderquad"])
def test_generate_interp_depth_matches_reference(name, tmp_path):
    """IntDepth 1 (InterpDepth with Interp and FindNearest, main/grid.c:
    3686-3733, main/util.c): scattered soundings, some exactly on Voronoi
    points -- one of them twice, with different depths, where the
    reference takes the second -- and scaledepth."""
    src = os.path.join(EXAMPLES, EXAMPLE_GRIDS[name])
    probe = tmp_path / "probe"
    probe.mkdir()
    copy_dir(src, probe)
    xv = generate(str(probe)).get("xv"), generate(str(probe)).get("yv")
    xs, ys = xv
    rng = np.random.default_rng(17)
    n = 3 * len(xs)
    x = rng.uniform(xs.min() - 0.05 * np.ptp(xs), xs.max() + 0.05 * np.ptp(xs), n)
    y = rng.uniform(ys.min() - 0.05 * np.ptp(ys), ys.max() + 0.05 * np.ptp(ys), n)
    z = -rng.uniform(5.0, 50.0, n)
    hits = rng.choice(len(xs), 20, replace=False)
    x = np.r_[x, xs[hits], xs[hits[0]]]
    y = np.r_[y, ys[hits], ys[hits[0]]]
    z = np.r_[z, -rng.uniform(5.0, 50.0, 21)]
    lines = "".join(f"{a!r} {b!r} {c!r}\n" for a, b, c in zip(x, y, z))

    def edit(line):
        for key, val in (("IntDepth", "1"), ("depth", "soundings.dat")):
            if line.split()[:1] == [key]:
                return f"{key}\t{val}\t# edited"
        return line

    def prepare(d):
        (d / "soundings.dat").write_text(lines)
        text = (d / "suntans.dat").read_text().splitlines()
        text = [edit(l) for l in text]
        text.insert(1, "scaledepth\t1\t# edited")
        text.insert(1, "scaledepthfactor\t1.5\t# edited")
        (d / "suntans.dat").write_text("\n".join(text) + "\n")
    both_ways(name, tmp_path, prepare=prepare)


@NEEDS_REF
def test_generate_large_matches_reference(tmp_path):
    """The 55,870-cell Monterey Bay planform (examples/tides/datasets/
    big3dgrid with the example's suntans.dat): about 15 s in the
    reference, whose partition is quadratic."""
    src = os.path.join(EXAMPLES, "tides", "datasets", "big3dgrid")
    if not os.path.isdir(src):
        pytest.skip("big3dgrid is not in this copy of the examples")

    def prepare(d):
        copy_dir(src, d)
        shutil.copyfile(os.path.join(EXAMPLES, "tides", "rundata", "suntans.dat"), d / "suntans.dat")
    both_ways("tides", tmp_path, prepare=prepare)


def both_ways(name, tmp_path, edit=None, prepare=None):
    src = os.path.join(EXAMPLES, EXAMPLE_GRIDS[name])
    a, b = tmp_path / "ref", tmp_path / "ours"
    a.mkdir()
    b.mkdir()
    copy_dir(src, a)
    copy_dir(src, b)
    if edit:
        for d in (a, b):
            lines = (d / "suntans.dat").read_text().splitlines()
            (d / "suntans.dat").write_text("\n".join(edit(l) for l in lines) + "\n")
    if prepare:
        prepare(a)
        prepare(b)
    code = ("import ctypes as C, sys; L = C.CDLL(sys.argv[1]); L.ref_set_verbose(0); "
            "L.ref_grid_generate.argtypes = [C.c_char_p]; L.ref_grid_generate(sys.argv[2].encode())")
    subprocess.run([sys.executable, "-c", code, REF_SO, str(a)], check=True, cwd=a, capture_output=True)
    generate(str(b))
    outs = [f for f in os.listdir(a) if f in RUN_FILES or f.endswith("-voro")]
    assert len(outs) >= len(RUN_FILES)
    for f in outs:
        assert filecmp.cmp(a / f, b / f, shallow=False), f

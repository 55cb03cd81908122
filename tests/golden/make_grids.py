"""Write the run-phase grid fixtures from the reference's own example grids.

Run where the reference was compiled from its own sources (oracle/_ref):

    python tests/golden/make_grids.py            # every fixture
    python tests/golden/make_grids.py tides      # the named ones

For each example the data directory is copied to a scratch directory, the
reference's GetGrid (main/grid.c:110, one process) writes its run-phase
files there, and this script keeps cells.dat.0, edges.dat.0,
celldata.dat.0, edgedata.dat.0, nodes.dat.0, topology.dat.0 and
vertspace.dat as written, with a suntans.dat holding the keys a run-phase
reader needs.  planform/ keeps the example's own input -- points, edges and
cells files, the "<depth>-voro" file where IntDepth is 2, and a suntans.dat
with the keys grid generation reads -- so a generator can be checked against
the files above.  Data only: the inputs are the example's, the rest the
reference's output.
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsunref.so")
EXAMPLES = "/root/reference/examples"

# fixture name -> example data directory
GRIDS = {
    "tides": "tides/rundata",    # Monterey Bay: 1534 cells, stair-stepped bathymetry, open boundaries
    "iwaves": "iwaves/rundata",  # 200 cells on 100 stretched levels
}
KEYS = ("Nkmax", "maxFaces", "stairstep", "fixdzz", "dzsmall", "smoothbot")
# what GetGrid reads (main/grid.c:110-212, GetDepth, VertGrid, GetDZ)
GEN_KEYS = ("maxFaces", "Nkmax", "stairstep", "rstretch", "CorrectVoronoi", "VoronoiRatio", "vertgridcorrect",
            "IntDepth", "minimum_depth", "dzsmall", "fixdzz", "smoothbot")
FILE_KEYS = ("points", "edges", "cells", "depth", "celldata", "edgedata", "nodes", "topology", "vertspace")


def generate(src: str, dst: str) -> None:
    """The reference's GetGrid into dst, in a child process (the reference
    exits the process on a fatal grid error)."""
    for f in os.listdir(src):
        if os.path.isfile(os.path.join(src, f)):
            shutil.copy(os.path.join(src, f), dst)
    code = ("import ctypes as C, sys; L = C.CDLL(sys.argv[1]); L.ref_set_verbose(0); "
            "L.ref_grid_generate.argtypes = [C.c_char_p]; L.ref_grid_generate(sys.argv[2].encode())")
    subprocess.run([sys.executable, "-c", code, REF_SO, dst], check=True, cwd=dst, capture_output=True)


def dat_lines(path: str, keys=KEYS) -> dict:
    out = {}
    with open(path) as f:
        for line in f:
            w = line.split()
            if len(w) >= 2 and w[0] in keys and w[0] not in out:
                out[w[0]] = w[1]
    return out


def make(name: str) -> None:
    dst = os.path.join(HERE, "grids", name)
    os.makedirs(dst, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        generate(os.path.join(EXAMPLES, GRIDS[name]), tmp)
        for f in ("cells.dat.0", "edges.dat.0", "celldata.dat.0", "edgedata.dat.0", "nodes.dat.0",
                  "topology.dat.0", "vertspace.dat"):
            shutil.copy(os.path.join(tmp, f), dst)
        keys = dat_lines(os.path.join(tmp, "suntans.dat"))
        gen = dat_lines(os.path.join(tmp, "suntans.dat"), GEN_KEYS + FILE_KEYS)
        src = os.path.join(EXAMPLES, GRIDS[name])
        plan = os.path.join(dst, "planform")
        os.makedirs(plan, exist_ok=True)
        inputs = [gen["points"], gen["edges"], gen["cells"]]
        if int(float(gen["IntDepth"])) == 2:
            inputs.append(gen["depth"] + "-voro")
        for f in inputs:
            shutil.copyfile(os.path.join(src, f), os.path.join(plan, f))
        with open(os.path.join(plan, "suntans.dat"), "w") as f:
            f.write(f"# grid-generation keys of examples/{GRIDS[name]}\n")
            for k in GEN_KEYS + FILE_KEYS:
                if k in gen:
                    f.write(f"{k}\t\t{gen[k]}\t# from the example\n")
    with open(os.path.join(dst, "suntans.dat"), "w") as f:
        f.write(f"# run-phase grid of examples/{GRIDS[name]}, written by the reference's GetGrid\n")
        for k in KEYS:
            if k in keys:
                f.write(f"{k}\t\t{keys[k]}\t# from the example\n")
        for k, v in (("celldata", "celldata.dat"), ("edgedata", "edgedata.dat"), ("nodes", "nodes.dat"),
                     ("topology", "topology.dat"), ("vertspace", "vertspace.dat")):
            f.write(f"{k}\t\t{v}\t# run-phase file\n")
    print(f"{name}: {sorted(os.listdir(dst))}")


if __name__ == "__main__":
    for n in sys.argv[1:] or list(GRIDS):
        make(n)

"""The polygon planform builder (sun_grid_polygons) against closed forms and
against the triangulated-rectangle builder on the same triangles."""
import numpy as np
import pytest

from suntans_amd import native, planform
from suntans_amd.state import State

KINDS = [("quad", 0, None), ("quad", 3, None), ("hex", 0, None), ("hex", 4, 7), ("hex", 3, 8)]


def _state(kind, split, mf, nx=7, ny=5, Lx=700.0, Ly=500.0):
    return State(nx, ny, Lx, Ly, 4, 1.0, 4.0, planform=kind, split_every=split, maxfaces=mf)


def _tables(st):
    mf = st.scalar("maxfaces")
    return (mf, st.ints("nfaces"), st.ints("face").reshape(-1, mf), st.ints("neigh").reshape(-1, mf),
            st.ints("normal").reshape(-1, mf), st.ints("grad").reshape(-1, 2))


@pytest.mark.parametrize("kind,split,mf", KINDS)
def test_closed_forms(kind, split, mf):
    nx, ny, Lx, Ly = 7, 5, 700.0, 500.0
    if kind == "hex":  # regular hexagons: half width = sqrt(3) * half the circumradius
        Lx = (2 * nx + 1) * np.sqrt(3.0) * Ly / (3 * ny + 1)
    st = _state(kind, split, mf, nx, ny, Lx, Ly)
    if kind == "quad":
        area = Lx * Ly
    else:  # each hexagon covers 6 lattice cells of (Lx/(2nx+1)) x (Ly/(3ny+1))
        area = nx * ny * 6 * (Lx / (2 * nx + 1)) * (Ly / (3 * ny + 1))
    assert np.isclose(st.get("Ac").sum(), area, rtol=1e-13)
    # Euler: points - edges + cells = 1 for a simply connected planform
    assert st.scalar("Np") - st.Ne + st.Nc == 1
    nc_expected = nx * ny + (sum(1 for j in range(ny) for i in range(nx) if split and (i + j) % split == 0))
    assert st.Nc == nc_expected
    m, nfaces, face, neigh, normal, grad = _tables(st)
    assert m == (mf if mf else (4 if kind == "quad" else 6))
    # every edge appears once per adjacent cell, with opposite normals
    seen = np.zeros(st.Ne, int)
    nsum = np.zeros(st.Ne, int)
    for i in range(st.Nc):
        for f in range(m):
            if f >= nfaces[i]:
                assert face[i, f] == -1 and neigh[i, f] == -1
                continue
            e = face[i, f]
            seen[e] += 1
            nsum[e] += normal[i, f]
            other = grad[e, 1] if grad[e, 0] == i else grad[e, 0]
            assert neigh[i, f] == other
            assert normal[i, f] == (-1 if grad[e, 0] == i else 1)
    boundary = grad[:, 1] < 0
    assert np.all(seen[boundary] == 1) and np.all(seen[~boundary] == 2)
    assert np.all(nsum[~boundary] == 0)
    assert np.array_equal(st.ints("mark"), boundary.astype(np.int32))
    # outward: normal * n points from the cell's Voronoi point to the edge
    xv, yv, xe, ye = st.get("xv"), st.get("yv"), st.get("xe"), st.get("ye")
    n1, n2 = st.get("n1"), st.get("n2")
    for i in range(st.Nc):
        for f in range(nfaces[i]):
            e = face[i, f]
            assert normal[i, f] * (n1[e] * (xe[e] - xv[i]) + n2[e] * (ye[e] - yv[i])) > 0
    # unsplit regular grids are orthogonal: the faces' outward normals close
    if split == 0:
        df = st.get("df")
        for i in range(st.Nc):
            sx = sum(normal[i, f] * df[face[i, f]] * n1[face[i, f]] for f in range(nfaces[i]))
            sy = sum(normal[i, f] * df[face[i, f]] * n2[face[i, f]] for f in range(nfaces[i]))
            assert abs(sx) < 1e-9 * Lx and abs(sy) < 1e-9 * Ly


def test_matches_triangle_builder():
    """The triangulated rectangle rebuilt from its own cell list gives the
    same cells, areas and, edge by edge, the same metrics."""
    t = State(6, 4, 600.0, 400.0, 4, 1.0, 4.0)
    cells = np.ascontiguousarray(t.ints("cells"))
    xp = np.array([ix * 100.0 for iy in range(5) for ix in range(7)])
    yp = np.array([iy * 100.0 for iy in range(5) for ix in range(7)])
    L = native.state_lib()
    xv, yv = t.get("xv"), t.get("yv")
    g = L.sun_grid_polygons(len(xp), native.dptr(xp), native.dptr(yp), t.Nc, 3,
                            native.iptr(np.full(t.Nc, 3, np.int32)), native.iptr(cells), native.dptr(xv),
                            native.dptr(yv))
    assert g
    p = State.__new__(State)
    p._L, p.g, p.p, p.prop, p._owns_grid = L, g, None, None, True
    try:
        assert p.Nc == t.Nc and p.Ne == t.Ne
        assert np.allclose(p.get("Ac"), t.get("Ac"), rtol=1e-15)
        ed_t = t.ints("edges").reshape(-1, 3)
        ed_p = p.ints("edges").reshape(-1, 3)
        key = lambda ed: [tuple(sorted(r[:2])) for r in ed]
        pos = {k: j for j, k in enumerate(key(ed_p))}
        perm = np.array([pos[k] for k in key(ed_t)])
        for f in ("df", "dg"):  # symmetric in the edge's end points and sides
            assert np.array_equal(p.get(f)[perm], t.get(f)), f
        for f in ("xe", "ye"):  # projected from the other end point, maybe
            assert np.allclose(p.get(f)[perm], t.get(f), rtol=1e-14, atol=1e-12), f
        # n1, n2 flip with the edge's side order
        gt, gp = t.ints("grad").reshape(-1, 2), p.ints("grad").reshape(-1, 2)[perm]
        sgn = np.where(gt[:, 0] == gp[:, 0], 1.0, -1.0)
        assert np.allclose(p.get("n1")[perm] * sgn, t.get("n1"), rtol=0, atol=1e-15)
        assert np.allclose(p.get("n2")[perm] * sgn, t.get("n2"), rtol=0, atol=1e-15)
        assert np.array_equal(np.sort(p.ints("neigh").reshape(-1, 3), 1), np.sort(t.ints("neigh").reshape(-1, 3), 1))
    finally:
        p.p = None


def test_rejects_malformed():
    L = native.state_lib()
    xp = np.array([0.0, 1.0, 1.0, 0.0, 2.0])
    yp = np.array([0.0, 0.0, 1.0, 1.0, 0.5])
    nf = np.array([3, 3, 3], np.int32)
    # three triangles on the side (1, 2): not a planform
    cells = np.array([0, 1, 2, 1, 4, 2, 2, 1, 3], np.int32)
    xv = np.array([0.6, 1.3, 0.9])
    yv = np.array([0.3, 0.5, 0.6])
    assert not L.sun_grid_polygons(5, native.dptr(xp), native.dptr(yp), 3, 3, native.iptr(nf), native.iptr(cells),
                                   native.dptr(xv), native.dptr(yv))
    bad = np.array([0, 1, 9], np.int32)  # corner out of range
    assert not L.sun_grid_polygons(5, native.dptr(xp), native.dptr(yp), 1, 3, native.iptr(nf[:1]),
                                   native.iptr(bad), native.dptr(xv), native.dptr(yv))
    with pytest.raises(ValueError):
        planform.rectangles(2, 2, 1.0, 1.0, maxfaces=3)

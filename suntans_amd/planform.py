"""Planforms other than the triangulated rectangle, as polygon lists for
``sun_grid_polygons``: rectangles (optionally with some split into two
triangles) and hexagons (optionally with some split into two quadrilaterals).

These are the grid families the reference runs besides triangles
(examples/cylinderquad, examples/heatflux: maxFaces 4;
examples/headland/grids/hex, examples/riverplume/grids/hybrid: maxFaces 7).
Voronoi points are the polygon centroids, so every dg is positive.
"""
from __future__ import annotations

import numpy as np


class Polygons:
    """Points, cells (padded to maxfaces with -1) and Voronoi points."""

    def __init__(self, xp, yp, cells, maxfaces):
        nf = np.array([len(c) for c in cells], np.int32)
        if maxfaces is None:
            maxfaces = int(nf.max())
        if nf.max() > maxfaces:
            raise ValueError(f"a cell has {nf.max()} corners, more than maxfaces={maxfaces}")
        tab = np.full((len(cells), maxfaces), -1, np.int32)
        for i, c in enumerate(cells):
            tab[i, :len(c)] = c
        self.xp = np.ascontiguousarray(xp, np.float64)
        self.yp = np.ascontiguousarray(yp, np.float64)
        self.nfaces = nf
        self.cells = tab
        self.maxfaces = int(maxfaces)
        self.xv, self.yv = self._centroids(cells)

    def _centroids(self, cells):
        xv = np.empty(len(cells))
        yv = np.empty(len(cells))
        for i, c in enumerate(cells):
            x, y = self.xp[c], self.yp[c]
            x1, y1 = np.roll(x, -1), np.roll(y, -1)
            cr = x * y1 - x1 * y
            A = 0.5 * cr.sum()
            xv[i] = ((x + x1) * cr).sum() / (6 * A)
            yv[i] = ((y + y1) * cr).sum() / (6 * A)
        return xv, yv

    @property
    def Nc(self) -> int:
        return len(self.nfaces)


def rectangles(nx: int, ny: int, Lx: float, Ly: float, split_every: int = 0, maxfaces=None) -> Polygons:
    """nx*ny rectangles; with split_every > 0, those with (ix + iy) divisible
    by it are split into two triangles along a diagonal (a hybrid grid)."""
    def pt(ix, iy):
        return iy * (nx + 1) + ix
    X, Y = np.meshgrid(np.arange(nx + 1) * (Lx / nx), np.arange(ny + 1) * (Ly / ny))
    cells = []
    for iy in range(ny):
        for ix in range(nx):
            a, b, c, d = pt(ix, iy), pt(ix + 1, iy), pt(ix + 1, iy + 1), pt(ix, iy + 1)
            if split_every > 0 and (ix + iy) % split_every == 0:
                if (ix // split_every) % 2:
                    cells += [[a, b, c], [a, c, d]]
                else:
                    cells += [[a, b, d], [b, c, d]]
            else:
                cells.append([a, b, c, d])
    return Polygons(X.ravel(), Y.ravel(), cells, maxfaces)


def hexagons(nx: int, ny: int, Lx: float, Ly: float, split_every: int = 0, maxfaces=None) -> Polygons:
    """nx*ny pointy-top hexagons in offset rows filling about [0,Lx]x[0,Ly];
    with split_every > 0, those with (i + j) divisible by it are split into
    two quadrilaterals through their top and bottom corners."""
    # corners on an integer lattice: x in half hexagon widths, y in half
    # circumradii; hexagon (i, j) is centred at (2i + j%2 + 1, 3j + 2)
    hx = Lx / (2 * nx + 1)
    hy = Ly / (3 * ny + 1)
    ids = {}
    xs, ys = [], []

    def pt(X, Y):
        k = (X, Y)
        if k not in ids:
            ids[k] = len(xs)
            xs.append(X * hx)
            ys.append(Y * hy)
        return ids[k]

    cells = []
    for j in range(ny):
        for i in range(nx):
            X, Y = 2 * i + (j % 2) + 1, 3 * j + 2
            ring = [(X + 1, Y - 1), (X + 1, Y + 1), (X, Y + 2), (X - 1, Y + 1), (X - 1, Y - 1), (X, Y - 2)]
            if split_every > 0 and (i + j) % split_every == 0:
                cells.append([pt(*c) for c in (ring[5], ring[0], ring[1], ring[2])])
                cells.append([pt(*c) for c in (ring[2], ring[3], ring[4], ring[5])])
            else:
                cells.append([pt(*c) for c in ring])
    return Polygons(np.array(xs), np.array(ys), cells, maxfaces)


def make(kind: str, nx: int, ny: int, Lx: float, Ly: float, split_every: int = 0, maxfaces=None) -> Polygons:
    if kind == "quad":
        return rectangles(nx, ny, Lx, Ly, split_every, maxfaces)
    if kind == "hex":
        return hexagons(nx, ny, Lx, Ly, split_every, maxfaces)
    raise ValueError(f"unknown planform {kind!r}")

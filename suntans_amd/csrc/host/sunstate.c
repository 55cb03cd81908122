/*
 * sunstate.c -- host construction of the state the drop-in solvers consume.
 * See include/sunstate.h for the contract of each function.
 */
#include <ctype.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <unistd.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include "sunstate.h"

#define DRY_HEIGHT 1e-10 /* main/suntans.h:27 (DRYCELLHEIGHT) */

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "sunstate: out of host memory\n"); exit(1); }
  return p;
}

/* One contiguous block carved into rows of len[i] (+extra) values. */
static REAL **rows_alloc(int nrows, const int *len, int extra) {
  size_t total = 0;
  for (int i = 0; i < nrows; i++) total += (size_t)(len[i] + extra);
  REAL **r = (REAL **)xcalloc((size_t)nrows, sizeof(REAL *));
  REAL *blk = (REAL *)xcalloc(total, sizeof(REAL));
  size_t off = 0;
  for (int i = 0; i < nrows; i++) { r[i] = blk + off; off += (size_t)(len[i] + extra); }
  if (!nrows) r[0] = blk;  /* the one slot xcalloc gives keeps the block for rows_free */
  return r;
}
static void rows_free(REAL **r) { if (r) { free(r[0]); free(r); } }
static size_t rows_total(int nrows, const int *len, int extra) {
  size_t t = 0;
  for (int i = 0; i < nrows; i++) t += (size_t)(len[i] + extra);
  return t;
}

/* ------------------------------------------------------------------------ */
/* planform                                                                  */

/* Cell, edge and point arrays of a planform with g->Nc, g->Ne, g->Np and
 * g->maxfaces set (shapes as main/gridio.c:86-150). */
static void planform_alloc(gridT *g) {
  const size_t Nc = (size_t)g->Nc, Ne = (size_t)g->Ne, mf = (size_t)g->maxfaces;
  g->cells = (int *)xcalloc(mf * Nc, sizeof(int));
  g->face = (int *)xcalloc(mf * Nc, sizeof(int));
  g->neigh = (int *)xcalloc(mf * Nc, sizeof(int));
  g->normal = (int *)xcalloc(mf * Nc, sizeof(int));
  g->nfaces = (int *)xcalloc(Nc, sizeof(int));
  g->xv = (REAL *)xcalloc(Nc, sizeof(REAL));
  g->yv = (REAL *)xcalloc(Nc, sizeof(REAL));
  g->Ac = (REAL *)xcalloc(Nc, sizeof(REAL));
  g->def = (REAL *)xcalloc(mf * Nc, sizeof(REAL));
  g->dv = (REAL *)xcalloc(Nc, sizeof(REAL));
  g->edges = (int *)xcalloc(3 * Ne, sizeof(int));
  g->grad = (int *)xcalloc(2 * Ne, sizeof(int));
  g->gradf = (int *)xcalloc(2 * Ne, sizeof(int));
  g->mark = (int *)xcalloc(Ne, sizeof(int));
  g->eneigh = (int *)xcalloc(2 * (mf - 1) * Ne, sizeof(int));
  g->df = (REAL *)xcalloc(Ne, sizeof(REAL));
  g->dg = (REAL *)xcalloc(Ne, sizeof(REAL));
  g->n1 = (REAL *)xcalloc(Ne, sizeof(REAL));
  g->n2 = (REAL *)xcalloc(Ne, sizeof(REAL));
  g->xe = (REAL *)xcalloc(Ne, sizeof(REAL));
  g->ye = (REAL *)xcalloc(Ne, sizeof(REAL));
}

/* Computational ordering of a single process (main/grid.c:1652-1679):
 * computational cells first, then the type-3 cells (celldist[1..2)), each
 * group in index order; edges sorted by mark 0 (and 5), 1, 2, 3, 4 into
 * edgedist[0..5].  The builders take a cell with a mark-3 face as type 3
 * (the reference reads the cell's type from its cells file). */
static void computational_order(gridT *g) {
  const int Nc = g->Nc, Ne = g->Ne, mf = g->maxfaces;
  int n = 0;
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) g->celldist[1] = n;
    for (int i = 0; i < Nc; i++) {
      int t3 = 0;
      for (int nf = 0; nf < g->nfaces[i]; nf++) t3 |= g->mark[g->face[i * mf + nf]] == 3;
      if (t3 == pass) g->cellp[n++] = i;
    }
  }
  g->celldist[0] = 0; g->celldist[2] = Nc;
  int k = 0;
  for (int e = 0; e < Ne; e++) if (g->mark[e] == 0 || g->mark[e] == 5) g->edgep[k++] = e;
  g->edgedist[0] = 0; g->edgedist[1] = k;
  for (int m = 1; m < 5; m++) {
    for (int e = 0; e < Ne; e++) if (g->mark[e] == m) g->edgep[k++] = e;
    g->edgedist[m + 1] = k;
  }
}

int sun_grid_open(gridT *g, const int *mark) {
  for (int e = 0; e < g->Ne; e++) {
    if (mark[e] == g->mark[e]) continue;
    if (g->grad[2 * e + 1] != -1 || mark[e] < 1 || mark[e] > 4) return -1;
  }
  for (int e = 0; e < g->Ne; e++) g->mark[e] = g->edges[3 * e + 2] = mark[e];
  computational_order(g);
  return 0;
}

/* Edge metrics, face normals and the single-process computational ordering
 * of a planform whose points, cells (nfaces, face, neigh), Voronoi points,
 * areas and edges (end points, grad, mark) are set. */
static void planform_finish(gridT *g) {
  const int Nc = g->Nc, Ne = g->Ne, mf = g->maxfaces;
  /* metrics, main/grid.c:3457-3533 conventions */
  for (int e = 0; e < Ne; e++) {
    int p1 = g->edges[3 * e], p2 = g->edges[3 * e + 1];
    double tx = g->xp[p2] - g->xp[p1], ty = g->yp[p2] - g->yp[p1];
    double tm = sqrt(tx * tx + ty * ty);
    g->df[e] = tm;
    tx /= tm; ty /= tm;
    int c = g->grad[2 * e] >= 0 ? g->grad[2 * e] : g->grad[2 * e + 1];
    double s = (g->xv[c] - g->xp[p1]) * tx + (g->yv[c] - g->yp[p1]) * ty;
    g->xe[e] = g->xp[p1] + s * tx; g->ye[e] = g->yp[p1] + s * ty;
    int a = g->grad[2 * e], b = g->grad[2 * e + 1];
    double nx_, ny_;
    if (a >= 0 && b >= 0) {
      nx_ = g->xv[a] - g->xv[b]; ny_ = g->yv[a] - g->yv[b];
      g->dg[e] = sqrt(nx_ * nx_ + ny_ * ny_);
      g->n1[e] = nx_ / g->dg[e]; g->n2[e] = ny_ / g->dg[e];
    } else {
      int in = a >= 0 ? a : b;
      nx_ = a >= 0 ? g->xv[in] - g->xe[e] : g->xe[e] - g->xv[in];
      ny_ = a >= 0 ? g->yv[in] - g->ye[e] : g->ye[e] - g->yv[in];
      double d = sqrt(nx_ * nx_ + ny_ * ny_);
      g->n1[e] = nx_ / d; g->n2[e] = ny_ / d; g->dg[e] = 2.0 * d;
    }
  }
  for (int i = 0; i < Nc; i++)
    for (int nf = 0; nf < g->nfaces[i]; nf++) {
      int e = g->face[mf * i + nf];
      g->normal[mf * i + nf] = (i == g->grad[2 * e]) ? -1 : 1;
      if (g->grad[2 * e] == i) g->gradf[2 * e] = nf; else g->gradf[2 * e + 1] = nf;
      int p1 = g->edges[3 * e];
      g->def[mf * i + nf] = -((g->xv[i] - g->xp[p1]) * g->n1[e] + (g->yv[i] - g->yp[p1]) * g->n2[e]) *
                           g->normal[mf * i + nf];
    }

  g->cellp = (int *)xcalloc(Nc, sizeof(int));
  g->celldist = (int *)xcalloc(3, sizeof(int));
  g->edgep = (int *)xcalloc(Ne, sizeof(int));
  g->edgedist = (int *)xcalloc(6, sizeof(int));
  computational_order(g);
  g->mnptr = (int *)xcalloc(Nc, sizeof(int));
  g->eptr = (int *)xcalloc(Ne, sizeof(int));
  for (int i = 0; i < Nc; i++) g->mnptr[i] = i;
  for (int e = 0; e < Ne; e++) g->eptr[e] = e;
  g->Nneighs = 0;
  g->myneighs = (int *)xcalloc(1, sizeof(int));
}

gridT *sun_grid_tri(int nx, int ny, double Lx, double Ly) {
  gridT *g = (gridT *)xcalloc(1, sizeof(gridT));
  const int Nc = 2 * nx * ny, Np = (nx + 1) * (ny + 1);
  const int Ne = 3 * nx * ny + nx + ny;
  const int rowE = 3 * nx + 1;
  const double dx = Lx / nx, dy = Ly / ny;
  g->Nc = g->Nc_comp = Nc; g->Ne = g->Ne_comp = Ne; g->Np = Np;
  g->maxfaces = 3; g->Nge = 0;

#define PT(ix, iy) ((iy) * (nx + 1) + (ix))
#define T0(ix, iy) (2 * ((iy) * nx + (ix)))
#define T1(ix, iy) (2 * ((iy) * nx + (ix)) + 1)
#define EH(ix, iy) ((iy) < ny ? (iy) * rowE + 3 * (ix) : ny * rowE + (ix))
#define EV(ix, iy) ((ix) < nx ? (iy) * rowE + 3 * (ix) + 1 : (iy) * rowE + 3 * nx)
#define ED(ix, iy) ((iy) * rowE + 3 * (ix) + 2)

  g->xp = (REAL *)xcalloc(Np, sizeof(REAL));
  g->yp = (REAL *)xcalloc(Np, sizeof(REAL));
  for (int iy = 0; iy <= ny; iy++)
    for (int ix = 0; ix <= nx; ix++) { g->xp[PT(ix, iy)] = ix * dx; g->yp[PT(ix, iy)] = iy * dy; }

  planform_alloc(g);

  for (int iy = 0; iy < ny; iy++)
    for (int ix = 0; ix < nx; ix++) {
      int c0 = T0(ix, iy), c1 = T1(ix, iy);
      int *v0 = g->cells + 3 * c0, *v1 = g->cells + 3 * c1;
      v0[0] = PT(ix, iy); v0[1] = PT(ix + 1, iy); v0[2] = PT(ix + 1, iy + 1);
      v1[0] = PT(ix, iy); v1[1] = PT(ix + 1, iy + 1); v1[2] = PT(ix, iy + 1);
      /* face nf joins corner nf to corner nf+1 */
      int *f0 = g->face + 3 * c0, *f1 = g->face + 3 * c1;
      int *n0 = g->neigh + 3 * c0, *n1 = g->neigh + 3 * c1;
      f0[0] = EH(ix, iy);     n0[0] = iy > 0 ? T1(ix, iy - 1) : -1;
      f0[1] = EV(ix + 1, iy); n0[1] = ix + 1 < nx ? T1(ix + 1, iy) : -1;
      f0[2] = ED(ix, iy);     n0[2] = c1;
      f1[0] = ED(ix, iy);     n1[0] = c0;
      f1[1] = EH(ix, iy + 1); n1[1] = iy + 1 < ny ? T0(ix, iy + 1) : -1;
      f1[2] = EV(ix, iy);     n1[2] = ix > 0 ? T0(ix - 1, iy) : -1;
      g->nfaces[c0] = g->nfaces[c1] = 3;
      g->xv[c0] = (ix + 2.0 / 3.0) * dx; g->yv[c0] = (iy + 1.0 / 3.0) * dy;
      g->xv[c1] = (ix + 1.0 / 3.0) * dx; g->yv[c1] = (iy + 2.0 / 3.0) * dy;
      g->Ac[c0] = g->Ac[c1] = 0.5 * dx * dy;
    }

  /* edges: end points and the cells on either side (second is -1 on the
   * boundary); ordering of the two sides is below/above, left/right, T0/T1 */
  for (int iy = 0; iy <= ny; iy++)
    for (int ix = 0; ix < nx; ix++) {
      int e = EH(ix, iy);
      int below = iy > 0 ? T1(ix, iy - 1) : -1, above = iy < ny ? T0(ix, iy) : -1;
      g->edges[3 * e] = PT(ix, iy); g->edges[3 * e + 1] = PT(ix + 1, iy);
      if (below >= 0 && above >= 0) { g->grad[2 * e] = below; g->grad[2 * e + 1] = above; }
      else { g->grad[2 * e] = below >= 0 ? below : above; g->grad[2 * e + 1] = -1; g->mark[e] = 1; }
    }
  for (int iy = 0; iy < ny; iy++)
    for (int ix = 0; ix <= nx; ix++) {
      int e = EV(ix, iy);
      int left = ix > 0 ? T0(ix - 1, iy) : -1, right = ix < nx ? T1(ix, iy) : -1;
      g->edges[3 * e] = PT(ix, iy); g->edges[3 * e + 1] = PT(ix, iy + 1);
      if (left >= 0 && right >= 0) { g->grad[2 * e] = left; g->grad[2 * e + 1] = right; }
      else { g->grad[2 * e] = left >= 0 ? left : right; g->grad[2 * e + 1] = -1; g->mark[e] = 1; }
    }
  for (int iy = 0; iy < ny; iy++)
    for (int ix = 0; ix < nx; ix++) {
      int e = ED(ix, iy);
      g->edges[3 * e] = PT(ix, iy); g->edges[3 * e + 1] = PT(ix + 1, iy + 1);
      g->grad[2 * e] = T0(ix, iy); g->grad[2 * e + 1] = T1(ix, iy);
    }
  for (int e = 0; e < Ne; e++) g->edges[3 * e + 2] = g->mark[e];

  planform_finish(g);
  return g;
#undef PT
#undef T0
#undef T1
#undef EH
#undef EV
#undef ED
}

/* One side of a cell: its end points (lower index first) and where it is. */
struct side { int lo, hi, cell, nf; };
static int side_by_points(const void *a, const void *b) {
  const struct side *x = (const struct side *)a, *y = (const struct side *)b;
  if (x->lo != y->lo) return x->lo < y->lo ? -1 : 1;
  if (x->hi != y->hi) return x->hi < y->hi ? -1 : 1;
  if (x->cell != y->cell) return x->cell < y->cell ? -1 : 1;
  return x->nf < y->nf ? -1 : (x->nf > y->nf);
}

gridT *sun_grid_polygons(int Np, const double *xp, const double *yp, int Nc, int maxfaces, const int *nfaces,
                         const int *cells, const double *xv, const double *yv) {
  if (Np < 3 || Nc < 1 || maxfaces < 3) return NULL;
  for (int i = 0; i < Nc; i++) {
    if (nfaces[i] < 3 || nfaces[i] > maxfaces) return NULL;
    for (int v = 0; v < nfaces[i]; v++) {
      const int pt = cells[(size_t)i * maxfaces + v];
      if (pt < 0 || pt >= Np) return NULL;
    }
  }
  /* sides grouped by their end points; each group is one edge */
  size_t ns = 0;
  for (int i = 0; i < Nc; i++) ns += (size_t)nfaces[i];
  struct side *sd = (struct side *)xcalloc(ns, sizeof(struct side));
  size_t k = 0;
  for (int i = 0; i < Nc; i++)
    for (int nf = 0; nf < nfaces[i]; nf++) {
      const int a = cells[(size_t)i * maxfaces + nf], b = cells[(size_t)i * maxfaces + (nf + 1) % nfaces[i]];
      if (a == b) { free(sd); return NULL; }
      sd[k].lo = a < b ? a : b; sd[k].hi = a < b ? b : a; sd[k].cell = i; sd[k].nf = nf; k++;
    }
  qsort(sd, ns, sizeof(struct side), side_by_points);
  /* edges are numbered in the order their first side appears in the cell
   * list: rank the groups by (cell, nf) of their first member */
  int ne = 0;
  int *first = (int *)xcalloc(ns, sizeof(int)); /* group -> index of its first side in sd */
  for (size_t j = 0; j < ns; j++) {
    if (j > 0 && sd[j].lo == sd[j - 1].lo && sd[j].hi == sd[j - 1].hi) {
      if (j > 1 && sd[j].lo == sd[j - 2].lo && sd[j].hi == sd[j - 2].hi) { free(sd); free(first); return NULL; }
      continue;
    }
    first[ne++] = (int)j;
  }
  int *order = (int *)xcalloc((size_t)Nc * maxfaces, sizeof(int)); /* (cell, nf) -> group, -1 */
  for (size_t j = 0; j < (size_t)Nc * maxfaces; j++) order[j] = -1;
  for (int e = 0; e < ne; e++) {
    const struct side *f = &sd[first[e]];
    order[(size_t)f->cell * maxfaces + f->nf] = e;
  }
  int *edge_of_group = (int *)xcalloc((size_t)ne, sizeof(int));
  int next = 0;
  for (size_t j = 0; j < (size_t)Nc * maxfaces; j++)
    if (order[j] >= 0) edge_of_group[order[j]] = next++;

  gridT *g = (gridT *)xcalloc(1, sizeof(gridT));
  g->Nc = g->Nc_comp = Nc; g->Ne = g->Ne_comp = ne; g->Np = Np;
  g->maxfaces = maxfaces; g->Nge = 0;
  g->xp = (REAL *)xcalloc((size_t)Np, sizeof(REAL));
  g->yp = (REAL *)xcalloc((size_t)Np, sizeof(REAL));
  memcpy(g->xp, xp, sizeof(REAL) * (size_t)Np);
  memcpy(g->yp, yp, sizeof(REAL) * (size_t)Np);
  planform_alloc(g);
  for (size_t j = 0; j < (size_t)Nc * maxfaces; j++) {
    g->cells[j] = cells[j];
    g->face[j] = -1;
    g->neigh[j] = -1;
  }
  for (int i = 0; i < Nc; i++) {
    g->nfaces[i] = nfaces[i];
    g->xv[i] = xv[i];
    g->yv[i] = yv[i];
    /* shoelace area of the (counterclockwise or clockwise) polygon */
    double A = 0;
    for (int v = 0; v < nfaces[i]; v++) {
      const int a = cells[(size_t)i * maxfaces + v], b = cells[(size_t)i * maxfaces + (v + 1) % nfaces[i]];
      A += xp[a] * yp[b] - xp[b] * yp[a];
    }
    g->Ac[i] = 0.5 * fabs(A);
  }
  for (int grp = 0; grp < ne; grp++) {
    const int e = edge_of_group[grp];
    const struct side *f = &sd[first[grp]];
    const int two = first[grp] + 1 < (int)ns && sd[first[grp] + 1].lo == f->lo && sd[first[grp] + 1].hi == f->hi;
    /* first side in cell order gives the end-point order and grad[2e] */
    const struct side *s0 = f, *s1 = two ? f + 1 : NULL;
    if (s1 && (s1->cell < s0->cell || (s1->cell == s0->cell && s1->nf < s0->nf))) { s0 = f + 1; s1 = f; }
    g->edges[3 * e] = cells[(size_t)s0->cell * maxfaces + s0->nf];
    g->edges[3 * e + 1] = cells[(size_t)s0->cell * maxfaces + (s0->nf + 1) % nfaces[s0->cell]];
    g->grad[2 * e] = s0->cell;
    g->grad[2 * e + 1] = s1 ? s1->cell : -1;
    g->mark[e] = s1 ? 0 : 1;
    g->edges[3 * e + 2] = g->mark[e];
    g->face[(size_t)s0->cell * maxfaces + s0->nf] = e;
    g->neigh[(size_t)s0->cell * maxfaces + s0->nf] = s1 ? s1->cell : -1;
    if (s1) {
      g->face[(size_t)s1->cell * maxfaces + s1->nf] = e;
      g->neigh[(size_t)s1->cell * maxfaces + s1->nf] = s0->cell;
    }
  }
  free(sd); free(first); free(order); free(edge_of_group);
  planform_finish(g);
  return g;
}

int sun_grid_vertical(gridT *g, int Nkmax, const double *dz, const double *dv, int stairstep) {
  const int Nc = g->Nc, Ne = g->Ne;
  g->Nkmax = Nkmax;
  g->stairstep = stairstep;
  g->dz = (REAL *)xcalloc(Nkmax, sizeof(REAL));
  memcpy(g->dz, dz, sizeof(REAL) * (size_t)Nkmax);
  memcpy(g->dv, dv, sizeof(REAL) * (size_t)Nc);
  double dmax = 0;
  for (int i = 0; i < Nc; i++) if (dv[i] > dmax) dmax = dv[i];
  g->Nk = (int *)xcalloc(Nc, sizeof(int));
  /* layer count: the first k whose cumulative thickness reaches the depth */
  for (int i = 0; i < Nc; i++) {
    if (Nkmax <= 1 || g->dv[i] == dmax) { g->Nk[i] = Nkmax; continue; }
    double z = 0;
    for (int k = 0; k < Nkmax; k++) {
      z += g->dz[k];
      g->Nk[i] = k + 1;
      if (z >= g->dv[i]) break;
    }
  }
  if (stairstep)
    for (int i = 0; i < Nc; i++) {
      double z = 0;
      for (int k = 0; k < g->Nk[i]; k++) z += g->dz[k];
      g->dv[i] = z;
    }
  g->ctop = (int *)xcalloc(Nc, sizeof(int));
  g->ctopold = (int *)xcalloc(Nc, sizeof(int));
  g->dztop = (REAL *)xcalloc(Nc, sizeof(REAL));
  g->dzbot = (REAL *)xcalloc(Nc, sizeof(REAL));
  for (int i = 0; i < Nc; i++) g->dztop[i] = g->dz[0];
  g->etop = (int *)xcalloc(Ne, sizeof(int));
  g->etopold = (int *)xcalloc(Ne, sizeof(int));
  g->Nke = (int *)xcalloc(Ne, sizeof(int));
  g->Nkc = (int *)xcalloc(Ne, sizeof(int));
  for (int e = 0; e < Ne; e++) {
    int a = g->grad[2 * e], b = g->grad[2 * e + 1];
    if (a < 0 || b < 0) {
      int c = a >= 0 ? a : b;
      g->Nke[e] = g->Nkc[e] = g->Nk[c];
    } else {
      int na = g->Nk[a], nb = g->Nk[b];
      g->Nke[e] = na < nb ? na : nb;
      g->Nkc[e] = na < nb ? nb : na;
    }
  }
  g->Nkp = (int *)xcalloc(g->Np, sizeof(int));
  for (int p = 0; p < g->Np; p++) g->Nkp[p] = Nkmax;
  g->dzz = rows_alloc(Nc, g->Nk, 0);
  g->dzzold = rows_alloc(Nc, g->Nk, 0);
  for (int i = 0; i < Nc; i++)
    for (int k = 0; k < g->Nk[i]; k++) g->dzz[i][k] = g->dzzold[i][k] = g->dz[k];
  g->dzf = rows_alloc(Ne, g->Nkc, 0);
  g->dzfB = (REAL *)xcalloc(Ne, sizeof(REAL));
  return 0;
}

/* ------------------------------------------------------------------------ */
/* physical state                                                            */

struct fdesc { const char *name; int kind; size_t off; };
/* kinds: 1 cell rows (Nk), 2 w rows (Nk+1), 3 edge rows (Nkc), 4 cell 2-D,
 * 5 edge 2-D, 6 column scratch (Nkmax+1), 8 open-boundary rows (Nke+1, one
 * per edge of edgedist[2..5)), 9 open-boundary rows (Nke); grid-owned
 * vectors also use 7 (Nkmax), 10 (Nc*maxfaces) and 11 (Np) */
#define ROWS_KIND(k) ((k) <= 3 || (k) == 8 || (k) == 9)
#define PF(f, k) {#f, k, offsetof(physT, f)}
static const struct fdesc phys_fields[] = {
  PF(uc, 1), PF(vc, 1), PF(wc, 1), PF(uold, 1), PF(vold, 1), PF(q, 1), PF(qc, 1),
  PF(s, 1), PF(T, 1), PF(s0, 1), PF(rho, 1), PF(stmp, 1), PF(stmp2, 1), PF(stmp3, 1),
  PF(nu_tv, 1), PF(kappa_tv, 1), PF(nu_lax, 1), PF(Cn_R, 1), PF(Cn_T, 1),
  PF(gradSx, 1), PF(gradSy, 1), PF(Ttmp, 1),
  PF(w, 2), PF(wnew, 2), PF(wtmp, 2), PF(wtmp2, 2), PF(Cn_W, 2), PF(Cn_W2, 2),
  PF(u, 3), PF(utmp, 3), PF(utmp2, 3), PF(ut, 3), PF(Cn_U, 3), PF(Cn_U2, 3),
  PF(SfHp, 3), PF(SfHm, 3),
  PF(h, 4), PF(hcorr, 4), PF(hold, 4), PF(htmp, 4), PF(htmp2, 4), PF(htmp3, 4),
  PF(hcoef, 4), PF(Tsurf, 4), PF(dT, 4),
  PF(D, 5), PF(tau_T, 5), PF(tau_B, 5), PF(CdT, 5), PF(CdB, 5),
  PF(ap, 6), PF(am, 6), PF(bp, 6), PF(bm, 6), PF(a, 6), PF(b, 6), PF(c, 6), PF(d, 6),
  PF(Cp, 6), PF(Cm, 6), PF(rp, 6), PF(rm, 6), PF(wp, 6), PF(wm, 6),
  PF(boundary_w, 8), PF(boundary_tmp, 8), PF(boundary_s, 9), PF(boundary_T, 9), PF(boundary_u, 9),
  PF(boundary_v, 9), PF(boundary_rho, 9),
};
#undef PF
#define NPF (int)(sizeof(phys_fields) / sizeof(phys_fields[0]))

physT *sun_phys_alloc(const gridT *g) {
  physT *p = (physT *)xcalloc(1, sizeof(physT));
  const int Nc = g->Nc, Ne = g->Ne, mf = g->maxfaces;
  for (int f = 0; f < NPF; f++) {
    void *slot = (char *)p + phys_fields[f].off;
    switch (phys_fields[f].kind) {
      case 1: *(REAL ***)slot = rows_alloc(Nc, g->Nk, 0); break;
      case 2: *(REAL ***)slot = rows_alloc(Nc, g->Nk, 1); break;
      case 3: *(REAL ***)slot = rows_alloc(Ne, g->Nkc, 0); break;
      case 4: *(REAL **)slot = (REAL *)xcalloc(Nc, sizeof(REAL)); break;
      case 5: *(REAL **)slot = (REAL *)xcalloc(Ne, sizeof(REAL)); break;
      case 6: *(REAL **)slot = (REAL *)xcalloc((size_t)g->Nkmax + 1, sizeof(REAL)); break;
    }
  }
  /* htmp is 10*Nc long in the reference (main/phys.c:177) */
  free(p->htmp);
  p->htmp = (REAL *)xcalloc(10 * (size_t)Nc, sizeof(REAL));
  p->hfcoef = (REAL *)xcalloc((size_t)mf * Nc, sizeof(REAL));
  /* qtmp: one Nk row per cell face (main/phys.c:194,261-262) */
  {
    int *len = (int *)xcalloc((size_t)mf * Nc, sizeof(int));
    for (int i = 0; i < Nc; i++)
      for (int nf = 0; nf < g->nfaces[i]; nf++) len[i * mf + nf] = g->Nk[i];
    p->qtmp = rows_alloc(mf * Nc, len, 0);
    free(len);
  }
  p->active = (unsigned char *)xcalloc(Nc, 1);
  for (int i = 0; i < Nc; i++) p->active[i] = 1;
  /* open-boundary rows: one per edge of marks 2..4 (main/phys.c:285-305) */
  {
    int nb = g->edgedist[5] - g->edgedist[2];
    int *len = (int *)xcalloc(nb, sizeof(int));
    for (int j = 0; j < nb; j++) len[j] = g->Nke[g->edgep[g->edgedist[2] + j]];
    p->boundary_s = rows_alloc(nb, len, 0);
    p->boundary_T = rows_alloc(nb, len, 0);
    p->boundary_u = rows_alloc(nb, len, 0);
    p->boundary_v = rows_alloc(nb, len, 0);
    p->boundary_rho = rows_alloc(nb, len, 0);
    p->boundary_w = rows_alloc(nb, len, 1);
    p->boundary_tmp = rows_alloc(nb, len, 1);
    p->boundary_h = (REAL *)xcalloc(nb, sizeof(REAL));
    p->boundary_flag = (REAL *)xcalloc(nb, sizeof(REAL));
    free(len);
  }
  return p;
}

void sun_phys_free(const gridT *g, physT *p) {
  (void)g;
  if (!p) return;
  for (int f = 0; f < NPF; f++) {
    void *slot = (char *)p + phys_fields[f].off;
    if (ROWS_KIND(phys_fields[f].kind)) rows_free(*(REAL ***)slot); else free(*(REAL **)slot);
  }
  free(p->hfcoef);
  rows_free(p->qtmp);
  free(p->active);
  free(p->boundary_h); free(p->boundary_flag);
  free(p);
}

static size_t kind_len(const gridT *g, int kind) {
  switch (kind) {
    case 1: return rows_total(g->Nc, g->Nk, 0);
    case 2: return rows_total(g->Nc, g->Nk, 1);
    case 3: return rows_total(g->Ne, g->Nkc, 0);
    case 4: return (size_t)g->Nc;
    case 5: return (size_t)g->Ne;
    case 6: return (size_t)g->Nkmax + 1;
    case 10: return (size_t)g->Nc * (size_t)g->maxfaces;
    case 11: return (size_t)g->Np;
    case 8: case 9: {
      size_t n = 0;
      for (int j = g->edgedist[2]; j < g->edgedist[5]; j++) n += (size_t)(g->Nke[g->edgep[j]] + (kind == 8));
      return n;
    }
  }
  return 0;
}

void sun_phys_copy(const gridT *g, const physT *from, physT *to) {
  for (int f = 0; f < NPF; f++) {
    size_t off = phys_fields[f].off, n = kind_len(g, phys_fields[f].kind);
    if (!n) continue;
    if (ROWS_KIND(phys_fields[f].kind))
      memcpy((*(REAL ***)((char *)to + off))[0], (*(REAL ***)((char *)from + off))[0], n * sizeof(REAL));
    else
      memcpy(*(REAL **)((char *)to + off), *(REAL **)((char *)from + off), n * sizeof(REAL));
  }
  memcpy(to->htmp, from->htmp, 10 * (size_t)g->Nc * sizeof(REAL));
  memcpy(to->hfcoef, from->hfcoef, (size_t)g->maxfaces * g->Nc * sizeof(REAL));
  memcpy(to->active, from->active, (size_t)g->Nc);
  to->mass = from->mass; to->volume = from->volume;
}

physT *sun_phys_clone(const gridT *g, const physT *p) {
  physT *q = sun_phys_alloc(g);
  sun_phys_copy(g, p, q);
  return q;
}

propT *sun_prop_default(void) {
  propT *pr = (propT *)xcalloc(1, sizeof(propT));
  /* examples/lockexchange/rundata/suntans.dat values where the path reads them */
  pr->dt = 5.0; pr->theta = 0.55; pr->thetaS = 0.55; pr->thetaB = 0.55;
  pr->nu = 0.025; pr->nu_H = 0.016; pr->CdB = 0.025; pr->grav = 9.81;
  pr->epsilon = 1e-10; pr->qepsilon = 1e-10; pr->resnorm = 0;
  pr->beta = 1.0; pr->kappa_s = 0; pr->kappa_sH = 0;
  pr->nonhydrostatic = 1; pr->cgsolver = 1; pr->maxiters = 1000; pr->qmaxiters = 2000;
  pr->qprecond = 2; pr->hprecond = 1; pr->nonlinear = 2; pr->TVDsalt = 4;
  pr->n = 1; pr->nstart = 0; pr->nsteps = 1;
  return pr;
}
void sun_prop_free(propT *prop) { free(prop); }

void sun_grid_free(gridT *g) {
  if (!g) return;
  REAL *r1[] = {g->xp, g->yp, g->xv, g->yv, g->xe, g->ye, g->dv, g->dz, g->dzfB, g->dztop,
                g->dzbot, g->df, g->dg, g->def, g->Ac, g->n1, g->n2};
  for (size_t i = 0; i < sizeof(r1) / sizeof(r1[0]); i++) free(r1[i]);
  int *i1[] = {g->face, g->edges, g->cells, g->neigh, g->eneigh, g->grad, g->gradf, g->mark,
               g->normal, g->celldist, g->edgedist, g->cellp, g->edgep, g->mnptr, g->eptr,
               g->myneighs, g->Nk, g->ctop, g->ctopold, g->etop, g->etopold, g->Nkp, g->Nke,
               g->Nkc, g->nfaces, g->edge_id};
  for (size_t i = 0; i < sizeof(i1) / sizeof(i1[0]); i++) free(i1[i]);
  rows_free(g->dzz); rows_free(g->dzzold); rows_free(g->dzf);
  free(g);
}

/* ------------------------------------------------------------------------ */
/* vertical grid updates                                                     */

void sun_update_dz(gridT *g, physT *p, const propT *prop, int option) {
  const int Nc = g->Nc, Ne = g->Ne;
  if (prop->linearFS) return;
  if (!option) {
    for (int j = 0; j < Ne; j++) g->etopold[j] = g->etop[j];
    for (int i = 0; i < Nc; i++) {
      g->ctopold[i] = g->ctop[i];
      for (int k = 0; k < g->Nk[i]; k++) g->dzzold[i][k] = k < g->ctop[i] ? 0 : g->dzz[i][k];
    }
  }
  if (option == -1) for (int i = 0; i < Nc; i++) p->h[i] = 0;
  /* partial-step bottom layer */
  for (int i = 0; i < Nc; i++) {
    double z = 0;
    for (int k = 0; k < g->Nk[i]; k++) z -= g->dz[k];
    g->dzz[i][g->Nk[i] - 1] = g->dz[g->Nk[i] - 1] + g->dv[i] + z;
  }
  /* layer thicknesses cut by the free surface */
  if (g->Nkmax > 1) {
    for (int i = 0; i < Nc; i++) {
      double z = 0;
      int found = 0;
      const int last = g->Nk[i] - 1;
      for (int k = 0; k <= last; k++) {
        z -= g->dz[k];
        if (p->h[i] < z) { g->dzz[i][k] = 0; continue; }
        if (!found) {
          found = 1;
          if (k == last) { g->dzz[i][k] = p->h[i] + g->dv[i]; g->ctop[i] = k; }
          else if (p->h[i] == z) { g->dzz[i][k] = 0; g->ctop[i] = k + 1; }
          else { g->dzz[i][k] = p->h[i] - z; g->ctop[i] = k; }
        } else if (k == last) {
          g->dzz[i][k] = g->dz[k] + g->dv[i] + z;
        } else {
          g->dzz[i][k] = (z < -g->dv[i]) ? 0 : g->dz[k];
        }
      }
    }
  } else {
    for (int i = 0; i < Nc; i++) g->dzz[i][0] = g->dv[i] + p->h[i];
  }
  for (int j = 0; j < Ne; j++) {
    int a = g->grad[2 * j], b = g->grad[2 * j + 1];
    if (a == -1) g->etop[j] = g->ctop[b];
    else if (b == -1) g->etop[j] = g->ctop[a];
    else g->etop[j] = g->ctop[a] < g->ctop[b] ? g->ctop[a] : g->ctop[b];
  }
  if (option == -1) {
    for (int i = 0; i < Nc; i++) {
      int k = g->Nk[i] - 1;
      g->dzbot[i] = g->dzz[i][k];
      if (!g->stairstep && g->fixdzz && g->dzz[i][k] < g->dz[k] * g->dzsmall)
        g->dv[i] += g->dz[k] * g->dzsmall - g->dzz[i][k];
    }
  }
  if (option == 1) {
    for (int j = 0; j < Ne; j++) g->etopold[j] = g->etop[j];
    for (int i = 0; i < Nc; i++) {
      g->ctopold[i] = g->ctop[i];
      for (int k = 0; k < g->Nk[i]; k++) g->dzzold[i][k] = g->dzz[i][k];
    }
    for (int j = 0; j < Ne; j++) {
      int a = g->grad[2 * j], b = g->grad[2 * j + 1];
      if (a == -1) a = b;
      if (b == -1) b = a;
      double d1 = g->dzz[a][g->Nk[a] - 1], d2 = g->dzz[b][g->Nk[b] - 1], z = 0;
      for (int k = 0; k < g->Nke[j] - 1; k++) z -= g->dz[k];
      if (p->h[a] < z) d1 = d1 - z + p->h[a];
      if (p->h[b] < z) d2 = d2 - z + p->h[b];
      g->dzfB[j] = d1 < d2 ? d1 : d2;
    }
  }
}

/* flux height chosen by the sign of u: downstream-of-flow cell's thickness
 * (main/util.c:219-232) */
static double upwind(double u, double a, double b) {
  if (fabs(u) > 0) return u > 0 ? b : a;
  return b > a ? b : a;
}

void sun_set_flux_height(gridT *g, physT *p, const propT *prop) {
  (void)prop;
  for (int j = 0; j < g->Ne; j++)
    for (int k = 0; k < g->Nkc[j]; k++) g->dzf[j][k] = 0;
  if (g->smoothbot)
    for (int i = 0; i < g->Nc; i++) {
      double lo = g->smoothbot * g->dz[g->Nk[i] - 1];
      g->dzz[i][g->Nk[i] - 1] = g->dzbot[i] > lo ? g->dzbot[i] : lo;
    }
  for (int j = 0; j < g->Ne; j++) {
    int a = g->grad[2 * j], b = g->grad[2 * j + 1];
    if (a == -1) a = b;
    if (b == -1) b = a;
    for (int k = g->etop[j]; k < g->Nke[j]; k++)
      g->dzf[j][k] = upwind(p->u[j][k], g->dzz[a][k], g->dzz[b][k]);
    int kb = g->Nke[j] - 1;
    if (g->etop[j] == kb) {
      double dmin = g->dv[a] < g->dv[b] ? g->dv[a] : g->dv[b];
      double v = upwind(p->u[j][kb], p->h[a], p->h[b]) + dmin;
      g->dzf[j][kb] = v > 0 ? v : 0;
    } else {
      g->dzf[j][kb] = g->dzz[a][kb] < g->dzz[b][kb] ? g->dzz[a][kb] : g->dzz[b][kb];
    }
    for (int k = g->etop[j]; k < g->Nke[j]; k++)
      if (g->dzf[j][k] <= DRY_HEIGHT) g->dzf[j][k] = 0;
  }
  if (g->smoothbot)
    for (int i = 0; i < g->Nc; i++) {
      double lo = g->dzsmall * g->dz[g->Nk[i] - 1];
      double *bot = &g->dzz[i][g->Nk[i] - 1];
      if (lo > *bot) *bot = lo;
    }
}

void sun_compute_qsource(REAL **src, gridT *g, physT *p, const propT *prop) {
  const double tf = (1 - prop->theta) / prop->theta;
  for (int i = 0; i < g->Nc; i++)
    for (int k = g->ctop[i]; k < g->Nk[i]; k++) src[i][k] = 0;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip];
    const double Ac = g->Ac[i];
    for (int k = g->ctop[i]; k < g->Nk[i]; k++)
      src[i][k] = Ac * (p->w[i][k] - p->w[i][k + 1]) + tf * Ac * (p->wtmp2[i][k] - p->wtmp2[i][k + 1]);
    for (int nf = 0; nf < g->nfaces[i]; nf++) {
      const int ne = g->face[i * g->maxfaces + nf];
      const double nrm = g->normal[i * g->maxfaces + nf];
      for (int k = g->ctop[i]; k < g->Nke[ne]; k++)
        src[i][k] += (p->u[ne][k] + tf * p->utmp2[ne][k]) * g->dzf[ne][k] * nrm * g->df[ne];
    }
    for (int k = g->ctop[i]; k < g->Nk[i]; k++) src[i][k] /= prop->dt;
  }
  for (int j = 0; j < g->Ne; j++) p->D[j] = g->df[j] / g->dg[j];
}

/* ------------------------------------------------------------------------ */
/* flat access                                                               */

static const struct fdesc *find_phys(const char *name) {
  for (int f = 0; f < NPF; f++)
    if (!strcmp(phys_fields[f].name, name)) return &phys_fields[f];
  return NULL;
}

/* grid-owned real fields */
static int grid_real(const gridT *g, const char *name, int *kind, REAL ***rows, REAL **vec) {
  *rows = NULL; *vec = NULL;
  if (!strcmp(name, "dzz")) { *kind = 1; *rows = g->dzz; }
  else if (!strcmp(name, "dzzold")) { *kind = 1; *rows = g->dzzold; }
  else if (!strcmp(name, "dzf")) { *kind = 3; *rows = g->dzf; }
  else if (!strcmp(name, "Ac")) { *kind = 4; *vec = g->Ac; }
  else if (!strcmp(name, "dv")) { *kind = 4; *vec = g->dv; }
  else if (!strcmp(name, "xv")) { *kind = 4; *vec = g->xv; }
  else if (!strcmp(name, "yv")) { *kind = 4; *vec = g->yv; }
  else if (!strcmp(name, "df")) { *kind = 5; *vec = g->df; }
  else if (!strcmp(name, "dg")) { *kind = 5; *vec = g->dg; }
  else if (!strcmp(name, "n1")) { *kind = 5; *vec = g->n1; }
  else if (!strcmp(name, "n2")) { *kind = 5; *vec = g->n2; }
  else if (!strcmp(name, "xe")) { *kind = 5; *vec = g->xe; }
  else if (!strcmp(name, "ye")) { *kind = 5; *vec = g->ye; }
  else if (!strcmp(name, "dzfB")) { *kind = 5; *vec = g->dzfB; }
  else if (!strcmp(name, "dzbot")) { *kind = 4; *vec = g->dzbot; }
  else if (!strcmp(name, "dz")) { *kind = 7; *vec = g->dz; }
  else if (!strcmp(name, "def")) { *kind = 10; *vec = g->def; }
  else if (!strcmp(name, "xp")) { *kind = 11; *vec = g->xp; }
  else if (!strcmp(name, "yp")) { *kind = 11; *vec = g->yp; }
  else return 0;
  return 1;
}

static long xfer(const gridT *g, int kind, REAL **rows, REAL *vec, double *flat, int to_flat) {
  size_t n = 0;
  if (kind == 8 || kind == 9) {
    for (int j = g->edgedist[2]; j < g->edgedist[5]; j++) {
      size_t m = (size_t)(g->Nke[g->edgep[j]] + (kind == 8));
      REAL *row = rows[j - g->edgedist[2]];
      if (to_flat) memcpy(flat + n, row, m * sizeof(REAL));
      else memcpy(row, flat + n, m * sizeof(REAL));
      n += m;
    }
  } else if (kind <= 3) {
    const int nr = kind == 3 ? g->Ne : g->Nc;
    const int *len = kind == 3 ? g->Nkc : g->Nk;
    const int extra = kind == 2;
    for (int i = 0; i < nr; i++) {
      size_t m = (size_t)(len[i] + extra);
      if (to_flat) memcpy(flat + n, rows[i], m * sizeof(REAL));
      else memcpy(rows[i], flat + n, m * sizeof(REAL));
      n += m;
    }
  } else {
    n = kind == 7 ? (size_t)g->Nkmax : kind_len(g, kind);
    if (to_flat) memcpy(flat, vec, n * sizeof(REAL)); else memcpy(vec, flat, n * sizeof(REAL));
  }
  return (long)n;
}

long sun_field_len(const gridT *g, const char *name) {
  const struct fdesc *f = find_phys(name);
  if (f) return (long)kind_len(g, f->kind);
  if (!strcmp(name, "hfcoef")) return (long)g->maxfaces * g->Nc;
  int kind; REAL **rows; REAL *vec;
  if (grid_real(g, name, &kind, &rows, &vec)) return kind == 7 ? g->Nkmax : (long)kind_len(g, kind);
  return -1;
}

long sun_field_get(const gridT *g, const physT *p, const char *name, double *out) {
  const struct fdesc *f = find_phys(name);
  if (f) {
    void *slot = (char *)p + f->off;
    return xfer(g, f->kind, ROWS_KIND(f->kind) ? *(REAL ***)slot : NULL, ROWS_KIND(f->kind) ? NULL : *(REAL **)slot,
                out, 1);
  }
  if (!strcmp(name, "hfcoef")) {
    memcpy(out, p->hfcoef, sizeof(REAL) * (size_t)g->maxfaces * g->Nc);
    return (long)g->maxfaces * g->Nc;
  }
  int kind; REAL **rows; REAL *vec;
  if (grid_real(g, name, &kind, &rows, &vec)) return xfer(g, kind, rows, vec, out, 1);
  return -1;
}

long sun_field_set(gridT *g, physT *p, const char *name, const double *in) {
  const struct fdesc *f = find_phys(name);
  if (f) {
    void *slot = (char *)p + f->off;
    return xfer(g, f->kind, ROWS_KIND(f->kind) ? *(REAL ***)slot : NULL, ROWS_KIND(f->kind) ? NULL : *(REAL **)slot,
                (double *)in, 0);
  }
  int kind; REAL **rows; REAL *vec;
  if (grid_real(g, name, &kind, &rows, &vec)) return xfer(g, kind, rows, vec, (double *)in, 0);
  return -1;
}

static int *grid_ivec(const gridT *g, const char *name, long *n) {
  const long Nc = g->Nc, Ne = g->Ne, mf = g->maxfaces;
  struct { const char *nm; int *p; long n; } t[] = {
    {"Nk", g->Nk, Nc}, {"ctop", g->ctop, Nc}, {"ctopold", g->ctopold, Nc}, {"nfaces", g->nfaces, Nc},
    {"cellp", g->cellp, Nc}, {"neigh", g->neigh, Nc * mf}, {"face", g->face, Nc * mf},
    {"normal", g->normal, Nc * mf}, {"cells", g->cells, Nc * mf}, {"Nke", g->Nke, Ne},
    {"Nkc", g->Nkc, Ne}, {"etop", g->etop, Ne}, {"etopold", g->etopold, Ne}, {"mark", g->mark, Ne},
    {"edgep", g->edgep, Ne}, {"grad", g->grad, 2 * Ne}, {"gradf", g->gradf, 2 * Ne},
    {"edges", g->edges, 3 * Ne}, {"celldist", g->celldist, 3}, {"edgedist", g->edgedist, 6},
    {"mnptr", g->mnptr, Nc}, {"eptr", g->eptr, Ne}, {"edge_id", g->edge_id, Ne}, {"Nkp", g->Nkp, g->Np},
    {"eneigh", g->eneigh, 2 * (mf - 1) * Ne},
  };
  for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); i++)
    if (!strcmp(t[i].nm, name)) { *n = t[i].n; return t[i].p; }
  return NULL;
}

long sun_grid_int(const gridT *g, const char *name, int *out) {
  long n; int *v = grid_ivec(g, name, &n);
  if (!v) return -1;
  if (out) memcpy(out, v, sizeof(int) * (size_t)n);
  return n;
}

long sun_grid_set_int(gridT *g, const char *name, const int *in) {
  long n; int *v = grid_ivec(g, name, &n);
  if (!v) return -1;
  memcpy(v, in, sizeof(int) * (size_t)n);
  return n;
}

void *sun_phys_ptr(const physT *p, const char *name) {
  const struct fdesc *f = find_phys(name);
  if (f) return *(void **)((char *)p + f->off);
  if (!strcmp(name, "hfcoef")) return p->hfcoef;
  if (!strcmp(name, "qtmp")) return p->qtmp;
  if (!strcmp(name, "boundary_s")) return p->boundary_s;
  if (!strcmp(name, "active")) return p->active;
  return NULL;
}

void *sun_grid_ptr(const gridT *g, const char *name) {
  if (!strcmp(name, "dzz")) return g->dzz;
  if (!strcmp(name, "dzzold")) return g->dzzold;
  if (!strcmp(name, "dzf")) return g->dzf;
  return NULL;
}

size_t sun_prop_size(void) { return sizeof(propT); }
void sun_prop_copy(const propT *from, propT *to) { memcpy(to, from, sizeof(propT)); }

int sun_grid_set_param(gridT *g, const char *name, double value) {
  if (!strcmp(name, "smoothbot")) g->smoothbot = value;
  else if (!strcmp(name, "dzsmall")) g->dzsmall = value;
  else if (!strcmp(name, "fixdzz")) g->fixdzz = (int)value;
  else if (!strcmp(name, "stairstep")) g->stairstep = (int)value;
  else return -1;
  return 0;
}

double sun_grid_param(const gridT *g, const char *name) {
  if (!strcmp(name, "smoothbot")) return g->smoothbot;
  if (!strcmp(name, "dzsmall")) return g->dzsmall;
  if (!strcmp(name, "fixdzz")) return g->fixdzz;
  if (!strcmp(name, "stairstep")) return g->stairstep;
  return 0;
}

long sun_grid_scalar(const gridT *g, const char *name) {
  if (!strcmp(name, "Nc")) return g->Nc;
  if (!strcmp(name, "Ne")) return g->Ne;
  if (!strcmp(name, "Np")) return g->Np;
  if (!strcmp(name, "Nkmax")) return g->Nkmax;
  if (!strcmp(name, "maxfaces")) return g->maxfaces;
  if (!strcmp(name, "N")) {
    long n = 0;
    for (int i = 0; i < g->Nc; i++) n += g->Nk[i] - g->ctop[i];
    return n;
  }
  if (!strcmp(name, "Ek")) {
    long n = 0;
    for (int j = 0; j < g->Ne; j++) {
      int a = g->grad[2 * j], b = g->grad[2 * j + 1];
      if (a < 0 || b < 0) continue;
      int kt = g->ctop[a] > g->ctop[b] ? g->ctop[a] : g->ctop[b];
      if (g->Nke[j] > kt) n += g->Nke[j] - kt;
    }
    return n;
  }
  return -1;
}

#define PROP_D(f) if (!strcmp(name, #f)) { if (set) prop->f = value; return prop->f; }
#define PROP_I(f) if (!strcmp(name, #f)) { if (set) prop->f = (int)value; return prop->f; }
static double prop_access(propT *prop, const char *name, double value, int set, int *ok) {
  *ok = 1;
  PROP_D(dt) PROP_D(theta) PROP_D(thetaS) PROP_D(thetaB) PROP_D(epsilon) PROP_D(qepsilon)
  PROP_D(resnorm) PROP_D(grav) PROP_D(beta) PROP_D(kappa_s) PROP_D(kappa_sH) PROP_D(gamma)
  PROP_D(nu) PROP_D(nu_H) PROP_D(CdB) PROP_D(CdW) PROP_D(thetaM) PROP_D(laxWendroff_Vertical) PROP_D(rtime)
  PROP_I(n) PROP_I(nstart) PROP_I(nsteps) PROP_I(maxiters) PROP_I(qmaxiters) PROP_I(hprecond)
  PROP_I(qprecond) PROP_I(TVD) PROP_I(horiTVD) PROP_I(vertTVD) PROP_I(TVDsalt) PROP_I(TVDtemp)
  PROP_I(nonhydrostatic) PROP_I(cgsolver) PROP_I(linearFS) PROP_I(stairstep) PROP_I(wetdry)
  PROP_I(nonlinear) PROP_I(thetaramptime) PROP_I(netcdfBdy) PROP_I(AB) PROP_I(conserveMomentum)
  PROP_I(TVDmomentum)
  *ok = 0;
  return 0;
}
#undef PROP_D
#undef PROP_I

int sun_prop_set(propT *prop, const char *name, double value) {
  int ok; prop_access(prop, name, value, 1, &ok);
  return ok ? 0 : -1;
}
double sun_prop_get(const propT *prop, const char *name) {
  int ok; return prop_access((propT *)prop, name, 0, 0, &ok);
}

/* ------------------------------------------------------------------------ */
/* run-phase grid files                                                      */

/* main/grid.h:24-25, main/suntans.h:21 */
#define GRID_MAXBCTYPES 4
#define GRID_MAXMARKS 7
#define GRID_EDGECOLUMNS 3

/* The value of key in a suntans.dat-style file (main/fileio.c:237-267,
 * 462-487): the first line whose first word is key; the search stops at
 * the first empty line, and a value that is the line's last character is
 * not taken.  Returns 1 with the value's first word in val, else 0. */
static int dat_string(const char *file, const char *key, char *val, size_t n) {
  FILE *f = fopen(file, "r");
  if (!f) return 0;
  char line[4096];
  int found = 0;
  while (fgets(line, sizeof line, f)) {
    size_t len = strcspn(line, "\n");
    line[len] = 0;
    if (len == 0) break;
    size_t kl = 0;
    while (kl < len && !isspace((unsigned char)line[kl])) kl++;
    if (kl != strlen(key) || strncmp(line, key, kl)) continue;
    size_t i = kl;
    while (i < len && isspace((unsigned char)line[i])) i++;
    if (i + 1 == len) break;
    size_t vl = 0;
    while (i + vl < len && !isspace((unsigned char)line[i + vl]) && vl + 1 < n) vl++;
    memcpy(val, line + i, vl);
    val[vl] = 0;
    found = 1;
    break;
  }
  fclose(f);
  return found;
}

/* A file named by key, relative to the data directory unless absolute
 * (main/mympi.c:94-102), with suffix appended. */
static int dat_file(const char *dir, const char *datafile, const char *key, const char *suffix, char *out,
                    size_t n) {
  char v[1024];
  if (!dat_string(datafile, key, v, sizeof v)) return 0;
  if (v[0] == '/') snprintf(out, n, "%s%s", v, suffix);
  else snprintf(out, n, "%s/%s%s", dir, v, suffix);
  return 1;
}

/* Rows of a text file (main/fileio.c:183-196). */
static long file_lines(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  long n = 0;
  int c;
  while ((c = getc_unlocked(f)) != EOF)
    if (c == '\n') n++;
  fclose(f);
  return n;
}

/* The next number of a grid file (main/fileio.c:137-174): words split at
 * space, tab, CR and LF; a word with any character other than digits, '.',
 * '+', '-' and 'e' is skipped; 0 at the end of the file. */
static double next_field(FILE *f) {
  char w[256];
  for (;;) {
    int c, i = 0;
    while ((c = getc_unlocked(f)) == ' ' || c == '\t' || c == '\n' || c == '\r') {}
    if (c == EOF) return 0;
    while (c != EOF && c != ' ' && c != '\t' && c != '\n' && c != '\r') {
      if (i < 255) w[i++] = (char)c;
      c = getc_unlocked(f);
    }
    w[i] = 0;
    int ok = 1;
    for (int k = 0; k < i; k++)
      if (!isdigit((unsigned char)w[k]) && w[k] != '.' && w[k] != '-' && w[k] != '+' && w[k] != 'e') { ok = 0; break; }
    if (ok) return strtod(w, NULL);
  }
}

static void grid_vertical_rows(gridT *g, const char *datafile);

gridT *sun_grid_load(const char *dir) {
  char datafile[1024], v[256];
  snprintf(datafile, sizeof datafile, "%s/suntans.dat", dir);
  if (!dat_string(datafile, "Nkmax", v, sizeof v)) return NULL;
  gridT *g = (gridT *)xcalloc(1, sizeof(gridT));
  g->Nkmax = (int)strtod(v, NULL);
  g->maxfaces = dat_string(datafile, "maxFaces", v, sizeof v) ? (int)strtod(v, NULL) : 3;
  const char *keys[5] = {"celldata", "edgedata", "nodes", "topology", "vertspace"};
  char files[5][1100];
  for (int i = 0; i < 5; i++)
    if (!dat_file(dir, datafile, keys[i], i < 4 ? ".0" : "", files[i], sizeof files[i])) { free(g); return NULL; }
  const long Nc = file_lines(files[0]), Ne = file_lines(files[1]), Np = file_lines(files[2]);
  if (Nc <= 0 || Ne <= 0 || Np <= 0) { free(g); return NULL; }
  g->Nc = (int)Nc; g->Ne = (int)Ne; g->Np = (int)Np;
  const int mf = g->maxfaces;

  /* topology (main/gridio.c:365-425): one process, no neighbours */
  FILE *f = fopen(files[3], "r");
  if (!f) { free(g); return NULL; }
  const int nprocs = (int)next_field(f);
  g->Nneighs = (int)next_field(f);
  if (nprocs != 1 || g->Nneighs != 0) { fclose(f); free(g); return NULL; }
  g->myneighs = (int *)xcalloc(1, sizeof(int));
  g->celldist = (int *)xcalloc(GRID_MAXBCTYPES - 1, sizeof(int));
  g->edgedist = (int *)xcalloc(GRID_MAXMARKS - 1, sizeof(int));
  g->cellp = (int *)xcalloc(Nc, sizeof(int));
  g->edgep = (int *)xcalloc(Ne, sizeof(int));
  for (int n = 0; n < GRID_MAXBCTYPES - 1; n++) g->celldist[n] = (int)next_field(f);
  for (int n = 0; n < GRID_MAXMARKS - 1; n++) g->edgedist[n] = (int)next_field(f);
  for (int n = 0; n < Nc; n++) g->cellp[n] = (int)next_field(f);
  for (int n = 0; n < Ne; n++) g->edgep[n] = (int)next_field(f);
  fclose(f);

  /* cell-centred data (main/gridio.c:631-662) */
  planform_alloc(g);
  free(g->eneigh); g->eneigh = NULL;
  g->Nk = (int *)xcalloc(Nc, sizeof(int));
  g->mnptr = (int *)xcalloc(Nc, sizeof(int));
  f = fopen(files[0], "r");
  if (!f) { sun_grid_free(g); return NULL; }
  for (int n = 0; n < Nc; n++) {
    g->nfaces[n] = (int)next_field(f);
    if (g->nfaces[n] > mf || g->nfaces[n] < 0) { fclose(f); sun_grid_free(g); return NULL; }
    g->xv[n] = next_field(f);
    g->yv[n] = next_field(f);
    g->Ac[n] = next_field(f);
    g->dv[n] = next_field(f);
    g->Nk[n] = (int)next_field(f);
    for (int nf = 0; nf < g->nfaces[n]; nf++) g->face[mf * n + nf] = (int)next_field(f);
    for (int nf = 0; nf < g->nfaces[n]; nf++) g->neigh[mf * n + nf] = (int)next_field(f);
    for (int nf = 0; nf < g->nfaces[n]; nf++) g->normal[mf * n + nf] = (int)next_field(f);
    for (int nf = 0; nf < g->nfaces[n]; nf++) g->def[mf * n + nf] = next_field(f);
    for (int nf = 0; nf < g->nfaces[n]; nf++) g->cells[mf * n + nf] = (int)next_field(f);
    g->mnptr[n] = (int)next_field(f);
  }
  fclose(f);

  /* edge-centred data (main/gridio.c:687-711) */
  g->Nke = (int *)xcalloc(Ne, sizeof(int));
  g->Nkc = (int *)xcalloc(Ne, sizeof(int));
  g->edge_id = (int *)xcalloc(Ne, sizeof(int));
  g->eptr = (int *)xcalloc(Ne, sizeof(int));
  f = fopen(files[1], "r");
  if (!f) { sun_grid_free(g); return NULL; }
  for (int n = 0; n < Ne; n++) {
    g->df[n] = next_field(f);
    g->dg[n] = next_field(f);
    g->n1[n] = next_field(f);
    g->n2[n] = next_field(f);
    g->xe[n] = next_field(f);
    g->ye[n] = next_field(f);
    g->Nke[n] = (int)next_field(f);
    g->Nkc[n] = (int)next_field(f);
    g->grad[2 * n] = (int)next_field(f);
    g->grad[2 * n + 1] = (int)next_field(f);
    g->gradf[2 * n] = (int)next_field(f);
    g->gradf[2 * n + 1] = (int)next_field(f);
    g->mark[n] = (int)next_field(f);
    g->edges[GRID_EDGECOLUMNS * n] = (int)next_field(f);
    g->edges[GRID_EDGECOLUMNS * n + 1] = (int)next_field(f);
    g->edge_id[n] = (int)next_field(f);
    g->eptr[n] = (int)next_field(f);
  }
  fclose(f);

  /* nodes (main/gridio.c:728-783): coordinates; the neighbour lists and
   * per-layer areas are skipped (nothing on the path reads them) */
  g->xp = (REAL *)xcalloc(Np, sizeof(REAL));
  g->yp = (REAL *)xcalloc(Np, sizeof(REAL));
  g->Nkp = (int *)xcalloc(Np, sizeof(int));
  f = fopen(files[2], "r");
  if (!f) { sun_grid_free(g); return NULL; }
  for (int n = 0; n < Np; n++) {
    next_field(f);
    g->xp[n] = next_field(f);
    g->yp[n] = next_field(f);
    for (int list = 0; list < 3; list++) {
      const int m = (int)next_field(f);
      for (int j = 0; j < m; j++) next_field(f);
    }
    g->Nkp[n] = (int)next_field(f);
    for (int j = 0; j < g->Nkp[n]; j++) next_field(f);
  }
  fclose(f);

  /* vertical spacing (main/gridio.c:496-504) and the vertical grid of
   * InitializeVerticalGrid (main/phys.c:765-799) */
  g->dz = (REAL *)xcalloc(g->Nkmax, sizeof(REAL));
  f = fopen(files[4], "r");
  if (!f) { sun_grid_free(g); return NULL; }
  for (int k = 0; k < g->Nkmax; k++) g->dz[k] = next_field(f);
  fclose(f);
  grid_vertical_rows(g, datafile);
  return g;
}

/* The vertical state InitializeVerticalGrid starts from (main/phys.c:765-799)
 * on a grid whose Nk, Nkc and dz are set, with the suntans.dat switches it
 * reads. */
static void grid_vertical_rows(gridT *g, const char *datafile) {
  char v[256];
  const int Nc = g->Nc, Ne = g->Ne;
  g->stairstep = dat_string(datafile, "stairstep", v, sizeof v) ? (int)strtod(v, NULL) : 0;
  g->fixdzz = dat_string(datafile, "fixdzz", v, sizeof v) ? (int)strtod(v, NULL) : 0;
  g->dzsmall = dat_string(datafile, "dzsmall", v, sizeof v) ? strtod(v, NULL) : 0;
  g->smoothbot = dat_string(datafile, "smoothbot", v, sizeof v) ? strtod(v, NULL) : 0;
  g->ctop = (int *)xcalloc(Nc, sizeof(int));
  g->ctopold = (int *)xcalloc(Nc, sizeof(int));
  g->etop = (int *)xcalloc(Ne, sizeof(int));
  g->etopold = (int *)xcalloc(Ne, sizeof(int));
  g->dztop = (REAL *)xcalloc(Nc, sizeof(REAL));
  g->dzbot = (REAL *)xcalloc(Nc, sizeof(REAL));
  g->dzfB = (REAL *)xcalloc(Ne, sizeof(REAL));
  g->dzz = rows_alloc((int)Nc, g->Nk, 0);
  g->dzzold = rows_alloc((int)Nc, g->Nk, 0);
  for (int i = 0; i < Nc; i++)
    for (int k = 0; k < g->Nk[i]; k++) g->dzz[i][k] = g->dzzold[i][k] = g->dz[k];
  g->dzf = rows_alloc((int)Ne, g->Nkc, 0);
}

/* ------------------------------------------------------------------------ */
/* grid generation                                                           */

/* main/suntans.h:23 (the reference's value of pi, which sets the angle
 * test of the Voronoi-point correction) */
#define GRID_PI 3.141592654

/* A number from suntans.dat as MPI_GetValue reads it (main/mympi.c:55-77):
 * the file's value, else the default of main/fileio.c:277 where one exists
 * (has_default), else an error. */
static int gen_value(const char *datafile, const char *key, int has_default, double dflt, double *out) {
  char v[256];
  if (dat_string(datafile, key, v, sizeof v)) { *out = strtod(v, NULL); return 1; }
  if (has_default) { *out = dflt; return 1; }
  fprintf(stderr, "sun_grid_generate: %s: no value for %s\n", datafile, key);
  return 0;
}

/* Columns of a grid file's first line (main/fileio.c:205-228): a blank or
 * tab after a word starts a column; a blank just before the line's end
 * takes one back. */
static int gen_columns(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return 0;
  int n = 1, c, prev = 0;
  while ((c = fgetc(f)) != EOF) {
    if ((c == ' ' || c == '\t') && !(prev == ' ' || prev == '\t' || prev == 0)) n++;
    if (c == '\n') { if (prev == ' ') n--; break; }
    prev = c;
  }
  fclose(f);
  return n;
}

/* The vertical spacing of GetDZ (main/initialization.c:27-56): uniform for
 * rstretch 1, geometric growing downward for 1 < r <= 1.1 and upward for
 * -1.1 < r < -1.  Returns 0 for any other r. */
static int gen_dz(double *dz, double depth, int Nkmax, double r) {
  if (r == 1) {
    for (int k = 0; k < Nkmax; k++) dz[k] = depth / Nkmax;
  } else if (r > 1 && r <= 1.1) {
    dz[0] = depth * (r - 1) / (pow(r, Nkmax) - 1);
    for (int k = 1; k < Nkmax; k++) dz[k] = r * dz[k - 1];
  } else if (r > -1.1 && r < -1) {
    r = fabs(r);
    dz[Nkmax - 1] = depth * (r - 1) / (pow(r, Nkmax) - 1);
    for (int k = Nkmax - 2; k >= 0; k--) dz[k] = r * dz[k + 1];
  } else {
    return 0;
  }
  return 1;
}

/* Area of a polygon as GetArea computes it (main/grid.c:3647-3678): a
 * triangle from its base and height, a polygon as the fan from its first
 * point. */
static double gen_tri_area(const double *x, const double *y) {
  const double b = sqrt(pow(x[1] - x[0], 2) + pow(y[1] - y[0], 2));
  const double r2 = sqrt(pow(x[2] - x[1], 2) + pow(y[2] - y[1], 2));
  const double l = ((x[0] - x[1]) * (x[2] - x[1]) + (y[0] - y[1]) * (y[2] - y[1])) / b;
  const double h = sqrt(r2 * r2 - l * l);
  return 0.5 * b * h;
}
static double gen_area(const double *x, const double *y, int nf) {
  if (nf < 3) return 0;
  double a = gen_tri_area(x, y);
  for (int i = 0; i < nf - 3; i++) {
    const double tx[3] = {x[i + 2], x[i + 3], x[0]}, ty[3] = {y[i + 2], y[i + 3], y[0]};
    a += gen_tri_area(tx, ty);
  }
  return a;
}

/* face and gradf of a planform from grad and neigh (main/grid.c:654-713).
 * First each interior edge takes the slot of its neighbour in both cells,
 * scanning the first cell's nfaces slots in both (so a slot of the second
 * cell past the first cell's count is not seen here); then every cell side
 * still without its edge takes the first free slot, edges in index order. */
static void gen_faces(int *grad, int *gradf, const int *neigh, int *face, const int *nfaces, int mf, int Nc,
                      int Ne) {
  for (int n = 0; n < Nc; n++)
    for (int nf = 0; nf < nfaces[n]; nf++) face[n * mf + nf] = -1;
  for (int e = 0; e < 2 * Ne; e++) gradf[e] = -1;
  for (int e = 0; e < Ne; e++) {
    const int c1 = grad[2 * e], c2 = grad[2 * e + 1];
    if (c1 == -1 || c2 == -1) continue;
    for (int nf = 0; nf < nfaces[c1]; nf++) {
      if (neigh[c1 * mf + nf] == c2) { face[c1 * mf + nf] = e; gradf[2 * e] = nf; }
      if (neigh[c2 * mf + nf] == c1) { face[c2 * mf + nf] = e; gradf[2 * e + 1] = nf; }
    }
  }
  for (int e = 0; e < Ne; e++)
    for (int j = 0; j < 2; j++) {
      const int c = grad[2 * e + j];
      if (c == -1) continue;
      int have = 0;
      for (int nf = 0; nf < nfaces[c]; nf++) have |= face[c * mf + nf] == e;
      if (have) continue;
      for (int nf = 0; nf < nfaces[c]; nf++)
        if (face[c * mf + nf] == -1) { face[c * mf + nf] = e; gradf[2 * e + j] = nf; break; }
    }
}

/* Triangle corners wound with the faces (main/grid.c:754-832): corner nf+1
 * is the point faces nf and nf+1 share, corner 0 the point of the last and
 * first faces (the first of the earlier face's two points found in the
 * other).  Returns the number of sides with no shared point. */
static int gen_shared(const int *e1, const int *e2) {
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      if (e1[i] == e2[j]) return e1[i];
  return -1;
}
static int gen_wind_triangles(const int *face, const int *edges, int *cells, const int *nfaces, int mf, int Nc) {
  int bad = 0;
  for (int c = 0; c < Nc; c++) {
    const int nff = nfaces[c];
    for (int nf = 0; nf < nff - 1; nf++) {
      const int s = gen_shared(edges + 3 * face[c * mf + nf], edges + 3 * face[c * mf + nf + 1]);
      if (s == -1) bad++;
      else cells[mf * c + nf + 1] = s;
    }
    const int s = gen_shared(edges + 3 * face[c * mf + nff - 1], edges + 3 * face[c * mf]);
    if (s == -1) bad++;
    else cells[mf * c] = s;
  }
  return bad;
}

static int gen_fail(const char *what, const char *detail) {
  fprintf(stderr, "sun_grid_generate: %s%s%s\n", what, detail ? ": " : "", detail ? detail : "");
  return 0;
}

static int gen_write(const char *path, FILE **f) {
  *f = fopen(path, "w");
  if (!*f) return gen_fail("cannot write", path);
  return 1;
}

/* Text output formatted in parallel: lines [lo, hi) of a file go to one
 * thread's buffer through the C library's own printf, so the bytes are
 * those of a serial fprintf; the buffers are written in line order. */
struct gen_buf { char *p; size_t n, cap; };
static void gb_printf(struct gen_buf *b, const char *fmt, ...) {
  for (;;) {
    va_list ap;
    va_start(ap, fmt);
    const int m = vsnprintf(b->p + b->n, b->cap - b->n, fmt, ap);
    va_end(ap);
    if (m < 0) return;
    if ((size_t)m < b->cap - b->n) { b->n += (size_t)m; return; }
    b->cap = 2 * b->cap + (size_t)m + 1;
    b->p = (char *)realloc(b->p, b->cap);
    if (!b->p) { fprintf(stderr, "sun_grid_generate: out of memory\n"); abort(); }
  }
}
/* what the per-line writers read */
struct gen_out {
  const gridT *g;
  const int *pc_off, *pc, *pe_off, *pe, *pp, *npp;
};
typedef void (*gen_line_fn)(struct gen_buf *b, long i, const struct gen_out *o);
struct gen_job { gen_line_fn fn; const struct gen_out *o; long lo, hi; struct gen_buf *b; };
static void *gen_run(void *arg) {
  struct gen_job *j = (struct gen_job *)arg;
  j->b->n = 0;
  for (long i = j->lo; i < j->hi; i++) j->fn(j->b, i, j->o);
  return NULL;
}
static int gen_emit(const char *path, long n, gen_line_fn fn, const struct gen_out *o) {
  FILE *f = fopen(path, "w");
  if (!f) return gen_fail("cannot write", path);
  enum { GEN_TMAX = 16, GEN_CHUNK = 4096 };
  long nt = sysconf(_SC_NPROCESSORS_ONLN);
  if (nt < 1) nt = 1;
  if (nt > GEN_TMAX) nt = GEN_TMAX;
  struct gen_buf b[GEN_TMAX];
  for (int t = 0; t < nt; t++) { b[t].cap = 1 << 20; b[t].n = 0; b[t].p = (char *)xcalloc(b[t].cap, 1); }
  for (long base = 0; base < n; base += nt * GEN_CHUNK) {
    pthread_t th[GEN_TMAX];
    struct gen_job jb[GEN_TMAX];
    int started[GEN_TMAX];
    for (int t = 0; t < nt; t++) {
      const long lo = base + t * GEN_CHUNK, hi = lo + GEN_CHUNK < n ? lo + GEN_CHUNK : n;
      jb[t] = (struct gen_job){fn, o, lo, lo < hi ? hi : lo, &b[t]};
      started[t] = t > 0 && lo < hi && pthread_create(&th[t], NULL, gen_run, &jb[t]) == 0;
    }
    for (int t = 0; t < nt; t++) if (!started[t]) gen_run(&jb[t]);
    for (int t = 0; t < nt; t++) {
      if (started[t]) pthread_join(th[t], NULL);
      fwrite(b[t].p, 1, b[t].n, f);
    }
  }
  for (int t = 0; t < nt; t++) free(b[t].p);
  const int bad = ferror(f);
  fclose(f);
  if (bad) return gen_fail("write failed", path);
  return 1;
}

/* one line of each run-phase file (main/gridio.c:608-627, 562-578,
 * 631-662, 687-711, 728-783) */
static void line_cells0(struct gen_buf *b, long j, const struct gen_out *o) {
  const gridT *g = o->g;
  const int mf = g->maxfaces;
  gb_printf(b, "%d ", g->nfaces[j]);
  gb_printf(b, "%e %e ", g->xv[j], g->yv[j]);
  for (int nf = 0; nf < g->nfaces[j]; nf++) gb_printf(b, "%d ", g->cells[j * mf + nf]);
  for (int nf = 0; nf < g->nfaces[j]; nf++) gb_printf(b, "%d ", g->neigh[j * mf + nf]);
  gb_printf(b, "\n");
}
static void line_edges0(struct gen_buf *b, long j, const struct gen_out *o) {
  const gridT *g = o->g;
  gb_printf(b, "%d %d %d %d %d %d \n", g->edges[3 * j], g->edges[3 * j + 1], g->mark[j], g->grad[2 * j],
            g->grad[2 * j + 1], g->edge_id[j]);
}
static void line_celldata(struct gen_buf *b, long n, const struct gen_out *o) {
  const gridT *g = o->g;
  const int mf = g->maxfaces;
  gb_printf(b, "%d %e %e %e %e %d ", g->nfaces[n], g->xv[n], g->yv[n], g->Ac[n], g->dv[n], g->Nk[n]);
  for (int nf = 0; nf < g->nfaces[n]; nf++) gb_printf(b, "%d ", g->face[mf * n + nf]);
  for (int nf = 0; nf < g->nfaces[n]; nf++) gb_printf(b, "%d ", g->neigh[mf * n + nf]);
  for (int nf = 0; nf < g->nfaces[n]; nf++) gb_printf(b, "%d ", g->normal[mf * n + nf]);
  for (int nf = 0; nf < g->nfaces[n]; nf++) gb_printf(b, "%e ", g->def[mf * n + nf]);
  for (int nf = 0; nf < g->nfaces[n]; nf++) gb_printf(b, "%d ", g->cells[mf * n + nf]);
  gb_printf(b, "%d ", g->mnptr[n]);
  gb_printf(b, "\n");
}
static void line_edgedata(struct gen_buf *b, long n, const struct gen_out *o) {
  const gridT *g = o->g;
  gb_printf(b, "%e %e %e %e %e %e %d %d %d %d %d %d %d %d %d %d %d\n", g->df[n], g->dg[n], g->n1[n], g->n2[n],
            g->xe[n], g->ye[n], g->Nke[n], g->Nkc[n], g->grad[2 * n], g->grad[2 * n + 1], g->gradf[2 * n],
            g->gradf[2 * n + 1], g->mark[n], g->edges[3 * n], g->edges[3 * n + 1], g->edge_id[n], g->eptr[n]);
}
static void line_nodes(struct gen_buf *b, long n, const struct gen_out *o) {
  const gridT *g = o->g;
  gb_printf(b, "%d %e %e ", (int)n, g->xp[n], g->yp[n]);
  gb_printf(b, "%d ", o->npp[n]);
  for (int i = 0; i < o->npp[n]; i++) gb_printf(b, "%d ", o->pp[o->pe_off[n] + i]);
  gb_printf(b, "%d ", o->pe_off[n + 1] - o->pe_off[n]);
  for (int i = o->pe_off[n]; i < o->pe_off[n + 1]; i++) gb_printf(b, "%d ", o->pe[i]);
  gb_printf(b, "%d ", o->pc_off[n + 1] - o->pc_off[n]);
  for (int i = o->pc_off[n]; i < o->pc_off[n + 1]; i++) gb_printf(b, "%d ", o->pc[i]);
  gb_printf(b, "%d ", g->Nkp[n]);
  /* the summed area of the point's cells in each layer */
  for (int k = 0; k < g->Nkp[n]; k++) {
    double a = 0;
    for (int i = o->pc_off[n]; i < o->pc_off[n + 1]; i++)
      if (k < g->Nk[o->pc[i]]) a += g->Ac[o->pc[i]];
    gb_printf(b, "%e ", a);
  }
  gb_printf(b, "\n");
}

/* Depth at each query point from scattered soundings (InterpDepth,
 * main/grid.c:3686-3733, with Interp and FindNearest, main/util.c): the
 * np = maxFaces + 1 soundings nearest to the point, in order of squared
 * distance and then index (FindNearest's repeated scans), weighted by
 * inverse squared distance; a point on a sounding takes that sounding's
 * depth, or the second's where several coincide with it (the scan returns at
 * the second zero distance it meets).  The reference scans every sounding
 * for every point; here soundings sit in a bucket grid searched in growing
 * rings until no unvisited bucket can hold a nearer sounding. */
struct gen_near { double d; int i; };
static int gen_near_before(const struct gen_near *a, const struct gen_near *b) {
  return a->d < b->d || (a->d == b->d && a->i < b->i);
}
static int gen_interp_depth(const double *xd, const double *yd, const double *zd, int Nd, const double *xq,
                            const double *yq, double *zq, long Nq, int np) {
  if (Nd < np) return 0;
  double x0 = xd[0], x1 = xd[0], y0 = yd[0], y1 = yd[0];
  for (int i = 1; i < Nd; i++) {
    if (xd[i] < x0) x0 = xd[i];
    if (xd[i] > x1) x1 = xd[i];
    if (yd[i] < y0) y0 = yd[i];
    if (yd[i] > y1) y1 = yd[i];
  }
  /* about two soundings per bucket */
  const double w = x1 - x0 > 0 ? x1 - x0 : 1, h = y1 - y0 > 0 ? y1 - y0 : 1;
  double side = sqrt(2.0 * w * h / Nd);
  if (!(side > 0)) side = w > h ? w : h;
  long nbx = (long)(w / side) + 1, nby = (long)(h / side) + 1;
  while (nbx * nby > 4L * Nd + 16) { side *= 1.5; nbx = (long)(w / side) + 1; nby = (long)(h / side) + 1; }
  int *start = (int *)xcalloc(nbx * nby + 1, sizeof(int)), *items = (int *)xcalloc(Nd, sizeof(int));
  int *fillc = (int *)xcalloc(nbx * nby, sizeof(int));
  long *bk = (long *)xcalloc(Nd, sizeof(long));
  for (int i = 0; i < Nd; i++) {
    long bx = (long)((xd[i] - x0) / side), by = (long)((yd[i] - y0) / side);
    if (bx >= nbx) bx = nbx - 1;
    if (by >= nby) by = nby - 1;
    bk[i] = by * nbx + bx;
    start[bk[i] + 1]++;
  }
  for (long b = 0; b < nbx * nby; b++) start[b + 1] += start[b];
  for (int i = 0; i < Nd; i++) items[start[bk[i]] + fillc[bk[i]]++] = i;
  struct gen_near best[65];
  for (long n = 0; n < Nq; n++) {
    const double qx = xq[n], qy = yq[n];
    long cx = (long)floor((qx - x0) / side), cy = (long)floor((qy - y0) / side);
    if (cx < 0) cx = 0;
    if (cx >= nbx) cx = nbx - 1;
    if (cy < 0) cy = 0;
    if (cy >= nby) cy = nby - 1;
    int m = 0, zeros = 0, zero1 = -1, zero2 = -1;
    for (long r = 0;; r++) {
      for (long by = cy - r; by <= cy + r; by++) {
        if (by < 0 || by >= nby) continue;
        for (long bx = cx - r; bx <= cx + r; bx++) {
          if (bx < 0 || bx >= nbx) continue;
          if (r > 0 && by != cy - r && by != cy + r && bx != cx - r && bx != cx + r) continue;
          for (int t = start[by * nbx + bx]; t < start[by * nbx + bx + 1]; t++) {
            const int i = items[t];
            const struct gen_near c = {(xd[i] - qx) * (xd[i] - qx) + (yd[i] - qy) * (yd[i] - qy), i};
            if (c.d == 0) {
              zeros++;
              if (zero1 < 0 || i < zero1) { zero2 = zero1; zero1 = i; }
              else if (zero2 < 0 || i < zero2) zero2 = i;
            }
            if (m < np || gen_near_before(&c, &best[np - 1])) {
              int j = m < np ? m++ : np - 1;
              while (j > 0 && gen_near_before(&c, &best[j - 1])) { best[j] = best[j - 1]; j--; }
              best[j] = c;
            }
          }
        }
      }
      if (cx - r <= 0 && cy - r <= 0 && cx + r >= nbx - 1 && cy + r >= nby - 1) break;
      if (m == np) {
        /* the nearest a sounding outside the searched square can be */
        const double lx = x0 + (cx - r) * side, hx = x0 + (cx + r + 1) * side;
        const double ly = y0 + (cy - r) * side, hy = y0 + (cy + r + 1) * side;
        if (qx > lx && qx < hx && qy > ly && qy < hy) {
          double g = qx - lx;
          if (hx - qx < g) g = hx - qx;
          if (qy - ly < g) g = qy - ly;
          if (hy - qy < g) g = hy - qy;
          if (g * g > best[np - 1].d * (1 + 1e-9) + 1e-300) break;
        }
      }
    }
    if (zeros) {
      zq[n] = zd[zeros >= 2 ? zero2 : zero1];
      continue;
    }
    double z = 0, rs = 0;
    for (int j = 0; j < np; j++) {
      const int i = best[j].i;
      const double dist = pow(xd[i] - qx, 2.0) + pow(yd[i] - qy, 2.0);
      const double r0 = 1.0 / dist;
      z += zd[i] * r0;
      rs += r0;
    }
    zq[n] = z / rs;
  }
  free(start); free(items); free(fillc); free(bk);
  return 1;
}

gridT *sun_grid_generate(const char *dir, double (*depth)(double x, double y)) {
  char datafile[1024];
  snprintf(datafile, sizeof datafile, "%s/suntans.dat", dir);
  double Nkmax_d, stair_d, intdepth_d, mindepth, dzsmall, maxfaces_d, correct_d, vgc_d, rstretch;
  if (!gen_value(datafile, "maxFaces", 1, 3, &maxfaces_d) || !gen_value(datafile, "CorrectVoronoi", 0, 0, &correct_d) ||
      !gen_value(datafile, "Nkmax", 0, 0, &Nkmax_d) || !gen_value(datafile, "stairstep", 0, 0, &stair_d) ||
      !gen_value(datafile, "IntDepth", 0, 0, &intdepth_d) || !gen_value(datafile, "minimum_depth", 1, 0, &mindepth) ||
      !gen_value(datafile, "dzsmall", 0, 0, &dzsmall) || !gen_value(datafile, "vertgridcorrect", 0, 0, &vgc_d) ||
      !gen_value(datafile, "rstretch", 0, 0, &rstretch))
    return NULL;
  const int mf = (int)maxfaces_d, correct = (int)correct_d, intdepth = (int)intdepth_d;
  int Nkmax = (int)Nkmax_d;
  const int stairstep = (int)stair_d, vgc = (int)vgc_d;
  (void)dzsmall;
  double vratio = 0;
  if (correct != 0 && !gen_value(datafile, "VoronoiRatio", 0, 0, &vratio)) return NULL;
  double scaledepth = 0, scalefactor = 1;
  if (intdepth == 1 && (!gen_value(datafile, "scaledepth", 0, 0, &scaledepth) ||
                        !gen_value(datafile, "scaledepthfactor", 0, 0, &scalefactor)))
    return NULL;
  if (mf > 64) { gen_fail("maxFaces above 64", NULL); return NULL; }
  if (intdepth != 2 && intdepth != 1 && !depth) { gen_fail("IntDepth 0 needs the depth function (ReturnDepth)", NULL); return NULL; }
  if (mf < 3 || Nkmax < 1) { gen_fail("maxFaces and Nkmax must be at least 3 and 1", NULL); return NULL; }

  char fpts[1100], fedg[1100], fcel[1100], fdep[1100], fcd[1100], fed[1100], fnod[1100], ftop[1100], fvs[1100];
  if (!dat_file(dir, datafile, "points", "", fpts, sizeof fpts) || !dat_file(dir, datafile, "edges", "", fedg, sizeof fedg) ||
      !dat_file(dir, datafile, "cells", "", fcel, sizeof fcel) || !dat_file(dir, datafile, "depth", "", fdep, sizeof fdep) ||
      !dat_file(dir, datafile, "celldata", ".0", fcd, sizeof fcd) ||
      !dat_file(dir, datafile, "edgedata", ".0", fed, sizeof fed) ||
      !dat_file(dir, datafile, "nodes", ".0", fnod, sizeof fnod) ||
      !dat_file(dir, datafile, "topology", ".0", ftop, sizeof ftop) ||
      !dat_file(dir, datafile, "vertspace", "", fvs, sizeof fvs)) {
    gen_fail("suntans.dat lacks a grid file name", datafile);
    return NULL;
  }
  const long Np = file_lines(fpts), Ne = file_lines(fedg), Nc = file_lines(fcel);
  if (Np <= 0 || Ne <= 0 || Nc <= 0) { gen_fail("empty or missing points, edges or cells file", dir); return NULL; }

  /* the planform as read (main/gridio.c:515-603); slots past a cell's
   * count start at zero */
  double *xp = (double *)xcalloc(Np, sizeof(double)), *yp = (double *)xcalloc(Np, sizeof(double));
  int *medges = (int *)xcalloc(3 * Ne, sizeof(int)), *mmark = (int *)xcalloc(Ne, sizeof(int));
  int *mgrad = (int *)xcalloc(2 * Ne, sizeof(int)), *mgradf = (int *)xcalloc(2 * Ne, sizeof(int));
  int *medge_id = (int *)xcalloc(Ne, sizeof(int));
  int *mnf = (int *)xcalloc(Nc, sizeof(int)), *mcells = (int *)xcalloc(mf * Nc, sizeof(int));
  int *mneigh = (int *)xcalloc(mf * Nc, sizeof(int)), *mface = (int *)xcalloc(mf * Nc, sizeof(int));
  double *mxv = (double *)xcalloc(Nc, sizeof(double)), *myv = (double *)xcalloc(Nc, sizeof(double));
  double *mdv = (double *)xcalloc(Nc, sizeof(double));
  int *mNk = (int *)xcalloc(Nc, sizeof(int)), *lcptr = (int *)xcalloc(Nc, sizeof(int));
  int *leptr = (int *)xcalloc(Ne, sizeof(int));
  int *pc_off = NULL, *pc = NULL, *fill = NULL, *pe_off = NULL, *pe = NULL, *pp = NULL, *npp = NULL;
  int *type3 = NULL;
  gridT *g = NULL;
  int ok = 0;
  FILE *f = NULL;
  do {
    if (!(f = fopen(fpts, "r"))) { gen_fail("cannot read", fpts); break; }
    for (long n = 0; n < Np; n++) { xp[n] = next_field(f); yp[n] = next_field(f); next_field(f); }
    fclose(f);
    const int ecols = gen_columns(fedg);
    if (!(f = fopen(fedg, "r"))) { gen_fail("cannot read", fedg); break; }
    for (long n = 0; n < Ne; n++) {
      medges[3 * n] = (int)next_field(f);
      medges[3 * n + 1] = (int)next_field(f);
      mmark[n] = (int)next_field(f);
      mgrad[2 * n] = (int)next_field(f);
      mgrad[2 * n + 1] = (int)next_field(f);
      medge_id[n] = ecols > 5 ? (int)next_field(f) : 0;
    }
    fclose(f);
    const int ccols = gen_columns(fcel);
    if (!(f = fopen(fcel, "r"))) { gen_fail("cannot read", fcel); break; }
    int bad = 0;
    for (long n = 0; n < Nc && !bad; n++) {
      mnf[n] = ccols > 8 ? (int)next_field(f) : 3;
      if (mnf[n] > mf || mnf[n] < 3) { bad = 1; break; }
      mxv[n] = next_field(f);
      myv[n] = next_field(f);
      for (int nf = 0; nf < mnf[n]; nf++) mcells[n * mf + nf] = (int)next_field(f);
      for (int nf = 0; nf < mnf[n]; nf++) mneigh[n * mf + nf] = (int)next_field(f);
    }
    fclose(f);
    if (bad) { gen_fail("a cell has more faces than maxFaces (or fewer than 3)", fcel); break; }
    for (long e = 0; e < Ne; e++)
      for (int j = 0; j < 2; j++)
        if (medges[3 * e + j] < 0 || medges[3 * e + j] >= Np || mgrad[2 * e + j] < -1 || mgrad[2 * e + j] >= Nc) bad = 1;
    for (long n = 0; n < Nc; n++)
      for (int nf = 0; nf < mnf[n]; nf++)
        if (mcells[n * mf + nf] < 0 || mcells[n * mf + nf] >= Np || mneigh[n * mf + nf] < -1 ||
            mneigh[n * mf + nf] >= Nc)
          bad = 1;
    if (bad) { gen_fail("an index in the edges or cells file is out of range", dir); break; }

    /* Voronoi points too close to their neighbours' moved toward the
     * centroids (CorrectVoronoi, main/grid.c:3735-3821): edge by edge in
     * index order, each test reading the points as earlier edges left them */
    if (correct > 0) {
      for (long e = 0; e < Ne; e++) {
        const int c1 = mgrad[2 * e], c2 = mgrad[2 * e + 1];
        double cx[2] = {0, 0}, cy[2] = {0, 0};
        const int cc[2] = {c1, c2};
        for (int j = 0; j < 2; j++) {
          if (cc[j] == -1) continue;
          for (int nf = 0; nf < mnf[cc[j]]; nf++) {
            cx[j] += xp[mcells[cc[j] * mf + nf]] / mnf[cc[j]];
            cy[j] += yp[mcells[cc[j] * mf + nf]] / mnf[cc[j]];
          }
        }
        if (c1 != -1 && c2 != -1 && mnf[c1] == 3 && mnf[c2] == 3) {
          const double xc = 0.5 * (cx[0] + cx[1]), yc = 0.5 * (cy[0] + cy[1]);
          const double dg0 = sqrt(pow(cx[1] - cx[0], 2) + pow(cy[1] - cy[0], 2));
          const double dg = sqrt(pow(mxv[c2] - mxv[c1], 2) + pow(myv[c2] - myv[c1], 2));
          if (dg < vratio * dg0) {
            mxv[c1] = xc + vratio * (cx[0] - xc);
            mxv[c2] = xc + vratio * (cx[1] - xc);
            myv[c1] = yc + vratio * (cy[0] - yc);
            myv[c2] = yc + vratio * (cy[1] - yc);
          }
        } else {
          /* a boundary edge, or an edge of a non-triangle: the cell on the
           * first side that exists against the edge's midpoint */
          const double xc = 0.5 * (xp[medges[3 * e]] + xp[medges[3 * e + 1]]);
          const double yc = 0.5 * (yp[medges[3 * e]] + yp[medges[3 * e + 1]]);
          const int j = c1 == -1 ? 1 : 0, c = cc[j];
          const double dg0 = sqrt(pow(cx[j] - xc, 2) + pow(cy[j] - yc, 2));
          const double dg = sqrt(pow(mxv[c] - xc, 2) + pow(myv[c] - yc, 2));
          if (dg < vratio * dg0) {
            mxv[c] = xc + vratio * (cx[j] - xc);
            myv[c] = yc + vratio * (cy[j] - yc);
          }
        }
      }
    }

    /* Voronoi points of obtuse triangles moved to the centroid
     * (CorrectAngles, main/grid.c:3823-3869) */
    if (correct < 0) {
      const double cosr = cos(vratio * GRID_PI / 180.0);
      for (long n = 0; n < Nc; n++) {
        if (mnf[n] != 3) continue;
        const int *pc = mcells + n * mf;
        const double x1 = xp[pc[0]], x2 = xp[pc[1]], x3 = xp[pc[2]];
        const double y1 = yp[pc[0]], y2 = yp[pc[1]], y3 = yp[pc[2]];
        const double m1 = sqrt(pow(x2 - x1, 2) + pow(y2 - y1, 2));
        const double m2 = sqrt(pow(x3 - x1, 2) + pow(y3 - y1, 2));
        const double m3 = sqrt(pow(x3 - x2, 2) + pow(y3 - y2, 2));
        const double d1 = (m1 * m1 + m2 * m2 - m3 * m3) / (2 * m1 * m2);
        const double d2 = (m2 * m2 + m3 * m3 - m1 * m1) / (2 * m2 * m3);
        const double d3 = (m3 * m3 + m1 * m1 - m2 * m2) / (2 * m3 * m1);
        if (d1 <= cosr || d2 <= cosr || d3 <= cosr) {
          const double sx = 0 + (x1 + x2 + x3) / 3, sy = 0 + (y1 + y2 + y3) / 3;
          mxv[n] = sx / (mnf[n] - 2);
          myv[n] = sy / (mnf[n] - 2);
        }
      }
    }

    /* depth (GetDepth, main/grid.c:429-535) */
    if (intdepth == 2) {
      char fv[1200];
      snprintf(fv, sizeof fv, "%s-voro", fdep);
      if (!(f = fopen(fv, "r"))) { gen_fail("cannot read", fv); break; }
      for (long n = 0; n < Nc; n++) { next_field(f); next_field(f); mdv[n] = next_field(f); }
      fclose(f);
    } else {
      if (intdepth == 1) {
        const long Nd = file_lines(fdep);
        if (Nd <= 0 || !(f = fopen(fdep, "r"))) { gen_fail("cannot read", fdep); break; }
        double *xd = (double *)xcalloc(Nd, sizeof(double)), *yd = (double *)xcalloc(Nd, sizeof(double));
        double *zd = (double *)xcalloc(Nd, sizeof(double));
        for (long n = 0; n < Nd; n++) { xd[n] = next_field(f); yd[n] = next_field(f); zd[n] = fabs(next_field(f)); }
        fclose(f);
        f = NULL;
        const int got = gen_interp_depth(xd, yd, zd, (int)Nd, mxv, myv, mdv, Nc, mf + 1);
        free(xd); free(yd); free(zd);
        if (!got) { gen_fail("fewer soundings than maxFaces + 1", fdep); break; }
        if ((int)scaledepth)
          for (long n = 0; n < Nc; n++) mdv[n] *= scalefactor;
      } else {
        for (long n = 0; n < Nc; n++) mdv[n] = depth(mxv[n], myv[n]);
      }
      char fv[1200];
      snprintf(fv, sizeof fv, "%s-voro", fdep);
      if (!gen_write(fv, &f)) break;
      for (long n = 0; n < Nc; n++) fprintf(f, "%f %f %f\n", mxv[n], myv[n], mdv[n]);
      fclose(f);
    }
    double maxd = 0;
    for (long n = 0; n < Nc; n++) if (mdv[n] > maxd) maxd = mdv[n];
    if (mindepth > 0) {
      for (long n = 0; n < Nc; n++) if (mdv[n] < mindepth) mdv[n] = mindepth;
    } else if (mindepth < 0) {
      double *dz0 = (double *)xcalloc(Nkmax, sizeof(double));
      if (!gen_dz(dz0, maxd, Nkmax, rstretch)) { free(dz0); gen_fail("rstretch out of range", NULL); break; }
      for (long n = 0; n < Nc; n++) if (mdv[n] < dz0[0]) mdv[n] = dz0[0];
      free(dz0);
    }

    /* the whole grid's faces and triangle corners (Connectivity,
     * main/grid.c:1026-1048) */
    gen_faces(mgrad, mgradf, mneigh, mface, mnf, mf, (int)Nc, (int)Ne);
    for (long n = 0; n < Nc; n++)
      for (int nf = 0; nf < mnf[n]; nf++) if (mface[n * mf + nf] < 0) bad = 1;
    if (bad) { gen_fail("a cell side has no edge (grad and neigh disagree)", dir); break; }
    if (mf == 3 && gen_wind_triangles(mface, medges, mcells, mnf, mf, (int)Nc))
      fprintf(stderr, "sun_grid_generate: a triangle's faces share no point\n");

    /* cells in the single process's order (TransferData,
     * main/grid.c:2670-2760): those without a type-3 face, then those with
     * one, each in index order */
    type3 = (int *)xcalloc(Nc, sizeof(int));
    for (long n = 0; n < Nc; n++)
      for (int nf = 0; nf < mnf[n]; nf++) type3[n] |= mmark[mface[n * mf + nf]] == 3;
    g = (gridT *)xcalloc(1, sizeof(gridT));
    g->Nc = (int)Nc; g->Ne = (int)Ne; g->Np = (int)Np; g->maxfaces = mf;
    planform_alloc(g);
    g->mnptr = (int *)xcalloc(Nc, sizeof(int));
    int k = 0;
    for (int pass = 0; pass < 2; pass++)
      for (long j = 0; j < Nc; j++)
        if (type3[j] == pass) {
          lcptr[j] = k;
          g->mnptr[k] = (int)j;
          g->xv[k] = mxv[j]; g->yv[k] = myv[j]; g->dv[k] = mdv[j]; g->nfaces[k] = mnf[j];
          for (int nf = 0; nf < mnf[j]; nf++) g->cells[k * mf + nf] = mcells[j * mf + nf];
          k++;
        }
    for (long i = 0; i < Nc; i++)
      for (int nf = 0; nf < g->nfaces[i]; nf++) {
        const int m = mneigh[g->mnptr[i] * mf + nf];
        g->neigh[i * mf + nf] = m >= 0 ? lcptr[m] : -1;
      }
    /* edges in order of first appearance around the cells (main/grid.c:2799-2831) */
    g->eptr = (int *)xcalloc(Ne, sizeof(int));
    g->edge_id = (int *)xcalloc(Ne, sizeof(int));
    for (long e = 0; e < Ne; e++) leptr[e] = -1;
    k = 0;
    for (long i = 0; i < Nc; i++)
      for (int nf = 0; nf < g->nfaces[i]; nf++) {
        const int ie = mface[g->mnptr[i] * mf + nf];
        if (leptr[ie] >= 0) continue;
        leptr[ie] = k;
        for (int j = 0; j < 3; j++) g->edges[3 * k + j] = medges[3 * ie + j];
        g->mark[k] = mmark[ie];
        g->eptr[k] = ie;
        g->edge_id[k] = medge_id[ie];
        k++;
      }
    if (k != Ne) { gen_fail("the cells do not reach every edge", dir); break; }
    for (long e = 0; e < Ne; e++)
      for (int j = 0; j < 2; j++) {
        const int c = mgrad[2 * g->eptr[e] + j];
        g->grad[2 * e + j] = c != -1 ? lcptr[c] : -1;
      }
    gen_faces(g->grad, g->gradf, g->neigh, g->face, g->nfaces, mf, (int)Nc, (int)Ne);
    if (mf == 3 && gen_wind_triangles(g->face, g->edges, g->cells, g->nfaces, mf, (int)Nc))
      fprintf(stderr, "sun_grid_generate: a triangle's faces share no point\n");
    for (long n = 0; n < Nc; n++)
      for (int nf = 0; nf < g->nfaces[n]; nf++)
        g->normal[n * mf + nf] = (n == g->grad[2 * g->face[mf * n + nf]]) ? -1 : 1;

    /* the other faces of each edge's cells (CreateMomentumCV,
     * main/grid.c:1071-1092) */
    for (long n = 0; n < Ne; n++) {
      int *en = g->eneigh + 2 * (mf - 1) * n, m = 0;
      for (int j = 0; j < 2 * (mf - 1); j++) en[j] = -1;
      for (int j = 0; j < 2; j++) {
        const int c = g->grad[2 * n + j];
        if (c == -1) continue;
        for (int nf = 0; nf < g->nfaces[c]; nf++)
          if (g->face[mf * c + nf] != n && m < 2 * (mf - 1)) en[m++] = g->face[mf * c + nf];
      }
    }

    /* each point's cells, edges and neighbour points (CreateNodeArray,
     * main/grid.c:834-990), as offsets into flat lists */
    pc_off = (int *)xcalloc(Np + 1, sizeof(int));
    for (long c = 0; c < Nc; c++)
      for (int nf = 0; nf < g->nfaces[c]; nf++) pc_off[g->cells[c * mf + nf] + 1]++;
    for (long p = 0; p < Np; p++) pc_off[p + 1] += pc_off[p];
    pc = (int *)xcalloc(pc_off[Np] + 1, sizeof(int));
    fill = (int *)xcalloc(Np, sizeof(int));
    for (long c = 0; c < Nc; c++)
      for (int nf = 0; nf < g->nfaces[c]; nf++) {
        const int p = g->cells[c * mf + nf];
        pc[pc_off[p] + fill[p]++] = (int)c;
      }
    pe_off = (int *)xcalloc(Np + 1, sizeof(int));
    for (long p = 0; p < Np; p++) {
      int m = 0;
      for (int i = pc_off[p]; i < pc_off[p + 1]; i++)
        for (int nf = 0; nf < g->nfaces[pc[i]]; nf++) {
          const int ie = g->face[mf * pc[i] + nf];
          m += (g->edges[3 * ie] == p) + (g->edges[3 * ie + 1] == p);
        }
      pe_off[p + 1] = pe_off[p] + m;
    }
    pe = (int *)xcalloc(pe_off[Np] + 1, sizeof(int));
    pp = (int *)xcalloc(pe_off[Np] + 1, sizeof(int));
    npp = (int *)xcalloc(Np, sizeof(int));
    for (long p = 0; p < Np; p++) {
      int m = pe_off[p];
      for (int i = pc_off[p]; i < pc_off[p + 1]; i++)
        for (int nf = 0; nf < g->nfaces[pc[i]]; nf++) {
          const int ie = g->face[mf * pc[i] + nf], a = g->edges[3 * ie], b = g->edges[3 * ie + 1];
          if (a == p) { pe[m] = ie; pp[m++] = b; }
          if (b == p) { pe[m] = ie; pp[m++] = a; }
        }
      /* neighbour points without repeats, first appearances kept */
      int u = 0;
      for (int i = pe_off[p]; i < pe_off[p + 1]; i++) {
        int seen = 0;
        for (int j = pe_off[p]; j < pe_off[p] + u; j++) seen |= pp[j] == pp[i];
        if (!seen) pp[pe_off[p] + u++] = pp[i];
      }
      npp[p] = u;
    }

    /* vertical grid (VertGrid, main/grid.c:1387-1585) */
    double dmax = mdv[0], dmin = mdv[0];
    for (long i = 0; i < Nc; i++) { if (mdv[i] > dmax) dmax = mdv[i]; if (mdv[i] < dmin) dmin = mdv[i]; }
    double dz0 = dmax / (double)Nkmax;
    if (dmin < dz0 && vgc) { dz0 = dmin; Nkmax = 1 + (int)(dmax / dz0); }
    g->Nkmax = Nkmax;
    g->dz = (REAL *)xcalloc(Nkmax, sizeof(REAL));
    if (!gen_dz(g->dz, dmax, Nkmax, rstretch)) { gen_fail("rstretch out of range", NULL); break; }
    for (long i = 0; i < Nc; i++) {
      if (Nkmax > 1 && mdv[i] != dmax) {
        double s = 0;
        for (int kk = 0; kk < Nkmax; kk++) {
          s += g->dz[kk];
          mNk[i] = kk + 1;
          if (s >= mdv[i]) break;
        }
      } else {
        mNk[i] = Nkmax;
      }
    }
    g->Nk = (int *)xcalloc(Nc, sizeof(int));
    for (long i = 0; i < Nc; i++) g->Nk[i] = mNk[g->mnptr[i]];
    if (stairstep)
      for (long i = 0; i < Nc; i++) {
        double s = 0;
        for (int kk = 0; kk < g->Nk[i]; kk++) s += g->dz[kk];
        g->dv[i] = s;
      }
    g->Nke = (int *)xcalloc(Ne, sizeof(int));
    g->Nkc = (int *)xcalloc(Ne, sizeof(int));
    for (long j = 0; j < Ne; j++) {
      const int c1 = mgrad[2 * g->eptr[j]], c2 = mgrad[2 * g->eptr[j] + 1];
      if (c1 == -1) { g->Nke[j] = g->Nkc[j] = mNk[c2]; }
      else if (c2 == -1) { g->Nke[j] = g->Nkc[j] = mNk[c1]; }
      else {
        g->Nke[j] = mNk[c1] < mNk[c2] ? mNk[c1] : mNk[c2];
        g->Nkc[j] = mNk[c1] < mNk[c2] ? mNk[c2] : mNk[c1];
      }
    }
    g->Nkp = (int *)xcalloc(Np, sizeof(int));
    for (long p = 0; p < Np; p++) {
      int m = 0;
      for (int i = pc_off[p]; i < pc_off[p + 1]; i++) if (g->Nk[pc[i]] > m) m = g->Nk[pc[i]];
      g->Nkp[p] = m;
    }

    /* metrics (Geometry, main/grid.c:3278-3563) */
    double xt[64], yt[64];
    for (long n = 0; n < Nc; n++) {
      for (int nf = 0; nf < g->nfaces[n]; nf++) { xt[nf] = xp[g->cells[n * mf + nf]]; yt[nf] = yp[g->cells[n * mf + nf]]; }
      g->Ac[n] = gen_area(xt, yt, g->nfaces[n]);
    }
    for (long n = 0; n < Ne; n++) {
      const int p1 = g->edges[3 * n], p2 = g->edges[3 * n + 1];
      g->df[n] = sqrt(pow(xp[p1] - xp[p2], 2) + pow(yp[p1] - yp[p2], 2));
      double tx = xp[p2] - xp[p1], ty = yp[p2] - yp[p1];
      const double tm = sqrt(tx * tx + ty * ty);
      tx = tx / tm;
      ty = ty / tm;
      const int c = g->grad[2 * n] >= 0 ? g->grad[2 * n] : g->grad[2 * n + 1];
      const double xd = (g->xv[c] - xp[p1]) * tx + (g->yv[c] - yp[p1]) * ty;
      g->xe[n] = xp[p1] + xd * tx;
      g->ye[n] = yp[p1] + xd * ty;
    }
    for (long n = 0; n < Ne; n++) {
      const int c1 = g->grad[2 * n], c2 = g->grad[2 * n + 1];
      if (c1 != -1 && c2 != -1) {
        g->n1[n] = g->xv[c1] - g->xv[c2];
        g->n2[n] = g->yv[c1] - g->yv[c2];
        g->dg[n] = sqrt(pow(g->n1[n], 2) + pow(g->n2[n], 2));
        g->n1[n] = g->n1[n] / g->dg[n];
        g->n2[n] = g->n2[n] / g->dg[n];
      } else {
        if (c1 == -1) { g->n1[n] = g->xe[n] - g->xv[c2]; g->n2[n] = g->ye[n] - g->yv[c2]; }
        else { g->n1[n] = g->xv[c1] - g->xe[n]; g->n2[n] = g->yv[c1] - g->ye[n]; }
        g->dg[n] = sqrt(pow(g->n1[n], 2) + pow(g->n2[n], 2));
        g->n1[n] = g->n1[n] / g->dg[n];
        g->n2[n] = g->n2[n] / g->dg[n];
        g->dg[n] = 2.0 * g->dg[n];
      }
    }
    for (long n = 0; n < Nc; n++)
      for (int nf = 0; nf < g->nfaces[n]; nf++) {
        const int ne = g->face[n * mf + nf], p = g->edges[3 * ne];
        g->def[n * mf + nf] =
            -((g->xv[n] - xp[p]) * g->n1[ne] + (g->yv[n] - yp[p]) * g->n2[ne]) * g->normal[n * mf + nf];
      }

    /* computational order (MakePointers, main/grid.c:1628-1679) */
    g->celldist = (int *)xcalloc(GRID_MAXBCTYPES - 1, sizeof(int));
    g->edgedist = (int *)xcalloc(GRID_MAXMARKS - 1, sizeof(int));
    g->cellp = (int *)xcalloc(Nc, sizeof(int));
    g->edgep = (int *)xcalloc(Ne, sizeof(int));
    g->myneighs = (int *)xcalloc(1, sizeof(int));
    computational_order(g);
    g->xp = xp; g->yp = yp; xp = yp = NULL;

    /* the run-phase files (OutputGridData, main/gridio.c:269-330) */
    char fc0[1200], fe0[1200];
    snprintf(fc0, sizeof fc0, "%s.0", fcel);
    snprintf(fe0, sizeof fe0, "%s.0", fedg);
    const struct gen_out o = {g, pc_off, pc, pe_off, pe, pp, npp};
    if (!gen_emit(fc0, Nc, line_cells0, &o) || !gen_emit(fe0, Ne, line_edges0, &o) ||
        !gen_emit(fcd, Nc, line_celldata, &o) || !gen_emit(fed, Ne, line_edgedata, &o) ||
        !gen_emit(fnod, Np, line_nodes, &o))
      break;
    if (!gen_write(ftop, &f)) break;
    fprintf(f, "%d %d\n", 1, 0);
    fprintf(f, "\n");
    for (int n = 0; n < GRID_MAXBCTYPES - 1; n++) fprintf(f, "%d ", g->celldist[n]);
    fprintf(f, "\n");
    for (int n = 0; n < GRID_MAXMARKS - 1; n++) fprintf(f, "%d ", g->edgedist[n]);
    fprintf(f, "\n");
    for (long n = 0; n < Nc; n++) fprintf(f, "%d ", g->cellp[n]);
    fprintf(f, "\n");
    for (long n = 0; n < Ne; n++) fprintf(f, "%d ", g->edgep[n]);
    fprintf(f, "\n");
    fclose(f);
    if (!gen_write(fvs, &f)) break;
    for (int n = 0; n < Nkmax; n++) fprintf(f, "%e\n", g->dz[n]);
    fclose(f);
    f = NULL;
    grid_vertical_rows(g, datafile);
    ok = 1;
  } while (0);
  free(xp); free(yp); free(medges); free(mmark); free(mgrad); free(mgradf); free(medge_id); free(mnf);
  free(mcells); free(mneigh); free(mface); free(mxv); free(myv); free(mdv); free(mNk); free(lcptr); free(leptr);
  free(pc_off); free(pc); free(fill); free(pe_off); free(pe); free(pp); free(npp); free(type3);
  if (f) fclose(f);
  if (!ok) { sun_grid_free(g); return NULL; }
  return g;
}

/* ------------------------------------------------------------------------ */
/* restart files                                                             */

/* The rows of one restart record in file order (main/physio.c:345-387),
 * with their lengths: 'e' Nke per edge, 'c' Nk per cell, 'w' Nk+1. */
struct restart_part { REAL **rows; char len; };
static int restart_parts(const physT *p, int turbmodel, struct restart_part *out) {
  struct restart_part a[] = {
    {p->Cn_U, 'e'}, {p->Cn_U2, 'e'}, {p->Cn_W, 'c'}, {p->Cn_W2, 'c'}, {p->Cn_R, 'c'}, {p->Cn_T, 'c'},
    {p->Cn_q, 'c'}, {p->Cn_l, 'c'}, {p->qT, 'c'}, {p->lT, 'c'},
    {p->nu_tv, 'c'}, {p->kappa_tv, 'c'}, {p->u, 'e'}, {p->w, 'w'}, {p->q, 'c'}, {p->qc, 'c'},
    {p->s, 'c'}, {p->T, 'c'}, {p->s0, 'c'}};
  int n = 0;
  for (int i = 0; i < (int)(sizeof a / sizeof a[0]); i++) {
    if (i >= 6 && i < 10 && turbmodel < 1) continue;  /* turbulence quantities only with a model */
    if (!a[i].rows) return -1;
    out[n++] = a[i];
  }
  return n;
}

static size_t row_len(const gridT *g, char len, int i) {
  return len == 'e' ? (size_t)g->Nke[i] : (size_t)g->Nk[i] + (len == 'w');
}

int sun_restart_write(const char *path, const gridT *g, const physT *p, int step, int turbmodel) {
  struct restart_part parts[19];
  const int np = restart_parts(p, turbmodel, parts);
  if (np < 0) return -1;
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  int ok = fwrite(&step, sizeof(int), 1, f) == 1 && fwrite(p->h, sizeof(REAL), g->Nc, f) == (size_t)g->Nc;
  for (int k = 0; ok && k < np; k++) {
    const int nr = parts[k].len == 'e' ? g->Ne : g->Nc;
    for (int i = 0; ok && i < nr; i++) {
      const size_t m = row_len(g, parts[k].len, i);
      ok = fwrite(parts[k].rows[i], sizeof(REAL), m, f) == m;
    }
  }
  return (fclose(f) == 0 && ok) ? 0 : -1;
}

int sun_restart_read(const char *path, const gridT *g, physT *p, int *step, int turbmodel) {
  struct restart_part parts[19];
  const int np = restart_parts(p, turbmodel, parts);
  if (np < 0) return -1;
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  int n = 0;
  int ok = fread(&n, sizeof(int), 1, f) == 1 && fread(p->h, sizeof(REAL), g->Nc, f) == (size_t)g->Nc;
  for (int k = 0; ok && k < np; k++) {
    const int nr = parts[k].len == 'e' ? g->Ne : g->Nc;
    for (int i = 0; ok && i < nr; i++) {
      const size_t m = row_len(g, parts[k].len, i);
      ok = fread(parts[k].rows[i], sizeof(REAL), m, f) == m;
    }
  }
  /* the record must end here: a longer file was written for another grid or model */
  if (ok && fgetc(f) != EOF) ok = 0;
  fclose(f);
  if (!ok) return -1;
  if (step) *step = n;
  return 0;
}

/*
 * sunstate.h -- host-side construction of SUNTANS grid/physics/property
 * state for the drop-in solvers: planform builders (a triangulated
 * rectangle, and any polygon list) with the reference's geometry
 * conventions, allocation of the jagged per-cell and per-edge rows in the
 * reference's shapes, the vertical-grid and flux-height updates the solvers
 * consume, and flat (column-compact) access to fields.
 *
 * These are set-up utilities that produce the solvers' inputs; they are not
 * part of the accelerated path.  Every function here is plain C.
 */
#ifndef SUNSTATE_H
#define SUNSTATE_H

#include "sun_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Rectangle [0,Lx]x[0,Ly] of nx*ny quads, each split into two right
 * triangles along its lower-left/upper-right diagonal.  Voronoi points are
 * centroids (with circumcentres the two triangles of a quad would share a
 * point and dg would vanish).  All boundary edges get mark 1 (closed).
 * Cell, edge and face conventions follow main/grid.c:654-720 (face/neigh),
 * :992-1007 (normal), :3493-3533 (n1, n2, dg), :1628-1685 (cellp/edgep,
 * celldist/edgedist). */
gridT *sun_grid_tri(int nx, int ny, double Lx, double Ly);

/* General unstructured planform: Np points, Nc polygons of nfaces[i] corners
 * (cells[i*maxfaces + v], either orientation, 3 <= nfaces[i] <= maxfaces)
 * with Voronoi points (xv, yv).  Edges are the polygon sides, numbered in
 * the order they first appear in the cell list; an edge's end points and
 * grad[2e] come from that first cell, grad[2e+1] is the cell across or -1
 * (boundary, mark 1).  Face nf of cell i joins corners nf and nf+1; unused
 * face slots hold -1.  Ac is the polygon area.  Metrics, normals and the
 * computational ordering follow the same conventions as sun_grid_tri.
 * Returns NULL for a malformed cell list (a corner out of range, a
 * degenerate side, or a side shared by more than two cells). */
gridT *sun_grid_polygons(int Np, const double *xp, const double *yp, int Nc, int maxfaces, const int *nfaces,
                         const int *cells, const double *xv, const double *yv);

/* Open boundaries: mark[Ne] gives every edge its mark; only boundary edges
 * (grad[2e+1] == -1) may change, to 1..4 (2 velocity, 3 specified-value
 * cells behind it, 4 no-slip).  Cells with a mark-3 face become type-3
 * cells; cellp/celldist and edgep/edgedist are rebuilt.  Call before
 * sun_phys_alloc (the open-boundary rows follow edgedist).  Returns -1 and
 * changes nothing for a mark the rule does not allow. */
int sun_grid_open(gridT *g, const int *mark);

/* A grid as the reference's run phase reads it for one process
 * (ReadGrid, main/gridio.c:54-187, then InitializeVerticalGrid,
 * main/phys.c:765-799): dir holds suntans.dat, whose celldata, edgedata,
 * nodes, topology and vertspace keys name the files GetGrid wrote (".0"
 * appended to the first four), maxFaces (default 3), Nkmax, stairstep,
 * fixdzz, dzsmall and smoothbot.  Numbers are parsed as the reference's
 * getfield parses them (main/fileio.c:137-174), so every array is the
 * reference's bit for bit.  The nodes' neighbour lists are skipped.
 * Returns NULL when a file is missing or written for more than one
 * process. */
gridT *sun_grid_load(const char *dir);

/* The reference's grid generation for one process (GetGrid with -g,
 * main/grid.c:110-212, TRIANGULATE off), in time linear in the grid: reads
 * the points, edges and cells files suntans.dat names in dir (main/gridio.c:
 * 515-603), corrects the Voronoi points (CorrectVoronoi > 0: toward the
 * centroids by VoronoiRatio, < 0: obtuse triangles to the centroid;
 * main/grid.c:3735-3869), takes the depth from the "<depth>-voro" file
 * for IntDepth 2, interpolates it from the depth file's soundings for
 * IntDepth 1 (inverse squared distance over the maxFaces + 1 nearest, as
 * InterpDepth, found through a bucket grid), or takes it from depth(x, y) -- the case's ReturnDepth
 * (main/initialization.c:78) -- otherwise (then writing "<depth>-voro" as
 * GetDepth does), applies minimum_depth, builds the faces, the single
 * process's cell and edge order, the point lists, the vertical grid
 * (rstretch as GetDZ, main/initialization.c:27-56, vertgridcorrect,
 * stairstep) and the metrics, and writes cells.dat.0, edges.dat.0 and the
 * run-phase files (celldata, edgedata, nodes, topology with ".0", and
 * vertspace) as OutputGridData does (main/gridio.c:269-330): the same bytes
 * as the reference's.  Returns the grid as GetGrid leaves it in memory
 * (full-precision metrics; sun_grid_load of the files gives the run phase's
 * rounded copy), or NULL with a message on stderr for a missing file or
 * key, or an inconsistent grid. */
gridT *sun_grid_generate(const char *dir, double (*depth)(double x, double y));

/* Restart records in the reference's binary format (OutputData's store
 * block and ReadPhysicalVariables, main/physio.c:337-394, 407-495): the step
 * number, h, then Cn_U, Cn_U2 (Nke per edge), Cn_W, Cn_W2, Cn_R, Cn_T (Nk
 * per cell), with turbmodel >= 1 Cn_q, Cn_l, qT, lT, then nu_tv, kappa_tv,
 * u (Nke), w (Nk + 1), q, qc, s, T, s0, in native byte order.  The reader
 * fills only these arrays (the reference's reader also runs UpdateDZ with
 * options -1 before and 0 after, and recomputes uc, vc and rho); it returns
 * -1 for a short or a longer file.  Both return -1 when a needed array is
 * not allocated. */
int sun_restart_write(const char *path, const gridT *g, const physT *p, int step, int turbmodel);
int sun_restart_read(const char *path, const gridT *g, physT *p, int *step, int turbmodel);

/* Vertical grid: dz[Nkmax] reference layer thicknesses and dv[Nc] depths.
 * Sets Nk (main/grid.c:1466-1498), Nke/Nkc (:1530-1563), ctop/etop = 0, and
 * allocates dzz/dzzold (Nk) and dzf (Nkc) rows as main/phys.c:765-799. */
int sun_grid_vertical(gridT *g, int Nkmax, const double *dz, const double *dv, int stairstep);

/* Every array the path and its callers read, in the reference's shapes
 * (main/phys.c:115-355): cell rows Nk long, w-type rows Nk+1, edge rows Nkc,
 * column scratch Nkmax+1.  All zero. */
physT *sun_phys_alloc(const gridT *g);
propT *sun_prop_default(void);

/* UpdateDZ (main/phys.c:816-968) and SetFluxHeight (main/phys.c:4338-4435)
 * semantics for the linear free surface state in phys->h and phys->u. */
void sun_update_dz(gridT *g, physT *p, const propT *prop, int option);
void sun_set_flux_height(gridT *g, physT *p, const propT *prop);

/* ComputeQSource (main/phys.c:2345-2395): RHS of the pressure Poisson
 * equation from phys->u/w/utmp2/wtmp2; also sets D[j] = df/dg. */
void sun_compute_qsource(REAL **src, gridT *g, physT *p, const propT *prop);

physT *sun_phys_clone(const gridT *g, const physT *p);
void sun_phys_copy(const gridT *g, const physT *from, physT *to);
void sun_grid_free(gridT *g);
void sun_phys_free(const gridT *g, physT *p);
void sun_prop_free(propT *prop);

/* Flat access.  Field names are the physT/gridT member names.  Cell fields
 * are column-compact (cell i occupies Nk[i] values, w-type fields Nk[i]+1),
 * edge fields occupy Nkc[j] values per edge, 2-D fields one value per cell
 * or edge.  Returns the number of values, or -1 for an unknown name. */
long sun_field_len(const gridT *g, const char *name);
long sun_field_get(const gridT *g, const physT *p, const char *name, double *out);
long sun_field_set(gridT *g, physT *p, const char *name, const double *in);
long sun_grid_int(const gridT *g, const char *name, int *out);
long sun_grid_set_int(gridT *g, const char *name, const int *in);
/* Scalars: Nc, Ne, Np, Nkmax, maxfaces, N (= sum Nk), Ek (edge-layers with two
 * wet neighbours). */
long sun_grid_scalar(const gridT *g, const char *name);
/* gridT's vertical-grid parameters smoothbot, dzsmall, fixdzz, stairstep. */
int sun_grid_set_param(gridT *g, const char *name, double value);
double sun_grid_param(const gridT *g, const char *name);
/* The member itself (a REAL** or REAL* row table) of physT / gridT by name,
 * for passing to functions with the reference's signatures. */
void *sun_phys_ptr(const physT *p, const char *name);
void *sun_grid_ptr(const gridT *g, const char *name);
size_t sun_prop_size(void);
void sun_prop_copy(const propT *from, propT *to);
/* propT members by name (doubles and ints). */
int sun_prop_set(propT *prop, const char *name, double value);
double sun_prop_get(const propT *prop, const char *name);

#ifdef __cplusplus
}
#endif
#endif

/*
 * sunhip.h -- the thin C ABI of the gfx950 kernels (libsunhip.so).
 *
 * Plain pointers and sizes only.  A grid context owns the device copy of one
 * SUNTANS grid (numprocs=1, the whole grid on one GPU) in a tiled,
 * layer-contiguous layout: 64 spatially ordered columns per tile (one
 * wavefront, one lane per column) and, inside a tile, level k of all 64
 * columns contiguous.  Edge rows are tiled the same way.
 *
 * Host arrays crossing this ABI are "flat": a per-cell field stores cell i's
 * column (Nk[i] values, Nk[i]+1 for w-type fields) contiguously, cells in
 * order; a per-edge field stores Nkc[j] values per edge; 2-D fields one value
 * per cell or edge.  Every function returns 0 on success and a negative code
 * on failure (sunhip_error() describes the last failure); none exits.
 *
 * The reference-signature drop-ins in suntans_hip.h are written on top of
 * this ABI.
 */
#ifndef SUNHIP_H
#define SUNHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sunhip_grid sunhip_grid;

/* Planform topology and metrics.  Arrays are the reference's gridT members
 * with the same meaning (main/grid.h:32-149). */
typedef struct {
  int Nc, Ne, maxfaces, Nkmax;
  const int *nfaces;   /* [Nc] */
  const int *neigh;    /* [Nc*maxfaces], -1 on the boundary */
  const int *face;     /* [Nc*maxfaces] */
  const int *normal;   /* [Nc*maxfaces], +-1 */
  const int *grad;     /* [2*Ne], -1 on the boundary side */
  const int *mark;     /* [Ne] */
  const int *Nk;       /* [Nc] */
  const int *Nkc;      /* [Ne] */
  const int *cellp;    /* [Nc] */
  const int *celldist; /* [3] */
  const int *edgep;    /* [Ne] */
  const int *edgedist; /* [6] */
  const double *Ac;    /* [Nc] */
  const double *df;    /* [Ne] */
  const double *dg;    /* [Ne] */
  const double *xv, *yv; /* [Nc], used only to order columns spatially */
} sunhip_topology;

/* Vertical extents that move with the free surface (UpdateDZ output). */
typedef struct {
  const int *ctop;     /* [Nc] */
  const int *ctopold;  /* [Nc] */
  const int *Nke;      /* [Ne] */
  const int *etop;     /* [Ne] */
} sunhip_extents;

/* Solve controls (propT members of the same names). */
typedef struct {
  int maxiters;        /* qmaxiters or maxiters */
  int precond;         /* qprecond (0 none, 2 vertical line) or hprecond (0/1) */
  int resnorm;
  double epsilon;      /* qepsilon or epsilon */
  int check_every;     /* iterations between host convergence polls (0: auto) */
  int qc_alias;        /* qprecond=1 only: 1 when the caller's c rows are its residual rows, as in
                          Solve (main/phys.c:1280 passes phys->stmp3, CGSolveQ's r), so the final
                          x /= c divides by the final residual (main/phys.c:2539-2546); 0 divides
                          by the conditioning c */
} sunhip_cg_params;

typedef struct {
  int niters;          /* iterations run (the reference's loop counter n) */
  int converged;
  double resid;        /* sqrt(eps/eps0) at exit */
  double eps0;
  double device_ms;    /* device time of the iteration loop (HIP events) */
  double setup_ms;     /* device time of coefficients + initial residual */
  int outer;           /* mixed forms: refinement steps (0 otherwise) */
  int inner;           /* mixed forms: FP32 PCG iterations (0 otherwise) */
} sunhip_cg_stats;

const char *sunhip_error(void);
int sunhip_device_count(void);

/* Checks a topology, and the extents X unless NULL, without touching a
 * device: every index the kernels follow in range (neigh and grad in
 * [-1,Nc), face in [0,Ne), cellp and edgep permutations, celldist and
 * edgedist nondecreasing splits, 1 <= Nk <= Nkmax, 0 <= Nkc <= Nkmax;
 * ctop and ctopold in [0,Nk], Nke and etop in [0,Nkc]).  0, or -1 with
 * sunhip_error() naming the first bad entry.  sunhip_grid_create and
 * sunhip_set_extents run the same checks first. */
int sunhip_validate(const sunhip_topology *topo, const sunhip_extents *ext);
int sunhip_grid_create(sunhip_grid **out, const sunhip_topology *topo, int device);
int sunhip_grid_destroy(sunhip_grid *g);
int sunhip_set_extents(sunhip_grid *g, const sunhip_extents *ext);
/* Device bytes held by the context. */
int64_t sunhip_grid_bytes(const sunhip_grid *g);
/* Sizes: N cell-layers (sum Nk), N flat w values, edge-layer values (sum Nkc),
 * padded cell-layer slots and tiles. */
int sunhip_grid_sizes(const sunhip_grid *g, int64_t *out6);

/* Named device fields; the name fixes the layout.  Cell fields (any other
 * name): "q","src","dzz","dzzold","s","kappa_tv","Cn","gradSx","gradSy",
 * "stmp"...; w-type fields take Nk+1 values per column (names starting
 * with 'w'); edge fields: "dzf","u","utmp2","u0","SfHp","SfHm","bscal" (the
 * open-boundary scalar, Nke values per type-2 edge, zero elsewhere); 2-D
 * cell fields: "h","htmp","htmp0","hcoef","active","Ftop","Fbot"...; 2-D edge fields:
 * "D","Dq","Dh".  Upload/download move between a flat host array and the
 * device layout. */
int sunhip_upload(sunhip_grid *g, const char *name, const double *flat);
int sunhip_download(sunhip_grid *g, const char *name, double *flat);
/* Pinned host buffer slot (0..SUNHIP_PIN_SLOTS-1) of the context, at least
 * n doubles, for transfers at the full host-link rate.  sunhip_upload from a
 * slot returns once the copy is enqueued (no wait); the next sunhip_pinned of
 * that slot waits until the device has read it, so a caller can fill one
 * slot while another is in flight.  Valid until the next call for the same
 * slot or sunhip_grid_destroy.  No reference counterpart. */
#define SUNHIP_PIN_SLOTS 4
int sunhip_pinned(sunhip_grid *g, int slot, int64_t n, double **out);
/* Device-side copy of one field into another of the same kind. */
int sunhip_copy_field(sunhip_grid *g, const char *dst, const char *src);
/* dst = a*src element-wise on the device (fields of the same kind);
 * enqueued, no host wait.  No reference counterpart: builds a time step's
 * forcing on the device from a resident base field. */
int sunhip_field_scale(sunhip_grid *g, const char *dst, const char *src, double a);

/* ---- nonhydrostatic pressure (main/phys.c:2411-2550) ------------------- */
/* Solve on the device-resident fields: "q" (warm start in, solution out),
 * "src" (RHS, clobbered as the reference clobbers it), "dzz", "dzf", "D".
 * Requires sunhip_set_extents.  theta/dt are not needed by this solve.
 * qprecond (prm->precond) 0, 1 or 2.  With 1 the conditioning c of
 * ConditionQ is formed in cell field "qcond" on the computational columns;
 * its other columns are read as the caller left them (the reference reads a
 * boundary neighbour's c as it finds it), and prm->qc_alias selects the
 * final rescaling. */
int sunhip_q_solve(sunhip_grid *g, const sunhip_cg_params *prm, sunhip_cg_stats *st);
/* The reference layer thicknesses grid->dz (Nkmax values, host memory).  No
 * reference counterpart: with them the pressure solve finds, per column and
 * face, the levels where QCoefficients' coef and the face heights dzf are
 * what dz gives (the z-level grid's full cells) and computes them there
 * instead of reading them; elsewhere it reads them.  Results are the same
 * bits either way (tuning "zreg" 0 switches it off).  sunhip_update_dz and
 * sunhip_set_flux_height set them too. */
int sunhip_set_dz(sunhip_grid *g, const double *dz);
/* y = OperatorQ(x) on named device fields (QCoefficients from "dzz"). */
int sunhip_q_operator(sunhip_grid *g, const char *x, const char *y);
/* xc = Preconditioner(x) (vertical-line Thomas solve per column). */
int sunhip_q_precondition(sunhip_grid *g, const char *x, const char *xc);
/* Run exactly n iterations regardless of the tolerance (parity at a fixed
 * iteration count). */
int sunhip_q_solve_fixed(sunhip_grid *g, const sunhip_cg_params *prm, int n, sunhip_cg_stats *st);

/* No reference counterpart: the mixed-precision form.  FP64 true residual
 * r = src - L q and its norm on the device; while sqrt(r.r/r0.r0) >=
 * prm->epsilon, solve L e = r with an FP32 vertical-line PCG (dots
 * accumulated in FP64) to inner_tol relative reduction and add e to q in FP64.
 * st->niters counts refinement steps; *inner_iters the FP32 iterations. */
int sunhip_q_solve_mixed(sunhip_grid *g, const sunhip_cg_params *prm, double inner_tol, int max_outer,
                         sunhip_cg_stats *st, int *inner_iters);
/* QCoefficients from the device field "dzz" into "coef". */
int sunhip_q_coefficients(sunhip_grid *g);
/* Kernel tuning: "unroll" (levels per load group of the operator kernel:
 * 1, 2, 4, 8), "prec_unroll" (the same for the preconditioner), "reduce"
 * (0: last block finishes the dot products, 1: a finisher kernel), "tiles"
 * (tiles of 64 columns per block of the iteration's operator kernel: 4, 8,
 * 16), "ntstore" (1: its results stored with the non-temporal hint), "rzid"
 * (1: the preconditioned residual dot r.M^-1 r taken as sum d'^2/b' from
 * the forward Thomas sweep, so the backward sweep does not re-read r;
 * M^-1 r itself is unchanged), "minwaves" (register budget of the
 * operator kernel: 3 or 4 waves per SIMD; 0 the compiler's choice),
 * "blocklds" (1: the operator's neighbour values staged in LDS per block of
 * four tiles walking the levels together), "band" (levels per workgroup of
 * the operator kernel, 8 or more; 0 whole columns), "prec_head" (top levels
 * of each column whose Thomas pivots and forward values the preconditioner
 * keeps in registers between its sweeps: 0, 8, 16, 24, 32; -1, the
 * default: 24 from 1024 tiles of 64 columns up in FP64, 16 from 2048 tiles
 * up in FP32, else 0), "zreg" (1, the default: the
 * z-level windows of sunhip_set_dz; 0 off), "order" (1, the default: blocks
 * of four tiles dispatched deepest first within each XCD's run; 0 in tile
 * order), "xchg" (the iteration's operator kernel exchanges neighbour values
 * inside each wave through LDS, gathering only cross-tile neighbours from
 * memory: 0 off, 1 on, 2 on with a register budget of 4 waves per SIMD, 3 on
 * with 2 levels per load group, 4-6 the pipelined form; -1, the default: 2
 * from 4096 tiles up, else 0), "tvd_group" (levels per load group of the
 * scalar step's column update: 1, 2 (the default) or 4), "prec_lds" (levels
 * below the preconditioner's register head of 16 or 24 that keep their
 * Thomas rows in LDS between the sweeps: 0 or 12, the default), "tvd_head" (Thomas
 * rows below the top row that the TVD column update keeps in registers
 * between its sweeps, with load groups of 2: 0, 8 (the default) or 16),
 * "tvd_sums_group" (the same as "tvd_group" for the TVD inflow sums: 1, 2 or
 * 4, the default), "h_blocks"
 * (most blocks of the free-surface iteration kernels, each adding all the
 * others' sums: 256, 512, 1024 (the default) or 2048).  Only "tiles",
 * "reduce", "rzid", "band", "order" and "h_blocks" change the rounding of a
 * dot product;
 * every other knob gives bit-identical results. */
int sunhip_set_tuning(sunhip_grid *g, const char *key, int value);
/* Per-kernel device time: sunhip_profile(g, 1) starts recording an event
 * pair around every solver kernel launch; sunhip_kernel_times fills ms[] and
 * n[] (SUNHIP_NPROF entries each) with the totals (ms) and launch counts per
 * kernel id and returns the number of ids: q apply, q precond, h apply,
 * h precond, mixed apply, mixed precond, mixed residual, other, TVD face
 * sums, TVD face values, scalar column update, Poisson source, Corrector,
 * Continuity, WPredictor cell source, WPredictor column (its face values
 * count as other). */
#define SUNHIP_NPROF 16
int sunhip_profile(sunhip_grid *g, int on);
int sunhip_kernel_times(sunhip_grid *g, double *ms, int *n);
/* A plain device copy of nbytes, reps times (HBM byte-counter calibration);
 * *avg_ms is the mean duration. */
int sunhip_copy_probe(sunhip_grid *g, int64_t nbytes, int reps, double *avg_ms);
/* No reference counterpart: transfers in parts, so that a caller packing or
 * unpacking host rows overlaps that work with the copies.
 *
 * sunhip_upload_part copies values [off, off + n) of the field's flat array
 * from flat (the whole array's base) into the device staging buffer; parts
 * come in order, each starting where the last ended, and the part with
 * last = 1 must end at the field's length: it then scatters the array into
 * the field.  flat must not change until the last part has been called and,
 * for a pinned slot (sunhip_pinned), until the slot is handed out again.
 *
 * sunhip_download_parts gathers the field and starts its copy into flat in
 * nparts (1..SUNHIP_MAXPARTS) equal parts without waiting;
 * sunhip_download_wait returns once values [0, upto) of the oldest download
 * in flight are on the host (upto at or beyond the length: all of them, and
 * the next one in flight becomes the oldest).  Two downloads in parts may be
 * in flight, into different host buffers: the second's gather and copies
 * then overlap the caller's work on the first.  Uploads may be issued while
 * downloads are in flight (they take other staging buffers and another copy
 * stream); whole downloads need none in flight.
 *
 * The copies run on streams of the grid's own, one per direction, through
 * device staging buffers taken in turn (two per direction), so one field's
 * copy overlaps the scatter or gather kernel of the field before it, and a
 * download overlaps the uploads issued after it. */
#define SUNHIP_MAXPARTS 16
int sunhip_upload_part(sunhip_grid *g, const char *name, const double *flat, int64_t off, int64_t n, int last);
int sunhip_download_parts(sunhip_grid *g, const char *name, double *flat, int nparts);
int sunhip_download_wait(sunhip_grid *g, int64_t upto);
/* No reference counterpart: host-to-device and device-to-host copy times (ms
 * per copy) of nbytes between pinned host memory and HBM, averaged over reps. */
int sunhip_pcie_probe(sunhip_grid *g, int64_t nbytes, int reps, double *h2d_ms, double *d2h_ms);

/* ---- free surface (main/phys.c:3086-3217) ------------------------------ */
/* Fields: "h" (in: boundary values; out: solution), "htmp" (RHS, clobbered),
 * "D".  grav, theta, dt enter through HCoefficients. */
int sunhip_h_solve(sunhip_grid *g, const sunhip_cg_params *prm, double grav, double theta, double dt,
                   sunhip_cg_stats *st);
/* No counterpart in the reference: sunhip_h_solve for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It allocates nothing: "h",
 * "htmp" and "D" must already be on the device, and the coefficients and CG
 * vectors live in a static device array of the library, 32 MiB per device,
 * which grids of up to 32 MiB / ((6 + maxfaces) x 8 B) cell slots fit
 * (about 466,000 triangles); larger grids get -1 and an error.  Same kernels
 * in the same order as sunhip_h_solve, so the same result.  Every call on a
 * device shares that device's array: calls that overlap on one device
 * (other threads, other grids' streams) race. */
int sunhip_h_solve_noalloc(sunhip_grid *g, const sunhip_cg_params *prm, double grav, double theta, double dt,
                           sunhip_cg_stats *st);
int sunhip_h_operator(sunhip_grid *g, double grav, double theta, double dt, const char *x, const char *y);
/* No reference counterpart: the mixed-precision free surface.  FP64 true
 * residual r = htmp - L h (h = 0 on computational cells first, which lifts
 * the boundary values like the reference), FP32 PCG on L e = r (dots in
 * FP64) to inner_tol relative reduction, h += e in FP64, until
 * sqrt(r.r/r0.r0) < prm->epsilon.  st->niters counts refinement steps. */
int sunhip_h_solve_mixed(sunhip_grid *g, const sunhip_cg_params *prm, double grav, double theta, double dt,
                         double inner_tol, int max_outer, sunhip_cg_stats *st, int *inner_iters);
/* OperatorH with caller-supplied coefficients: "hcoef" (2-D cell field) and
 * the face coefficients flat as [Nc*maxfaces] in the reference's indexing. */
int sunhip_h_operator_coef(sunhip_grid *g, const double *hfcoef_flat, const char *x, const char *y);

/* ---- TVD scalar transport (main/tvd.c:36-120, main/scalars.c:35-325) ---- */
typedef struct {
  double theta, dt, kappa;
  int TVD;             /* 0 none (upwind), 1..4 Psi limiter */
  int step;            /* prop->n */
  int alpha_top, alpha_bot;
} sunhip_scalar_params;
/* SfHp/SfHm from "s", "u", "utmp2", "dzf" (gradSx/gradSy receive sumQ/sumQC). */
int sunhip_face_scalars(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal);
/* One UpdateScalars step of field scal with "wnew","wtmp2","dzz","dzzold",
 * "kappa_tv","u","utmp2","dzf"; writes scal and "Cn". */
int sunhip_update_scalars(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal);
/* The same step with UpdateScalars' source terms (main/scalars.c:151-183,
 * 213-242): src1 and src2 name cell fields (the implicit and explicit
 * sources, main/phys.c:1207-1209 and 1241-1243 pass uold and wtmp), ftop and
 * fbot 2-D cell fields ("Ftop", "Fbot"); NULL leaves a term out.  Where the
 * surface rose (ctopold > ctop) the reference reads the scalar at a stale
 * level index past the row in the src1 term and writes src2's Cn below the
 * row; the device uses the top cell's scalar and leaves those writes out.
 * bcell (first-order upwind, TVD 0, with open boundaries) names a cell field
 * holding the rows the reference builds in stmp2 (main/scalars.c:84-96: the
 * boundary values in the rows of the cells next to edges of edgedist[2..5),
 * the last such edge's where a cell has several); faces of type-2 and type-3
 * edges read their outside value there (:260-263).  NULL: they read the
 * cell's own value, as the reference does without boundary_scal. */
int sunhip_update_scalars_src(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal, const char *src1,
                              const char *src2, const char *ftop, const char *fbot, const char *bcell);
/* No reference counterpart: sunhip_update_scalars (TVD >= 1, no source
 * terms) in two calls, so that a caller can move data between them.
 * sunhip_update_scalars_faces runs the faces pass (HorizontalFaceScalars:
 * "SfHp", "SfHm" and the column's flux terms) from scal, "u", "utmp2", "dzf"
 * and "bscal", enqueued without a wait; sunhip_update_scalars_column then
 * runs the column update, which reads "wnew", "wtmp2", "dzz", "dzzold",
 * "kappa_tv", "active" and the flux terms, also enqueued without a wait (a
 * download or sunhip_sync waits for it).  The pair gives
 * the result of sunhip_update_scalars bit for bit provided scal, "u",
 * "utmp2", "dzf" and "bscal" do not change between the two calls. */
int sunhip_update_scalars_faces(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal);
int sunhip_update_scalars_column(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal);

/* ---- the vertical grid that moves with the free surface -------------------
 * UpdateDZ (main/phys.c:816-968) and SetFluxHeight (:4338-4435) on the
 * device; the per-face level ranges are rebuilt from the new extents there,
 * so a resident simulation needs no sunhip_set_extents between steps.
 * Fields: cell "dzz", "dzzold"; cell 2-D "h", "dv", "dzbot"; edge "u",
 * "dzf"; edge 2-D "dzfB".  Bit-identical to the reference. */
typedef struct {
  const double *dz;        /* [Nkmax] reference layer thicknesses (grid->dz), host memory */
  double dzsmall, smoothbot;
  int stairstep, fixdzz, linearFS;
} sunhip_vgrid_params;
/* option 0 (each step), -1 (initialisation from h = 0) or 1 (after it),
 * as the reference's UpdateDZ; linearFS makes it a no-op. */
int sunhip_update_dz(sunhip_grid *g, const sunhip_vgrid_params *vp, int option);
int sunhip_set_flux_height(sunhip_grid *g, const sunhip_vgrid_params *vp);
/* The device's current extents in the reference's cell/edge order (any
 * pointer may be NULL). */
int sunhip_get_extents(sunhip_grid *g, int *ctop, int *ctopold, int *etop, int *etopold);

/* ---- the horizontal-momentum predictor (UPredictor, main/phys.c:2579-3018) --
 * In two parts either side of the reference's open-boundary hooks
 * (BoundaryVelocities, OpenBoundaryFluxes, user code of each case).  Fields:
 * edge "u", "utmp" (in: the horizontal source; out: U***), "utmp2", "ue"
 * (E = A^-1 1, the reference's phys->ut), "dzf"; edge 2-D "D", "tau_T",
 * "CdT", "CdB"; cell "dzz", "nu_tv", "nu_lax", "s"; w-type "w", "wtmp2";
 * cell 2-D "h", "htmp", "active", "hcorr", "dv" (depth); scratch "up_*".
 * Bit-identical to the reference.  On a one-layer edge with no-slip drag
 * (CdB or CdT = -1) the reference reads a coefficient left from the
 * previous edge; this build uses the edge's own, 0. */
typedef struct {
  double dt, theta, grav;
  double nu;             /* prop->nu */
  double lax_vertical;   /* prop->laxWendroff_Vertical */
  double thetaM;         /* implicitness of vertical momentum advection; < 0 none */
  int nonlinear;         /* prop->nonlinear (0: no momentum advection) */
} sunhip_upred_params;
/* main/phys.c:2596-2942: wtmp2 = w, utmp2 = u, D = 0; type-2 edges utmp = u;
 * computational edges: explicit free-surface gradient, wind stress, the
 * vertical system of viscosity, drag (implicit) and vertical advection
 * solved for U*** and E, D = dzf^T E.  Where the reference exits it returns
 * -5 (an edge level whose two cells have zero height) or -6 (U*** = nan). */
int sunhip_u_predict(sunhip_grid *g, const sunhip_upred_params *p);
/* main/phys.c:2951-3018: htmp = Ac h - dt sum of theta-weighted face fluxes,
 * the free-surface PCG (as sunhip_h_solve, from h = 0), cells drying below
 * 1e-10 above the bed clamped (h, active = 0, bottom "s" = 0), and
 * u = U*** - g theta dt E (h1 - h2)/dg on computational edges. */
int sunhip_u_solve(sunhip_grid *g, const sunhip_upred_params *p, const sunhip_cg_params *prm, sunhip_cg_stats *st);

/* ---- the vertical-momentum predictor (WPredictor, main/phys.c:2025-2290) ----
 * With the face values of the cell-centred w it advects (GetMomentumFaceValues
 * main/phys.c:4879-4992, UFaceFlux :4265-4289).  Fields: w-type "w" (in/out),
 * "wtmp" (out: the right-hand side as the reference's TriSolve leaves it),
 * "Cn_W", "Cn_W2" (in/out); cell "wc" (out), "stmp" (out: the cell source),
 * "qp" (the nonhydrostatic pressure phys->q, which sunhip_corrector
 * accumulates), "dzz", "nu_tv", "nu_lax"; edge "ue" (the reference's phys->ut: face
 * values out on the edges the reference writes), "utmp2", "bw" (type-2
 * boundary rows of w, Nke values per edge); for scheme 5 also edge "u",
 * "dzf", "SfHp", "SfHm" and cell "gradSx", "gradSy" (HorizontalFaceScalars
 * of wc); scratch "wp_b", "wp_c".  Bit-identical to the reference.
 * Refused (-1): scheme 4 (its vertical Lax-Wendroff term reads dzz one past
 * the column bottom), conserveMomentum (the top-cell fill reads scratch left
 * from the previous column), grids with type-4 edges (the type-4 term runs
 * to Nke of a cell index). */
typedef struct {
  double dt, theta;
  double nu, nu_H, CdW;  /* prop->nu, nu_H, CdW */
  double lax_vertical;   /* prop->laxWendroff_Vertical */
  int nonlinear;         /* 0 none, 1 upwind, 2 central, 5 TVD (TVDmomentum) */
  int step;              /* prop->n (1: Cn_W and Cn_W2 start from zero) */
  int AB;                /* prop->AB: 2 or 3 (Adams-Bashforth order from the third step) */
  int TVDmomentum;       /* limiter of scheme 5 */
  int conserveMomentum;  /* must be 0 */
} sunhip_wpred_params;
/* Distances from each edge's first cell centre to the edge midpoint, [Ne]
 * host (main/phys.c:4272-4273); central differencing (scheme 2) needs them. */
int sunhip_set_face_distances(sunhip_grid *g, const double *def1);
int sunhip_w_predict(sunhip_grid *g, const sunhip_wpred_params *p);

/* ---- the stencils either side of the pressure solve ------------------------
 * Solve's nonhydrostatic sequence is ComputeQSource, CGSolveQ, Corrector,
 * Continuity (main/phys.c:1270-1300).  Bit-identical to the reference. */
/* ComputeQSource (main/phys.c:2345-2395): the Poisson source into cell field
 * src from "u", "utmp2", "w", "wtmp2" and "dzf"; "D" <- df/dg on every edge. */
int sunhip_compute_qsource(sunhip_grid *g, double theta, double dt, const char *src);
/* Corrector (main/phys.c:2301-2332): with the pressure correction in cell
 * field qc, "u" and "w" are projected (reading "D" and "dzz") and cell field
 * q, the nonhydrostatic pressure, accumulates qc. */
int sunhip_corrector(sunhip_grid *g, double dt, const char *qc, const char *q);
/* Continuity (main/phys.c:3814-3844): w-type field w from "u", "utmp2",
 * "wtmp2" and "dzf". */
int sunhip_continuity(sunhip_grid *g, double theta, const char *w);

/* Synchronise the context's stream. */
int sunhip_sync(sunhip_grid *g);

/* ---- batched solves of many grids ----------------------------------------
 * No reference counterpart (the reference solves one grid per process).  A
 * batch owns one stream; grids attached to it enqueue all their work there.
 * The PCG iterations of every active solve of every attached grid run in the
 * same launches (a member table and a (member, tile) work list walked by
 * persistent kernels, replayed from a hipGraph), each member with its own
 * scalars and a fixed-order reduction over its own tiles, so a member's
 * result is the same bit for bit alone or in any mix.  Solves join and leave
 * between iterations (continuous batching).  Pressure solves use
 * qprecond=2. */
typedef struct sunhip_batch sunhip_batch;
typedef struct { double grav, theta, dt; } sunhip_h_phys;

int sunhip_batch_create(sunhip_batch **b, int device);
int sunhip_batch_destroy(sunhip_batch *b);
/* Route a grid's work to the batch's stream (implicit in the calls below);
 * detach restores the grid's own stream.  Destroying an attached grid
 * detaches it first. */
int sunhip_batch_attach(sunhip_batch *b, sunhip_grid *g);
int sunhip_batch_detach(sunhip_batch *b, sunhip_grid *g);
/* Start a FP64 pressure solve of g (fields as for sunhip_q_solve: "q" is the
 * warm start) or a free-surface solve (as sunhip_h_solve); tag is returned by
 * sunhip_batch_poll when it completes. */
int sunhip_batch_begin_q(sunhip_batch *b, sunhip_grid *g, const sunhip_cg_params *prm, int64_t tag);
int sunhip_batch_begin_h(sunhip_batch *b, sunhip_grid *g, const sunhip_cg_params *prm, const sunhip_h_phys *ph,
                         int64_t tag);
/* Start a mixed-precision pressure solve of g (as sunhip_q_solve_mixed): its
 * refinement steps advance with every poll, its FP64 residuals and FP32
 * inner iterations sharing launches with the other members'. */
int sunhip_batch_begin_q_mixed(sunhip_batch *b, sunhip_grid *g, const sunhip_cg_params *prm, double inner_tol,
                               int max_outer, int64_t tag);
/* Enqueue at least iters iterations of every active solve (whole graph
 * chunks); returns without waiting. */
int sunhip_batch_advance(sunhip_batch *b, int iters);
/* Wait for the enqueued work, retire up to max converged (or exhausted)
 * solves: finish their x, return their tags, kinds (0 pressure FP64,
 * 1 free surface, 2 pressure FP32 inner, 3 pressure mixed precision: niters
 * and inner = FP32 iterations, outer = refinement steps) and stats
 * (device_ms = time since begin).  Returns the number retired. */
int sunhip_batch_poll(sunhip_batch *b, int64_t *tags, int *kinds, sunhip_cg_stats *st, int max);
int sunhip_batch_active(const sunhip_batch *b);
int sunhip_batch_sync(sunhip_batch *b);
/* Blocking forms over n grids (the batch must be idle); st[i] per grid. */
int sunhip_batch_q_solve(sunhip_batch *b, sunhip_grid **g, int n, const sunhip_cg_params *prm, sunhip_cg_stats *st);
int sunhip_batch_h_solve(sunhip_batch *b, sunhip_grid **g, int n, const sunhip_cg_params *prm,
                         const sunhip_h_phys *ph, sunhip_cg_stats *st);
/* Mixed precision over n grids: one FP64 residual launch per refinement step
 * for all unconverged members, their FP32 inner solves batched, FP64
 * updates; per member the same stopping rule as sunhip_q_solve_mixed. */
int sunhip_batch_q_solve_mixed(sunhip_batch *b, sunhip_grid **g, int n, const sunhip_cg_params *prm,
                               double inner_tol, int max_outer, sunhip_cg_stats *st, int *inner_iters);
/* One scalar step (as sunhip_update_scalars) of each of n grids, sp[i]. */
int sunhip_batch_update_scalars(sunhip_batch *b, sunhip_grid **g, int n, const sunhip_scalar_params *sp,
                                const char *scal);
/* The stencils of each of n grids in one launch per stencil (theta[i],
 * dt[i]): the Poisson source (as sunhip_compute_qsource), or Corrector then
 * Continuity (as sunhip_corrector followed by sunhip_continuity). */
int sunhip_batch_compute_qsource(sunhip_batch *b, sunhip_grid **g, int n, const double *theta, const double *dt,
                                 const char *src);
int sunhip_batch_correct(sunhip_batch *b, sunhip_grid **g, int n, const double *theta, const double *dt,
                         const char *qc, const char *q, const char *w);
/* Launch configuration: unroll (1, 2, 4; 0 = autotune), blocks per CU of the
 * persistent launch (0 = one work item per wavefront), iterations per graph
 * chunk (0 = automatic).  sunhip_batch_tuning reports the configuration in
 * use for the current batch shape and returns the number of tuned shapes. */
int sunhip_batch_set_tuning(sunhip_batch *b, int unroll, int blocks_per_cu, int chunk);
int sunhip_batch_tuning(const sunhip_batch *b, int *unroll, int *blocks_per_cu, int *chunk);

#ifdef __cplusplus
}
#endif
#endif

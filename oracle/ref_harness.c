/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into
 * oracle/_ref/libsunref.so from the unmodified reference sources, to check
 * the oracle restatement against the real SUNTANS code.
 *
 * The elliptic solvers are file-static in main/phys.c (main/phys.c:39-73),
 * so this unit includes main/phys.c itself and exports thin wrappers with a
 * ref_ prefix.  The wrappers pass numprocs=1 and comm=0 (the serial MPI
 * stand-ins of main/no-mpi.c).  Every other symbol is hidden so the library
 * cannot interpose on the product's exported names.
 */
#include "phys.c"
#include "tvd.h"
#include "scalars.h"

#define EXPORT __attribute__((visibility("default")))

EXPORT void ref_set_verbose(int v) { VERBOSE = v; WARNING = 0; }

EXPORT void ref_CGSolveQ(REAL **q, REAL **src, REAL **c, gridT *g, physT *p, propT *pr) {
  CGSolveQ(q, src, c, g, p, pr, 0, 1, 0);
}
EXPORT void ref_CGSolve(gridT *g, physT *p, propT *pr) { CGSolve(g, p, pr, 0, 1, 0); }
EXPORT void ref_QCoefficients(REAL **coef, REAL **fcoef, REAL **c, gridT *g, physT *p, propT *pr) {
  QCoefficients(coef, fcoef, c, g, p, pr);
}
EXPORT void ref_OperatorQ(REAL **coef, REAL **x, REAL **y, REAL **c, gridT *g, physT *p, propT *pr) {
  OperatorQ(coef, x, y, c, g, p, pr);
}
EXPORT void ref_OperatorQC(REAL **coef, REAL **fcoef, REAL **x, REAL **y, REAL **c, gridT *g, physT *p,
                           propT *pr) {
  OperatorQC(coef, fcoef, x, y, c, g, p, pr);
}
EXPORT void ref_ConditionQ(REAL **x, gridT *g, physT *p, propT *pr) { ConditionQ(x, g, p, pr, 0, 0); }
EXPORT void ref_Preconditioner(REAL **x, REAL **xc, REAL **coef, gridT *g, physT *p, propT *pr) {
  Preconditioner(x, xc, coef, g, p, pr);
}
EXPORT double ref_InnerProduct3(REAL **x, REAL **y, gridT *g) { return InnerProduct3(x, y, g, 0, 1, 0); }
EXPORT void ref_HCoefficients(REAL *coef, REAL *fcoef, gridT *g, physT *p, propT *pr) {
  HCoefficients(coef, fcoef, g, p, pr);
}
EXPORT void ref_OperatorH(REAL *x, REAL *y, REAL *coef, REAL *fcoef, gridT *g, physT *p, propT *pr) {
  OperatorH(x, y, coef, fcoef, g, p, pr);
}
EXPORT void ref_ComputeQSource(REAL **src, gridT *g, physT *p, propT *pr) {
  ComputeQSource(src, g, p, pr, 0, 1);
}
EXPORT void ref_Corrector(REAL **qc, gridT *g, physT *p, propT *pr) { Corrector(qc, g, p, pr, 0, 1, 0); }
EXPORT void ref_Continuity(REAL **w, gridT *g, physT *p, propT *pr) { Continuity(w, g, p, pr); }
EXPORT void ref_UpdateDZ(gridT *g, physT *p, propT *pr, int option) { UpdateDZ(g, p, pr, option); }
EXPORT void ref_SetFluxHeight(gridT *g, physT *p, propT *pr) { SetFluxHeight(g, p, pr); }
EXPORT void ref_UPredictor(gridT *g, physT *p, propT *pr) { UPredictor(g, p, pr, 0, 1, 0); }
EXPORT void ref_WPredictor(gridT *g, physT *p, propT *pr) { WPredictor(g, p, pr, 0, 1, 0); }
EXPORT void ref_GetMomentumFaceValues(gridT *g, physT *p, propT *pr) {
  GetMomentumFaceValues(p->ut, p->wc, p->boundary_w, p->utmp2, g, p, pr, 0, 0, pr->nonlinear);
}
EXPORT void ref_HorizontalFaceScalars(gridT *g, physT *p, propT *pr, REAL **scal, REAL **bscal, int TVD) {
  HorizontalFaceScalars(g, p, pr, scal, bscal, TVD, 0, 0);
}
EXPORT void ref_GetApAm(REAL *ap, REAL *am, REAL *wp, REAL *wm, REAL *Cp, REAL *Cm, REAL *rp, REAL *rm,
                        REAL **w, REAL **dzz, REAL **scal, int i, int Nk, int ktop, REAL dt, int TVD) {
  GetApAm(ap, am, wp, wm, Cp, Cm, rp, rm, w, dzz, scal, i, Nk, ktop, dt, TVD);
}
EXPORT void ref_TriSolve(REAL *a, REAL *b, REAL *c, REAL *d, REAL *u, int N) { TriSolve(a, b, c, d, u, N); }
EXPORT void ref_update_salinity(gridT *g, physT *p, propT *pr) {
  UpdateScalars(g, p, pr, p->wnew, p->s, p->boundary_s, p->Cn_R, pr->kappa_s, pr->kappa_sH, p->kappa_tv,
                pr->theta, NULL, NULL, NULL, NULL, 0, 0, 0, 0, 1, pr->TVDsalt);
}
/* The salinity update with the source terms of the reference's own call
 * sites (main/phys.c:1207-1209, 1241-1243: src1 = uold, src2 = wtmp) and the
 * flux arguments: flags bit 1 src1, 2 src2, 4 Ftop = htmp, 8 Fbot = htmp2. */
EXPORT void ref_update_salinity_src(gridT *g, physT *p, propT *pr, int flags, int alpha_top, int alpha_bot) {
  UpdateScalars(g, p, pr, p->wnew, p->s, p->boundary_s, p->Cn_R, pr->kappa_s, pr->kappa_sH, p->kappa_tv,
                pr->theta, (flags & 1) ? p->uold : NULL, (flags & 2) ? p->wtmp : NULL, (flags & 4) ? p->htmp : NULL,
                (flags & 8) ? p->htmp2 : NULL, alpha_top, alpha_bot, 0, 0, 1, pr->TVDsalt);
}
EXPORT void ref_solve_q(gridT *g, physT *p, propT *pr) {
  CGSolveQ(p->qc, p->stmp, p->stmp3, g, p, pr, 0, 1, 0);
}

/* Grid generation and the run-phase reader, for the grid-file tests: the
 * data directory dir holds suntans.dat and the planform files it names;
 * GetGrid (main/grid.c:110) writes the run-phase files there for one
 * process, ReadGrid (main/gridio.c:54) reads them back. */
EXPORT void ref_grid_generate(const char *dir) {
  snprintf(DATADIR, sizeof DATADIR, "%s", dir);
  snprintf(DATAFILE, sizeof DATAFILE, "%s/suntans.dat", dir);
  TRIANGULATE = 0;
  gridT *grid = NULL;
  GetGrid(&grid, 0, 1, 0);
}
/* ReadPhysicalVariables (main/physio.c:407-495) from the restart file path. */
EXPORT int ref_read_restart(gridT *g, physT *p, propT *pr, const char *path) {
  pr->StartFID = fopen(path, "rb");
  if (!pr->StartFID) return -1;
  ReadPhysicalVariables(g, p, pr, 0, 0);
  return 0;
}
EXPORT gridT *ref_grid_read(const char *dir) {
  snprintf(DATADIR, sizeof DATADIR, "%s", dir);
  snprintf(DATAFILE, sizeof DATAFILE, "%s/suntans.dat", dir);
  gridT *grid = NULL;
  ReadGrid(&grid, 0, 1, 0);
  return grid;
}

/*
 * oracle.h -- CPU FP64 restatement of the SUNTANS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP path is
 * compared against; nothing in the product links it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function restates one reference function operation for operation
 * (same loop order, same floating-point expression order) so that, built
 * with gcc -O2 on x86-64 (no FMA contraction), it reproduces the reference
 * bit for bit.  tests/test_oracle_vs_ref.py checks that against the real
 * reference compiled from its sources (oracle/_ref); the golden fixtures in
 * tests/golden/ and the closed-form/brute-force checks in
 * tests/test_oracle_math.py pin it without the reference.
 */
#ifndef SUN_ORACLE_H
#define SUN_ORACLE_H
#include "sun_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Iteration count and final sqrt(eps/eps0) of the last CG solve. */
typedef struct { int niters; double resid; double eps0; } oracle_cg_stats;
extern oracle_cg_stats oracle_last_q, oracle_last_h;

void oracle_TriSolve(REAL *a, REAL *b, REAL *c, REAL *d, REAL *u, int N);
double oracle_Psi(double r, int TVD);
void oracle_QCoefficients(REAL **coef, REAL **fcoef, REAL **c, gridT *g, physT *p, propT *prop);
void oracle_OperatorQ(REAL **coef, REAL **x, REAL **y, REAL **c, gridT *g, physT *p, propT *prop);
void oracle_OperatorQC(REAL **coef, REAL **fcoef, REAL **x, REAL **y, REAL **c, gridT *g, physT *p,
                       propT *prop);
void oracle_ConditionQ(REAL **x, gridT *g, physT *p, propT *prop);
void oracle_Preconditioner(REAL **x, REAL **xc, REAL **coef, gridT *g, physT *p, propT *prop);
double oracle_InnerProduct3(REAL **x, REAL **y, gridT *g);
double oracle_InnerProduct(REAL *x, REAL *y, gridT *g);
void oracle_CGSolveQ(REAL **q, REAL **src, REAL **c, gridT *g, physT *p, propT *prop);
void oracle_HCoefficients(REAL *coef, REAL *fcoef, gridT *g, physT *p, propT *prop);
void oracle_OperatorH(REAL *x, REAL *y, REAL *coef, REAL *fcoef, gridT *g, physT *p, propT *prop);
void oracle_HPreconditioner(REAL *x, REAL *y, gridT *g, physT *p, propT *prop);
void oracle_CGSolve(gridT *g, physT *p, propT *prop);
void oracle_ComputeQSource(REAL **src, gridT *g, physT *p, propT *prop);
void oracle_Corrector(REAL **qc, gridT *g, physT *p, propT *prop);
void oracle_Continuity(REAL **w, gridT *g, physT *p, propT *prop);
int oracle_UPredictor_begin(gridT *g, physT *p, propT *prop);
void oracle_UpdateDZ(gridT *g, physT *p, propT *prop, int option);
void oracle_SetFluxHeight(gridT *g, physT *p, propT *prop);
int oracle_UPredictor_solve(gridT *g, physT *p, propT *prop);
void oracle_UPredictor_rhs(gridT *g, physT *p, propT *prop);
void oracle_UPredictor_finish(gridT *g, physT *p, propT *prop);
double oracle_face_def1(gridT *g, int j);
void oracle_MomentumFaceValues(REAL **uf, REAL **ui, REAL **bui, REAL **U, gridT *g, physT *p, propT *prop);
void oracle_WPredictor(gridT *g, physT *p, propT *prop);
void oracle_HorizontalFaceScalars(gridT *g, physT *p, propT *prop, REAL **scal, REAL **boundary_scal,
                                  int TVD);
void oracle_GetApAm(REAL *ap, REAL *am, REAL *wp, REAL *wm, REAL *Cp, REAL *Cm, REAL *rp, REAL *rm,
                    REAL **w, REAL **dzz, REAL **scal, int i, int Nk, int ktop, REAL dt, int TVD);
void oracle_UpdateScalars(gridT *g, physT *p, propT *prop, REAL **wnew, REAL **scal,
                          REAL **boundary_scal, REAL **Cn, REAL kappa, REAL kappaH, REAL **kappa_tv,
                          REAL theta, REAL **src1, REAL **src2, REAL *Ftop, REAL *Fbot, int alpha_top,
                          int alpha_bot, int checkflag, int TVDscheme);

/* Convenience wrappers with the argument pattern of the reference's call
 * sites (main/phys.c:1280, main/phys.c:2989, main/phys.c:1249-1251). */
void oracle_solve_q(gridT *g, physT *p, propT *prop);
void oracle_solve_h(gridT *g, physT *p, propT *prop);
void oracle_update_salinity(gridT *g, physT *p, propT *prop);
void oracle_update_salinity_src(gridT *g, physT *p, propT *prop, int flags, int alpha_top, int alpha_bot);

#ifdef __cplusplus
}
#endif
#endif

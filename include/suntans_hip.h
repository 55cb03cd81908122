/*
 * suntans_hip.h -- drop-in replacements for the SUNTANS elliptic solvers and
 * TVD scalar transport, with the reference's own C signatures, running on one
 * MI355X (numprocs=1, the whole grid on one GPU).  Link libsuntans.so in
 * place of the reference's definitions of these functions.
 *
 * Semantics follow the reference function each declaration replaces.  The
 * caller owns every array; results are written to the same rows the
 * reference writes (computational cells / edges).  ComputeQSource,
 * Corrector and Continuity, the stencils either side of CGSolveQ, run on the
 * device too.  Scratch rows the
 * reference clobbers as a side effect (src, stmp2, stmp3, uold, wtmp, qtmp,
 * a..d, hold, htmp*, hcoef, hfcoef, and CGSolveQ's c with qprecond=1) are
 * left untouched: the reference reads none of them afterwards.
 * Unsupported configurations (numprocs>1, first-order upwind scalars with
 * open boundaries, and the branches named at each function) print an error
 * and exit, as the reference does on fatal conditions; nonconvergence is not
 * an error (query it with sunhip_last_solve).
 */
#ifndef SUNTANS_HIP_H
#define SUNTANS_HIP_H

#include "sun_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* main/phys.c:2411 (static CGSolveQ; prototype main/phys.c:50-51).  qprecond
 * 0, 1 or 2.  With 1, when c's rows are phys->stmp3 (Solve's call,
 * main/phys.c:1280) they are also the residual rows, and the result is the
 * reference's: the scaled solution divided by the final residual
 * (main/phys.c:2539-2546). */
void CGSolveQ(REAL **q, REAL **src, REAL **c, gridT *grid, physT *phys, propT *prop, int myproc,
              int numprocs, MPI_Comm comm);
/* main/phys.c:3086 (static CGSolve; prototype main/phys.c:45-46) */
void CGSolve(gridT *grid, physT *phys, propT *prop, int myproc, int numprocs, MPI_Comm comm);
/* main/phys.c:3486 (static OperatorQ; prototype main/phys.c:71-72) */
void OperatorQ(REAL **coef, REAL **x, REAL **y, REAL **c, gridT *grid, physT *phys, propT *prop);
/* main/phys.c:3577 (static Preconditioner; prototype main/phys.c:55-56) */
void Preconditioner(REAL **x, REAL **xc, REAL **coef, gridT *grid, physT *phys, propT *prop);
/* main/phys.c:3335 (static OperatorH; prototype main/phys.c:65-66) */
void OperatorH(REAL *x, REAL *y, REAL *coef, REAL *fcoef, gridT *grid, physT *phys, propT *prop);
/* main/phys.c:2345 (static ComputeQSource; prototype main/phys.c:43-44) */
void ComputeQSource(REAL **src, gridT *grid, physT *phys, propT *prop, int myproc, int numprocs);
/* main/phys.c:2301 (static Corrector; prototype main/phys.c:41-42) */
void Corrector(REAL **qc, gridT *grid, physT *phys, propT *prop, int myproc, int numprocs, MPI_Comm comm);
/* main/phys.c:3814 (Continuity; prototypes main/phys.c:73-74) */
void Continuity(REAL **w, gridT *grid, physT *phys, propT *prop);
/* UPredictor (static, main/phys.c:2579-3074) in the two parts either side
 * of its open-boundary hooks; the caller runs BoundaryVelocities and
 * OpenBoundaryFluxes between them and UpdateDZ plus the boundary-flux tail
 * (main/phys.c:3023-3073) after, as the reference's UPredictor does.
 * _begin: main/phys.c:2596-2942 (U*** in utmp, E in ut, D; wtmp2, utmp2);
 * _solve: main/phys.c:2951-3018 (htmp, CGSolve, drying, u).  No counterpart
 * by these names: the split is where a device call cannot reach the hooks. */
void sunhip_UPredictor_begin(gridT *grid, physT *phys, propT *prop, int myproc, int numprocs, MPI_Comm comm);
void sunhip_UPredictor_solve(gridT *grid, physT *phys, propT *prop, int myproc, int numprocs, MPI_Comm comm);
/* WPredictor (static, main/phys.c:2025-2290; called at main/phys.c:1271)
 * under this name, same signature.  Exits through the drop-ins' error path
 * for nonlinear=4, conserveMomentum or type-4 edges (not built for the
 * device). */
void sunhip_WPredictor(gridT *grid, physT *phys, propT *prop, int myproc, int numprocs, MPI_Comm comm);
/* main/phys.c:816 (UpdateDZ; prototype main/phys.h:220) */
void UpdateDZ(gridT *grid, physT *phys, propT *prop, int option);
/* main/phys.c:4338 (SetFluxHeight; prototype main/phys.c:103) */
void SetFluxHeight(gridT *grid, physT *phys, propT *prop);
/* main/tvd.h:27-28 (definition main/tvd.c:36) */
void HorizontalFaceScalars(gridT *grid, physT *phys, propT *prop, REAL **scal, REAL **boundary_scal,
                           int TVD, MPI_Comm comm, int myproc);
/* main/scalars.h:19-22 (definition main/scalars.c:35).  The source terms
 * src1, src2, Ftop and Fbot are applied as the reference applies them; where
 * the surface rose (ctopold > ctop) and src1 or src2 is given, the reference
 * reads the scalar at a stale level index past the row and writes src2's Cn
 * below the row (main/scalars.c:153-154, 213-214): there the top cell's scalar
 * is used and those writes are left out. */
void UpdateScalars(gridT *grid, physT *phys, propT *prop, REAL **wnew, REAL **scal, REAL **boundary_scal,
                   REAL **Cn, REAL kappa, REAL kappaH, REAL **kappa_tv, REAL theta, REAL **src1,
                   REAL **src2, REAL *Ftop, REAL *Fbot, int alpha_top, int alpha_bot, MPI_Comm comm,
                   int myproc, int checkflag, int TVDscheme);

/* No counterpart in the reference: its solvers report convergence only on
 * stdout (main/phys.c:2527-2536, :3204-3213).  which = 0 for the last
 * CGSolveQ, 1 for the last CGSolve.  niters is the reference's loop counter
 * n at exit; resid is sqrt(eps/eps0).  Returns 0, or -1 before any solve. */
int sunhip_last_solve(int which, int *niters, double *resid, double *device_ms);
/* No counterpart in the reference: drop the cached device copies of all
 * grids (call after freeing or rebuilding a gridT). */
void sunhip_dropin_reset(void);
/* No counterpart in the reference: select the HIP device used by the
 * drop-ins (default 0). */
void sunhip_dropin_set_mode(int device);
/* No counterpart in the reference: mixed = 1 makes CGSolveQ and CGSolve run
 * the mixed-precision form (FP32 inner PCG inside FP64 iterative refinement,
 * FP64 true-residual test against the same qepsilon/epsilon); inner_tol is
 * the inner relative tolerance (<= 0 keeps the current one, default 1e-4).
 * sunhip_last_solve then reports the FP32 iterations.  mixed = 0 (default)
 * is the FP64 form whose iterates follow the reference's. */
void sunhip_dropin_set_precision(int mixed, double inner_tol);

#ifdef __cplusplus
}
#endif
#endif

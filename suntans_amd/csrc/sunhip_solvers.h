// sunhip_solvers.h -- host-side pieces of the single-grid solvers that the
// batched engine (sunhip_batch.hip) reuses per member: set-up of one PCG
// (coefficients, pivots, initial residual, first preconditioned residual) and
// the final x update.  All of them only enqueue work on g->stream.
#pragma once
#include "sunhip_qtile.h"
#include "sunhip_htile.h"
#include "sunhip_tvdtile.h"
#include "sunhip_steptile.h"

namespace sunhip {

// One pressure PCG instance: storage type, its scalars and its buffers.
template <typename T>
struct QSolver {
  CGScal *sc;
  const T *coef, *dzf, *D;
  T *x, *r, *z, *rtmp, *p0, *p1, *bq;
  const T *fcoef;  // qprecond=1: the scaled face coefficients
  ZReg Z;          // z-level windows of this solve (cwin null: none)
};

// One free-surface PCG instance.
template <typename T>
struct HSolver {
  CGScal *sc;
  const T *coef, *fcoef;
  T *x, *r, *z, *rt, *p0, *p1;
};

// The layer thicknesses dz[0..Nkmax) the z-level windows use (host array).
int set_zdz(sunhip_grid *g, const double *dz);
// The z-level windows of the pressure operator for the coefficients coef
// (FP64, from "dzz") and the face heights dzf now on the device; an empty
// ZReg when the layer thicknesses are unknown or the windows are switched
// off ("zreg" 0).
int q_windows(sunhip_grid *g, const double *coef, const double *dzf, ZReg *Z);

// Upload fresh CG scalars (tolerance, limits) to d from the pinned host slot
// stage (g->h_sc when null).  The slot must not be rewritten before the
// stream has executed the copy.
int upload_scal(sunhip_grid *g, CGScal *d, const sunhip_cg_params *prm, CGScal *stage = nullptr);

// Pressure solve, FP64, qprecond=2 or 0: everything up to the first iteration
// (main/phys.c:2441-2475) into the scalars at sc (g->d_qs when null).
int q_begin64(sunhip_grid *g, const sunhip_cg_params *prm, QSolver<double> *S, CGScal *stage, CGScal *sc = nullptr);
// x += nu*p of the last iteration.
int q_end64(sunhip_grid *g, const QSolver<double> &S);
int q_end32(sunhip_grid *g, const QSolver<float> &S);

// Mixed form: FP64 coefficients, FP32 copies of coef/dzf/D, FP32 pivots;
// S32 gets the FP32 PCG buffers (x = the correction e, scalars g->d_ms) and
// R the FP64 residual operands (out = the FP32 residual).
int q_mixed_prepare(sunhip_grid *g, const sunhip_cg_params *prm, QSolver<float> *S32, QVec<double, float> *R);
// Inner FP32 PCG from e = 0 on the residual already in S32.r.
int q_inner_begin(sunhip_grid *g, const sunhip_cg_params *inner, QSolver<float> &S32, CGScal *stage);
// x64 += e32.
int q_mix_update(sunhip_grid *g, double *x, const float *e);

// Free surface, FP64: coefficients, x = 0 on computational cells,
// r = b - L x, first preconditioned residual (main/phys.c:3108-3158).
int h_begin64(sunhip_grid *g, const sunhip_cg_params *prm, double grav, double theta, double dt,
              HSolver<double> *S, CGScal *stage);
int h_end64(sunhip_grid *g, const HSolver<double> &S);

// Arguments of one grid's scalar step (fields created on demand); no launch.
int tvd_args(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal, TvdArgs *T, ColArgs *C);

// Arguments of one grid's source / Corrector / Continuity stencils: qc, q and
// w name the correction, pressure and vertical velocity fields, src the
// source (fields created on demand); no launch.
int step_args(sunhip_grid *g, double theta, double dt, const char *qc, const char *q, const char *w,
              const char *src, StepArgs *out);

}  // namespace sunhip

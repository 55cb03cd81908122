// sunhip_htile.h -- the per-cell arithmetic of the free-surface solve.
//
// Reference: CGSolve main/phys.c:3086-3217, OperatorH :3335-3349,
// HPreconditioner :3227-3235.  Shared by the single-grid kernels
// (sunhip_h.hip) and the batched kernels over many grids (sunhip_batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "sunhip_common.h"

namespace sunhip {

enum { HM_PLAIN = 0, HM_RESID = 1, HM_FUSED = 2, HM_RESID_DOT = 3 };
enum { HOP_APPLY = 1, HOP_SETUP = 2, HOP_ITER = 3, HOP_STORE = 4 };

template <typename T, typename TO = T>
struct HVec {
  const T *coef, *fcoef;   // hcoef [NS], hfcoef [MAXF*NS]
  const T *vin, *src;
  const T *rtmp, *pold;
  T *pnew, *x;
  TO *out;
};

__device__ inline void h_op(CGScal *sc, int op, const double *tot) {
  const int pc = sc->precond;
  switch (op) {
    case HOP_APPLY:
      sc->pz = tot[0];
      sc->muinv = 1 / sc->alpha;
      sc->nu = sc->alpha / tot[0];
      break;
    case HOP_SETUP:  // main/phys.c:3137-3158
      sc->alpha = sc->alpha0 = pc == 1 ? tot[0] : tot[1];
      if (!sc->resnorm) sc->alpha0 = 1;
      sc->eps = sc->eps0 = pc == 1 ? tot[1] : sc->alpha0;
      sc->mu = 0; sc->nu = 0; sc->iter = 0; sc->nloop = 0; sc->converged = 0;
      sc->done = (sc->eps == 0 || sc->alpha == 0 || sc->maxit <= 0);
      sc->resid = sqrt(sc->eps / sc->eps0);
      break;
    case HOP_ITER: {  // main/phys.c:3175-3203
      const double alpha = pc == 1 ? tot[0] : tot[1];
      sc->mu = sc->muinv * alpha;
      sc->alpha = alpha;
      sc->eps = pc == 1 ? tot[1] : alpha;
      sc->iter += 1;
      sc->resid = sqrt(sc->eps / sc->eps0);
      if (sc->resid < sc->tol) { sc->done = 1; sc->converged = 1; sc->nloop = sc->iter - 1; }
      else if (sc->iter >= sc->maxit || sc->eps == 0 || sc->alpha == 0) { sc->done = 1; sc->nloop = sc->iter; }
      break;
    }
    case HOP_STORE:
      sc->pz = tot[0];
      break;
  }
}

// y = OperatorH(v) at slot s in one of the modes (see k_happly); returns the
// slot's share of the dot product.
template <int MODE, typename T, typename TO>
__device__ __forceinline__ double happly_cell(const Geo &G, const HVec<T, TO> &V, T mu, T nu, int s) {
  double dot = 0;
  if (s < G.NS && G.nk[s] > 0 && G.comp[s] == 1) {
    T po0 = 0, v0;
    if (MODE == HM_FUSED) { po0 = V.pold[s]; v0 = V.rtmp[s] + mu * po0; } else v0 = V.vin[s];
    T y = V.coef[s] * v0;
    for (int f = 0; f < G.nf[s]; f++) {
      const int nb = G.nbs[f * G.NS + s];
      if (nb < 0) continue;
      const T vn = MODE == HM_FUSED ? V.rtmp[nb] + mu * V.pold[nb] : V.vin[nb];
      y -= V.fcoef[f * G.NS + s] * vn;
    }
    if (MODE == HM_FUSED) {
      V.pnew[s] = v0;
      V.x[s] = V.x[s] + nu * po0;
      V.out[s] = y;
      dot = (double)v0 * (double)y;
    } else if (MODE == HM_RESID) {
      // p = p - z ; r = p (main/phys.c:3122-3128)
      V.out[s] = V.src[s] - y;
    } else if (MODE == HM_RESID_DOT) {
      const T rk = V.src[s] - y;
      V.out[s] = (TO)rk;
      dot = (double)rk * (double)rk;
    } else {
      V.out[s] = y;
    }
  }
  return dot;
}

// (update) r -= nu*z ; rt = r/hcoef (hprecond=1) or r ; rz, rr collect r.rt, r.r
template <typename T>
__device__ __forceinline__ void hprec_cell(const Geo &G, const T *__restrict__ coef, const T *__restrict__ z,
                                           T *__restrict__ r, T *__restrict__ rt, T nu, int pc, bool update, int s,
                                           double &rz, double &rr) {
  if (s < G.NS && G.nk[s] > 0 && G.comp[s] == 1) {
    T rk = r[s];
    if (update) { rk = rk - nu * z[s]; r[s] = rk; }
    const T u = pc == 1 ? rk / coef[s] : rk;
    rt[s] = u;
    rz += (double)rk * (double)u;
    rr += (double)rk * (double)rk;
  }
}

}  // namespace sunhip

/*
 * oracle.c -- CPU FP64 restatement of the SUNTANS elliptic solves and TVD
 * scalar transport.  TEST INFRASTRUCTURE ONLY (see oracle.h): this file is
 * the checker for the HIP path and is never linked into the product.
 *
 * Every expression keeps the reference's evaluation order, so that with
 * gcc -O2 on x86-64 the results are bit-identical to the reference.  Where
 * the reference reads outside a row (documented per function) the oracle
 * reads 0 instead; those cases are excluded from comparisons with the real
 * reference.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define TVD_EPS 1e-12 /* main/tvd.c:21 */

oracle_cg_stats oracle_last_q, oracle_last_h;

static double fmax_ref(double x1, double x2) { return x2 > x1 ? x2 : x1; } /* main/util.c:212 */
static double fmin_ref(double x, double y) { return x < y ? x : y; }       /* main/util.c:240 */
static double upwind_ref(double u, double a, double b) {                    /* main/util.c:219-232 */
  if (fabs(u) > 0) return u > 0 ? b : a;
  return fmax_ref(a, b);
}
static int kmin_face(const gridT *g, int i, int nc) {
  return g->ctop[nc] > g->ctop[i] ? g->ctop[nc] : g->ctop[i];
}

/* main/util.c:185-203: Thomas elimination without pivoting; b and d are
 * overwritten by the forward sweep. */
void oracle_TriSolve(REAL *a, REAL *b, REAL *c, REAL *d, REAL *u, int N) {
  for (int k = 1; k < N; k++) {
    b[k] -= a[k] * c[k - 1] / b[k - 1];
    d[k] -= a[k] * d[k - 1] / b[k - 1];
  }
  u[N - 1] = d[N - 1] / b[N - 1];
  for (int k = N - 2; k >= 0; k--) u[k] = d[k] / b[k] - c[k] * u[k + 1] / b[k];
}

/* main/tvd.c:130-151.  Case 4 is minmod (min(r,1) for r>0), whatever its
 * comment says. */
double oracle_Psi(double r, int TVD) {
  if (TVD == 2) return 1;
  if (TVD == 3) return fmax_ref(0, fmax_ref(fmin_ref(2 * (r), 1), fmin_ref(r, 2)));
  if (TVD == 4) return r > 0 ? fmin_ref(r, 1) : 0;
  return 0;
}

/* main/phys.c:3420-3471 */
void oracle_QCoefficients(REAL **coef, REAL **fcoef, REAL **c, gridT *g, physT *p, propT *prop) {
  const int mf = g->maxfaces;
  if (prop->qprecond == 1) {
    for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
      const int i = g->cellp[ip], kt = g->ctop[i];
      coef[i][kt] = g->Ac[i] / g->dzz[i][kt] / c[i][kt];
      for (int k = kt + 1; k < g->Nk[i]; k++)
        coef[i][k] = 2 * g->Ac[i] / (g->dzz[i][k] + g->dzz[i][k - 1]) / (c[i][k] * c[i][k - 1]);
      for (int nf = 0; nf < g->nfaces[i]; nf++) {
        const int nc = g->neigh[i * mf + nf];
        if (nc == -1) continue;
        const int ne = g->face[i * mf + nf];
        for (int k = kmin_face(g, i, nc); k < g->Nke[ne]; k++)
          fcoef[i * mf + nf][k] = g->dzz[i][k] * p->D[ne] / (c[i][k] * c[nc][k]);
      }
    }
    return;
  }
  /* unpreconditioned coefficients are built for every cell, not just cellp */
  for (int i = 0; i < g->Nc; i++) {
    const int kt = g->ctop[i];
    coef[i][kt] = g->Ac[i] / g->dzz[i][kt];
    for (int k = kt + 1; k < g->Nk[i]; k++) coef[i][k] = 2 * g->Ac[i] / (g->dzz[i][k] + g->dzz[i][k - 1]);
  }
}

/* main/phys.c:3486-3536: y = L x, horizontal faces first (face order), then
 * the vertical three-point part with q=0 above the top (ghost -x) and
 * dq/dz=0 below the bottom. */
void oracle_OperatorQ(REAL **coef, REAL **x, REAL **y, REAL **c, gridT *g, physT *p, propT *prop) {
  (void)c; (void)prop;
  const int mf = g->maxfaces;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip], kt = g->ctop[i], nk = g->Nk[i];
    REAL *yi = y[i], *xi = x[i], *ci = coef[i];
    for (int k = kt; k < nk; k++) yi[k] = 0;
    for (int nf = 0; nf < g->nfaces[i]; nf++) {
      const int nc = g->neigh[i * mf + nf];
      if (nc == -1) continue;
      const int ne = g->face[i * mf + nf];
      for (int k = kmin_face(g, i, nc); k < g->Nke[ne]; k++)
        yi[k] += (x[nc][k] - xi[k]) * g->dzf[ne][k] * p->D[ne];
    }
    for (int k = kt + 1; k < nk - 1; k++)
      yi[k] += ci[k] * xi[k - 1] - (ci[k] + ci[k + 1]) * xi[k] + ci[k + 1] * xi[k + 1];
    if (kt < nk - 1) {
      yi[kt] += (-2 * ci[kt] - ci[kt + 1]) * xi[kt] + ci[kt + 1] * xi[kt + 1];
      yi[nk - 1] += ci[nk - 1] * xi[nk - 2] - ci[nk - 1] * xi[nk - 1];
    } else {
      yi[kt] -= 2.0 * ci[kt] * xi[kt];
    }
  }
}

/* main/phys.c:3364-3409 (qprecond=1: symmetrically scaled operator with unit
 * diagonal) */
void oracle_OperatorQC(REAL **coef, REAL **fcoef, REAL **x, REAL **y, REAL **c, gridT *g, physT *p,
                       propT *prop) {
  (void)c; (void)p; (void)prop;
  const int mf = g->maxfaces;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip], kt = g->ctop[i], nk = g->Nk[i];
    for (int k = kt; k < nk; k++) y[i][k] = -x[i][k];
    for (int nf = 0; nf < g->nfaces[i]; nf++) {
      const int nc = g->neigh[i * mf + nf];
      if (nc == -1) continue;
      const int ne = g->face[i * mf + nf];
      for (int k = kmin_face(g, i, nc); k < g->Nke[ne]; k++) y[i][k] += x[nc][k] * fcoef[i * mf + nf][k];
    }
    for (int k = kt + 1; k < nk - 1; k++) y[i][k] += coef[i][k] * x[i][k - 1] + coef[i][k + 1] * x[i][k + 1];
    if (kt < nk - 1) {
      y[i][kt] += coef[i][kt + 1] * x[i][kt + 1];
      y[i][nk - 1] += coef[i][nk - 1] * x[i][nk - 2];
    }
  }
}

/* main/phys.c:3629-3687: c = sqrt(|diag L|), 1 where the diagonal is not
 * positive. */
void oracle_ConditionQ(REAL **x, gridT *g, physT *p, propT *prop) {
  (void)prop;
  const int mf = g->maxfaces;
  REAL *a = p->a;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip], kt = g->ctop[i], nk = g->Nk[i];
    for (int k = kt; k < nk; k++) x[i][k] = 0;
    for (int nf = 0; nf < g->nfaces[i]; nf++) {
      const int nc = g->neigh[i * mf + nf];
      if (nc == -1) continue;
      const int ne = g->face[i * mf + nf];
      for (int k = kmin_face(g, i, nc); k < g->Nke[ne]; k++) x[i][k] += g->dzz[i][k] * p->D[ne];
    }
    a[kt] = g->Ac[i] / g->dzz[i][kt];
    for (int k = kt + 1; k < nk; k++) a[k] = 2 * g->Ac[i] / (g->dzz[i][k] + g->dzz[i][k - 1]);
    for (int k = kt + 1; k < nk - 1; k++) x[i][k] += (a[k] + a[k + 1]);
    if (kt < nk - 1) {
      x[i][kt] += 2 * a[kt] + a[kt + 1];
      x[i][nk - 1] += a[nk - 1];
    }
  }
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip];
    for (int k = g->ctop[i]; k < g->Nk[i]; k++) {
      if (x[i][k] <= 0) x[i][k] = 1;
      x[i][k] = sqrt(x[i][k]);
    }
  }
}

/* main/phys.c:3577-3619: per column, invert the vertical part of L with the
 * Thomas algorithm (scratch rows phys->a..d). */
void oracle_Preconditioner(REAL **x, REAL **xc, REAL **coef, gridT *g, physT *p, propT *prop) {
  (void)prop;
  REAL *a = p->a, *b = p->b, *c = p->c, *d = p->d;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip], kt = g->ctop[i], nk = g->Nk[i];
    const REAL *ci = coef[i];
    if (kt >= nk - 1) {
      xc[i][kt] = -0.5 * x[i][kt] / ci[kt];
      continue;
    }
    for (int k = kt + 1; k < nk - 1; k++) {
      a[k] = ci[k];
      b[k] = -ci[k] - ci[k + 1];
      c[k] = ci[k + 1];
      d[k] = x[i][k];
    }
    b[kt] = -2 * ci[kt] - ci[kt + 1];
    c[kt] = ci[kt + 1];
    d[kt] = x[i][kt];
    a[nk - 1] = ci[nk - 1];
    b[nk - 1] = -ci[nk - 1];
    d[nk - 1] = x[i][nk - 1];
    oracle_TriSolve(a + kt, b + kt, c + kt, d + kt, xc[i] + kt, nk - kt);
  }
}

/* main/phys.c:3304-3319 (numprocs=1: the reduce is a copy) */
double oracle_InnerProduct3(REAL **x, REAL **y, gridT *g) {
  double s = 0;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip];
    for (int k = g->ctop[i]; k < g->Nk[i]; k++) s += x[i][k] * y[i][k];
  }
  return s;
}

/* main/phys.c:3280-3294 */
double oracle_InnerProduct(REAL *x, REAL *y, gridT *g) {
  double s = 0;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip];
    s += x[i] * y[i];
  }
  return s;
}

#define FOR_CELLS_K(g, i, k)                                        \
  for (int ip_ = (g)->celldist[0]; ip_ < (g)->celldist[1]; ip_++)   \
    for (int i = (g)->cellp[ip_], k = (g)->ctop[i]; k < (g)->Nk[i]; k++)

/* main/phys.c:2411-2550.  x aliases q (warm start), p aliases src, r is
 * phys->stmp3, z phys->stmp2, rtmp phys->uold.  With qprecond=1 the caller
 * in Solve passes c=phys->stmp3, the same rows as r, so the closing x/=c
 * divides by the final residual; that aliasing is reproduced because the
 * same rows are used. */
void oracle_CGSolveQ(REAL **q, REAL **src, REAL **c, gridT *g, physT *p, propT *prop) {
  REAL **z = p->stmp2, **x = q, **r = p->stmp3, **rtmp = p->uold, **pp = src;
  double mu, nu, alpha, alpha0, eps, eps0;
  const int pc = prop->qprecond;
  if (pc == 1) {
    oracle_ConditionQ(c, g, p, prop);
    FOR_CELLS_K(g, i, k) { pp[i][k] /= c[i][k]; x[i][k] *= c[i][k]; }
  }
  const int niters = prop->qmaxiters;
  oracle_QCoefficients(p->wtmp, p->qtmp, c, g, p, prop);
  if (pc == 1) oracle_OperatorQC(p->wtmp, p->qtmp, x, z, c, g, p, prop);
  else oracle_OperatorQ(p->wtmp, x, z, c, g, p, prop);
  FOR_CELLS_K(g, i, k) r[i][k] = pp[i][k] - z[i][k];
  if (pc == 2) {
    oracle_Preconditioner(r, rtmp, p->wtmp, g, p, prop);
    FOR_CELLS_K(g, i, k) pp[i][k] = rtmp[i][k];
    alpha = alpha0 = oracle_InnerProduct3(r, rtmp, g);
  } else {
    FOR_CELLS_K(g, i, k) pp[i][k] = r[i][k];
    alpha = alpha0 = oracle_InnerProduct3(r, r, g);
  }
  if (!prop->resnorm) alpha0 = 1;
  if (pc == 2) eps = eps0 = oracle_InnerProduct3(r, r, g);
  else eps = eps0 = alpha0;

  int n;
  for (n = 0; n < niters && eps != 0; n++) {
    if (pc == 1) oracle_OperatorQC(p->wtmp, p->qtmp, pp, z, c, g, p, prop);
    else oracle_OperatorQ(p->wtmp, pp, z, c, g, p, prop);
    mu = 1 / alpha;
    nu = alpha / oracle_InnerProduct3(pp, z, g);
    FOR_CELLS_K(g, i, k) { x[i][k] += nu * pp[i][k]; r[i][k] -= nu * z[i][k]; }
    if (pc == 2) {
      oracle_Preconditioner(r, rtmp, p->wtmp, g, p, prop);
      alpha = oracle_InnerProduct3(r, rtmp, g);
      mu *= alpha;
      FOR_CELLS_K(g, i, k) pp[i][k] = rtmp[i][k] + mu * pp[i][k];
    } else {
      alpha = oracle_InnerProduct3(r, r, g);
      mu *= alpha;
      FOR_CELLS_K(g, i, k) pp[i][k] = r[i][k] + mu * pp[i][k];
    }
    eps = (pc == 2) ? oracle_InnerProduct3(r, r, g) : alpha;
    if (sqrt(eps / eps0) < prop->qepsilon) break;
  }
  oracle_last_q.niters = n;
  oracle_last_q.resid = sqrt(eps / eps0);
  oracle_last_q.eps0 = eps0;
  if (pc == 1) FOR_CELLS_K(g, i, k) x[i][k] /= c[i][k];
}

/* main/phys.c:3252-3269 */
void oracle_HCoefficients(REAL *coef, REAL *fcoef, gridT *g, physT *p, propT *prop) {
  const int mf = g->maxfaces;
  const double tmp = prop->grav * pow(prop->theta * prop->dt, 2);
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip];
    coef[i] = g->Ac[i];
    for (int nf = 0; nf < g->nfaces[i]; nf++) {
      if (g->neigh[i * mf + nf] == -1) continue;
      const int ne = g->face[i * mf + nf];
      fcoef[i * mf + nf] = tmp * p->D[ne] * g->df[ne] / g->dg[ne];
      coef[i] += fcoef[i * mf + nf];
    }
  }
}

/* main/phys.c:3335-3349 */
void oracle_OperatorH(REAL *x, REAL *y, REAL *coef, REAL *fcoef, gridT *g, physT *p, propT *prop) {
  (void)p; (void)prop;
  const int mf = g->maxfaces;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip];
    y[i] = coef[i] * x[i];
    for (int nf = 0; nf < g->nfaces[i]; nf++) {
      const int nb = g->neigh[i * mf + nf];
      if (nb != -1) y[i] -= fcoef[i * mf + nf] * x[nb];
    }
  }
}

/* main/phys.c:3227-3235 */
void oracle_HPreconditioner(REAL *x, REAL *y, gridT *g, physT *p, propT *prop) {
  (void)prop;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip];
    y[i] = x[i] / p->hcoef[i];
  }
}

#define FOR_CELLS(g, i, lo, hi) \
  for (int ip_ = (g)->celldist[lo], i; ip_ < (g)->celldist[hi] && ((i = (g)->cellp[ip_]), 1); ip_++)

/* main/phys.c:3086-3217.  x=h (zeroed on computational cells, boundary
 * values lifted into the RHS), r=hold, rtmp=htmp2, z=htmp3, p=htmp. */
void oracle_CGSolve(gridT *g, physT *p, propT *prop) {
  REAL *x = p->h, *r = p->hold, *rtmp = p->htmp2, *z = p->htmp3, *pp = p->htmp;
  double mu, nu, eps, eps0, alpha, alpha0;
  const int niters = prop->maxiters;
  oracle_HCoefficients(p->hcoef, p->hfcoef, g, p, prop);
  FOR_CELLS(g, i, 0, 1) x[i] = 0;
  oracle_OperatorH(x, z, p->hcoef, p->hfcoef, g, p, prop);
  FOR_CELLS(g, i, 0, 1) { pp[i] = pp[i] - z[i]; r[i] = pp[i]; x[i] = 0; }
  FOR_CELLS(g, i, 1, 2) pp[i] = 0;
  if (prop->hprecond == 1) {
    oracle_HPreconditioner(r, rtmp, g, p, prop);
    FOR_CELLS(g, i, 0, 1) pp[i] = rtmp[i];
    alpha = alpha0 = oracle_InnerProduct(r, rtmp, g);
  } else {
    FOR_CELLS(g, i, 0, 1) pp[i] = r[i];
    alpha = alpha0 = oracle_InnerProduct(r, r, g);
  }
  if (!prop->resnorm) alpha0 = 1;
  if (prop->hprecond == 1) eps = eps0 = oracle_InnerProduct(r, r, g);
  else eps = eps0 = alpha0;
  int n;
  for (n = 0; n < niters && eps != 0 && alpha != 0; n++) {
    oracle_OperatorH(pp, z, p->hcoef, p->hfcoef, g, p, prop);
    mu = 1 / alpha;
    nu = alpha / oracle_InnerProduct(pp, z, g);
    FOR_CELLS(g, i, 0, 1) { x[i] += nu * pp[i]; r[i] -= nu * z[i]; }
    if (prop->hprecond == 1) {
      oracle_HPreconditioner(r, rtmp, g, p, prop);
      alpha = oracle_InnerProduct(r, rtmp, g);
      mu *= alpha;
      FOR_CELLS(g, i, 0, 1) pp[i] = rtmp[i] + mu * pp[i];
    } else {
      alpha = oracle_InnerProduct(r, r, g);
      mu *= alpha;
      FOR_CELLS(g, i, 0, 1) pp[i] = r[i] + mu * pp[i];
    }
    eps = (prop->hprecond == 1) ? oracle_InnerProduct(r, r, g) : alpha;
    if (sqrt(eps / eps0) < prop->epsilon) break;
  }
  oracle_last_h.niters = n;
  oracle_last_h.resid = sqrt(eps / eps0);
  oracle_last_h.eps0 = eps0;
}

/* main/phys.c:2345-2395 */
void oracle_ComputeQSource(REAL **src, gridT *g, physT *p, propT *prop) {
  const double tf = (1 - prop->theta) / prop->theta;
  const int mf = g->maxfaces;
  for (int i = 0; i < g->Nc; i++)
    for (int k = g->ctop[i]; k < g->Nk[i]; k++) src[i][k] = 0;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip];
    for (int k = g->ctop[i]; k < g->Nk[i]; k++)
      src[i][k] = g->Ac[i] * (p->w[i][k] - p->w[i][k + 1]) + tf * g->Ac[i] * (p->wtmp2[i][k] - p->wtmp2[i][k + 1]);
    for (int nf = 0; nf < g->nfaces[i]; nf++) {
      const int ne = g->face[i * mf + nf];
      for (int k = g->ctop[i]; k < g->Nke[ne]; k++)
        src[i][k] += (p->u[ne][k] + tf * p->utmp2[ne][k]) * g->dzf[ne][k] * g->normal[i * mf + nf] * g->df[ne];
    }
    for (int k = g->ctop[i]; k < g->Nk[i]; k++) src[i][k] /= prop->dt;
  }
  for (int j = 0; j < g->Ne; j++) p->D[j] = g->df[j] / g->dg[j];
}

/* main/phys.c:2301-2332.  The horizontal velocity of computational interior
 * edges with D != 0 and more than one layer takes the gradient of the
 * pressure correction; the vertical velocity its vertical gradient (with the
 * surface ghost value -qc[ctop]); the pressure of multi-layer columns
 * accumulates the correction. */
void oracle_Corrector(REAL **qc, gridT *g, physT *p, propT *prop) {
  const double dt = prop->dt;
  for (int jp = g->edgedist[0]; jp < g->edgedist[1]; jp++) {
    const int j = g->edgep[jp];
    if (p->D[j] == 0 || g->etop[j] >= g->Nke[j] - 1) continue;
    const int a = g->grad[2 * j], b = g->grad[2 * j + 1];
    for (int k = g->etop[j]; k < g->Nke[j]; k++) p->u[j][k] -= dt / g->dg[j] * (qc[a][k] - qc[b][k]);
  }
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip], kt = g->ctop[i];
    for (int k = kt + 1; k < g->Nk[i]; k++)
      p->w[i][k] -= 2.0 * dt / (g->dzz[i][k - 1] + g->dzz[i][k]) * (qc[i][k - 1] - qc[i][k]);
    p->w[i][kt] += 2.0 * dt / g->dzz[i][kt] * qc[i][kt];
  }
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip];
    if (g->ctop[i] >= g->Nk[i] - 1) continue;
    for (int k = g->ctop[i]; k < g->Nk[i]; k++) p->q[i][k] += qc[i][k];
  }
}

/* main/phys.c:3814-3844.  Per computational column, from the bottom (w = 0
 * at Nk) up to ctop: the theta-weighted discrete continuity equation, with
 * the old vertical velocity wtmp2 and the face fluxes of u (new) and utmp2
 * (old); every other level of the column is zero. */
void oracle_Continuity(REAL **w, gridT *g, physT *p, propT *prop) {
  const double th = prop->theta;
  const int mf = g->maxfaces;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip];
    for (int k = 0; k <= g->Nk[i]; k++) w[i][k] = 0;
    for (int k = g->Nk[i] - 1; k >= g->ctop[i]; k--) {
      w[i][k] = w[i][k + 1] - (1 - th) / th * (p->wtmp2[i][k] - p->wtmp2[i][k + 1]);
      for (int f = 0; f < g->nfaces[i]; f++) {
        const int ne = g->face[i * mf + f];
        if (k < g->Nke[ne])
          w[i][k] -= (th * p->u[ne][k] + (1 - th) * p->utmp2[ne][k]) * g->df[ne] * g->normal[i * mf + f] /
                     g->Ac[i] / th * g->dzf[ne][k];
      }
    }
  }
}

/* main/phys.c:2596-2942, the part of UPredictor before the open-boundary
 * hooks: wtmp2 = w, utmp2 = u, D = 0; on type-2 edges utmp = u; on
 * computational edges the explicit free-surface gradient, the wind stress,
 * and per edge the vertical system of viscosity, drag (implicit, theta = 1)
 * and vertical momentum advection (thetaM), solved for U*** (into utmp) and
 * E = A^-1 1 (into ut), with D = dzf^T E.  Returns 1 where the reference
 * exits: a level of an edge whose two cells both have zero height
 * (main/phys.c:2900-2904) or U*** = nan (:2945-2950).
 * On a single-layer edge with no-slip drag (CdB or CdT = -1) the reference
 * reads a viscosity coefficient left over from the previous edge
 * (main/phys.c:2803-2815, :2880-2889, index k after its loops); the
 * restatement uses this edge's own value, 0. */
int oracle_UPredictor_begin(gridT *g, physT *p, propT *prop) {
  const double dt = prop->dt, theta = prop->theta, th = 1.0, alpha = 1.0;
  const double nu = prop->nu, lw = prop->laxWendroff_Vertical, tM = prop->thetaM;
  const int n1 = g->Nkmax + 1;
  double *a = (double *)calloc(n1, sizeof(double)), *b = (double *)calloc(n1, sizeof(double));
  double *c = (double *)calloc(n1, sizeof(double)), *d = (double *)calloc(n1, sizeof(double));
  double *a0 = (double *)calloc(n1, sizeof(double)), *b0 = (double *)calloc(n1, sizeof(double));
  double *c0 = (double *)calloc(n1, sizeof(double)), *dw = (double *)calloc(n1, sizeof(double));
  double *bp = (double *)calloc(n1, sizeof(double)), *dp = (double *)calloc(n1, sizeof(double));
  double *ep = (double *)calloc(n1, sizeof(double));
  int bad = 0;
  for (int i = 0; i < g->Nc; i++)
    for (int k = 0; k <= g->Nk[i]; k++) p->wtmp2[i][k] = p->w[i][k];
  for (int j = 0; j < g->Ne; j++) {
    p->D[j] = 0;
    for (int k = 0; k < g->Nke[j]; k++) p->utmp2[j][k] = p->u[j][k];
  }
  for (int jp = g->edgedist[2]; jp < g->edgedist[3]; jp++) {
    const int j = g->edgep[jp];
    for (int k = g->etop[j]; k < g->Nke[j]; k++) p->utmp[j][k] = p->u[j][k];
  }
  for (int jp = g->edgedist[0]; jp < g->edgedist[1]; jp++) {
    const int j = g->edgep[jp];
    const int i1 = g->grad[2 * j], i2 = g->grad[2 * j + 1];
    for (int k = g->etop[j]; k < g->Nke[j]; k++)
      p->utmp[j][k] -= prop->grav * (1 - theta) * dt * (p->h[i1] - p->h[i2]) / g->dg[j];
  }
  for (int jp = g->edgedist[0]; jp < g->edgedist[1] && !bad; jp++) {
    const int j = g->edgep[jp];
    int i1 = g->grad[2 * j], i2 = g->grad[2 * j + 1];
    if (i1 == -1) i1 = i2;
    if (i2 == -1) i2 = i1;
    const int et = g->etop[j], ne = g->Nke[j];
    double *u = p->u[j], *ut = p->utmp[j];
    REAL *z1 = g->dzz[i1], *z2 = g->dzz[i2];
    ut[et] += 2.0 * dt * p->tau_T[j] / (z1[et] + z2[et]);
    if (z1[et] == 0 && z2[et] == 0) continue;
    for (int k = et; k < ne; k++) a[k] = b[k] = c[k] = d[k] = 0;
    for (int k = et + 1; k < ne; k++)
      c[k] = 0.25 * (p->nu_tv[i1][k - 1] + p->nu_tv[i2][k - 1] + p->nu_tv[i1][k] + p->nu_tv[i2][k] +
                     lw * (p->nu_lax[i1][k - 1] + p->nu_lax[i2][k - 1] + p->nu_lax[i1][k] + p->nu_lax[i2][k]));
    for (int k = et + 1; k < ne; k++)
      a[k] = 2.0 * (nu + c[k]) / (0.25 * (z1[k] + z2[k]) * (z1[k - 1] + z2[k - 1] + z1[k] + z2[k]));
    for (int k = et; k < ne - 1; k++)
      b[k] = 2.0 * (nu + c[k + 1]) / (0.25 * (z1[k] + z2[k]) * (z1[k] + z2[k] + z1[k + 1] + z2[k + 1]));
    const int multi = ne - et > 1, adv = prop->nonlinear && tM >= 0 && multi;
    if (adv) {
      const int lo = g->ctop[i1] > g->ctop[i2] ? i2 : i1, hi = g->ctop[i1] > g->ctop[i2] ? i1 : i2;
      for (int k = 0; k < g->ctop[hi]; k++) dw[k] = 0;
      for (int k = g->ctop[hi]; k < ne; k++) dw[k] = 0.5 * (p->w[lo][k] + p->w[hi][k]);
      dw[ne] = 0;
      for (int k = et; k < ne; k++) {
        const double hz = 0.5 * (z1[k] + z2[k]);
        a0[k] = (alpha * 0.5 * (dw[k] - fabs(dw[k])) + 0.5 * (1 - alpha) * dw[k]) / hz;
        b0[k] = (alpha * 0.5 * (dw[k] + fabs(dw[k]) - dw[k + 1] + fabs(dw[k + 1])) +
                 0.5 * (1 - alpha) * (dw[k] - dw[k + 1])) / hz;
        c0[k] = -(alpha * 0.5 * (dw[k + 1] + fabs(dw[k + 1])) + 0.5 * (1 - alpha) * dw[k + 1]) / hz;
      }
    }
    const double CdT = p->CdT[j], CdB = p->CdB[j];
    if (multi) {
      for (int k = et + 1; k < ne - 1; k++)
        ut[k] += dt * (1 - th) * (a[k] * u[k - 1] - (a[k] + b[k]) * u[k] + b[k] * u[k + 1]);
      if (CdT == -1)
        ut[et] += dt * (1 - th) * (a[et] * -u[et] - (a[et] + b[et]) * u[et] + b[et] * u[et + 1]);
      else
        ut[et] += dt * (1 - th) * (-(b[et] + 2.0 * CdT * fabs(u[et]) / (z1[et] + z2[et])) * u[et] + b[et] * u[et + 1]);
      const int kb = ne - 1;
      if (CdB == -1)
        ut[kb] += dt * (1 - th) * (a[kb] * u[kb - 1] - (a[kb] + b[kb]) * u[kb] + b[kb] * -u[kb]);
      else
        ut[kb] += dt * (1 - th) * (a[kb] * u[kb - 1] - (a[kb] + 2.0 * CdB * fabs(u[kb]) / (z1[kb] + z2[kb])) * u[kb]);
    } else {
      const double s = z1[et] + z2[et];
      if (CdB == -1) ut[et] -= 2.0 * dt * (1 - th) * (2.0 * (2.0 * (nu + c[et])) * u[et] / (s * s));
      else ut[et] -= 2.0 * dt * (1 - th) * (CdB) / s * fabs(u[et]) * u[et];
      if (CdT == -1) ut[et] -= 2.0 * dt * (1 - th) * (2.0 * (2.0 * (nu + c[et])) * u[et] / (s * s));
      else ut[et] -= 2.0 * dt * (1 - th) * (CdT) / s * fabs(u[et]) * u[et];
    }
    if (adv) {
      for (int k = et + 1; k < ne - 1; k++)
        ut[k] -= dt * (1 - tM) * (a0[k] * u[k - 1] + b0[k] * u[k] + c0[k] * u[k + 1]);
      ut[et] -= dt * (1 - tM) * ((a0[et] + b0[et]) * u[et] + c0[et] * u[et + 1]);
      const int kb = ne - 1;
      ut[kb] -= dt * (1 - tM) * (a0[kb] * u[kb - 1] + (b0[kb] + c0[kb]) * u[kb]);
    }
    if (multi) {
      const int kb = ne - 1;
      c[et] = -th * dt * b[et];
      if (CdT == -1) b[et] = 1.0 + th * dt * (a[et] + a[et + 1] + b[et]);
      else b[et] = 1.0 + th * dt * (b[et] + 2.0 * CdT * fabs(u[et]) / (z1[et] + z2[et]));
      a[et] = 0;
      c[kb] = 0;
      if (CdB == -1) b[kb] = 1.0 + th * dt * (a[kb] + b[kb] + b[kb - 1]);
      else b[kb] = 1.0 + th * dt * (a[kb] + 2.0 * CdB * fabs(u[kb]) / (z1[kb] + z2[kb]));
      a[kb] = -th * dt * a[kb];
      for (int k = et + 1; k < kb; k++) {
        c[k] = -th * dt * b[k];
        b[k] = 1.0 + th * dt * (a[k] + b[k]);
        a[k] = -th * dt * a[k];
      }
    } else {
      const double s = z1[et] + z2[et];
      b[et] = 1.0;
      if (CdB == -1) b[et] += 4.0 * th * dt * 2.0 * (nu + c[et]) / (s * s);
      else b[et] += 2.0 * th * dt * fabs(ut[et]) / s * (CdB);
      if (CdT == -1) b[et] += 4.0 * th * dt * 2.0 * (nu + c[et]) / (s * s);
      else b[et] += 2.0 * th * dt * fabs(ut[et]) / s * (CdT);
    }
    if (adv) {
      const int kb = ne - 1;
      for (int k = et + 1; k < kb; k++) {
        a[k] += dt * tM * a0[k];
        b[k] += dt * tM * b0[k];
        c[k] += dt * tM * c0[k];
      }
      b[et] += dt * tM * (a0[et] + b0[et]);
      c[et] += dt * tM * c0[et];
      a[kb] += dt * tM * a0[kb];
      b[kb] += dt * tM * (b0[kb] + c0[kb]);
    }
    for (int k = et; k < ne; k++)
      if (z1[k] == 0 && z2[k] == 0) bad = 1;
    if (bad) break;
    double *E = p->ut[j];
    if (multi) {
      /* two Thomas solves with the same matrix (main/util.c:185-203):
       * U*** from U** and E from ones; the pivots are the same for both */
      bp[et] = b[et]; dp[et] = ut[et]; ep[et] = 1.0;
      for (int k = et + 1; k < ne; k++) {
        bp[k] = b[k] - a[k] * c[k - 1] / bp[k - 1];
        dp[k] = ut[k] - a[k] * dp[k - 1] / bp[k - 1];
        ep[k] = 1.0 - a[k] * ep[k - 1] / bp[k - 1];
      }
      ut[ne - 1] = dp[ne - 1] / bp[ne - 1];
      E[ne - 1] = ep[ne - 1] / bp[ne - 1];
      for (int k = ne - 2; k >= et; k--) {
        ut[k] = dp[k] / bp[k] - c[k] * ut[k + 1] / bp[k];
        E[k] = ep[k] / bp[k] - c[k] * E[k + 1] / bp[k];
      }
    } else {
      ut[et] /= b[et];
      E[et] = 1.0 / b[et];
    }
    p->D[j] = 0;
    for (int k = et; k < ne; k++) p->D[j] += E[k] * g->dzf[j][k];
  }
  for (int j = 0; j < g->Ne && !bad; j++)
    for (int k = g->etop[j]; k < g->Nke[j]; k++)
      if (p->utmp[j][k] != p->utmp[j][k]) bad = 1;
  free(a); free(b); free(c); free(d); free(a0); free(b0); free(c0); free(dw); free(bp); free(dp); free(ep);
  return bad;
}

/* main/phys.c:2956-2971, UPredictor's free-surface source: htmp = Ac h -
 * dt times the sum of the theta-weighted face fluxes of each computational
 * cell. */
void oracle_UPredictor_rhs(gridT *g, physT *p, propT *prop) {
  const double dt = prop->dt, theta = prop->theta;
  const int mf = g->maxfaces;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip];
    double sum = 0;
    for (int f = 0; f < g->nfaces[i]; f++) {
      const int ne = g->face[i * mf + f], nrm = g->normal[i * mf + f];
      for (int k = g->etop[ne]; k < g->Nke[ne]; k++)
        sum += ((1 - theta) * p->u[ne][k] + theta * p->utmp[ne][k]) * g->df[ne] * nrm * g->dzf[ne][k];
    }
    p->htmp[i] = g->Ac[i] * p->h[i] - dt * sum;
  }
}

/* main/phys.c:2991-3018, UPredictor after its CGSolve: cells that dry below
 * DRYCELLHEIGHT clamped (h, active, bottom scalar), and u = U*** - g theta
 * dt E grad h on computational edges (zero on a dry one-layer edge). */
void oracle_UPredictor_finish(gridT *g, physT *p, propT *prop) {
  const double dt = prop->dt, theta = prop->theta;
  for (int i = 0; i < g->Nc; i++)
    if (p->h[i] <= -g->dv[i] + 1e-10) {
      p->h[i] = -g->dv[i] + 1e-10;
      p->active[i] = 0;
      p->s[i][g->Nk[i] - 1] = 0;
    } else {
      p->hcorr[i] = 0;
      p->active[i] = 1;
    }
  for (int jp = g->edgedist[0]; jp < g->edgedist[1]; jp++) {
    const int j = g->edgep[jp];
    const int i1 = g->grad[2 * j], i2 = g->grad[2 * j + 1];
    for (int k = g->etop[j]; k < g->Nke[j]; k++)
      p->u[j][k] = p->utmp[j][k] - prop->grav * theta * dt * p->ut[j][k] * (p->h[i1] - p->h[i2]) / g->dg[j];
    if (g->etop[j] == g->Nke[j] - 1 && g->dzz[i1][g->etop[j]] == 0 && g->dzz[i2][g->etop[j]] == 0)
      p->u[j][g->etop[j]] = 0;
  }
}

/* main/phys.c:2951-3018, the part of UPredictor after the open-boundary
 * hooks: the free-surface source, CGSolve (cgsolver = 1), then the drying
 * clamp and the velocity update.  Returns 1 for cgsolver != 1 (GSSolve is
 * not restated). */
int oracle_UPredictor_solve(gridT *g, physT *p, propT *prop) {
  if (prop->cgsolver != 1) return 1;
  oracle_UPredictor_rhs(g, p, prop);
  oracle_CGSolve(g, p, prop);
  oracle_UPredictor_finish(g, p, prop);
  return 0;
}

static double ref_max(double x1, double x2) { return x2 > x1 ? x2 : x1; }  /* main/util.c:212-217 */
static double ref_min(double x, double y) { return x < y ? x : y; }      /* main/util.c:241-246 */
static double ref_upwind(double u, double a, double b) {                 /* main/util.c:219-232 */
  if (fabs(u) > 0) return u > 0 ? b : a;
  return ref_max(a, b);
}

/* main/phys.c:816-968.  The layer heights dzz and the top cell ctop of every
 * column from the free surface h (a column's cells are full below the
 * surface cell, the bottom cell reaches the bed dv, cells above the surface
 * are empty), and etop = the smaller ctop of an edge's cells.  option 0
 * first keeps the old extents and heights (etopold, ctopold, dzzold);
 * option -1 (initialisation) starts from h = 0 and records the bottom cell
 * height (dzbot), deepening dv where fixdzz asks; option 1 afterwards copies
 * the new state into the old one and sets the bottom flux height dzfB. */
void oracle_UpdateDZ(gridT *g, physT *p, propT *prop, int option) {
  const int Nc = g->Nc, Ne = g->Ne;
  if (prop->linearFS) return;
  if (option == 0) {
    for (int j = 0; j < Ne; j++) g->etopold[j] = g->etop[j];
    for (int i = 0; i < Nc; i++) {
      g->ctopold[i] = g->ctop[i];
      for (int k = 0; k < g->ctop[i]; k++) g->dzzold[i][k] = 0;
      for (int k = g->ctop[i]; k < g->Nk[i]; k++) g->dzzold[i][k] = g->dzz[i][k];
    }
  }
  if (option == -1)
    for (int i = 0; i < Nc; i++) p->h[i] = 0.0;
  for (int i = 0; i < Nc; i++) {
    double z = 0;
    for (int k = 0; k < g->Nk[i]; k++) z -= g->dz[k];
    g->dzz[i][g->Nk[i] - 1] = g->dz[g->Nk[i] - 1] + g->dv[i] + z;
  }
  if (g->Nkmax > 1) {
    for (int i = 0; i < Nc; i++) {
      const int n = g->Nk[i];
      double z = 0;
      int found = 0;
      for (int k = 0; k < n; k++) {
        z -= g->dz[k];
        if (p->h[i] >= z) {
          if (!found) {
            if (k == n - 1) { g->dzz[i][k] = p->h[i] + g->dv[i]; g->ctop[i] = k; }
            else if (p->h[i] == z) { g->dzz[i][k] = 0; g->ctop[i] = k + 1; }
            else { g->dzz[i][k] = p->h[i] - z; g->ctop[i] = k; }
            found = 1;
          } else if (k == n - 1) {
            g->dzz[i][k] = g->dz[k] + g->dv[i] + z;
          } else {
            g->dzz[i][k] = z < -g->dv[i] ? 0 : g->dz[k];
          }
        } else {
          g->dzz[i][k] = 0;
        }
      }
    }
  } else {
    for (int i = 0; i < Nc; i++) g->dzz[i][0] = g->dv[i] + p->h[i];
  }
  for (int j = 0; j < Ne; j++) {
    const int a = g->grad[2 * j], b = g->grad[2 * j + 1];
    if (a == -1) g->etop[j] = g->ctop[b];
    else if (b == -1) g->etop[j] = g->ctop[a];
    else g->etop[j] = g->ctop[a] < g->ctop[b] ? g->ctop[a] : g->ctop[b];
  }
  if (option == -1)
    for (int i = 0; i < Nc; i++) {
      const int k = g->Nk[i] - 1;
      g->dzbot[i] = g->dzz[i][k];
      if (!g->stairstep && g->fixdzz && g->dzz[i][k] < g->dz[k] * g->dzsmall)
        g->dv[i] += (g->dz[k] * g->dzsmall - g->dzz[i][k]);
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
      double z1 = g->dzz[a][g->Nk[a] - 1], z2 = g->dzz[b][g->Nk[b] - 1], z = 0;
      for (int k = 0; k < g->Nke[j] - 1; k++) z -= g->dz[k];
      if (p->h[a] < z) z1 = z1 - z + p->h[a];
      if (p->h[b] < z) z2 = z2 - z + p->h[b];
      g->dzfB[j] = ref_min(z1, z2);
    }
  }
}

/* main/phys.c:4338-4435.  Flux-face heights: zero everywhere, then on the
 * wet layers of each edge the upwind cell height by the sign of u (the
 * larger for u = 0); the bottom wet layer takes the smaller cell height, or
 * for a one-layer edge the upwind free surface over the shallower bed
 * (not below 0); heights at or below DRYCELLHEIGHT are zeroed.  smoothbot
 * bounds the bottom cell heights from below before and after. */
void oracle_SetFluxHeight(gridT *g, physT *p, propT *prop) {
  (void)prop;
  for (int j = 0; j < g->Ne; j++)
    for (int k = 0; k < g->Nkc[j]; k++) g->dzf[j][k] = 0;
  if (g->smoothbot)
    for (int i = 0; i < g->Nc; i++)
      g->dzz[i][g->Nk[i] - 1] = ref_max(g->dzbot[i], g->smoothbot * g->dz[g->Nk[i] - 1]);
  for (int j = 0; j < g->Ne; j++) {
    int a = g->grad[2 * j], b = g->grad[2 * j + 1];
    if (a == -1) a = b;
    if (b == -1) b = a;
    for (int k = 0; k < g->etop[j]; k++) g->dzf[j][k] = 0;
    for (int k = g->etop[j]; k < g->Nke[j]; k++) g->dzf[j][k] = ref_upwind(p->u[j][k], g->dzz[a][k], g->dzz[b][k]);
    const int k = g->Nke[j] - 1;
    if (g->etop[j] == k)
      g->dzf[j][k] = ref_max(0, ref_upwind(p->u[j][k], p->h[a], p->h[b]) + ref_min(g->dv[a], g->dv[b]));
    else
      g->dzf[j][k] = ref_min(g->dzz[a][k], g->dzz[b][k]);
    for (int kk = g->etop[j]; kk < g->Nke[j]; kk++)
      if (g->dzf[j][kk] <= 1e-10) g->dzf[j][kk] = 0;
  }
  if (g->smoothbot)
    for (int i = 0; i < g->Nc; i++)
      g->dzz[i][g->Nk[i] - 1] = ref_max(g->dzz[i][g->Nk[i] - 1], g->dzsmall * g->dz[g->Nk[i] - 1]);
}

/* main/tvd.c:36-120.  Pass 1 builds, per cell-layer, the inflow volume flux
 * sumQ and the inflow-weighted scalar jump sumQC (in gradSx/gradSy); pass 2
 * limits the face values of each wet interior edge-layer by Psi of the
 * upwind cell's ratio; type-2 edges take the boundary value.  The reference
 * reads scal[neigh][k] below the neighbour's bottom (multiplied by zero);
 * the oracle skips that term. */
void oracle_HorizontalFaceScalars(gridT *g, physT *p, propT *prop, REAL **scal, REAL **boundary_scal,
                                  int TVD) {
  REAL **sumQ = p->gradSx, **sumQC = p->gradSy;
  const int mf = g->maxfaces;
  const double th = prop->theta;
  for (int ip = g->celldist[0]; ip < g->celldist[2]; ip++) {
    const int i = g->cellp[ip];
    for (int k = g->ctop[i]; k < g->Nk[i]; k++) {
      sumQ[i][k] = 0;
      sumQC[i][k] = 0;
      for (int f = 0; f < g->nfaces[i]; f++) {
        const int ne = g->face[i * mf + f], nb = g->neigh[i * mf + f], nrm = g->normal[i * mf + f];
        const double un = nrm * (th * p->u[ne][k] + (1 - th) * p->utmp2[ne][k]);
        const double Qm = 0.5 * g->dzf[ne][k] * g->df[ne] * fabs(un - fabs(un));
        if (nb != -1 && k < g->Nk[nb]) sumQC[i][k] += Qm * (scal[i][k] - scal[nb][k]);
        sumQ[i][k] += Qm;
      }
    }
  }
  for (int jp = g->edgedist[0]; jp < g->edgedist[1]; jp++) {
    const int j = g->edgep[jp], nc1 = g->grad[2 * j], nc2 = g->grad[2 * j + 1];
    for (int k = 0; k < g->etop[j]; k++) p->SfHp[j][k] = p->SfHm[j][k] = 0;
    for (int k = g->etop[j]; k < g->Nke[j]; k++) {
      const double un = th * p->u[j][k] + (1 - th) * p->utmp2[j][k];
      const int up = un > 0 ? nc2 : nc1;
      const double si = un > 0 ? scal[nc2][k] : scal[nc1][k];
      const double sm = un > 0 ? scal[nc1][k] : scal[nc2][k];
      double r = 0;
      if (sumQ[up][k] != 0 && sm != si) r = sumQC[up][k] / (sumQ[up][k] * (sm - si));
      p->SfHp[j][k] = scal[nc2][k] + 0.5 * oracle_Psi(r, TVD) * (scal[nc1][k] - scal[nc2][k]);
      p->SfHm[j][k] = scal[nc1][k] - 0.5 * oracle_Psi(r, TVD) * (scal[nc1][k] - scal[nc2][k]);
    }
  }
  for (int jp = g->edgedist[2]; jp < g->edgedist[3]; jp++) {
    const int j = g->edgep[jp], ib = g->grad[2 * j];
    for (int k = 0; k < g->etop[j]; k++) p->SfHp[j][k] = p->SfHm[j][k] = 0;
    for (int k = g->etop[j]; k < g->Nke[j]; k++) {
      p->SfHp[j][k] = boundary_scal[jp - g->edgedist[2]][k];
      p->SfHm[j][k] = scal[ib][k];
    }
  }
}

/* Column read with the out-of-row cases of GetApAm mapped to 0. */
static double col(REAL **s, int i, int k, int nk) { return (k >= 0 && k < nk) ? s[i][k] : 0.0; }

/* main/tvd.c:161-204: vertical TVD flux coefficients for one column.  For
 * columns with fewer than three wet layers the reference reads outside the
 * row; the oracle reads 0 there. */
void oracle_GetApAm(REAL *ap, REAL *am, REAL *wp, REAL *wm, REAL *Cp, REAL *Cm, REAL *rp, REAL *rm,
                    REAL **w, REAL **dzz, REAL **scal, int i, int Nk, int ktop, REAL dt, int TVD) {
  for (int k = 0; k < Nk + 1; k++) {
    wp[k] = 0.5 * (w[i][k] + fabs(w[i][k]));
    wm[k] = 0.5 * (w[i][k] - fabs(w[i][k]));
  }
  for (int k = 1; k < Nk; k++) {
    Cp[k] = 2 * wp[k] * dt / (dzz[i][k] + dzz[i][k - 1]);
    Cm[k] = 2 * wm[k] * dt / (dzz[i][k] + dzz[i][k - 1]);
  }
  Cp[Nk] = wp[Nk] * dt / dzz[i][Nk - 1];
  Cm[Nk] = wm[Nk] * dt / dzz[i][Nk - 1];
#define S(k) col(scal, i, (k), Nk)
  for (int k = ktop + 2; k < Nk - 1; k++) {
    rp[k - ktop] = (S(k) - S(k + 1) + TVD_EPS) / (S(k - 1) - S(k) + TVD_EPS);
    rm[k - ktop] = (S(k - 2) - S(k - 1) + TVD_EPS) / (S(k - 1) - S(k) + TVD_EPS);
  }
  rp[1] = (S(ktop + 1) - S(ktop + 2) + TVD_EPS) / (S(ktop) - S(ktop + 1) + TVD_EPS);
  rm[1] = TVD_EPS / (S(ktop) - S(ktop + 1) + TVD_EPS);
  {
    const int k = Nk - 1;
    rp[k - ktop] = TVD_EPS / (S(k - 1) - S(k) + TVD_EPS);
    rm[k - ktop] = (S(k - 2) - S(k - 1) + TVD_EPS) / (S(k - 1) - S(k) + TVD_EPS);
  }
  rp[Nk - ktop] = 1;
  rm[Nk - ktop] = (S(Nk - 2) - S(Nk - 1) + TVD_EPS) / TVD_EPS;
#undef S
  for (int k = ktop + 1; k < Nk + 1; k++) {
    am[k] = 0.5 * wp[k] * oracle_Psi(rp[k - ktop], TVD) * (1 - Cp[k]) +
            wm[k] * (1 - 0.5 * oracle_Psi(rm[k - ktop], TVD) * (1 + Cm[k]));
    ap[k] = wp[k] * (1 - 0.5 * oracle_Psi(rp[k - ktop], TVD) * (1 - Cp[k])) +
            0.5 * wm[k] * oracle_Psi(rm[k - ktop], TVD) * (1 + Cm[k]);
  }
}

/* main/scalars.c:35-325 with CHECKCONSISTENCY=0 (main/suntans.h:30).
 * Adams-Bashforth is disabled in the reference (fab=1, Cn zeroed first). */
void oracle_UpdateScalars(gridT *g, physT *p, propT *prop, REAL **wnew, REAL **scal,
                          REAL **boundary_scal, REAL **Cn, REAL kappa, REAL kappaH, REAL **kappa_tv,
                          REAL theta, REAL **src1, REAL **src2, REAL *Ftop, REAL *Fbot, int alpha_top,
                          int alpha_bot, int checkflag, int TVDscheme) {
  (void)kappaH; (void)checkflag;
  const int mf = g->maxfaces;
  const double dt = prop->dt, fab = 1;
  REAL *ap = p->ap, *am = p->am, *bd = p->bp, *temp = p->bm;
  REAL *a = p->a, *b = p->b, *c = p->c, *d = p->d;
  prop->TVD = TVDscheme;
  prop->horiTVD = 1;
  prop->vertTVD = 1;
  const int tvd_h = prop->TVD && prop->horiTVD, tvd_v = prop->TVD && prop->vertTVD;

  for (int i = 0; i < g->Nc; i++)
    for (int k = 0; k < g->Nk[i]; k++) Cn[i][k] = 0;
  for (int i = 0; i < g->Nc; i++)
    for (int k = 0; k < g->Nk[i]; k++) p->stmp[i][k] = scal[i][k];
  FOR_CELLS(g, i, 0, 1) for (int k = g->ctop[i]; k < g->Nk[i]; k++) p->stmp2[i][k] = 0;
  if (boundary_scal)
    for (int jp = g->edgedist[2]; jp < g->edgedist[5]; jp++) {
      const int j = g->edgep[jp], ib = g->grad[2 * j];
      for (int k = g->ctop[ib]; k < g->Nk[ib]; k++) p->stmp2[ib][k] = boundary_scal[jp - g->edgedist[2]][k];
    }
  if (tvd_h) oracle_HorizontalFaceScalars(g, p, prop, scal, boundary_scal, prop->TVD);

  FOR_CELLS(g, i, 0, 1) {
    const double Ac = g->Ac[i];
    const int nk = g->Nk[i], ct = g->ctop[i], cto = g->ctopold[i];
    int ktop;
    double dznew;
    if (ct >= cto) {
      ktop = ct;
      dznew = g->dzz[i][ktop];
    } else {
      ktop = cto;
      dznew = 0;
      for (int k = ct; k <= cto; k++) dznew += g->dzz[i][k];
    }
    const int nb = nk - 1 - ktop; /* row of the bottom cell */

    /* implicit (new time level) vertical advection */
    if (!tvd_v)
      for (int k = 0; k < nk + 1; k++) {
        ap[k] = 0.5 * (wnew[i][k] + fabs(wnew[i][k]));
        am[k] = 0.5 * (wnew[i][k] - fabs(wnew[i][k]));
      }
    else
      oracle_GetApAm(ap, am, p->wp, p->wm, p->Cp, p->Cm, p->rp, p->rm, wnew, g->dzz, scal, i, nk, ktop, dt,
                     prop->TVD);
    for (int k = ktop + 1; k < nk; k++) {
      a[k - ktop] = theta * dt * am[k];
      b[k - ktop] = g->dzz[i][k] + theta * dt * (ap[k] - am[k + 1]);
      c[k - ktop] = -theta * dt * ap[k + 1];
    }
    a[0] = 0;
    b[0] = dznew - theta * dt * am[ktop + 1];
    c[0] = -theta * dt * ap[ktop + 1];
    b[nb] += c[nb];

    /* implicit vertical diffusion */
    for (int k = ktop + 1; k < nk; k++)
      bd[k] = (2.0 * kappa + kappa_tv[i][k - 1] + kappa_tv[i][k]) / (g->dzz[i][k - 1] + g->dzz[i][k]);
    for (int k = ktop + 1; k < nk - 1; k++) {
      a[k - ktop] -= theta * dt * bd[k];
      b[k - ktop] += theta * dt * (bd[k] + bd[k + 1]);
      c[k - ktop] -= theta * dt * bd[k + 1];
    }
    if (src1)
      for (int k = ktop; k < nk; k++) b[k - ktop] += g->dzz[i][k] * src1[i][k] * theta * dt;
    if (ktop < nk - 1) {
      b[0] += theta * dt * (bd[ktop + 1] + 2 * alpha_top * bd[ktop + 1]);
      c[0] -= theta * dt * bd[ktop + 1];
      a[nb] -= theta * dt * bd[nk - 1];
      b[nb] += theta * dt * (bd[nk - 1] + 2 * alpha_bot * bd[nk - 1]);
    }

    /* explicit (old time level) part */
    for (int k = ktop + 1; k < nk; k++) d[k - ktop] = g->dzzold[i][k] * p->stmp[i][k];
    if (src1)
      for (int k = ktop + 1; k < nk; k++)
        d[k - ktop] -= src1[i][k] * (1 - theta) * dt * g->dzzold[i][k] * p->stmp[i][k];
    d[0] = 0;
    if (cto <= ct) {
      for (int k = cto; k <= ct; k++) d[0] += g->dzzold[i][k] * p->stmp[i][k];
      if (src1)
        for (int k = cto; k <= ct; k++) d[0] -= src1[i][k] * (1 - theta) * dt * g->dzzold[i][k] * p->stmp[i][k];
    } else {
      d[0] = g->dzzold[i][ktop] * p->stmp[i][ktop];
      /* the reference multiplies by stmp at a stale loop index here; the
       * oracle uses the top cell */
      if (src1) d[0] -= src1[i][ktop] * (1 - theta) * dt * g->dzzold[i][ktop] * p->stmp[i][ktop];
    }
    if (!tvd_v)
      for (int k = 0; k < nk + 1; k++) {
        ap[k] = 0.5 * (p->wtmp2[i][k] + fabs(p->wtmp2[i][k]));
        am[k] = 0.5 * (p->wtmp2[i][k] - fabs(p->wtmp2[i][k]));
      }
    else
      oracle_GetApAm(ap, am, p->wp, p->wm, p->Cp, p->Cm, p->rp, p->rm, p->wtmp2, g->dzzold, p->stmp, i, nk,
                     ktop, dt, prop->TVD);
    REAL *s = p->stmp[i];
    for (int k = ktop + 1; k < nk - 1; k++)
      d[k - ktop] -= (1 - theta) * dt * (am[k] * s[k - 1] + (ap[k] - am[k + 1]) * s[k] - ap[k + 1] * s[k + 1]) -
                     (1 - theta) * dt * (bd[k] * s[k - 1] - (bd[k] + bd[k + 1]) * s[k] + bd[k + 1] * s[k + 1]);
    if (ktop < nk - 1) {
      int k = ktop;
      d[0] = d[0] - (1 - theta) * dt * (-am[k + 1] * s[k] - ap[k + 1] * s[k + 1]) +
             (1 - theta) * dt * (-(2 * alpha_top * bd[k + 1] + bd[k + 1]) * s[k] + bd[k + 1] * s[k + 1]);
      if (Ftop) d[0] += dt * (1 - alpha_top + 2 * alpha_top * bd[k + 1]) * Ftop[i];
      k = nk - 1;
      d[k - ktop] -= (1 - theta) * dt * (am[k] * s[k - 1] + ap[k] * s[k]) -
                     (1 - theta) * dt * (bd[k] * s[k - 1] - (bd[k] + 2 * alpha_bot * bd[k]) * s[k]);
      if (Fbot) d[k - ktop] += dt * (-1 + alpha_bot + 2 * alpha_bot * bd[k]) * Fbot[i];
    }

    /* previous-step AB term: weight 1-fab = 0 */
    if (ct <= cto) {
      for (int k = ct; k <= cto; k++) d[0] += (1 - fab) * Cn[i][cto] / (1 + abs(ct - cto));
      for (int k = cto + 1; k < nk; k++) d[k - cto] += (1 - fab) * Cn[i][k];
    } else {
      for (int k = cto; k <= ct; k++) d[0] += (1 - fab) * Cn[i][k];
      for (int k = ct + 1; k < nk; k++) d[k - ct] += (1 - fab) * Cn[i][k];
    }
    for (int k = 0; k < ct; k++) Cn[i][k] = 0;
    /* with src2 the reference writes Cn[i][k-ktop] from k = ctop: below row 0
     * when ctop < ktop (the surface rose), outside the row; those writes
     * are left out here */
    if (src2) {
      for (int k = ct; k < nk; k++)
        if (k >= ktop) Cn[i][k - ktop] = dt * src2[i][k] * g->dzzold[i][k];
    } else {
      for (int k = ct; k < nk; k++) Cn[i][k] = 0;
    }

    /* horizontal flux divergence with the TVD face values */
    for (int k = 0; k < nk; k++) ap[k] = 0;
    for (int nf = 0; nf < g->nfaces[i]; nf++) {
      const int ne = g->face[i * mf + nf], nrm = g->normal[i * mf + nf];
      const double df = g->df[ne];
      int nc1 = g->grad[2 * ne], nc2 = g->grad[2 * ne + 1];
      REAL *sp;
      if (nc1 == -1) nc1 = nc2;
      if (nc2 == -1) {
        nc2 = nc1;
        sp = (boundary_scal && (g->mark[ne] == 2 || g->mark[ne] == 3)) ? p->stmp2[nc1] : p->stmp[nc1];
      } else {
        sp = p->stmp[nc2];
      }
      if (!tvd_h)
        for (int k = 0; k < g->Nke[ne]; k++) temp[k] = upwind_ref(p->utmp2[ne][k], p->stmp[nc1][k], sp[k]);
      else
        for (int k = 0; k < g->Nke[ne]; k++) temp[k] = p->utmp2[ne][k] > 0 ? p->SfHp[ne][k] : p->SfHm[ne][k];
      for (int k = 0; k < g->Nke[ne]; k++)
        ap[k] += dt * df * nrm / Ac * (theta * p->u[ne][k] + (1 - theta) * p->utmp2[ne][k]) * temp[k] *
                 g->dzf[ne][k];
    }
    for (int k = ktop + 1; k < nk; k++) Cn[i][k - ktop] -= ap[k];
    for (int k = 0; k <= ktop; k++) Cn[i][0] -= ap[k];
    for (int k = 0; k < nk - ktop; k++) d[k] += fab * Cn[i][k];
    for (int k = ktop; k < nk; k++) ap[k] = Cn[i][k - ktop];
    for (int k = 0; k <= ktop; k++) Cn[i][k] = 0;
    for (int k = ktop + 1; k < nk; k++) Cn[i][k] = ap[k];
    for (int k = ct; k <= ktop; k++) Cn[i][k] = ap[ktop] / (1 + abs(ct - ktop));

    if (nk - ktop > 1) {
      oracle_TriSolve(a, b, c, d, scal[i] + ktop, nk - ktop);
    } else if (prop->n > 1) {
      scal[i][ktop] = (b[0] > 0 && p->active[i]) ? d[0] / b[0] : 0;
    }
    for (int k = 0; k < ct; k++) scal[i][k] = 0;
    for (int k = ct; k < cto; k++) scal[i][k] = scal[i][ktop];
  }
}

void oracle_solve_q(gridT *g, physT *p, propT *prop) {
  oracle_CGSolveQ(p->qc, p->stmp, p->stmp3, g, p, prop);
}
void oracle_solve_h(gridT *g, physT *p, propT *prop) { oracle_CGSolve(g, p, prop); }
void oracle_update_salinity(gridT *g, physT *p, propT *prop) {
  oracle_UpdateScalars(g, p, prop, p->wnew, p->s, p->boundary_s, p->Cn_R, prop->kappa_s, prop->kappa_sH,
                       p->kappa_tv, prop->theta, NULL, NULL, NULL, NULL, 0, 0, 1, prop->TVDsalt);
}
/* the salinity update with source terms; flags as ref_update_salinity_src in
 * oracle/ref_harness.c (1 src1 = uold, 2 src2 = wtmp, 4 Ftop = htmp, 8 Fbot
 * = htmp2) */
void oracle_update_salinity_src(gridT *g, physT *p, propT *prop, int flags, int alpha_top, int alpha_bot) {
  oracle_UpdateScalars(g, p, prop, p->wnew, p->s, p->boundary_s, p->Cn_R, prop->kappa_s, prop->kappa_sH,
                       p->kappa_tv, prop->theta, (flags & 1) ? p->uold : NULL, (flags & 2) ? p->wtmp : NULL,
                       (flags & 4) ? p->htmp : NULL, (flags & 8) ? p->htmp2 : NULL, alpha_top, alpha_bot, 1,
                       prop->TVDsalt);
}

/* ------------------------------------------------------------------------ */
/* WPredictor (main/phys.c:2025-2290) and the face values it advects with   */

/* Distance from the centre of the edge's first cell to the edge midpoint,
 * as main/phys.c:4272-4273 computes it. */
double oracle_face_def1(gridT *g, int j) {
  int nc1 = g->grad[2 * j];
  if (nc1 == -1) nc1 = g->grad[2 * j + 1];
  return sqrt(pow(g->xv[nc1] - g->xe[j], 2) + pow(g->yv[nc1] - g->ye[j], 2));
}

/* main/phys.c:4265-4289: distance-weighted interpolation of phi to face j
 * (method 2), less the Lax-Wendroff correction (method 4); upwind where a
 * cell centre lies on the face. */
static double face_interp(gridT *g, int j, int k, REAL **phi, REAL **U, double dt, int method) {
  int nc1 = g->grad[2 * j], nc2 = g->grad[2 * j + 1];
  if (nc1 == -1) nc1 = nc2;
  if (nc2 == -1) nc2 = nc1;
  const double dg = g->dg[j], d1 = oracle_face_def1(g, j), d2 = dg - d1;
  double C = 0;
  if (method == 4) C = U[j][k] * dt / dg;
  if (d1 == 0 || d2 == 0) return upwind_ref(U[j][k], phi[nc1][k], phi[nc2][k]);
  return (phi[nc1][k] * d2 + phi[nc2][k] * d1) / (d1 + d2) - C / 2 * (phi[nc1][k] - phi[nc2][k]);
}

/* main/phys.c:4879-4992: face values uf of the cell field ui advected by the
 * edge velocity U.  Type-2 edges take the boundary row where U > 0 (inflow)
 * and the inner cell otherwise; type-4 edges the boundary row; each
 * computational edge is zero above the deeper of its cells' tops and below
 * that takes the scheme's value (1 upwind, 2/4 face_interp, 5 the TVD face
 * values of HorizontalFaceScalars); faces between a type-3 cell and a
 * neighbour are upwind whatever the scheme.  Other edges keep their values. */
void oracle_MomentumFaceValues(REAL **uf, REAL **ui, REAL **bui, REAL **U, gridT *g, physT *p, propT *prop) {
  const int mf = g->maxfaces, b0 = g->edgedist[2], scheme = prop->nonlinear;
  for (int jp = g->edgedist[2]; jp < g->edgedist[3]; jp++) {
    const int j = g->edgep[jp], i = g->grad[2 * j];
    for (int k = g->etop[j]; k < g->Nke[j]; k++) uf[j][k] = U[j][k] > 0 ? bui[jp - b0][k] : ui[i][k];
  }
  for (int jp = g->edgedist[4]; jp < g->edgedist[5]; jp++) {
    const int j = g->edgep[jp];
    for (int k = g->etop[j]; k < g->Nke[j]; k++) uf[j][k] = bui[jp - b0][k];
  }
  if (scheme == 5) oracle_HorizontalFaceScalars(g, p, prop, ui, bui, prop->TVDmomentum);
  for (int jp = g->edgedist[0]; jp < g->edgedist[1]; jp++) {
    const int j = g->edgep[jp], nc1 = g->grad[2 * j], nc2 = g->grad[2 * j + 1];
    const int kmin = g->ctop[nc1] > g->ctop[nc2] ? g->ctop[nc1] : g->ctop[nc2];
    for (int k = 0; k < kmin; k++) uf[j][k] = 0;
    for (int k = kmin; k < g->Nke[j]; k++) {
      const int up = U[j][k] > 0 ? nc2 : nc1;
      if (scheme == 2 || scheme == 4) uf[j][k] = face_interp(g, j, k, ui, U, prop->dt, scheme);
      else if (scheme == 5) uf[j][k] = U[j][k] > 0 ? p->SfHp[j][k] : p->SfHm[j][k];
      else uf[j][k] = ui[up][k];
    }
  }
  for (int ip = g->celldist[1]; ip < g->celldist[2]; ip++) {
    const int i = g->cellp[ip];
    for (int f = 0; f < g->nfaces[i]; f++) {
      if (g->neigh[i * mf + f] == -1) continue;
      const int j = g->face[i * mf + f], nc1 = g->grad[2 * j], nc2 = g->grad[2 * j + 1];
      const int kmin = g->ctop[nc1] > g->ctop[nc2] ? g->ctop[nc1] : g->ctop[nc2];
      for (int k = 0; k < kmin; k++) uf[j][k] = 0;
      for (int k = kmin; k < g->Nke[j]; k++) uf[j][k] = ui[U[j][k] > 0 ? nc2 : nc1][k];
    }
  }
}

/* main/phys.c:2025-2290.  Adams-Bashforth weights from the step number and
 * AB (Cn_W/Cn_W2 cleared on the first step); per computational column
 * wtmp = w + f2 Cn_W + f3 Cn_W2 less the explicit nonhydrostatic pressure
 * gradient, Cn_W2 <- Cn_W.  The cell source stmp: horizontal advection of
 * the cell-centred w (wc) through the face values above (ut), vertical
 * advection of w^2 for schemes 1, 2 and 5, horizontal diffusion
 * (nu_H + nu_lax) and sidewall drag (CdW) scattered from computational
 * edges in edgep order, the no-slip term of type-4 edges.  Cn_W takes
 * -dt stmp averaged to the faces (and for scheme 4 the vertical
 * Lax-Wendroff flux), wtmp += f1 Cn_W, and the implicit vertical diffusion
 * of each column with more than one layer is solved into w (TriSolve),
 * a one-layer column copies wtmp.
 * Reads outside rows as the reference makes them: with conserveMomentum the
 * top-cell fill reads a[k] above ctop left over from the previous column
 * (the same scratch here, so the result matches); scheme 4 reads dzz one
 * past the column bottom (read as 0 here: parity unpinned for scheme 4's
 * vertical Lax-Wendroff term, which the device refuses); the type-4 term
 * runs to Nke[cell] of the inner cell's index and reads its boundary row
 * past Nke+1 (read as 0 here; pinned where it stays inside the row). */
void oracle_WPredictor(gridT *g, physT *p, propT *prop) {
  REAL *a = p->a, *b = p->b, *c = p->c;
  const int mf = g->maxfaces, scheme = prop->nonlinear;
  const double dt = prop->dt;
  double f1, f2, f3;
  if (prop->n == 1) {
    f1 = 1; f2 = f3 = 0;
    for (int i = 0; i < g->Nc; i++)
      for (int k = 0; k < g->Nk[i]; k++) p->Cn_W[i][k] = p->Cn_W2[i][k] = 0;
  } else if (prop->n == 2 || prop->AB == 2) {
    f1 = 3.0 / 2.0; f2 = -1.0 / 2.0; f3 = 0;
  } else {
    f1 = 23.0 / 12.0; f2 = -4.0 / 3.0; f3 = 5.0 / 12.0;
  }
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip], ct = g->ctop[i];
    for (int k = ct; k < g->Nk[i]; k++) {
      p->wtmp[i][k] = p->w[i][k] + f2 * p->Cn_W[i][k] + f3 * p->Cn_W2[i][k];
      p->Cn_W2[i][k] = p->Cn_W[i][k];
      p->Cn_W[i][k] = 0;
    }
    for (int k = ct + 1; k < g->Nk[i]; k++)
      p->wtmp[i][k] -= 2.0 * dt / (g->dzz[i][k - 1] + g->dzz[i][k]) * (p->q[i][k - 1] - p->q[i][k]);
    p->wtmp[i][ct] += 2.0 * dt / g->dzz[i][ct] * p->q[i][ct];
  }
  for (int i = 0; i < g->Nc; i++)
    for (int k = 0; k < g->Nk[i]; k++) p->stmp[i][k] = 0;

  if (scheme) {
    for (int i = 0; i < g->Nc; i++)
      for (int k = g->ctop[i]; k < g->Nk[i]; k++) p->wc[i][k] = 0.5 * (p->w[i][k] + p->w[i][k + 1]);
    oracle_MomentumFaceValues(p->ut, p->wc, p->boundary_w, p->utmp2, g, p, prop);
    if (prop->conserveMomentum)
      for (int jp = g->edgedist[0]; jp < g->edgedist[5]; jp++) {
        const int j = g->edgep[jp];
        for (int k = g->etop[j]; k < g->Nke[j]; k++) p->ut[j][k] *= g->dzf[j][k];
      }
    for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
      const int i = g->cellp[ip], ct = g->ctop[i];
      for (int k = ct; k < g->Nk[i]; k++) a[k] = prop->conserveMomentum ? g->dzz[i][k] : 1.0;
      for (int f = 0; f < g->nfaces[i]; f++) {
        const int ne = g->face[i * mf + f];
        for (int k = ct; k < g->Nk[i]; k++)
          p->stmp[i][k] += p->ut[ne][k] * p->utmp2[ne][k] * g->df[ne] * g->normal[i * mf + f] / (a[k] * g->Ac[i]);
        if (prop->conserveMomentum)
          for (int k = g->etop[ne]; k < ct; k++)
            p->stmp[i][ct] += p->ut[ne][k] * p->utmp2[ne][k] * g->df[ne] * g->normal[i * mf + f] / (a[k] * g->Ac[i]);
      }
      if (scheme == 1 || scheme == 2 || scheme == 5)
        for (int k = ct; k < g->Nk[i]; k++)
          p->stmp[i][k] += (pow(p->w[i][k], 2) - pow(p->w[i][k + 1], 2)) / g->dzz[i][k];
    }
  }

  for (int jp = g->edgedist[0]; jp < g->edgedist[1]; jp++) {
    const int j = g->edgep[jp], nc1 = g->grad[2 * j], nc2 = g->grad[2 * j + 1];
    const int kmin = g->ctop[nc1] > g->ctop[nc2] ? g->ctop[nc1] : g->ctop[nc2];
    for (int k = kmin; k < g->Nke[j]; k++) {
      a[k] = .5 * (prop->nu_H + 0.5 * (p->nu_lax[nc1][k] + p->nu_lax[nc2][k])) *
             (p->w[nc2][k] - p->w[nc1][k] + p->w[nc2][k + 1] - p->w[nc1][k + 1]) * g->df[j] / g->dg[j];
      p->stmp[nc1][k] -= a[k] / g->Ac[nc1];
      p->stmp[nc2][k] += a[k] / g->Ac[nc2];
    }
    const int side[2] = {nc1, nc2};
    for (int s = 0; s < 2; s++) {
      const int i = side[s];
      for (int k = g->Nke[j]; k < g->Nk[i]; k++)
        p->stmp[i][k] += 0.25 * prop->CdW * fabs(p->w[i][k] + p->w[i][k + 1]) * (p->w[i][k] + p->w[i][k + 1]) *
                         g->df[j] / g->Ac[i];
    }
  }
  for (int jp = g->edgedist[4]; jp < g->edgedist[5]; jp++) {
    const int j = g->edgep[jp];
    const int i = g->grad[2 * j] == -1 ? g->grad[2 * j + 1] : g->grad[2 * j];
    const REAL *bw = p->boundary_w[jp - g->edgedist[2]];
    for (int k = g->ctop[i]; k < g->Nke[i]; k++) {
      const double bk = k <= g->Nke[j] ? bw[k] : 0.0;
      p->stmp[i][k] += -2.0 * prop->nu_H * (bk - 0.5 * (p->w[i][k] + p->w[i][k + 1])) / g->dg[j] * g->df[j] /
                       g->Ac[i];
    }
  }

  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip], ct = g->ctop[i];
    for (int k = ct + 1; k < g->Nk[i]; k++)
      p->Cn_W[i][k] -= dt * (g->dzz[i][k - 1] * p->stmp[i][k - 1] + g->dzz[i][k] * p->stmp[i][k]) /
                       (g->dzz[i][k - 1] + g->dzz[i][k]);
    p->Cn_W[i][ct] -= dt * p->stmp[i][ct];
  }
  if (scheme == 4)
    for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
      const int i = g->cellp[ip], n = g->Nk[i];
      for (int k = g->ctop[i] + 1; k < n + 1; k++) {
        const double wm = 0.5 * (p->w[i][k - 1] + p->w[i][k]);
        const double Cz = wm * dt / g->dzz[i][k - 1];
        a[k] = wm * (wm - 0.5 * Cz * (p->w[i][k - 1] - p->w[i][k]));
      }
      for (int k = g->ctop[i] + 1; k < n; k++) {
        const double below = k + 1 < n ? g->dzz[i][k + 1] : 0.0;
        p->Cn_W[i][k] -= 2.0 * dt * (a[k] - a[k + 1]) / (g->dzz[i][k] + below);
      }
    }
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip];
    for (int k = g->ctop[i]; k < g->Nk[i]; k++) p->wtmp[i][k] += f1 * p->Cn_W[i][k];
  }

  const double nu = prop->nu, lw = prop->laxWendroff_Vertical, th = prop->theta;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++) {
    const int i = g->cellp[ip], ct = g->ctop[i], n = g->Nk[i];
    REAL *z = g->dzz[i], *w = p->w[i], *rhs = p->wtmp[i];
    if (n - ct <= 1) {
      for (int k = ct; k < n; k++) w[k] = rhs[k];
      continue;
    }
    for (int k = ct + 1; k < n; k++) {
      a[k] = 2 * (nu + lw * p->nu_lax[i][k - 1] + p->nu_tv[i][k - 1]) / z[k - 1] / (z[k] + z[k - 1]);
      b[k] = 2 * (nu + lw * p->nu_lax[i][k] + p->nu_tv[i][k]) / z[k] / (z[k] + z[k - 1]);
    }
    b[ct] = (nu + lw * p->nu_lax[i][ct] + p->nu_tv[i][ct]) / pow(z[ct], 2);
    a[ct] = b[ct];
    for (int k = ct + 1; k < n; k++)
      rhs[k] += dt * (1 - th) * (a[k] * w[k - 1] - (a[k] + b[k]) * w[k] + b[k] * w[k + 1]);
    rhs[ct] += dt * (1 - th) * (-(a[ct] + b[ct]) * w[ct] + (a[ct] + b[ct]) * w[ct + 1]);
    for (int k = ct; k < n; k++) {
      c[k] = 1 + dt * th * (a[k] + b[k]);
      a[k] *= (-dt * th);
      b[k] *= (-dt * th);
    }
    b[ct] += a[ct];
    oracle_TriSolve(a + ct, c + ct, b + ct, rhs + ct, w + ct, n - ct);
  }
}

// sunhip_h.hip -- free-surface PCG on gfx950.
//
// Reference: CGSolve main/phys.c:3086-3217, HCoefficients :3252-3269,
// OperatorH :3335-3349, HPreconditioner :3227-3235, InnerProduct :3280-3294.
// Same two-kernel iteration as the pressure solve, one thread per cell.
#include <hip/hip_runtime.h>

#include <cmath>

#include "sunhip_common.h"

using namespace sunhip;

enum { HM_PLAIN = 0, HM_RESID = 1, HM_FUSED = 2 };

struct HVec {
  const double *coef, *fcoef;   // hcoef [NS], hfcoef [MAXF*NS]
  const double *vin, *src;
  const double *rtmp, *pold;
  double *pnew, *x, *out;
};

// HCoefficients: fcoef = g(theta dt)^2 D df/dg on faces with a neighbour,
// coef = Ac + sum fcoef, added in face order.
__global__ __launch_bounds__(256) void k_hcoef(Geo G, const double *__restrict__ D, double tmp,
                                               double *__restrict__ coef, double *__restrict__ fcoef) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= G.NS || G.nk[s] == 0 || G.comp[s] != 1) return;
  double c = G.Ac[s];
  for (int f = 0; f < G.nf[s]; f++) {
    if (G.nbs[f * G.NS + s] < 0) continue;
    const int e = G.es[f * G.NS + s];
    const double fc = tmp * D[e] * G.df[e] / G.dg[e];
    fcoef[f * G.NS + s] = fc;
    c += fc;
  }
  coef[s] = c;
}

__global__ void k_hzero_comp(Geo G, double *__restrict__ x) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < G.NS && G.nk[s] > 0 && G.comp[s] == 1) x[s] = 0;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_happly(Geo G, HVec V, CGScal *sc, double *part, unsigned *cnt) {
  if (MODE == HM_FUSED && sc->done) return;
  const int s = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  double dot = 0;
  if (s < G.NS && G.nk[s] > 0 && G.comp[s] == 1) {
    const double mu = MODE == HM_FUSED ? sc->mu : 0.0, nu = MODE == HM_FUSED ? sc->nu : 0.0;
    double po0 = 0, v0;
    if (MODE == HM_FUSED) { po0 = V.pold[s]; v0 = V.rtmp[s] + mu * po0; } else v0 = V.vin[s];
    double y = V.coef[s] * v0;
    for (int f = 0; f < G.nf[s]; f++) {
      const int nb = G.nbs[f * G.NS + s];
      if (nb < 0) continue;
      const double vn = MODE == HM_FUSED ? V.rtmp[nb] + mu * V.pold[nb] : V.vin[nb];
      y -= V.fcoef[f * G.NS + s] * vn;
    }
    if (MODE == HM_FUSED) {
      V.pnew[s] = v0;
      V.x[s] = V.x[s] + nu * po0;
      V.out[s] = y;
      dot = v0 * y;
    } else if (MODE == HM_RESID) {
      // p = p - z ; r = p (main/phys.c:3122-3128)
      V.out[s] = V.src[s] - y;
    } else {
      V.out[s] = y;
    }
  }
  if (MODE == HM_FUSED) {
    double v[1] = {dot}, tot[1];
    if (grid_reduce<1, 256>(v, part, cnt, tot) && threadIdx.x == 0) {
      sc->pz = tot[0];
      sc->muinv = 1 / sc->alpha;
      sc->nu = sc->alpha / tot[0];
    }
  }
}

// r -= nu*z ; rtmp = r/hcoef (hprecond=1) or r ; dots; scalars.
__global__ __launch_bounds__(256) void k_hprec(Geo G, const double *__restrict__ coef, const double *__restrict__ z,
                                               double *__restrict__ r, double *__restrict__ rt, CGScal *sc,
                                               double *part, unsigned *cnt, int finish, int update) {
  if (update && sc->done) return;
  const int s = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  double rz = 0, rr = 0;
  const int pc = sc->precond;
  if (s < G.NS && G.nk[s] > 0 && G.comp[s] == 1) {
    double rk = r[s];
    if (update) { rk = rk - sc->nu * z[s]; r[s] = rk; }
    const double u = pc == 1 ? rk / coef[s] : rk;
    rt[s] = u;
    rz = rk * u;
    rr = rk * rk;
  }
  double v[2] = {rz, rr}, tot[2];
  if (grid_reduce<2, 256>(v, part, cnt, tot) && threadIdx.x == 0) {
    if (finish == 1) {
      sc->alpha = sc->alpha0 = pc == 1 ? tot[0] : tot[1];
      if (!sc->resnorm) sc->alpha0 = 1;
      sc->eps = sc->eps0 = pc == 1 ? tot[1] : sc->alpha0;
      sc->mu = 0; sc->nu = 0; sc->iter = 0; sc->nloop = 0; sc->converged = 0;
      sc->done = (sc->eps == 0 || sc->alpha == 0 || sc->maxit <= 0);
      sc->resid = sqrt(sc->eps / sc->eps0);
    } else {
      const double alpha = pc == 1 ? tot[0] : tot[1];
      sc->mu = sc->muinv * alpha;
      sc->alpha = alpha;
      sc->eps = pc == 1 ? tot[1] : alpha;
      sc->iter += 1;
      sc->resid = sqrt(sc->eps / sc->eps0);
      if (sc->resid < sc->tol) { sc->done = 1; sc->converged = 1; sc->nloop = sc->iter - 1; }
      else if (sc->iter >= sc->maxit || sc->eps == 0 || sc->alpha == 0) { sc->done = 1; sc->nloop = sc->iter; }
    }
  }
}

__global__ void k_hfinish(Geo G, double *__restrict__ x, const double *__restrict__ p0,
                          const double *__restrict__ p1, const CGScal *sc) {
  if (sc->iter == 0) return;
  const double *__restrict__ p = (sc->iter & 1) ? p1 : p0;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < G.NS && G.nk[s] > 0 && G.comp[s] == 1) x[s] = x[s] + sc->nu * p[s];
}

__global__ void k_hzero(double *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

namespace {

struct HBufs {
  Field *h, *src, *D, *coef, *r, *z, *rt, *p0, *p1;
  double *fcoef;
};


int hbufs(sunhip_grid *g, HBufs &B) {
  int rc = field(g, "h", &B.h) | field(g, "htmp", &B.src) | field(g, "D", &B.D) | field(g, "hcoef", &B.coef) |
           field(g, "hr", &B.r) | field(g, "hz", &B.z) | field(g, "hrt", &B.rt) | field(g, "hp0", &B.p0) |
           field(g, "hp1", &B.p1);
  if (rc) return rc;
  Field *fc;
  // face coefficients: MAXF rows of NS, kept as a cell field of its own size
  auto it = g->fields.find("#hfcoef");
  if (it == g->fields.end()) {
    Field f;
    f.kind = CELL2;
    f.n = (size_t)SUNHIP_MAXF * g->NS;
    if ((rc = dev_alloc(g, (void **)&f.d, f.n * sizeof(double)))) return rc;
    HIPCHK(hipMemsetAsync(f.d, 0, f.n * sizeof(double), g->stream));
    it = g->fields.emplace("#hfcoef", f).first;
  }
  fc = &it->second;
  B.fcoef = fc->d;
  return 0;
}

int h_coefficients(sunhip_grid *g, HBufs &B, double grav, double theta, double dt) {
  const double td = theta * dt;
  const double tmp = grav * (td * td);  // grav*pow(theta*dt,2)
  k_hcoef<<<(g->NS + 255) / 256, 256, 0, g->stream>>>(g->geo(), B.D->d, tmp, B.coef->d, B.fcoef);
  HIPCHK(hipGetLastError());
  return 0;
}

int h_iteration(sunhip_grid *g, HBufs &B, int it) {
  const Geo G = g->geo();
  HVec V{};
  V.coef = B.coef->d; V.fcoef = B.fcoef; V.rtmp = B.rt->d;
  V.pold = (it & 1) ? B.p1->d : B.p0->d;
  V.pnew = (it & 1) ? B.p0->d : B.p1->d;
  V.x = B.h->d; V.out = B.z->d;
  const int nb = (g->NS + 255) / 256;
  int ps = prof_begin(g, PK_HAPPLY);
  k_happly<HM_FUSED><<<nb, 256, 0, g->stream>>>(G, V, g->d_hs, g->d_part, g->d_count + 8);
  prof_end(g, ps);
  ps = prof_begin(g, PK_HPREC);
  k_hprec<<<nb, 256, 0, g->stream>>>(G, B.coef->d, B.z->d, B.r->d, B.rt->d, g->d_hs, g->d_part, g->d_count + 9, 2, 1);
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int sunhip_h_solve(sunhip_grid *g, const sunhip_cg_params *prm, double grav, double theta, double dt,
                   sunhip_cg_stats *st) {
  HIPCHK(hipSetDevice(g->device));
  if (prm->precond != 0 && prm->precond != 1) {
    set_error("hprecond=%d not supported (0 or 1)", prm->precond);
    return -1;
  }
  HBufs B;
  int rc = hbufs(g, B);
  if (rc) return rc;
  const Geo G = g->geo();
  const int nb = (g->NS + 255) / 256;
  CGScal h{};
  h.tol = prm->epsilon; h.maxit = prm->maxiters; h.precond = prm->precond; h.resnorm = prm->resnorm;
  *g->h_sc = h;
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  HIPCHK(hipMemcpyAsync(g->d_hs, g->h_sc, sizeof h, hipMemcpyHostToDevice, g->stream));
  if ((rc = h_coefficients(g, B, grav, theta, dt))) return rc;
  // x = 0 on computational cells; r = b - L x lifts type-3 boundary values
  k_hzero_comp<<<nb, 256, 0, g->stream>>>(G, B.h->d);
  HVec V{};
  V.coef = B.coef->d; V.fcoef = B.fcoef; V.vin = B.h->d; V.src = B.src->d; V.out = B.r->d;
  k_happly<HM_RESID><<<nb, 256, 0, g->stream>>>(G, V, g->d_hs, g->d_part, g->d_count + 8);
  k_hprec<<<nb, 256, 0, g->stream>>>(G, B.coef->d, nullptr, B.r->d, B.rt->d, g->d_hs, g->d_part, g->d_count + 9, 1, 0);
  k_hzero<<<nb, 256, 0, g->stream>>>(B.p0->d, g->NS);
  k_hzero<<<nb, 256, 0, g->stream>>>(B.p1->d, g->NS);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  int every = prm->check_every > 0 ? prm->check_every : 32;
  if (every & 1) every++;
  HIPCHK(hipMemcpyAsync(g->h_sc, g->d_hs, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  int it = 0;
  while (!g->h_sc->done && it < prm->maxiters) {
    for (int b = 0; b < every && it < prm->maxiters; b++, it++)
      if ((rc = h_iteration(g, B, it))) return rc;
    HIPCHK(hipMemcpyAsync(g->h_sc, g->d_hs, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
  }
  k_hfinish<<<nb, 256, 0, g->stream>>>(G, B.h->d, B.p0->d, B.p1->d, g->d_hs);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(g->h_sc, g->d_hs, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipEventRecord(g->ev2, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  if (st) {
    st->niters = g->h_sc->nloop;
    st->converged = g->h_sc->converged;
    st->resid = g->h_sc->resid;
    st->eps0 = g->h_sc->eps0;
    float ms = 0;
    hipEventElapsedTime(&ms, g->ev1, g->ev2);
    st->device_ms = ms;
    hipEventElapsedTime(&ms, g->ev0, g->ev1);
    st->setup_ms = ms;
  }
  return 0;
}

int sunhip_h_operator(sunhip_grid *g, double grav, double theta, double dt, const char *x, const char *y) {
  HIPCHK(hipSetDevice(g->device));
  HBufs B;
  int rc = hbufs(g, B);
  if (rc) return rc;
  Field *fx, *fy;
  if ((rc = field(g, x, &fx) | field(g, y, &fy))) return rc;
  if ((rc = h_coefficients(g, B, grav, theta, dt))) return rc;
  HVec V{};
  V.coef = B.coef->d; V.fcoef = B.fcoef; V.vin = fx->d; V.out = fy->d;
  k_happly<HM_PLAIN><<<(g->NS + 255) / 256, 256, 0, g->stream>>>(g->geo(), V, g->d_hs, g->d_part, g->d_count + 8);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

// OperatorH with caller-supplied coefficients: "hcoef" (2-D cell field) and
// face coefficients given flat as [Nc*maxfaces] in the reference's indexing.
int sunhip_h_operator_coef(sunhip_grid *g, const double *hfcoef_flat, const char *x, const char *y) {
  HIPCHK(hipSetDevice(g->device));
  HBufs B;
  int rc = hbufs(g, B);
  if (rc) return rc;
  Field *fx, *fy;
  if ((rc = field(g, x, &fx) | field(g, y, &fy))) return rc;
  std::vector<double> fc((size_t)SUNHIP_MAXF * g->NS, 0.0);
  for (int c = 0; c < g->Nc; c++) {
    const int s = g->slot_of_cell[c];
    for (int f = 0; f < g->nfaces[c]; f++) fc[(size_t)f * g->NS + s] = hfcoef_flat[(size_t)c * g->mf + f];
  }
  HIPCHK(hipMemcpyAsync(B.fcoef, fc.data(), fc.size() * sizeof(double), hipMemcpyHostToDevice, g->stream));
  HVec V{};
  V.coef = B.coef->d; V.fcoef = B.fcoef; V.vin = fx->d; V.out = fy->d;
  k_happly<HM_PLAIN><<<(g->NS + 255) / 256, 256, 0, g->stream>>>(g->geo(), V, g->d_hs, g->d_part, g->d_count + 8);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

}  // extern "C"

// sunhip_q.hip -- nonhydrostatic pressure Poisson PCG on gfx950.
//
// Reference: CGSolveQ main/phys.c:2411-2550, QCoefficients :3420-3471,
// OperatorQ :3486-3536, Preconditioner :3577-3619, TriSolve main/util.c:185-203,
// InnerProduct3 main/phys.c:3304-3319.
//
// One PCG iteration is two kernels, one lane per water column:
//   apply:   p = rtmp + mu*p_old ; x += nu_prev*p_old ; z = L p ; sum p.z
//   precond: r -= nu*z ; rtmp = M^-1 r (Thomas down and up the column) ;
//            sum r.rtmp, sum r.r ; mu, alpha, eps and the convergence test
// The scalars live on the device; the last block of each kernel finishes the
// reduction and updates them, so the host only polls a flag every few
// iterations.  Element-wise arithmetic keeps the reference's operation order
// and is built without FMA contraction, so OperatorQ and the preconditioner
// agree with the reference bit for bit; only the summation order of the dot
// products differs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "sunhip_common.h"

using namespace sunhip;

#define QAPPLY_THREADS 256
#define TILES_PER_APPLY (QAPPLY_THREADS / SUNHIP_TILE)

enum { QM_PLAIN = 0, QM_RESID = 1, QM_FUSED = 2 };

struct QVec {
  const double *coef, *dzf, *D;
  const double *vin;            // PLAIN / RESID: operand
  const double *src;            // RESID: right-hand side
  const double *rtmp, *pold;    // FUSED: p = rtmp + mu*pold
  double *pnew, *x;             // FUSED
  double *out;                  // L p, or src - L x
};

// QCoefficients, unpreconditioned branch (main/phys.c:3456-3469), every cell.
__global__ __launch_bounds__(256) void k_qcoef(Geo G, const double *__restrict__ dzz, double *__restrict__ coef) {
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0) return;
  const int kt = G.ctop[s];
  const double Ac = G.Ac[s];
  const int64_t base = G.toff[t] + l;
  double dprev = dzz[base + (int64_t)kt * 64];
  coef[base + (int64_t)kt * 64] = Ac / dprev;
  for (int k = kt + 1; k < n; k++) {
    const double dk = dzz[base + (int64_t)k * 64];
    coef[base + (int64_t)k * 64] = 2 * Ac / (dk + dprev);
    dprev = dk;
  }
}

// y = L v over the computational columns (main/phys.c:3486-3536).
template <int MODE>
__global__ __launch_bounds__(QAPPLY_THREADS) void k_qapply(Geo G, QVec V, CGScal *sc, double *part, unsigned *cnt) {
  if (MODE == QM_FUSED && sc->done) return;
  const int nblk = gridDim.x;
  const int t = xcd_block(blockIdx.x, nblk) * TILES_PER_APPLY + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  double dot = 0;
  const double mu = MODE == QM_FUSED ? sc->mu : 0.0;
  const double nu = MODE == QM_FUSED ? sc->nu : 0.0;
  const double *__restrict__ coef = V.coef;
  const double *__restrict__ dzf = V.dzf;
  const double *__restrict__ rt = V.rtmp;
  const double *__restrict__ po = V.pold;
  const double *__restrict__ vin = V.vin;
  if (t < G.Nt) {
    const int s = t * 64 + l;
    const int n = G.nk[s];
    if (n > 0 && G.comp[s] == 1) {
      const int kt = G.ctop[s];
      const int NS = G.NS;
      const int nf = G.nf[s];
      int nb[SUNHIP_MAXF], eb[SUNHIP_MAXF], kmn[SUNHIP_MAXF], kmx[SUNHIP_MAXF];
      double Df[SUNHIP_MAXF];
#pragma unroll
      for (int f = 0; f < SUNHIP_MAXF; f++) {
        nb[f] = -1; eb[f] = 0; kmn[f] = 0; kmx[f] = 0; Df[f] = 0;
        if (f < nf) {
          nb[f] = G.nbb[f * NS + s];
          if (nb[f] >= 0) {
            eb[f] = G.eb[f * NS + s];
            const int kr = G.kr[f * NS + s];
            kmn[f] = kr & 0xffff;
            kmx[f] = kr >> 16;
            Df[f] = V.D[G.es[f * NS + s]];
          }
        }
      }
      const int64_t base = G.toff[t] + l;
      auto val = [&](int64_t idx) -> double {
        if (MODE == QM_FUSED) return rt[idx] + mu * po[idx];
        return vin[idx];
      };
      double vm1 = 0.0, v0, vp1 = 0.0, c0, cp1 = 0.0, po0 = 0.0, po1 = 0.0;
      {
        const int64_t i0 = base + (int64_t)kt * 64;
        if (MODE == QM_FUSED) { po0 = po[i0]; v0 = rt[i0] + mu * po0; } else v0 = vin[i0];
        c0 = coef[i0];
      }
      for (int k = kt; k < n; k++) {
        const int64_t idx = base + (int64_t)k * 64;
        if (k + 1 < n) {
          if (MODE == QM_FUSED) { po1 = po[idx + 64]; vp1 = rt[idx + 64] + mu * po1; }
          else vp1 = vin[idx + 64];
          cp1 = coef[idx + 64];
        }
        double y = 0;
#pragma unroll
        for (int f = 0; f < SUNHIP_MAXF; f++)
          if (nb[f] >= 0 && k >= kmn[f] && k < kmx[f])
            y += (val(nb[f] + (int64_t)k * 64) - v0) * dzf[eb[f] + (int64_t)k * 64] * Df[f];
        if (kt == n - 1) y -= 2.0 * c0 * v0;
        else if (k == kt) y += (-2 * c0 - cp1) * v0 + cp1 * vp1;
        else if (k == n - 1) y += c0 * vm1 - c0 * v0;
        else y += c0 * vm1 - (c0 + cp1) * v0 + cp1 * vp1;
        if (MODE == QM_FUSED) {
          V.pnew[idx] = v0;
          V.x[idx] = V.x[idx] + nu * po0;
          dot += v0 * y;
          V.out[idx] = y;
        } else if (MODE == QM_RESID) {
          V.out[idx] = V.src[idx] - y;
        } else {
          V.out[idx] = y;
        }
        vm1 = v0; v0 = vp1; c0 = cp1; po0 = po1;
      }
    }
  }
  if (MODE == QM_FUSED) {
    double v[1] = {dot}, tot[1];
    if (grid_reduce<1, QAPPLY_THREADS>(v, part, cnt, tot) && threadIdx.x == 0) {
      sc->pz = tot[0];
      sc->muinv = 1 / sc->alpha;          // mu = 1/alpha  (main/phys.c:2484)
      sc->nu = sc->alpha / tot[0];        // nu = alpha/(p.z) (:2485)
    }
  }
}

// Thomas solve of one column's vertical operator (main/phys.c:3577-3619 with
// main/util.c:185-203), keeping the reference's operation order.  The forward
// sweep's modified diagonal b' and right-hand side d' of all levels stay in
// registers (the loops are fully unrolled over DC levels); the back
// substitution re-reads coef (for c) and r (for r.rtmp).
//   UPDATE: r -= nu*z first (main/phys.c:2490-2493), written back.
template <int DC, bool UPDATE>
__global__ __launch_bounds__(64) void k_qprec(Geo G, const double *__restrict__ coef,
                                              const double *__restrict__ z, double *__restrict__ r,
                                              double *__restrict__ out, const int *__restrict__ tiles,
                                              int ntiles, CGScal *sc, double *part, unsigned *cnt, int finish) {
  if (UPDATE && sc->done) return;
  const int bt = xcd_block(blockIdx.x, gridDim.x);
  const int t = tiles ? tiles[bt] : bt;
  const int l = threadIdx.x;
  const int s = t * 64 + l;
  const int n = G.nk[s];
  const int kt = G.ctop[s];
  const bool act = n > 0 && G.comp[s] == 1;
  const double nu = UPDATE ? sc->nu : 0.0;
  const int64_t base = G.toff[t] + l;
  double rz = 0, rr = 0;
  double yv[DC], bv[DC];
  // the tile's wet levels
  const int k0 = 0, kend = G.tdep[t] - 1;
  (void)ntiles;
  double cprev = 0, bprev = 1, dprev = 0;
  if (act && kt == n - 1) {
    const int64_t idx = base + (int64_t)kt * 64;
    double rk = r[idx];
    if (UPDATE) { rk = rk - nu * z[idx]; r[idx] = rk; }
    const double u = -0.5 * rk / coef[idx];
    out[idx] = u;
    rz += rk * u;
    rr += rk * rk;
  }
  const bool multi = act && kt < n - 1;
#pragma unroll
  for (int j = 0; j < DC; j++) {
    const int k = k0 + j;
    if (k < kend && multi && k >= kt && k < n) {
      const int64_t idx = base + (int64_t)k * 64;
      double rk = r[idx];
      if (UPDATE) { rk = rk - nu * z[idx]; r[idx] = rk; }
      rr += rk * rk;
      const double ck = coef[idx];
      double a = 0, b, c = 0;
      if (k == kt) {
        const double ck1 = coef[idx + 64];
        b = -2 * ck - ck1;
        c = ck1;
      } else if (k == n - 1) {
        a = ck;
        b = -ck;
      } else {
        const double ck1 = coef[idx + 64];
        a = ck;
        b = -ck - ck1;
        c = ck1;
      }
      double d = rk;
      if (k > kt) {
        b -= a * cprev / bprev;
        d -= a * dprev / bprev;
      }
      yv[j] = d;
      bv[j] = b;
      cprev = c; bprev = b; dprev = d;
    }
  }
  double unext = 0;
#pragma unroll
  for (int j = DC - 1; j >= 0; j--) {
    const int k = k0 + j;
    if (k < kend && multi && k >= kt && k < n) {
      const int64_t idx = base + (int64_t)k * 64;
      double u;
      if (k == n - 1) u = yv[j] / bv[j];
      else u = yv[j] / bv[j] - coef[idx + 64] * unext / bv[j];
      out[idx] = u;
      rz += r[idx] * u;
      unext = u;
    }
  }
  if (finish) {
    double v[2] = {rz, rr}, tot[2];
    if (grid_reduce<2, 64>(v, part, cnt, tot) && threadIdx.x == 0) {
      if (finish == 1) {
        // initial residual (main/phys.c:2452-2475)
        sc->alpha = sc->alpha0 = tot[0];
        if (!sc->resnorm) sc->alpha0 = 1;
        sc->eps = sc->eps0 = tot[1];
        sc->mu = 0; sc->nu = 0; sc->iter = 0; sc->nloop = 0; sc->converged = 0;
        sc->done = (sc->eps == 0 || sc->maxit <= 0);
        sc->resid = sqrt(sc->eps / sc->eps0);
      } else {
        // end of one iteration (main/phys.c:2496-2524)
        const double alpha = tot[0];
        sc->mu = sc->muinv * alpha;
        sc->alpha = alpha;
        sc->eps = tot[1];
        sc->iter += 1;
        sc->resid = sqrt(sc->eps / sc->eps0);
        if (sc->resid < sc->tol) { sc->done = 1; sc->converged = 1; sc->nloop = sc->iter - 1; }
        else if (sc->iter >= sc->maxit || sc->eps == 0) { sc->done = 1; sc->nloop = sc->iter; }
      }
    }
  }
}

// qprecond=0: r -= nu*z, rtmp = r, alpha = eps = r.r (main/phys.c:2505-2519).
__global__ __launch_bounds__(256) void k_qnoprec(Geo G, const double *__restrict__ z, double *__restrict__ r,
                                                 double *__restrict__ out, CGScal *sc, double *part,
                                                 unsigned *cnt, int finish, int update) {
  if (update && sc->done) return;
  const int t = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  double rr = 0;
  if (t < G.Nt) {
    const int s = t * 64 + l, n = G.nk[s];
    if (n > 0 && G.comp[s] == 1) {
      const double nu = update ? sc->nu : 0.0;
      const int64_t base = G.toff[t] + l;
      for (int k = G.ctop[s]; k < n; k++) {
        const int64_t idx = base + (int64_t)k * 64;
        double rk = r[idx];
        if (update) { rk = rk - nu * z[idx]; r[idx] = rk; }
        out[idx] = rk;
        rr += rk * rk;
      }
    }
  }
  double v[1] = {rr}, tot[1];
  if (grid_reduce<1, 256>(v, part, cnt, tot) && threadIdx.x == 0) {
    if (finish == 1) {
      sc->alpha = sc->alpha0 = tot[0];
      if (!sc->resnorm) sc->alpha0 = 1;
      sc->eps = sc->eps0 = sc->alpha0;
      sc->mu = 0; sc->nu = 0; sc->iter = 0; sc->nloop = 0; sc->converged = 0;
      sc->done = (sc->eps == 0 || sc->maxit <= 0);
      sc->resid = sqrt(sc->eps / sc->eps0);
    } else {
      const double alpha = tot[0];
      sc->mu = sc->muinv * alpha;
      sc->alpha = alpha;
      sc->eps = alpha;
      sc->iter += 1;
      sc->resid = sqrt(sc->eps / sc->eps0);
      if (sc->resid < sc->tol) { sc->done = 1; sc->converged = 1; sc->nloop = sc->iter - 1; }
      else if (sc->iter >= sc->maxit || sc->eps == 0) { sc->done = 1; sc->nloop = sc->iter; }
    }
  }
}

// x += nu*p for the last iteration (its x update is otherwise folded into the
// next apply).  p is the buffer the last apply wrote.
__global__ __launch_bounds__(256) void k_qfinish(Geo G, double *__restrict__ x, const double *__restrict__ p0,
                                                 const double *__restrict__ p1, const CGScal *sc) {
  if (sc->iter == 0) return;
  const double *__restrict__ p = (sc->iter & 1) ? p1 : p0;
  const double nu = sc->nu;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return;
  const int64_t base = G.toff[t] + l;
  for (int k = G.ctop[s]; k < n; k++) {
    const int64_t idx = base + (int64_t)k * 64;
    x[idx] = x[idx] + nu * p[idx];
  }
}

__global__ void k_zero(double *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// host side

namespace {

struct QBufs {
  Field *q, *src, *dzz, *dzf, *D, *coef, *r, *z, *rtmp, *p0, *p1;
};

int qbufs(sunhip_grid *g, QBufs &B) {
  int rc = 0;
  rc |= field(g, "q", &B.q);
  rc |= field(g, "src", &B.src);
  rc |= field(g, "dzz", &B.dzz);
  rc |= field(g, "dzf", &B.dzf);
  rc |= field(g, "D", &B.D);
  rc |= field(g, "coef", &B.coef);
  rc |= field(g, "r", &B.r);
  rc |= field(g, "z", &B.z);
  rc |= field(g, "rtmp", &B.rtmp);
  rc |= field(g, "p0", &B.p0);
  rc |= field(g, "p1", &B.p1);
  return rc;
}

// depth class of the register-resident Thomas kernel
int depth_class(int d) {
  static const int cls[] = {8, 16, 24, 32, 48, 64, 80, 96, 112, 128};
  for (int c : cls) if (d <= c) return c;
  return -1;
}

template <bool UPDATE>
int launch_prec(sunhip_grid *g, int dc, const double *coef, const double *z, double *r, double *out,
                CGScal *sc, int finish) {
  const Geo G = g->geo();
  const int nb = g->Nt;
  unsigned *cnt = g->d_count + 1;
#define PREC_CASE(D)                                                                                    \
  case D:                                                                                               \
    k_qprec<D, UPDATE><<<nb, 64, 0, g->stream>>>(G, coef, z, r, out, nullptr, nb, sc, g->d_part, cnt, finish); \
    break;
  switch (dc) {
    PREC_CASE(8) PREC_CASE(16) PREC_CASE(24) PREC_CASE(32) PREC_CASE(48) PREC_CASE(64) PREC_CASE(80)
    PREC_CASE(96) PREC_CASE(112) PREC_CASE(128)
    default:
      set_error("columns deeper than 128 layers are not supported by the preconditioner");
      return -1;
  }
#undef PREC_CASE
  HIPCHK(hipGetLastError());
  return 0;
}

int q_setup(sunhip_grid *g, const sunhip_cg_params *prm, QBufs &B, int &dc) {
  if (!g->have_extents) {
    sunhip::set_error("sunhip_q_*: call sunhip_set_extents first");
    return -1;
  }
  if (prm->precond != 0 && prm->precond != 2) {
    sunhip::set_error("qprecond=%d is not supported on the device (0 or 2)", prm->precond);
    return -1;
  }
  int rc = qbufs(g, B);
  if (rc) return rc;
  dc = depth_class(g->maxdepth);
  if (dc < 0) { set_error("column depth %d exceeds 128", g->maxdepth); return -1; }
  CGScal h{};
  h.tol = prm->epsilon;
  h.maxit = prm->maxiters;
  h.precond = prm->precond;
  h.resnorm = prm->resnorm;
  *g->h_sc = h;
  HIPCHK(hipMemcpyAsync(g->d_qs, g->h_sc, sizeof h, hipMemcpyHostToDevice, g->stream));
  const Geo G = g->geo();
  const int nb4 = (g->Nt + 3) / 4;
  k_qcoef<<<nb4, 256, 0, g->stream>>>(G, B.dzz->d, B.coef->d);
  // r = src - L x  (main/phys.c:2444-2451)
  QVec V{};
  V.coef = B.coef->d; V.dzf = B.dzf->d; V.D = B.D->d; V.vin = B.q->d; V.src = B.src->d; V.out = B.r->d;
  const int nba = (g->Nt + TILES_PER_APPLY - 1) / TILES_PER_APPLY;
  k_qapply<QM_RESID><<<nba, QAPPLY_THREADS, 0, g->stream>>>(G, V, g->d_qs, g->d_part, g->d_count);
  HIPCHK(hipGetLastError());
  // rtmp = M^-1 r, alpha0, eps0
  if (prm->precond == 2) {
    rc = launch_prec<false>(g, dc, B.coef->d, nullptr, B.r->d, B.rtmp->d, g->d_qs, 1);
  } else {
    k_qnoprec<<<nb4, 256, 0, g->stream>>>(G, nullptr, B.r->d, B.rtmp->d, g->d_qs, g->d_part, g->d_count + 2, 1, 0);
    HIPCHK(hipGetLastError());
  }
  if (rc) return rc;
  // p_old = 0 so the first apply reads p = rtmp exactly
  const int64_t n = g->NT;
  k_zero<<<(unsigned)((n + 255) / 256), 256, 0, g->stream>>>(B.p0->d, n);
  k_zero<<<(unsigned)((n + 255) / 256), 256, 0, g->stream>>>(B.p1->d, n);
  HIPCHK(hipGetLastError());
  return 0;
}

// One PCG iteration; it = iteration index (selects the p buffers).
int q_iteration(sunhip_grid *g, const sunhip_cg_params *prm, QBufs &B, int dc, int it) {
  const Geo G = g->geo();
  QVec V{};
  V.coef = B.coef->d; V.dzf = B.dzf->d; V.D = B.D->d;
  V.rtmp = B.rtmp->d;
  V.pold = (it & 1) ? B.p1->d : B.p0->d;
  V.pnew = (it & 1) ? B.p0->d : B.p1->d;
  V.x = B.q->d;
  V.out = B.z->d;
  const int nba = (g->Nt + TILES_PER_APPLY - 1) / TILES_PER_APPLY;
  int ps = prof_begin(g, PK_QAPPLY);
  k_qapply<QM_FUSED><<<nba, QAPPLY_THREADS, 0, g->stream>>>(G, V, g->d_qs, g->d_part, g->d_count);
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  ps = prof_begin(g, PK_QPREC);
  int rc = 0;
  if (prm->precond == 2) {
    rc = launch_prec<true>(g, dc, B.coef->d, B.z->d, B.r->d, B.rtmp->d, g->d_qs, 2);
  } else {
    k_qnoprec<<<(g->Nt + 3) / 4, 256, 0, g->stream>>>(G, B.z->d, B.r->d, B.rtmp->d, g->d_qs, g->d_part,
                                                       g->d_count + 2, 2, 1);
  }
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  return rc;
}

int q_finish(sunhip_grid *g, QBufs &B, sunhip_cg_stats *st) {
  const Geo G = g->geo();
  k_qfinish<<<(g->Nt + 3) / 4, 256, 0, g->stream>>>(G, B.q->d, B.p0->d, B.p1->d, g->d_qs);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(g->h_sc, g->d_qs, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
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

}  // namespace

extern "C" {

int sunhip_q_solve(sunhip_grid *g, const sunhip_cg_params *prm, sunhip_cg_stats *st) {
  HIPCHK(hipSetDevice(g->device));
  QBufs B;
  int dc, rc;
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  if ((rc = q_setup(g, prm, B, dc))) return rc;
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  int every = prm->check_every > 0 ? prm->check_every : (g->NT > (1 << 22) ? 2 : 16);
  if (every & 1) every++;
  int it = 0;
  // the host checks the device flag after every batch; kernels of a batch
  // that start after convergence return at once
  HIPCHK(hipMemcpyAsync(g->h_sc, g->d_qs, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  while (!g->h_sc->done && it < prm->maxiters) {
    for (int b = 0; b < every && it < prm->maxiters; b++, it++)
      if ((rc = q_iteration(g, prm, B, dc, it))) return rc;
    HIPCHK(hipMemcpyAsync(g->h_sc, g->d_qs, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
  }
  return q_finish(g, B, st);
}

int sunhip_q_solve_fixed(sunhip_grid *g, const sunhip_cg_params *prm, int n, sunhip_cg_stats *st) {
  HIPCHK(hipSetDevice(g->device));
  sunhip_cg_params p = *prm;
  p.epsilon = 0;       // never satisfied
  p.maxiters = n;
  QBufs B;
  int dc, rc;
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  if ((rc = q_setup(g, &p, B, dc))) return rc;
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  for (int it = 0; it < n; it++)
    if ((rc = q_iteration(g, &p, B, dc, it))) return rc;
  return q_finish(g, B, st);
}

int sunhip_q_operator(sunhip_grid *g, const char *x, const char *y) {
  HIPCHK(hipSetDevice(g->device));
  if (!g->have_extents) { set_error("sunhip_q_operator: call sunhip_set_extents first"); return -1; }
  Field *fx, *fy, *coef, *dzf, *D;
  int rc = field(g, x, &fx) | field(g, y, &fy) | field(g, "coef", &coef) | field(g, "dzf", &dzf) |
           field(g, "D", &D);
  if (rc) return rc;
  QVec V{};
  V.coef = coef->d; V.dzf = dzf->d; V.D = D->d; V.vin = fx->d; V.out = fy->d;
  const int nba = (g->Nt + TILES_PER_APPLY - 1) / TILES_PER_APPLY;
  k_qapply<QM_PLAIN><<<nba, QAPPLY_THREADS, 0, g->stream>>>(g->geo(), V, g->d_qs, g->d_part, g->d_count);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

// Coefficients from the device field "dzz" into "coef".
int sunhip_q_coefficients(sunhip_grid *g) {
  HIPCHK(hipSetDevice(g->device));
  Field *dzz, *coef;
  int rc = field(g, "dzz", &dzz) | field(g, "coef", &coef);
  if (rc) return rc;
  k_qcoef<<<(g->Nt + 3) / 4, 256, 0, g->stream>>>(g->geo(), dzz->d, coef->d);
  HIPCHK(hipGetLastError());
  return 0;
}

int sunhip_q_precondition(sunhip_grid *g, const char *x, const char *xc) {
  HIPCHK(hipSetDevice(g->device));
  if (!g->have_extents) { set_error("sunhip_q_precondition: call sunhip_set_extents first"); return -1; }
  Field *fx, *fxc, *coef;
  int rc = field(g, x, &fx) | field(g, xc, &fxc) | field(g, "coef", &coef);
  if (rc) return rc;
  const int dc = depth_class(g->maxdepth);
  if (dc < 0) { set_error("column depth %d exceeds 128", g->maxdepth); return -1; }
  if ((rc = launch_prec<false>(g, dc, coef->d, nullptr, fx->d, fxc->d, g->d_qs, 0))) return rc;
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

}  // extern "C"

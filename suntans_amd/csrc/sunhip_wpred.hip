// sunhip_wpred.hip -- the vertical-momentum predictor (WPredictor,
// main/phys.c:2025-2290) on gfx950, with the face values of the
// cell-centred vertical velocity it advects (GetMomentumFaceValues,
// main/phys.c:4879-4992, and UFaceFlux :4265-4289).
//
// Four launches:
//   centres (lane per column): wc = (w[k] + w[k+1]) / 2 on every column
//   faces   (lane per edge):   face values of wc into "ue" (the reference's
//                              phys->ut) by the advection scheme; for scheme 5
//                              the TVD face values of HorizontalFaceScalars
//                              run first
//   source  (lane per column): the cell source stmp, gathered where the
//                              reference scatters: each column visits its
//                              faces in the edgep order of their edges, so
//                              every stmp value receives its terms in the
//                              reference's order
//   column  (lane per column): Adams-Bashforth update, pressure gradient,
//                              Cn_W, and the implicit vertical diffusion
//                              (TriSolve, main/util.c:185-203) eliminated on
//                              the way down and substituted on the way up
// Every expression keeps the reference's operation order; results are the
// reference's bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>

#include "sunhip_solvers.h"
#include "sunhip_tvdtile.h"

using namespace sunhip;

namespace sunhip {

struct WPArgs {
  // cell fields
  double *w, *wtmp, *wc, *stmp, *Cn, *Cn2, *sb, *sc;  // sb, sc: per-level scratch of the vertical solve
  const double *q, *dzz, *nu_tv, *nu_lax;
  // edge fields
  double *uf;  // face values (the reference's phys->ut)
  const double *U, *bw, *SfHp, *SfHm;
  const double *def1;  // edge 2-D
  const int *ford;
  double dt, theta, nu, nu_H, CdW, lax, f1, f2, f3;
  int scheme, first;
};

__device__ __forceinline__ double wmax(double x1, double x2) { return x2 > x1 ? x2 : x1; }  // main/util.c:212-217
__device__ __forceinline__ double wupwind(double u, double a, double b) {                    // main/util.c:219-232
  if (fabs(u) > 0) return u > 0 ? b : a;
  return wmax(a, b);
}

}  // namespace sunhip

// wc on every column from its top cell down (main/phys.c:2096-2100).
__global__ __launch_bounds__(256) void k_wp_centres(Geo G, WPArgs A) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0) return;
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  double wk = at32(A.w, base + (uint32_t)G.ctop[s] * 64u);
  for (int k = G.ctop[s]; k < n; k++) {
    const double wn = at32(A.w, base + (uint32_t)(k + 1) * 64u);
    at32(A.wc, base + (uint32_t)k * 64u) = 0.5 * (wk + wn);
    wk = wn;
  }
}

// Face values of one edge slot (main/phys.c:4884-4990).  Faces between a
// type-3 cell and its neighbour are upwind whatever the scheme (the
// reference's last loop overwrites them); computational edges take the
// scheme's value below the deeper top; type-2 edges the boundary row on
// inflow.  Type-4 edges are refused by the host.  Other edges keep theirs.
__global__ __launch_bounds__(256) void k_wp_faces(Geo G, WPArgs A) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (t >= G.Net) return;
  const int s = t * 64 + l;
  if (G.e_nkc[s] == 0) return;  // padding
  const uint32_t base = (uint32_t)(G.eoff[t] + l);
  const int s1 = G.e_s1[s], s2 = G.e_s2[s], nke = G.e_nke[s], comp = G.e_comp[s];
  const uint32_t c1 = (uint32_t)G.e_c1[s], c2 = (uint32_t)G.e_c2[s];
  auto E = [&](const double *p, int k) { return at32(p, base + (uint32_t)k * 64u); };
  auto F = [&](int k) -> double & { return at32(A.uf, base + (uint32_t)k * 64u); };
  auto W1 = [&](int k) { return at32(A.wc, c1 + (uint32_t)k * 64u); };
  auto W2 = [&](int k) { return at32(A.wc, c2 + (uint32_t)k * 64u); };
  const bool type3 = s1 >= 0 && s2 >= 0 && (G.comp[s1] == 2 || G.comp[s2] == 2);
  if (type3 || comp == 1) {
    const int kmin = G.ctop[s1] > G.ctop[s2] ? G.ctop[s1] : G.ctop[s2];
    for (int k = 0; k < kmin; k++) F(k) = 0;
    const int scheme = type3 ? 1 : A.scheme;
    double d1 = 0, d2 = 0;
    if (scheme == 2) { d1 = A.def1[s]; d2 = G.dg[s] - d1; }
    for (int k = kmin; k < nke; k++) {
      const double U = E(A.U, k);
      double v;
      if (scheme == 2) {
        const double C = 0;  // the Lax-Wendroff term of scheme 4, not built here
        if (d1 == 0 || d2 == 0) v = wupwind(U, W1(k), W2(k));
        else v = (W1(k) * d2 + W2(k) * d1) / (d1 + d2) - C / 2 * (W1(k) - W2(k));
      } else if (scheme == 5) {
        v = U > 0 ? E(A.SfHp, k) : E(A.SfHm, k);
      } else {
        v = U > 0 ? W2(k) : W1(k);
      }
      F(k) = v;
    }
  } else if (comp == 2) {
    for (int k = G.e_etop[s]; k < nke; k++) F(k) = E(A.U, k) > 0 ? E(A.bw, k) : W1(k);
  }
}

// The cell source stmp of one column (main/phys.c:2093-2185): zero, then
// on computational columns the horizontal advection through every face
// (face order) and the vertical advection of w^2, then the horizontal
// diffusion and sidewall drag of each computational edge in edgep order.
template <int NF>
__global__ __launch_bounds__(256) void k_wp_source(Geo G, WPArgs A) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0) return;
  const int NS = G.NS, nf = G.nf[s], ct = G.ctop[s];
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  const double Ac = G.Ac[s];
  const bool adv = A.scheme != 0 && G.comp[s] == 1;
  const bool vadv = adv && (A.scheme == 1 || A.scheme == 2 || A.scheme == 5);
  // faces in the reference's face order (advection)
  uint32_t eb[NF];
  double dfn[NF], df[NF];
  // faces in edgep order (diffusion, drag): the edge's cells, ranges and metrics
  uint32_t oc1[NF], oc2[NF];
  int okmin[NF], onke[NF];
  bool ofirst[NF], ocomp[NF];
  double odf[NF], odg[NF], oAc1[NF], oAc2[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    eb[f] = 0; dfn[f] = 0; df[f] = 0;
    oc1[f] = oc2[f] = 0; okmin[f] = onke[f] = 0; ofirst[f] = ocomp[f] = false;
    odf[f] = odg[f] = oAc1[f] = oAc2[f] = 0;
    if (f < nf) {
      const int es = G.es[f * NS + s];
      eb[f] = (uint32_t)G.eb[f * NS + s];
      df[f] = G.df[es];
      dfn[f] = (double)G.nrm[f * NS + s];
      const int fo = A.ford[f * NS + s];
      const int eo = G.es[fo * NS + s];
      ocomp[f] = G.e_comp[eo] == 1;
      if (ocomp[f]) {
        const int e1 = G.e_s1[eo], e2 = G.e_s2[eo];
        oc1[f] = (uint32_t)G.e_c1[eo];
        oc2[f] = (uint32_t)G.e_c2[eo];
        okmin[f] = G.ctop[e1] > G.ctop[e2] ? G.ctop[e1] : G.ctop[e2];
        onke[f] = G.e_nke[eo];
        ofirst[f] = e1 == s;
        odf[f] = G.df[eo];
        odg[f] = G.dg[eo];
        oAc1[f] = G.Ac[e1];
        oAc2[f] = G.Ac[e2];
      }
    }
  }
  auto Wd = [&](uint32_t b, int k) { return at32(A.w, b + (uint32_t)k * 64u); };
  // w of this column at k and k+1, and of each face's other column at k and
  // k+1 over the face's levels, slide down the column in registers
  double wo0 = Wd(base, 0), wo1 = Wd(base, 1);
  double wn0[NF], wn1[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) { wn0[f] = 0; wn1[f] = 0; }
  for (int k = 0; k < n; k++) {
    double st = 0;
    const uint32_t kk = (uint32_t)k * 64u;
    if (adv && k >= ct) {
#pragma unroll
      for (int f = 0; f < NF; f++)
        if (f < nf) st += at32(A.uf, eb[f] + kk) * at32(A.U, eb[f] + kk) * df[f] * dfn[f] / (1.0 * Ac);
      if (vadv) st += (wo0 * wo0 - wo1 * wo1) / at32(A.dzz, base + kk);
    }
    const double nlo = at32(A.nu_lax, base + kk);
#pragma unroll
    for (int f = 0; f < NF; f++) {
      if (f >= nf || !ocomp[f]) continue;
      if (k >= okmin[f] && k < onke[f]) {
        const uint32_t other = ofirst[f] ? oc2[f] : oc1[f];
        if (k == okmin[f]) wn0[f] = Wd(other, k);
        else wn0[f] = wn1[f];
        wn1[f] = Wd(other, k + 1);
        const double nln = at32(A.nu_lax, other + kk);
        // the edge's first cell (grad[2j]) is c1
        const double w1k = ofirst[f] ? wo0 : wn0[f], w1k1 = ofirst[f] ? wo1 : wn1[f];
        const double w2k = ofirst[f] ? wn0[f] : wo0, w2k1 = ofirst[f] ? wn1[f] : wo1;
        const double l1 = ofirst[f] ? nlo : nln, l2 = ofirst[f] ? nln : nlo;
        const double a = .5 * (A.nu_H + 0.5 * (l1 + l2)) * (w2k - w1k + w2k1 - w1k1) * odf[f] / odg[f];
        if (ofirst[f]) st -= a / oAc1[f];
        else st += a / oAc2[f];
      } else if (k >= onke[f]) {
        const double sw = wo0 + wo1;
        st += 0.25 * A.CdW * fabs(sw) * sw * odf[f] / Ac;
      }
    }
    at32(A.stmp, base + kk) = st;
    wo0 = wo1;
    wo1 = k + 2 <= n ? Wd(base, k + 2) : 0.0;
  }
}

// The rest of one column (main/phys.c:2043-2090, 2211-2290): on the first
// step Cn_W and Cn_W2 start from zero on every column; on computational
// columns wtmp = w + f2 Cn_W + f3 Cn_W2 less the pressure gradient,
// Cn_W2 <- Cn_W, Cn_W <- -dt stmp averaged to the faces, wtmp += f1 Cn_W,
// then the theta-implicit vertical diffusion: its explicit part into wtmp,
// elimination top-down (wtmp keeps the eliminated right-hand side, as
// TriSolve leaves it) and substitution bottom-up into w.
__global__ __launch_bounds__(256) void k_wp_column(Geo G, WPArgs A) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0) return;
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  auto X = [&](double *p, int k) -> double & { return at32(p, base + (uint32_t)k * 64u); };
  auto R = [&](const double *p, int k) { return at32(p, base + (uint32_t)k * 64u); };
  if (A.first)
    for (int k = 0; k < n; k++) { X(A.Cn, k) = 0; X(A.Cn2, k) = 0; }
  if (G.comp[s] != 1) return;
  const int ct = G.ctop[s];
  const double dt = A.dt, th = A.theta, nu = A.nu, lw = A.lax;
  const bool multi = n - ct > 1;
  double zp = 0, sp = 0;       // dzz and stmp of the level above
  double cp = 0, bp = 0, dp = 0;  // eliminated diagonal, scaled upper diagonal and right-hand side above
  // w at k-1, k, k+1, q and the viscosity of the level above slide down the
  // column in registers: each is loaded once (the values and their order of
  // operations are those of loading them at every use)
  double wm = 0, w0 = R(A.w, ct), qp = 0, visp = 0;
  // the next level's loads are issued at the top of each level, before its
  // stores, so they are in flight while it computes
  double zN = R(A.dzz, ct), stN = R(A.stmp, ct), cnN = R(A.Cn, ct), c2N = R(A.Cn2, ct), qN = R(A.q, ct);
  double w1N = R(A.w, ct + 1), nlN = R(A.nu_lax, ct), ntN = R(A.nu_tv, ct);  // w-type rows hold n + 1 levels
  for (int k = ct; k < n; k++) {
    const double z = zN, st = stN, cn = cnN, c2 = c2N, qk = qN, w1 = w1N, nlk = nlN, ntk = ntN;
    if (k + 1 < n) {
      zN = R(A.dzz, k + 1); stN = R(A.stmp, k + 1); cnN = R(A.Cn, k + 1); c2N = R(A.Cn2, k + 1);
      qN = R(A.q, k + 1); w1N = R(A.w, k + 2); nlN = R(A.nu_lax, k + 1); ntN = R(A.nu_tv, k + 1);
    }
    double wt = w0 + A.f2 * cn + A.f3 * c2;
    X(A.Cn2, k) = cn;
    double c = 0;
    if (k > ct) wt -= 2.0 * dt / (zp + z) * (qp - qk);
    else wt += 2.0 * dt / z * qk;
    if (k > ct) c -= dt * (zp * sp + z * st) / (zp + z);
    else c -= dt * st;
    X(A.Cn, k) = c;
    wt += A.f1 * c;
    if (multi) {
      double a, b;
      const double vis = nu + lw * nlk + ntk;
      if (k > ct) {
        a = 2 * visp / zp / (z + zp);
        b = 2 * vis / z / (z + zp);
        wt += dt * (1 - th) * (a * wm - (a + b) * w0 + b * w1);
      } else {
        b = vis / (z * z);
        a = b;
        wt += dt * (1 - th) * (-(a + b) * w0 + (a + b) * w1);
      }
      visp = vis;
      double cd = 1 + dt * th * (a + b);
      const double as = a * (-dt * th);
      double bs = b * (-dt * th);
      if (k == ct) bs += as;
      if (k > ct) {
        cd -= as * bp / cp;
        wt -= as * dp / cp;
      }
      X(A.sc, k) = cd;
      X(A.sb, k) = bs;
      cp = cd; bp = bs; dp = wt;
    }
    X(A.wtmp, k) = wt;
    zp = z; sp = st;
    qp = qk; wm = w0; w0 = w1;
  }
  if (!multi) {
    for (int k = ct; k < n; k++) X(A.w, k) = R(A.wtmp, k);
    return;
  }
  double wn = R(A.wtmp, n - 1) / R(A.sc, n - 1);
  X(A.w, n - 1) = wn;
  for (int k = n - 2; k >= ct; k--) {
    const double cd = R(A.sc, k);
    wn = R(A.wtmp, k) / cd - R(A.sb, k) * wn / cd;
    X(A.w, k) = wn;
  }
}

namespace sunhip {

static int wpred_args(sunhip_grid *g, const sunhip_wpred_params *p, WPArgs *out) {
  if (!g->have_extents) { set_error("WPredictor: call sunhip_set_extents first"); return -1; }
  if (!(p->dt > 0)) { set_error("WPredictor: dt=%g must be positive", p->dt); return -1; }
  if (p->nonlinear != 0 && p->nonlinear != 1 && p->nonlinear != 2 && p->nonlinear != 5) {
    set_error("WPredictor: nonlinear=%d is not built for the device (0, 1, 2 and 5 are; scheme 4's vertical "
              "Lax-Wendroff term reads dzz one past the column bottom, main/phys.c:2230-2233)", p->nonlinear);
    return -1;
  }
  if (p->conserveMomentum) {
    set_error("WPredictor: conserveMomentum is not built for the device (its top-cell fill reads scratch "
              "left from the previous column, main/phys.c:2136-2138)");
    return -1;
  }
  if (g->n_type4) {
    set_error("WPredictor: type-4 (no-slip) edges are not built for the device (the reference's type-4 term "
              "runs to Nke of a cell index, main/phys.c:2199)");
    return -1;
  }
  if (!g->type2_ok) { set_error("WPredictor: a type-2 edge has no cell on its grad[2j] side"); return -1; }
  if (p->nonlinear == 2 && !g->have_def1) {
    set_error("WPredictor: central differencing needs the face distances (sunhip_set_face_distances)");
    return -1;
  }
  struct { const char *name; double **ptr; } f[] = {
      {"w", nullptr}, {"wtmp", nullptr}, {"wc", nullptr}, {"stmp", nullptr}, {"Cn_W", nullptr},
      {"Cn_W2", nullptr}, {"wp_b", nullptr}, {"wp_c", nullptr}, {"qp", nullptr}, {"dzz", nullptr},
      {"nu_tv", nullptr}, {"nu_lax", nullptr}, {"ue", nullptr}, {"utmp2", nullptr}, {"bw", nullptr},
      {"SfHp", nullptr}, {"SfHm", nullptr}, {"def1", nullptr}};
  WPArgs A{};
  double **dst[] = {&A.w, &A.wtmp, &A.wc, &A.stmp, &A.Cn, &A.Cn2, &A.sb, &A.sc, (double **)&A.q,
                    (double **)&A.dzz, (double **)&A.nu_tv, (double **)&A.nu_lax, &A.uf, (double **)&A.U,
                    (double **)&A.bw, (double **)&A.SfHp, (double **)&A.SfHm, (double **)&A.def1};
  for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); i++) {
    Field *fp;
    int rc = field(g, f[i].name, &fp);
    if (rc) return rc;
    *dst[i] = fp->d;
  }
  A.ford = g->d_ford;
  A.dt = p->dt; A.theta = p->theta; A.nu = p->nu; A.nu_H = p->nu_H; A.CdW = p->CdW; A.lax = p->lax_vertical;
  A.scheme = p->nonlinear;
  A.first = p->step == 1;
  // Adams-Bashforth weights (main/phys.c:2043-2066)
  if (p->step == 1) { A.f1 = 1; A.f2 = A.f3 = 0; }
  else if (p->step == 2 || p->AB == 2) { A.f1 = 3.0 / 2.0; A.f2 = -1.0 / 2.0; A.f3 = 0; }
  else { A.f1 = 23.0 / 12.0; A.f2 = -4.0 / 3.0; A.f3 = 5.0 / 12.0; }
  *out = A;
  return 0;
}

}  // namespace sunhip

extern "C" {

int sunhip_set_face_distances(sunhip_grid *g, const double *def1) {
  HIPCHK(hipSetDevice(g->device));
  Field *f;
  int rc = field(g, "def1", &f);
  if (rc) return rc;
  std::vector<double> slots((size_t)g->NES, 0.0);
  for (int s = 0; s < g->NES; s++) {
    const int e = g->edge_of_slot[s];
    if (e >= 0) slots[s] = def1[e];
  }
  HIPCHK(hipMemcpyAsync(f->d, slots.data(), sizeof(double) * slots.size(), hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  g->have_def1 = true;
  return 0;
}

int sunhip_w_predict(sunhip_grid *g, const sunhip_wpred_params *p) {
  HIPCHK(hipSetDevice(g->device));
  WPArgs A;
  int rc = wpred_args(g, p, &A);
  if (rc) return rc;
  const Geo G = g->geo();
  const int nc = (g->Nt + 3) / 4, ne = (g->Net + 3) / 4;
  if (A.scheme) {
    int ps = prof_begin(g, PK_OTHER);
    k_wp_centres<<<nc, 256, 0, g->stream>>>(G, A);
    prof_end(g, ps);
    HIPCHK(hipGetLastError());
    if (A.scheme == 5) {
      // HorizontalFaceScalars of wc with the boundary rows of w (main/phys.c:4906-4907)
      Field *u, *dzf, *sq, *sqc;
      if ((rc = field(g, "u", &u)) || (rc = field(g, "dzf", &dzf)) || (rc = field(g, "gradSx", &sq)) ||
          (rc = field(g, "gradSy", &sqc)))
        return rc;
      TvdArgs T{};
      T.s = A.wc; T.u = u->d; T.utmp2 = A.U; T.dzf = dzf->d; T.sumQ = sq->d; T.sumQC = sqc->d;
      T.oQ = sq->d; T.oQC = sqc->d; T.SfHp = const_cast<double *>(A.SfHp); T.SfHm = const_cast<double *>(A.SfHm);
      T.bscal = A.bw; T.theta = p->theta; T.TVD = p->TVDmomentum;
      if ((rc = face_scalars_launch(g, T))) return rc;
    }
    ps = prof_begin(g, PK_OTHER);
    k_wp_faces<<<ne, 256, 0, g->stream>>>(G, A);
    prof_end(g, ps);
    HIPCHK(hipGetLastError());
  }
  int ps = prof_begin(g, PK_WSOURCE);
  with_nf(g->mf, [&](auto nf) {
    k_wp_source<decltype(nf)::value><<<nc, 256, 0, g->stream>>>(G, A);
    return 0;
  });
  prof_end(g, ps);
  ps = prof_begin(g, PK_WCOLUMN);
  k_wp_column<<<nc, 256, 0, g->stream>>>(G, A);
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

}  // extern "C"

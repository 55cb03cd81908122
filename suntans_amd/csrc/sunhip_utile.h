// sunhip_utile.h -- per-tile bodies of the horizontal-momentum predictor
// (UPredictor, main/phys.c:2579-3018) on the device: the explicit
// free-surface term, the per-edge vertical system of viscosity, drag and
// vertical momentum advection solved for U*** and E = A^-1 1, D = dzf^T E,
// the free-surface source htmp, and after the free-surface solve the drying
// correction and the velocity update.
//
// One lane owns one edge (edge tiles) or one column (cell tiles).  The
// per-edge arrays the reference keeps in phys->a..d, am, bp, bm are edge
// fields here ("up_*"), walked by the lane level by level in the reference's
// loop order; every expression keeps the reference's operation order, so the
// results are the reference's bit for bit.
#pragma once
#include "sunhip_qtile.h"

namespace sunhip {

struct UPArgs {
  // edge fields
  double *u, *utmp, *utmp2, *E;
  const double *dzf;
  double *sa, *sb, *sc, *sa0, *sb0, *sc0, *sbp, *sep;  // per-edge scratch (up_*)
  // edge 2-D (by edge slot)
  double *D;
  const double *tau_T, *CdT, *CdB;
  // cell fields
  const double *dzz, *nu_tv, *nu_lax, *w;
  double *wtmp2, *s;
  // cell 2-D (by slot)
  double *h, *htmp, *active, *hcorr;
  const double *dv;
  double dt, theta, grav, nu, lax, thetaM;
  int nonlinear;
  int *err;  // [0]: a level of an edge with two zero-height cells; [1]: U*** = nan
};

// wtmp2 = w on every column (main/phys.c:2597-2599).
__device__ __forceinline__ void upred_cells_tile(const Geo &G, const UPArgs &A, int t, int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0) return;
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  for (int k = 0; k <= n; k++) at32(A.wtmp2, base + (uint32_t)k * 64u) = at32(A.w, base + (uint32_t)k * 64u);
}

// Everything of UPredictor before the open-boundary hooks for edge slot
// t*64+l (main/phys.c:2601-2942).
__device__ __forceinline__ void upred_edge_tile(const Geo &G, const UPArgs &A, int t, int l) {
  if (t >= G.Net) return;
  const int s = t * 64 + l;
  if (G.e_nkc[s] == 0) return;  // padding
  const uint32_t base = (uint32_t)(G.eoff[t] + l);
  const int ne = G.e_nke[s], et = G.e_etop[s], comp = G.e_comp[s];
  auto E = [&](double *p, int k) -> double & { return at32(p, base + (uint32_t)k * 64u); };
  // U*** = nan anywhere on the edge's wet layers (main/phys.c:2945-2950)
  auto nan_check = [&]() {
    for (int k = et; k < ne; k++)
      if (E(A.utmp, k) != E(A.utmp, k)) A.err[1] = 1;
  };
  A.D[s] = 0;
  for (int k = 0; k < ne; k++) E(A.utmp2, k) = E(A.u, k);
  if (comp == 2)
    for (int k = et; k < ne; k++) E(A.utmp, k) = E(A.u, k);
  if (comp != 1) { nan_check(); return; }
  const double dt = A.dt, theta = A.theta, th = 1.0, alpha = 1.0, nu = A.nu, lw = A.lax, tM = A.thetaM;
  int s1 = G.e_s1[s], s2 = G.e_s2[s];
  int b1 = G.e_c1[s], b2 = G.e_c2[s];
  {
    const double hd = A.h[s1] - A.h[s2], dg = G.dg[s];
    for (int k = et; k < ne; k++) E(A.utmp, k) -= A.grav * (1 - theta) * dt * hd / dg;
  }
  if (s1 < 0) { s1 = s2; b1 = b2; }
  if (s2 < 0) { s2 = s1; b2 = b1; }
  auto Z1 = [&](int k) { return at32(A.dzz, (uint32_t)b1 + (uint32_t)k * 64u); };
  auto Z2 = [&](int k) { return at32(A.dzz, (uint32_t)b2 + (uint32_t)k * 64u); };
  auto C1 = [&](const double *p, int k) { return at32(p, (uint32_t)b1 + (uint32_t)k * 64u); };
  auto C2 = [&](const double *p, int k) { return at32(p, (uint32_t)b2 + (uint32_t)k * 64u); };
  E(A.utmp, et) += 2.0 * dt * A.tau_T[s] / (Z1(et) + Z2(et));
  if (Z1(et) == 0 && Z2(et) == 0) { nan_check(); return; }
  for (int k = et; k < ne; k++) { E(A.sa, k) = 0; E(A.sb, k) = 0; E(A.sc, k) = 0; }
  for (int k = et + 1; k < ne; k++)
    E(A.sc, k) = 0.25 * (C1(A.nu_tv, k - 1) + C2(A.nu_tv, k - 1) + C1(A.nu_tv, k) + C2(A.nu_tv, k) +
                         lw * (C1(A.nu_lax, k - 1) + C2(A.nu_lax, k - 1) + C1(A.nu_lax, k) + C2(A.nu_lax, k)));
  for (int k = et + 1; k < ne; k++)
    E(A.sa, k) = 2.0 * (nu + E(A.sc, k)) / (0.25 * (Z1(k) + Z2(k)) * (Z1(k - 1) + Z2(k - 1) + Z1(k) + Z2(k)));
  for (int k = et; k < ne - 1; k++)
    E(A.sb, k) = 2.0 * (nu + E(A.sc, k + 1)) / (0.25 * (Z1(k) + Z2(k)) * (Z1(k) + Z2(k) + Z1(k + 1) + Z2(k + 1)));
  const bool multi = ne - et > 1, adv = A.nonlinear && tM >= 0 && multi;
  if (adv) {
    // w at the faces between U levels, from the cell whose top is lower
    // (n1) and the other (n0); zero above n1's top and at the bottom
    const int hi1 = G.ctop[s1] > G.ctop[s2];
    const int bl = hi1 ? b2 : b1, bh = hi1 ? b1 : b2, kth = hi1 ? G.ctop[s1] : G.ctop[s2];
    auto DW = [&](int k) -> double {
      if (k >= ne || k < kth) return 0.0;
      return 0.5 * (at32(A.w, (uint32_t)bl + (uint32_t)k * 64u) + at32(A.w, (uint32_t)bh + (uint32_t)k * 64u));
    };
    double dk = DW(et);
    for (int k = et; k < ne; k++) {
      const double dk1 = DW(k + 1);
      const double hz = 0.5 * (Z1(k) + Z2(k));
      E(A.sa0, k) = (alpha * 0.5 * (dk - fabs(dk)) + 0.5 * (1 - alpha) * dk) / hz;
      E(A.sb0, k) = (alpha * 0.5 * (dk + fabs(dk) - dk1 + fabs(dk1)) + 0.5 * (1 - alpha) * (dk - dk1)) / hz;
      E(A.sc0, k) = -(alpha * 0.5 * (dk1 + fabs(dk1)) + 0.5 * (1 - alpha) * dk1) / hz;
      dk = dk1;
    }
  }
  const double CdT = A.CdT[s], CdB = A.CdB[s];
  if (multi) {
    for (int k = et + 1; k < ne - 1; k++) {
      const double a = E(A.sa, k), b = E(A.sb, k);
      E(A.utmp, k) += dt * (1 - th) * (a * E(A.u, k - 1) - (a + b) * E(A.u, k) + b * E(A.u, k + 1));
    }
    {
      const double a = E(A.sa, et), b = E(A.sb, et), u0 = E(A.u, et), u1 = E(A.u, et + 1);
      if (CdT == -1) E(A.utmp, et) += dt * (1 - th) * (a * -u0 - (a + b) * u0 + b * u1);
      else E(A.utmp, et) += dt * (1 - th) * (-(b + 2.0 * CdT * fabs(u0) / (Z1(et) + Z2(et))) * u0 + b * u1);
    }
    {
      const int kb = ne - 1;
      const double a = E(A.sa, kb), b = E(A.sb, kb), um = E(A.u, kb - 1), u0 = E(A.u, kb);
      if (CdB == -1) E(A.utmp, kb) += dt * (1 - th) * (a * um - (a + b) * u0 + b * -u0);
      else E(A.utmp, kb) += dt * (1 - th) * (a * um - (a + 2.0 * CdB * fabs(u0) / (Z1(kb) + Z2(kb))) * u0);
    }
  } else {
    // the viscosity coefficient at the top of a one-layer edge is 0
    const double sz = Z1(et) + Z2(et), u0 = E(A.u, et), cv = E(A.sc, et);
    if (CdB == -1) E(A.utmp, et) -= 2.0 * dt * (1 - th) * (2.0 * (2.0 * (nu + cv)) * u0 / (sz * sz));
    else E(A.utmp, et) -= 2.0 * dt * (1 - th) * (CdB) / sz * fabs(u0) * u0;
    if (CdT == -1) E(A.utmp, et) -= 2.0 * dt * (1 - th) * (2.0 * (2.0 * (nu + cv)) * u0 / (sz * sz));
    else E(A.utmp, et) -= 2.0 * dt * (1 - th) * (CdT) / sz * fabs(u0) * u0;
  }
  if (adv) {
    for (int k = et + 1; k < ne - 1; k++)
      E(A.utmp, k) -= dt * (1 - tM) *
                      (E(A.sa0, k) * E(A.u, k - 1) + E(A.sb0, k) * E(A.u, k) + E(A.sc0, k) * E(A.u, k + 1));
    E(A.utmp, et) -= dt * (1 - tM) * ((E(A.sa0, et) + E(A.sb0, et)) * E(A.u, et) + E(A.sc0, et) * E(A.u, et + 1));
    const int kb = ne - 1;
    E(A.utmp, kb) -= dt * (1 - tM) * (E(A.sa0, kb) * E(A.u, kb - 1) + (E(A.sb0, kb) + E(A.sc0, kb)) * E(A.u, kb));
  }
  if (multi) {
    const int kb = ne - 1;
    E(A.sc, et) = -th * dt * E(A.sb, et);
    if (CdT == -1) E(A.sb, et) = 1.0 + th * dt * (E(A.sa, et) + E(A.sa, et + 1) + E(A.sb, et));
    else E(A.sb, et) = 1.0 + th * dt * (E(A.sb, et) + 2.0 * CdT * fabs(E(A.u, et)) / (Z1(et) + Z2(et)));
    E(A.sa, et) = 0;
    E(A.sc, kb) = 0;
    if (CdB == -1) E(A.sb, kb) = 1.0 + th * dt * (E(A.sa, kb) + E(A.sb, kb) + E(A.sb, kb - 1));
    else E(A.sb, kb) = 1.0 + th * dt * (E(A.sa, kb) + 2.0 * CdB * fabs(E(A.u, kb)) / (Z1(kb) + Z2(kb)));
    E(A.sa, kb) = -th * dt * E(A.sa, kb);
    for (int k = et + 1; k < kb; k++) {
      E(A.sc, k) = -th * dt * E(A.sb, k);
      E(A.sb, k) = 1.0 + th * dt * (E(A.sa, k) + E(A.sb, k));
      E(A.sa, k) = -th * dt * E(A.sa, k);
    }
  } else {
    const double sz = Z1(et) + Z2(et), cv = E(A.sc, et);
    double b = 1.0;
    if (CdB == -1) b += 4.0 * th * dt * 2.0 * (nu + cv) / (sz * sz);
    else b += 2.0 * th * dt * fabs(E(A.utmp, et)) / sz * (CdB);
    if (CdT == -1) b += 4.0 * th * dt * 2.0 * (nu + cv) / (sz * sz);
    else b += 2.0 * th * dt * fabs(E(A.utmp, et)) / sz * (CdT);
    E(A.sb, et) = b;
  }
  if (adv) {
    const int kb = ne - 1;
    for (int k = et + 1; k < kb; k++) {
      E(A.sa, k) += dt * tM * E(A.sa0, k);
      E(A.sb, k) += dt * tM * E(A.sb0, k);
      E(A.sc, k) += dt * tM * E(A.sc0, k);
    }
    E(A.sb, et) += dt * tM * (E(A.sa0, et) + E(A.sb0, et));
    E(A.sc, et) += dt * tM * E(A.sc0, et);
    E(A.sa, kb) += dt * tM * E(A.sa0, kb);
    E(A.sb, kb) += dt * tM * (E(A.sb0, kb) + E(A.sc0, kb));
  }
  for (int k = et; k < ne; k++)
    if (Z1(k) == 0 && Z2(k) == 0) { A.err[0] = 1; return; }
  if (multi) {
    // the reference's two TriSolve calls (main/util.c:185-203) share the
    // matrix: one forward sweep carries both right-hand sides
    double bprev = E(A.sb, et), dprev = E(A.utmp, et), eprev = 1.0, cprev = E(A.sc, et);
    E(A.sbp, et) = bprev;
    E(A.sep, et) = eprev;
    for (int k = et + 1; k < ne; k++) {
      const double a = E(A.sa, k);
      const double bk = E(A.sb, k) - a * cprev / bprev;
      const double dk = E(A.utmp, k) - a * dprev / bprev;
      const double ek = 1.0 - a * eprev / bprev;
      E(A.sbp, k) = bk;
      E(A.utmp, k) = dk;
      E(A.sep, k) = ek;
      bprev = bk; dprev = dk; eprev = ek; cprev = E(A.sc, k);
    }
    double un = dprev / bprev, en = eprev / bprev;
    E(A.utmp, ne - 1) = un;
    E(A.E, ne - 1) = en;
    for (int k = ne - 2; k >= et; k--) {
      const double bk = E(A.sbp, k), ck = E(A.sc, k);
      un = E(A.utmp, k) / bk - ck * un / bk;
      en = E(A.sep, k) / bk - ck * en / bk;
      E(A.utmp, k) = un;
      E(A.E, k) = en;
    }
  } else {
    E(A.utmp, et) /= E(A.sb, et);
    E(A.E, et) = 1.0 / E(A.sb, et);
  }
  double Dj = 0;
  for (int k = et; k < ne; k++) Dj += E(A.E, k) * at32(A.dzf, base + (uint32_t)k * 64u);
  A.D[s] = Dj;
  nan_check();
}

// htmp = Ac h - dt * sum of theta-weighted face fluxes (main/phys.c:2955-2971).
template <int NF>
__device__ __forceinline__ void upred_htmp_tile(const Geo &G, const UPArgs &A, int t, int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return;
  const int NS = G.NS, nf = G.nf[s];
  const double th = A.theta;
  double sum = 0;
#pragma unroll
  for (int f = 0; f < NF; f++) {
    if (f < nf) {
      const int es = G.es[f * NS + s];
      const uint32_t eb = (uint32_t)G.eb[f * NS + s];
      const double df = G.df[es], nrm = (double)G.nrm[f * NS + s];
      for (int k = G.e_etop[es]; k < G.e_nke[es]; k++) {
        const uint32_t ei = eb + (uint32_t)k * 64u;
        sum += ((1 - th) * at32(A.u, ei) + th * at32(A.utmp, ei)) * df * nrm * at32(A.dzf, ei);
      }
    }
  }
  A.htmp[s] = G.Ac[s] * A.h[s] - A.dt * sum;
}

// Cells that dry below 1e-10 above the bed (main/phys.c:2995-3004).
__device__ __forceinline__ void upred_dry_tile(const Geo &G, const UPArgs &A, int t, int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0) return;
  if (A.h[s] <= -A.dv[s] + 1e-10) {
    A.h[s] = -A.dv[s] + 1e-10;
    A.active[s] = 0;
    at32(A.s, (uint32_t)(G.toff[t] + l) + (uint32_t)(n - 1) * 64u) = 0;
  } else {
    A.hcorr[s] = 0;
    A.active[s] = 1;
  }
}

// u = U*** - g theta dt E (h1 - h2)/dg on computational edges, zero on a
// dry one-layer edge (main/phys.c:3008-3018).
__device__ __forceinline__ void upred_velocity_tile(const Geo &G, const UPArgs &A, int t, int l) {
  if (t >= G.Net) return;
  const int s = t * 64 + l;
  if (G.e_comp[s] != 1) return;
  const uint32_t base = (uint32_t)(G.eoff[t] + l);
  const int et = G.e_etop[s], ne = G.e_nke[s], s1 = G.e_s1[s], s2 = G.e_s2[s];
  const double hd = A.h[s1] - A.h[s2], dg = G.dg[s];
  for (int k = et; k < ne; k++) {
    const uint32_t ei = base + (uint32_t)k * 64u;
    at32(A.u, ei) = at32(A.utmp, ei) - A.grav * A.theta * A.dt * at32(A.E, ei) * hd / dg;
  }
  if (et == ne - 1 && at32(A.dzz, (uint32_t)G.e_c1[s] + (uint32_t)et * 64u) == 0 &&
      at32(A.dzz, (uint32_t)G.e_c2[s] + (uint32_t)et * 64u) == 0)
    at32(A.u, base + (uint32_t)et * 64u) = 0;
}

}  // namespace sunhip

// sunhip_utile.h -- per-tile bodies of the horizontal-momentum predictor
// (UPredictor, main/phys.c:2579-3018) on the device: the explicit
// free-surface term, the per-edge vertical system of viscosity, drag and
// vertical momentum advection solved for U*** and E = A^-1 1, D = dzf^T E,
// the free-surface source htmp, and after the free-surface solve the drying
// correction and the velocity update.
//
// One lane owns one edge (edge tiles) or one column (cell tiles).  Where the
// reference fills per-level arrays (phys->a..d, am, bp, bm) in separate
// passes, the lane forms each level's coefficients in registers in one
// top-down walk and keeps only the eliminated rows in edge fields ("up_*");
// every expression keeps the reference's operation order, so the results
// are the reference's bit for bit.
#pragma once
#include "sunhip_qtile.h"

namespace sunhip {

struct UPArgs {
  // edge fields
  double *u, *utmp, *utmp2, *E;
  const double *dzf;
  double *sc, *sbp, *sep;  // per-edge scratch (up_*): upper diagonal, eliminated diagonal and ones
  double *hterm;           // per edge-level: the free-surface source term, see upred_hterm_tile
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
  // streamed: utmp2 is read by later launches only
  for (int k = 0; k < ne; k++) __builtin_nontemporal_store(E(A.u, k), &E(A.utmp2, k));
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
  // The reference builds the edge's vertical system in passes over level
  // arrays (main/phys.c:2684-2920): viscosity coefficients, vertical
  // advection coefficients, the explicit part into U***, the implicit
  // matrix (with wall drag at the top and bottom), then TriSolve for U***
  // and for E = A^-1 1.  Every coefficient of level k depends only on levels
  // k-1..k+1, so here one top-down walk forms level k's final coefficients
  // and right-hand sides with the same expressions in the same order and
  // eliminates it at once; a bottom-up walk substitutes; a last top-down walk
  // sums D = dzf^T E in the reference's order.
  const bool multi = ne - et > 1, adv = A.nonlinear && tM >= 0 && multi;
  const double CdT = A.CdT[s], CdB = A.CdB[s];
  const int kb = ne - 1;
  // w at the faces between U levels, from the cell whose top is lower (n1)
  // and the other (n0); zero above n1's top and at the bottom
  const int hi1 = G.ctop[s1] > G.ctop[s2];
  const int bl = hi1 ? b2 : b1, bh = hi1 ? b1 : b2, kth = hi1 ? G.ctop[s1] : G.ctop[s2];
  auto DW = [&](int k) -> double {
    if (k >= ne || k < kth) return 0.0;
    return 0.5 * (at32(A.w, (uint32_t)bl + (uint32_t)k * 64u) + at32(A.w, (uint32_t)bh + (uint32_t)k * 64u));
  };
  // mean eddy and Lax-Wendroff viscosity at the face above level k (k > et)
  auto VIS = [&](int k) {
    return 0.25 * (C1(A.nu_tv, k - 1) + C2(A.nu_tv, k - 1) + C1(A.nu_tv, k) + C2(A.nu_tv, k) +
                   lw * (C1(A.nu_lax, k - 1) + C2(A.nu_lax, k - 1) + C1(A.nu_lax, k) + C2(A.nu_lax, k)));
  };
  const double f1 = dt * (1 - th), fm = dt * (1 - tM), ti = th * dt, ta = dt * tM;
  bool dry = false;
  // sliding window: the two cells' heights and u at k-1, k, k+1 and the
  // viscosity at the faces above k and k+1
  double h1m = 0, h2m = 0, h1k = Z1(et), h2k = Z2(et);
  double h1p = multi ? Z1(et + 1) : 0.0, h2p = multi ? Z2(et + 1) : 0.0;
  double um = 0, uk = E(A.u, et), up = multi ? E(A.u, et + 1) : 0.0;
  double sck = 0, scp = multi ? VIS(et + 1) : 0.0;
  double dk = adv ? DW(et) : 0.0;
  double sb_above = 0, btop = 0;  // level k-1's coefficient below it; the top row's diagonal
  double bprev = 0, dprev = 0, eprev = 0, cprev = 0;
  for (int k = et; k < ne; k++) {
    if (h1k == 0 && h2k == 0) dry = true;
    const double hs = h1k + h2k;
    // viscosity coefficients above (sa) and below (sb) level k
    const double sa = k > et ? 2.0 * (nu + sck) / (0.25 * hs * (h1m + h2m + h1k + h2k)) : 0.0;
    const double sb = k < kb ? 2.0 * (nu + scp) / (0.25 * hs * (h1k + h2k + h1p + h2p)) : 0.0;
    // vertical advection coefficients (thetaM)
    double sa0 = 0, sb0 = 0, sc0 = 0;
    if (adv) {
      const double dk1 = DW(k + 1);
      const double hz = 0.5 * hs;
      sa0 = (alpha * 0.5 * (dk - fabs(dk)) + 0.5 * (1 - alpha) * dk) / hz;
      sb0 = (alpha * 0.5 * (dk + fabs(dk) - dk1 + fabs(dk1)) + 0.5 * (1 - alpha) * (dk - dk1)) / hz;
      sc0 = -(alpha * 0.5 * (dk1 + fabs(dk1)) + 0.5 * (1 - alpha) * dk1) / hz;
      dk = dk1;
    }
    // explicit viscosity and drag, then explicit vertical advection, into U***
    double rhs = E(A.utmp, k);
    if (multi) {
      if (k == et) {
        if (CdT == -1) rhs += f1 * (sa * -uk - (sa + sb) * uk + sb * up);
        else rhs += f1 * (-(sb + 2.0 * CdT * fabs(uk) / hs) * uk + sb * up);
      } else if (k == kb) {
        if (CdB == -1) rhs += f1 * (sa * um - (sa + sb) * uk + sb * -uk);
        else rhs += f1 * (sa * um - (sa + 2.0 * CdB * fabs(uk) / hs) * uk);
      } else {
        rhs += f1 * (sa * um - (sa + sb) * uk + sb * up);
      }
    } else {
      // the viscosity coefficient at the top of a one-layer edge is 0
      const double cv = 0.0;
      if (CdB == -1) rhs -= 2.0 * dt * (1 - th) * (2.0 * (2.0 * (nu + cv)) * uk / (hs * hs));
      else rhs -= 2.0 * dt * (1 - th) * (CdB) / hs * fabs(uk) * uk;
      if (CdT == -1) rhs -= 2.0 * dt * (1 - th) * (2.0 * (2.0 * (nu + cv)) * uk / (hs * hs));
      else rhs -= 2.0 * dt * (1 - th) * (CdT) / hs * fabs(uk) * uk;
    }
    if (adv) {
      if (k == et) rhs -= fm * ((sa0 + sb0) * uk + sc0 * up);
      else if (k == kb) rhs -= fm * (sa0 * um + (sb0 + sc0) * uk);
      else rhs -= fm * (sa0 * um + sb0 * uk + sc0 * up);
    }
    // the implicit row: a (below the diagonal), b (diagonal), c (above)
    double a = 0, b, c = 0;
    if (multi) {
      if (k == et) {
        c = -th * dt * sb;
        if (CdT == -1) {
          // the coefficient above level et + 1
          const double sa_next = 2.0 * (nu + scp) / (0.25 * (h1p + h2p) * (h1k + h2k + h1p + h2p));
          b = 1.0 + ti * (sa + sa_next + sb);
        } else {
          b = 1.0 + ti * (sb + 2.0 * CdT * fabs(uk) / hs);
        }
        btop = b;
      } else if (k == kb) {
        // the reference reads level kb-1's coefficient after the top row
        // was formed, so on a two-layer edge it is the top diagonal
        const double sbm = kb - 1 == et ? btop : sb_above;
        if (CdB == -1) b = 1.0 + ti * (sa + sb + sbm);
        else b = 1.0 + ti * (sa + 2.0 * CdB * fabs(uk) / hs);
        a = -th * dt * sa;
      } else {
        c = -th * dt * sb;
        b = 1.0 + ti * (sa + sb);
        a = -th * dt * sa;
      }
      if (adv) {
        if (k == et) {
          b += ta * (sa0 + sb0);
          c += ta * sc0;
        } else if (k == kb) {
          a += ta * sa0;
          b += ta * (sb0 + sc0);
        } else {
          a += ta * sa0;
          b += ta * sb0;
          c += ta * sc0;
        }
      }
    } else {
      const double cv = 0.0;
      b = 1.0;
      if (CdB == -1) b += 4.0 * th * dt * 2.0 * (nu + cv) / (hs * hs);
      else b += 2.0 * th * dt * fabs(rhs) / hs * (CdB);
      if (CdT == -1) b += 4.0 * th * dt * 2.0 * (nu + cv) / (hs * hs);
      else b += 2.0 * th * dt * fabs(rhs) / hs * (CdT);
    }
    // elimination; the reference's two TriSolve calls (main/util.c:185-203)
    // share the matrix, so one sweep carries U*** and the ones of E
    double bk = b, dd = rhs, ek = 1.0;
    if (k > et) {
      bk = b - a * cprev / bprev;
      dd = rhs - a * dprev / bprev;
      ek = 1.0 - a * eprev / bprev;
    }
    E(A.sbp, k) = bk;
    E(A.sep, k) = ek;
    E(A.sc, k) = c;
    E(A.utmp, k) = dd;
    bprev = bk; dprev = dd; eprev = ek; cprev = c;
    sb_above = sb;
    // slide the window down one level
    h1m = h1k; h2m = h2k; h1k = h1p; h2k = h2p;
    um = uk; uk = up;
    sck = scp;
    if (k + 2 < ne) { h1p = Z1(k + 2); h2p = Z2(k + 2); up = E(A.u, k + 2); scp = VIS(k + 2); }
    else { h1p = 0; h2p = 0; up = 0; scp = 0; }
  }
  if (dry) { A.err[0] = 1; return; }
  // substitution, bottom up; it forms the final U*** of every wet level
  // (et..kb), so the nan test of U*** is taken on the values as they are
  // formed instead of reading them back
  double un = dprev / bprev, en = eprev / bprev;
  bool bad = un != un;
  E(A.utmp, kb) = un;
  E(A.E, kb) = en;
  for (int k = kb - 1; k >= et; k--) {
    const double bk = E(A.sbp, k), ck = E(A.sc, k);
    un = E(A.utmp, k) / bk - ck * un / bk;
    en = E(A.sep, k) / bk - ck * en / bk;
    bad |= un != un;
    __builtin_nontemporal_store(un, &E(A.utmp, k));
    E(A.E, k) = en;
  }
  double Dj = 0;
  for (int k = et; k < ne; k++) Dj += E(A.E, k) * at32(A.dzf, base + (uint32_t)k * 64u);
  A.D[s] = Dj;
  if (bad) A.err[1] = 1;
}

// The term of htmp's face sum at every level of edge slot s (main/phys.c:
// 2965-2967, ((1-theta)*u+theta*utmp)*df*normal*dzf) with the normal taken
// as +1: with -1 it is the same value negated, exactly, so a cell adds
// normal * hterm.  One lane per edge walks its levels (coalesced), and the
// cell sum reads one value per face and level instead of gathering u, utmp
// and dzf.
__device__ __forceinline__ void upred_hterm_tile(const Geo &G, const UPArgs &A, int t, int l) {
  if (t >= G.Net) return;
  const int s = t * 64 + l, ne = G.e_nke[s];
  if (ne == 0) return;
  const uint32_t base = (uint32_t)(G.eoff[t] + l);
  const double th = A.theta, df = G.df[s];
  for (int k = G.e_etop[s]; k < ne; k++) {
    const uint32_t ei = base + (uint32_t)k * 64u;
    at32(A.hterm, ei) = ((1 - th) * at32(A.u, ei) + th * at32(A.utmp, ei)) * df * at32(A.dzf, ei);
  }
}

// htmp = Ac h - dt * sum of theta-weighted face fluxes (main/phys.c:2955-2971),
// the faces' terms added in the reference's order (faces, then levels).
template <int NF>
__device__ __forceinline__ void upred_htmp_tile(const Geo &G, const UPArgs &A, int t, int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return;
  const int NS = G.NS, nf = G.nf[s];
  double sum = 0;
#pragma unroll
  for (int f = 0; f < NF; f++) {
    if (f < nf) {
      const int es = G.es[f * NS + s];
      const uint32_t eb = (uint32_t)G.eb[f * NS + s];
      const double nrm = (double)G.nrm[f * NS + s];
      for (int k = G.e_etop[es]; k < G.e_nke[es]; k++) sum += nrm * at32(A.hterm, eb + (uint32_t)k * 64u);
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

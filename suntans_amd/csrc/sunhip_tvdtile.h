// sunhip_tvdtile.h -- TVD scalar transport, one tile of columns (or of
// edges) at a time.
//
// Reference: HorizontalFaceScalars main/tvd.c:36-120, Psi :130-151,
// GetApAm :161-204, UpdateScalars main/scalars.c:35-325, TriSolve
// main/util.c:185-203, UpWind/Max/Min main/util.c:212-246.  Shared by the
// single-grid kernels (sunhip_tvd.hip) and the batched kernels over many
// grids (sunhip_batch.hip).  Operation order follows the reference so
// results agree bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "sunhip_qtile.h"

namespace sunhip {

#define TVD_EPS 1e-12

__device__ __forceinline__ double dmax_ref(double x1, double x2) { return x2 > x1 ? x2 : x1; }
__device__ __forceinline__ double dmin_ref(double x, double y) { return x < y ? x : y; }

__device__ __forceinline__ double psi(double r, int TVD) {
  if (TVD == 2) return 1;
  if (TVD == 3) return dmax_ref(0, dmax_ref(dmin_ref(2 * (r), 1), dmin_ref(r, 2)));
  if (TVD == 4) return r > 0 ? dmin_ref(r, 1) : 0;
  return 0;
}

struct TvdArgs {
  const double *s;       // scalar (cell field), old values
  const double *u, *utmp2, *dzf;   // edge fields
  const double *sumQ, *sumQC;      // cell fields (pass 1 output)
  double *oQ, *oQC;
  double *SfHp, *SfHm;             // edge fields
  const double *bscal;             // edge field, type-2 edges (may be null)
  // the column update's horizontal flux terms, one per edge-level and side
  // (null: not formed; see tvd_faces_tile)
  double *P1, *P2;
  double theta, dt;
  int TVD;
};

// pass 1 (main/tvd.c:46-65), cells of celldist[0..2).  The faces' level-
// invariant data (edge base, df, normal, neighbour base and depth) are read
// once per column.  NF: faces per cell at most.
template <int NF, int GRP = 1>
__device__ __forceinline__ void tvd_sums_tile(const Geo &G, const TvdArgs &A, int t, int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] == 0) return;
  const int NS = G.NS, nf = G.nf[s];
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  const double th = A.theta;
  uint32_t eb[NF], nbb[NF];
  double dfe[NF];
  int nrm[NF], nkn[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    if (f < nf) {
      const int nbs = G.nbs[f * NS + s];
      eb[f] = (uint32_t)G.eb[f * NS + s];
      dfe[f] = G.df[G.es[f * NS + s]];
      nrm[f] = G.nrm[f * NS + s];
      nkn[f] = nbs >= 0 ? G.nk[nbs] : 0;  // no neighbour: never below its bottom
      nbb[f] = nbs >= 0 ? (uint32_t)G.nbb[f * NS + s] : base;
    } else {
      eb[f] = 0; dfe[f] = 0; nrm[f] = 0; nkn[f] = 0; nbb[f] = base;
    }
  }
  // GRP levels at a time: every load of the group first (levels past the
  // column clamped to its last, the neighbour's scalar to its own bottom),
  // then the sums in face order as below
  const int k1 = G.ctop[s];
  for (int k0 = k1; k0 < n; k0 += GRP) {
    double vs[GRP], vu[GRP][NF], vt[GRP][NF], vz[GRP][NF], vn[GRP][NF];
#pragma unroll
    for (int j = 0; j < GRP; j++) {
      const int k = k0 + j < n ? k0 + j : n - 1;
      const uint32_t k64 = (uint32_t)k * 64u;
      vs[j] = at32(A.s, base + k64);
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const uint32_t ei = eb[f] + k64;
        vu[j][f] = at32(A.u, ei);
        vt[j][f] = at32(A.utmp2, ei);
        vz[j][f] = at32(A.dzf, ei);
        const int kn = k < nkn[f] ? k : (nkn[f] > 0 ? nkn[f] - 1 : 0);
        vn[j][f] = at32(A.s, nbb[f] + (uint32_t)kn * 64u);
      }
    }
#pragma unroll
    for (int j = 0; j < GRP; j++) {
      const int k = k0 + j;
      if (k >= n) break;
      const uint32_t idx = base + (uint32_t)k * 64u;
      const double si = vs[j];
      double sq = 0, sqc = 0;
#pragma unroll
      for (int f = 0; f < NF; f++) {
        if (f < nf) {
          const double un = nrm[f] * (th * vu[j][f] + (1 - th) * vt[j][f]);
          const double Qm = 0.5 * vz[j][f] * dfe[f] * fabs(un - fabs(un));
          if (k < nkn[f]) sqc += Qm * (si - vn[j][f]);
          sq += Qm;
        }
      }
      at32(A.oQ, idx) = sq;
      at32(A.oQC, idx) = sqc;
    }
  }
}

// The horizontal flux term of the scalar column update (main/scalars.c:
// 283-285, dt*df*normal/Ac*(theta*u+(1-theta)*utmp2)*temp*dzf) at level k
// of an edge for one of its cells, with the normal taken as +1: the product
// with normal -1 is the same value negated, exactly, so the column adds
// normal * P.  c = dt*df/Ac of that cell, w = theta*u+(1-theta)*utmp2, temp
// SfHp or SfHm by the sign of utmp2.  The faces pass stores it per edge-level
// for each side (P1: the grad[0] cell's, P2: the grad[1] cell's), and the
// column then reads one value per face and level instead of u, utmp2, dzf,
// SfHp and SfHm.
__device__ __forceinline__ double face_term(double c, double w, double temp, double fz) { return c * w * temp * fz; }

// Side constants of the flux terms of edge slot s (0 for a missing side).
__device__ __forceinline__ void face_term_consts(const Geo &G, const TvdArgs &A, int s, double &c1, double &c2) {
  const int s1 = G.e_s1[s], s2 = G.e_s2[s];
  const double dd = A.dt * G.df[s];
  c1 = s1 >= 0 ? dd / G.Ac[s1] : 0.0;
  c2 = s2 >= 0 ? dd / G.Ac[s2] : 0.0;
}

__device__ __forceinline__ void store_terms(const Geo &G, const TvdArgs &A, int s, int64_t ei, double c1, double c2,
                                            double w, double ut2, double hp, double hm) {
  const double temp = ut2 > 0 ? hp : hm, fz = A.dzf[ei];
  // streamed: the column kernel reads them after this launch
  if (G.e_s1[s] >= 0) __builtin_nontemporal_store(face_term(c1, w, temp, fz), &A.P1[ei]);
  if (G.e_s2[s] >= 0) __builtin_nontemporal_store(face_term(c2, w, temp, fz), &A.P2[ei]);
}

// pass 2 (main/tvd.c:69-119), one lane per edge; with TvdArgs::P1 set, the
// column's flux terms of each level too (store_terms), from the values in
// registers.  Edges outside the pass (closed boundaries) get theirs from the
// stored face values, which the column would otherwise read.
__device__ __forceinline__ void tvd_faces_tile(const Geo &G, const TvdArgs &A, int t, int l) {
  if (t >= G.Net) return;
  const int s = t * 64 + l, ec = G.e_comp[s];
  const int64_t base = G.eoff[t] + l;
  const int et = G.e_etop[s], nke = G.e_nke[s];
  const double th = A.theta;
  const bool terms = A.P1 != nullptr;
  double c1 = 0, c2 = 0;
  if (terms) face_term_consts(G, A, s, c1, c2);
  if (ec == 0) {
    if (terms)
      for (int k = 0; k < nke; k++) {
        const int64_t ei = base + (int64_t)k * 64;
        const double ut2 = A.utmp2[ei];
        store_terms(G, A, s, ei, c1, c2, th * A.u[ei] + (1 - th) * ut2, ut2, A.SfHp[ei], A.SfHm[ei]);
      }
    return;
  }
  for (int k = 0; k < et; k++) {
    const int64_t ei = base + (int64_t)k * 64;
    A.SfHp[ei] = A.SfHm[ei] = 0;
    if (terms) {
      const double ut2 = A.utmp2[ei];
      store_terms(G, A, s, ei, c1, c2, th * A.u[ei] + (1 - th) * ut2, ut2, 0.0, 0.0);
    }
  }
  if (ec == 1) {
    const int64_t e1 = G.e_c1[s], e2 = G.e_c2[s];
    for (int k = et; k < nke; k++) {
      const int64_t ei = base + (int64_t)k * 64;
      const double ut2 = A.utmp2[ei];
      const double un = th * A.u[ei] + (1 - th) * ut2;
      const double s1 = A.s[e1 + (int64_t)k * 64], s2 = A.s[e2 + (int64_t)k * 64];
      const int64_t up = (un > 0 ? e2 : e1) + (int64_t)k * 64;
      const double si = un > 0 ? s2 : s1, sm = un > 0 ? s1 : s2;
      double r = 0;
      const double q = A.sumQ[up];
      if (q != 0 && sm != si) r = A.sumQC[up] / (q * (sm - si));
      const double hp = s2 + 0.5 * psi(r, A.TVD) * (s1 - s2);
      const double hm = s1 - 0.5 * psi(r, A.TVD) * (s1 - s2);
      A.SfHp[ei] = hp;
      A.SfHm[ei] = hm;
      if (terms) store_terms(G, A, s, ei, c1, c2, un, ut2, hp, hm);
    }
  } else {
    const int64_t ib = G.e_c1[s];
    for (int k = et; k < nke; k++) {
      const int64_t ei = base + (int64_t)k * 64;
      const double hp = A.bscal ? A.bscal[ei] : 0.0;
      const double hm = A.s[ib + (int64_t)k * 64];
      A.SfHp[ei] = hp;
      A.SfHm[ei] = hm;
      if (terms) {
        const double ut2 = A.utmp2[ei];
        store_terms(G, A, s, ei, c1, c2, th * A.u[ei] + (1 - th) * ut2, ut2, hp, hm);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// UpdateScalars column kernel

struct ColArgs {
  const double *s_old;                   // stmp
  double *s;                             // scal (out)
  const double *wnew, *wtmp2;            // w-type
  const double *dzz, *dzzold, *kappa_tv;
  const double *u, *utmp2, *dzf, *SfHp, *SfHm;
  const double *active;                  // 2-D cell field (0/1)
  double *Cn;
  double *tb, *tc, *td;                  // Thomas scratch (cell fields)
  // source terms (null: absent): src1 (implicit, per cell-layer), src2
  // (explicit, per cell-layer), Ftop/Fbot (surface and bottom fluxes, per
  // column), main/scalars.c:151-183, 213-242
  const double *src1, *src2, *Ftop, *Fbot;
  // first-order upwind only: the reference's stmp2 rows (the boundary values
  // in the rows of cells next to open-boundary edges, main/scalars.c:84-96),
  // read on faces of type-2 and type-3 edges without a second cell; null:
  // those faces read the cell's own old value, as without boundary_scal
  const double *bcell;
  // the flux terms formed by the faces pass (TvdArgs::P1, P2), read instead
  // of u, utmp2, dzf, SfHp and SfHm when the column is instantiated with PRE
  const double *P1, *P2;
  double theta, dt, kappa;
  int TVD, step, alpha_top, alpha_bot;
};

// One column's view: levels are absolute k; reads outside [0,Nk) give 0, as
// in the oracle (the reference reads outside the row there).
struct Col {
  const double *p;
  int64_t base;
  int nk;
  __device__ double at(int k) const { return (k >= 0 && k < nk) ? p[base + (int64_t)k * 64] : 0.0; }
};

// ap[k], am[k] of GetApAm (main/tvd.c:161-204) for one k in [ktop+1, Nk]:
// the final values of rp/rm after the reference's sequence of assignments.
// Takes the values it reads: w at k, the layer heights at k and k-1, and the
// scalar at k-2..k+1 (0 outside the column).
__device__ __forceinline__ void apam(double wk, double dzk, double dzk1, double sm2, double sm1, double s0, double sp1,
                                     int nk, int ktop, int k, double dt, int TVD, double &ap, double &am) {
  const double wp = 0.5 * (wk + fabs(wk)), wm = 0.5 * (wk - fabs(wk));
  double Cp, Cm;
  if (k < nk) {
    const double den = dzk + dzk1;
    Cp = 2 * wp * dt / den;
    Cm = 2 * wm * dt / den;
  } else {
    Cp = wp * dt / dzk1;
    Cm = wm * dt / dzk1;
  }
  double rp, rm;
  if (k == nk) {
    rp = 1;
    rm = (sm2 - sm1 + TVD_EPS) / TVD_EPS;
  } else if (k == nk - 1) {
    rp = TVD_EPS / (sm1 - s0 + TVD_EPS);
    rm = (sm2 - sm1 + TVD_EPS) / (sm1 - s0 + TVD_EPS);
  } else if (k == ktop + 1) {
    rp = (s0 - sp1 + TVD_EPS) / (sm1 - s0 + TVD_EPS);
    rm = TVD_EPS / (sm1 - s0 + TVD_EPS);
  } else {
    rp = (s0 - sp1 + TVD_EPS) / (sm1 - s0 + TVD_EPS);
    rm = (sm2 - sm1 + TVD_EPS) / (sm1 - s0 + TVD_EPS);
  }
  am = 0.5 * wp * psi(rp, TVD) * (1 - Cp) + wm * (1 - 0.5 * psi(rm, TVD) * (1 + Cm));
  ap = wp * (1 - 0.5 * psi(rp, TVD) * (1 - Cp)) + 0.5 * wm * psi(rm, TVD) * (1 + Cm);
}

__device__ __forceinline__ void apam_plain(const double *w, int64_t base, int k, double &ap, double &am) {
  const double wk = w[base + (int64_t)k * 64];
  ap = 0.5 * (wk + fabs(wk));
  am = 0.5 * (wk - fabs(wk));
}

__device__ __forceinline__ double upwind_ref(double u, double a, double b) {
  if (fabs(u) > 0) return u > 0 ? b : a;
  return dmax_ref(a, b);
}

// A face's level-invariant part of the horizontal flux of main/scalars.c:
// 251-290: the edge's layer count and base, the upwind cell bases, and
// cf = dt*df*normal/Ac, the leading factors of the flux in the reference's
// order (so the product below rounds as the reference's does).  c2 < 0 marks
// an open-boundary face whose second value comes from ColArgs::bcell at
// base -c2 - 1.
struct FaceC {
  uint32_t eb;
  int nke;
  int64_t c1, c2;
  double cf;
  const double *pp;  // this cell's side of the precomputed flux terms
  double nd;         // the face's normal sign
};

template <int NF>
__device__ __forceinline__ void face_consts(const Geo &G, const ColArgs &A, int s, double Ac, FaceC (&fc)[NF]) {
  const int NS = G.NS, nf = G.nf[s];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    if (f < nf) {
      const int es = G.es[f * NS + s];
      fc[f].eb = (uint32_t)G.eb[f * NS + s];
      fc[f].nke = G.e_nke[es];
      int64_t c1 = G.e_c1[es], c2 = G.e_c2[es];
      if (c1 < 0) c1 = c2;
      if (c2 < 0) {
        const int mk = G.e_mark[es];
        c2 = A.bcell && (mk == 2 || mk == 3) ? -c1 - 1 : c1;
      }
      fc[f].c1 = c1;
      fc[f].c2 = c2;
      fc[f].cf = A.dt * G.df[es] * G.nrm[f * NS + s] / Ac;
      fc[f].pp = G.e_s1[es] == s ? A.P1 : A.P2;
      fc[f].nd = (double)G.nrm[f * NS + s];
    } else {
      fc[f].eb = 0; fc[f].nke = 0; fc[f].c1 = 0; fc[f].c2 = 0; fc[f].cf = 0;
      fc[f].pp = A.P1; fc[f].nd = 0;
    }
  }
}

// horizontal TVD flux divergence at level k (main/scalars.c:251-290)
template <int NF, bool PRE = false>
__device__ __forceinline__ double hflux(const ColArgs &A, const FaceC (&fc)[NF], int nf, int k) {
  double acc = 0;
  const uint32_t k64 = (uint32_t)k * 64u;
#pragma unroll
  for (int f = 0; f < NF; f++) {
    if (f >= nf || k >= fc[f].nke) continue;
    const uint32_t ei = fc[f].eb + k64;
    if (PRE) {
      acc += fc[f].nd * at32(fc[f].pp, ei);
      continue;
    }
    const double ut2 = at32(A.utmp2, ei);
    double temp;
    if (A.TVD) {
      temp = ut2 > 0 ? at32(A.SfHp, ei) : at32(A.SfHm, ei);
    } else {
      const int64_t c2 = fc[f].c2;
      const double sp = c2 >= 0 ? A.s_old[c2 + (int64_t)k * 64] : A.bcell[-c2 - 1 + (int64_t)k * 64];
      temp = upwind_ref(ut2, A.s_old[fc[f].c1 + (int64_t)k * 64], sp);
    }
    acc += fc[f].cf * (A.theta * at32(A.u, ei) + (1 - A.theta) * ut2) * temp * at32(A.dzf, ei);
  }
  return acc;
}

// The loads of one level k of the column walk (levels beyond the column
// clamped into it, so every load is in bounds and none waits on a branch):
// wnew and wtmp2 at k+1; the old scalar at k+3, dzz, dzzold and kappa_tv at
// k+2 (0 outside the column); per face at k the edge's u, utmp2, dzf and
// either SfHp and SfHm (TVD) or the old scalar of its two cells.
template <int NF>
struct LevelIn {
  double wn, wo, sn, zn, on, qn, g1, g2;  // g1, g2: src1 and src2 at k
  double fu[NF], ft[NF], fz[NF], fa[NF], fb[NF];
  double fp[NF];  // PRE: the precomputed flux term
};

template <int NF, bool SRC, bool PRE = false>
__device__ __forceinline__ void load_level(const ColArgs &A, const FaceC (&fc)[NF], int64_t base, int nk, int k,
                                           LevelIn<NF> &L) {
  const int kw = k < nk ? k : nk - 1;  // w-type rows hold nk + 1 levels
  L.wn = A.wnew[base + (int64_t)(kw + 1) * 64];
  L.wo = A.wtmp2[base + (int64_t)(kw + 1) * 64];
  auto cl = [&](const double *p, int kk) {
    const double v = p[base + (int64_t)(kk < nk ? kk : nk) * 64];  // a tile holds one level below its deepest column
    return kk < nk ? v : 0.0;
  };
  L.sn = cl(A.s_old, k + 3);
  L.zn = cl(A.dzz, k + 2);
  L.on = cl(A.dzzold, k + 2);
  L.qn = cl(A.kappa_tv, k + 2);
  L.g1 = SRC && A.src1 ? A.src1[base + (int64_t)k * 64] : 0.0;
  L.g2 = SRC && A.src2 ? A.src2[base + (int64_t)k * 64] : 0.0;
#pragma unroll
  for (int f = 0; f < NF; f++) {
    const int ne = fc[f].nke;
    const int ke = k < ne ? k : (ne > 0 ? ne - 1 : 0);
    const uint32_t ei = fc[f].eb + (uint32_t)ke * 64u;
    if (PRE) {
      L.fp[f] = at32(fc[f].pp, ei);
      continue;
    }
    L.fu[f] = at32(A.u, ei);
    L.ft[f] = at32(A.utmp2, ei);
    L.fz[f] = at32(A.dzf, ei);
    if (A.TVD) {
      L.fa[f] = at32(A.SfHp, ei);
      L.fb[f] = at32(A.SfHm, ei);
    } else {
      const int64_t c2 = fc[f].c2;
      const double *pb = c2 >= 0 ? A.s_old : A.bcell;
      L.fa[f] = A.s_old[fc[f].c1 + (int64_t)ke * 64];
      L.fb[f] = pb[(c2 >= 0 ? c2 : -c2 - 1) + (int64_t)ke * 64];
    }
  }
}

// hflux on the loaded values: the same products in the same order
template <int NF, bool PRE = false>
__device__ __forceinline__ double hflux_in(const ColArgs &A, const FaceC (&fc)[NF], int nf, int k,
                                           const LevelIn<NF> &L) {
  double acc = 0;
#pragma unroll
  for (int f = 0; f < NF; f++) {
    if (f >= nf || k >= fc[f].nke) continue;
    if (PRE) {
      acc += fc[f].nd * L.fp[f];
      continue;
    }
    const double ut2 = L.ft[f];
    const double temp = A.TVD ? (ut2 > 0 ? L.fa[f] : L.fb[f]) : upwind_ref(ut2, L.fa[f], L.fb[f]);
    acc += fc[f].cf * (A.theta * L.fu[f] + (1 - A.theta) * ut2) * temp * L.fz[f];
  }
  return acc;
}

// One column of UpdateScalars.  GRP > 1: the levels below the top row are
// walked GRP at a time, every load of a group issued before its arithmetic,
// and the back substitution reads its rows a group at a time; the values
// are those of GRP = 1 bit for bit.  SRC: the source terms of ColArgs are
// read (without it they are taken as absent, and cost no registers).  PRE:
// the horizontal flux terms are read from ColArgs::P1/P2 (TVD schemes, the
// faces pass having formed them), one value per face and level.
template <int NF, int GRP = 1, bool SRC = false, bool PRE = false, int KR = 0>
__device__ __forceinline__ void update_column_tile(const Geo &G, const ColArgs &A, int t, int l) {
  static_assert(KR == 0 || (GRP > 1 && KR % GRP == 0), "a register head needs load groups that divide it");
  if (t >= G.Nt) return;
  const int s = t * 64 + l, nk = G.nk[s];
  if (nk == 0 || G.comp[s] != 1) return;
  const int ct = G.ctop[s], cto = G.ctopold[s];
  const int64_t base = G.toff[t] + l;
  const double Ac = G.Ac[s], theta = A.theta, dt = A.dt;
  const Col dzz{A.dzz, base, nk}, dzo{A.dzzold, base, nk}, so{A.s_old, base, nk};
  FaceC fc[NF];
  face_consts(G, A, s, Ac, fc);
  const int nfs = G.nf[s];
  int ktop;
  double dznew;
  if (ct >= cto) {
    ktop = ct;
    dznew = dzz.at(ktop);
  } else {
    ktop = cto;
    dznew = 0;
    for (int k = ct; k <= cto; k++) dznew += dzz.at(k);
  }
  const int nb = nk - 1 - ktop;
  const bool tvdv = A.TVD != 0;
  // level windows, positioned for level k of the walk below: the old scalar
  // at k-1..k+2, the new and old layer heights and kappa_tv at k and k+1
  // (0 outside the column), so that every value is loaded once
  double s0 = so.at(ktop - 1), s1 = so.at(ktop), s2 = so.at(ktop + 1), s3 = so.at(ktop + 2);
  double z0 = dzz.at(ktop), z1 = dzz.at(ktop + 1), o0 = dzo.at(ktop), o1 = dzo.at(ktop + 1);
  auto kap = [&](int k) { return k < nk ? A.kappa_tv[base + (int64_t)k * 64] : 0.0; };
  double q0 = kap(ktop), q1 = kap(ktop + 1);
  // ap, am at level k+1 (its top face) from the windows at level k
  auto ap_am_new = [&](int k1, double &ap, double &am) {
    if (tvdv) apam(A.wnew[base + (int64_t)k1 * 64], z1, z0, s0, s1, s2, s3, nk, ktop, k1, dt, A.TVD, ap, am);
    else apam_plain(A.wnew, base, k1, ap, am);
  };
  auto ap_am_old = [&](int k1, double &ap, double &am) {
    if (tvdv) apam(A.wtmp2[base + (int64_t)k1 * 64], o1, o0, s0, s1, s2, s3, nk, ktop, k1, dt, A.TVD, ap, am);
    else apam_plain(A.wtmp2, base, k1, ap, am);
  };
  // vertical diffusion coefficient of the face above level k+1
  auto bd = [&]() { return (2.0 * A.kappa + q0 + q1) / (z0 + z1); };
  const double td = theta * dt, otd = (1 - theta) * dt;
  const bool hs1 = SRC && A.src1 != nullptr, hs2 = SRC && A.src2 != nullptr;
  const bool hft = SRC && A.Ftop != nullptr, hfb = SRC && A.Fbot != nullptr;
  auto S1 = [&](int k) { return A.src1[base + (int64_t)k * 64]; };
  // horizontal source of the top row: levels 0..ktop fold into row 0, after
  // the top cell's explicit source dt*src2*dzzold
  double cn0 = hs2 ? dt * A.src2[base + (int64_t)ktop * 64] * dzo.at(ktop) : 0.0;
  for (int k = 0; k <= ktop && k < nk; k++) cn0 -= hflux<NF, PRE>(A, fc, nfs, k);

  // forward: assemble row m (level k = ktop+m) and eliminate
  double bprev = 1, cprev = 0, dprev = 0;
  // new-time advection at the top face of the next level, carried
  double apN, amN, apO, amO;  // at k+1 (new and old velocities)
  ap_am_new(ktop + 1, apN, amN);
  ap_am_old(ktop + 1, apO, amO);
  double bd_k1 = (ktop + 1 < nk) ? bd() : 0.0;  // bd at k+1
  double apN_k = 0, amN_k = 0, apO_k = 0, amO_k = 0, bd_k = 0;
  // KR > 0: rows 0..KR (the top KR+1 levels) keep b, c and d in registers
  // between the sweeps, so they never go to tb/tc/td; the arithmetic and its
  // order are the same for every KR
  constexpr int KH = KR > 0 ? KR : 1;
  double hb[KH], hc[KH], hd[KH];
  double h0b = 1, h0c = 0, h0d = 0;
  for (int m = 0; m <= (GRP > 1 ? 0 : nb); m++) {
    const int k = ktop + m;
    // shift the level-(k+1) quantities to level k, compute level k+1
    if (m > 0) {
      apN_k = apN; amN_k = amN; apO_k = apO; amO_k = amO; bd_k = bd_k1;
      ap_am_new(k + 1, apN, amN);
      ap_am_old(k + 1, apO, amO);
      bd_k1 = (k + 1 < nk) ? bd() : 0.0;
    }
    double a, b, c, d;
    const double sk = s1, skm = s0, skp = s2;
    if (m == 0) {
      a = 0;
      b = dznew - td * amN;
      c = -td * apN;
      if (nb == 0) b += c;
      if (hs1) b += z0 * S1(ktop) * theta * dt;
      if (ktop < nk - 1) {
        b += td * (bd_k1 + 2 * A.alpha_top * bd_k1);
        c -= td * bd_k1;
      }
      if (cto <= ct) {
        d = 0;
        for (int kk = cto; kk <= ct; kk++) d += dzo.at(kk) * so.at(kk);
        if (hs1)
          for (int kk = cto; kk <= ct; kk++) d -= S1(kk) * (1 - theta) * dt * dzo.at(kk) * so.at(kk);
      } else {
        d = o0 * sk;
        // the reference multiplies by the scalar at a stale level index past
        // the row here (main/scalars.c:153-154); the top cell's, as the oracle
        if (hs1) d -= S1(ktop) * (1 - theta) * dt * o0 * sk;
      }
      if (ktop < nk - 1) {
        d = d - otd * (-amO * sk - apO * skp) + otd * (-(2 * A.alpha_top * bd_k1 + bd_k1) * sk + bd_k1 * skp);
        if (hft) d += dt * (1 - A.alpha_top + 2 * A.alpha_top * bd_k1) * A.Ftop[s];
      }
      d += cn0;
    } else {
      a = td * amN_k;
      b = z0 + td * (apN_k - amN);
      c = -td * apN;
      d = o0 * sk;
      if (hs1) d -= S1(k) * (1 - theta) * dt * o0 * sk;
      if (m == nb) {
        b += c;
        if (hs1) b += z0 * S1(k) * theta * dt;
        a -= td * bd_k;
        b += td * (bd_k + 2 * A.alpha_bot * bd_k);
        d -= otd * (amO_k * skm + apO_k * sk) - otd * (bd_k * skm - (bd_k + 2 * A.alpha_bot * bd_k) * sk);
        if (hfb) d += dt * (-1 + A.alpha_bot + 2 * A.alpha_bot * bd_k) * A.Fbot[s];
      } else {
        a -= td * bd_k;
        b += td * (bd_k + bd_k1);
        c -= td * bd_k1;
        if (hs1) b += z0 * S1(k) * theta * dt;
        d -= otd * (amO_k * skm + (apO_k - amO) * sk - apO * skp) -
             otd * (bd_k * skm - (bd_k + bd_k1) * sk + bd_k1 * skp);
      }
      // Cn[m] = dt*src2*dzzold (0 without src2) - ap[k]; d[m] += fab*Cn[m] with fab = 1
      const double cnk = hs2 ? dt * A.src2[base + (int64_t)k * 64] * o0 - hflux<NF, PRE>(A, fc, nfs, k)
                             : -hflux<NF, PRE>(A, fc, nfs, k);
      d += cnk;
      A.Cn[base + (int64_t)k * 64] = cnk;
    }
    if (m > 0) {
      b -= a * cprev / bprev;
      d -= a * dprev / bprev;
    }
    const int64_t idx = base + (int64_t)k * 64;
    if (KR > 0) {
      h0b = b; h0c = c; h0d = d;
    } else {
      A.tb[idx] = b;
      A.tc[idx] = c;
      A.td[idx] = d;
    }
    bprev = b; cprev = c; dprev = d;
    // slide the windows down one level
    s0 = s1; s1 = s2; s2 = s3; s3 = so.at(k + 3);
    z0 = z1; z1 = dzz.at(k + 2);
    o0 = o1; o1 = dzo.at(k + 2);
    q0 = q1; q1 = kap(k + 2);
  }
  if (GRP > 1) {
    // rows m0..m0+GRP-1 (up to nb) on loaded values (the body of the loop
    // above for m > 0); put(j, idx, b, c, d) keeps row m0+j
    auto group = [&](int m0, auto &&put) {
      LevelIn<NF> L[GRP];
#pragma unroll
      for (int j = 0; j < GRP; j++) {
        const int k = ktop + m0 + j;
        load_level<NF, SRC, PRE>(A, fc, base, nk, k < nk ? k : nk - 1, L[j]);
      }
#pragma unroll
      for (int j = 0; j < GRP; j++) {
        const int m = m0 + j;
        if (m > nb) break;
        const int k = ktop + m;
        apN_k = apN; amN_k = amN; apO_k = apO; amO_k = amO; bd_k = bd_k1;
        if (tvdv) {
          apam(L[j].wn, z1, z0, s0, s1, s2, s3, nk, ktop, k + 1, dt, A.TVD, apN, amN);
          apam(L[j].wo, o1, o0, s0, s1, s2, s3, nk, ktop, k + 1, dt, A.TVD, apO, amO);
        } else {
          apN = 0.5 * (L[j].wn + fabs(L[j].wn));
          amN = 0.5 * (L[j].wn - fabs(L[j].wn));
          apO = 0.5 * (L[j].wo + fabs(L[j].wo));
          amO = 0.5 * (L[j].wo - fabs(L[j].wo));
        }
        bd_k1 = (k + 1 < nk) ? bd() : 0.0;
        const double sk = s1, skm = s0, skp = s2;
        double a = td * amN_k;
        double b = z0 + td * (apN_k - amN);
        double c = -td * apN;
        double d = o0 * sk;
        if (hs1) d -= L[j].g1 * (1 - theta) * dt * o0 * sk;
        if (m == nb) {
          b += c;
          if (hs1) b += z0 * L[j].g1 * theta * dt;
          a -= td * bd_k;
          b += td * (bd_k + 2 * A.alpha_bot * bd_k);
          d -= otd * (amO_k * skm + apO_k * sk) - otd * (bd_k * skm - (bd_k + 2 * A.alpha_bot * bd_k) * sk);
          if (hfb) d += dt * (-1 + A.alpha_bot + 2 * A.alpha_bot * bd_k) * A.Fbot[s];
        } else {
          a -= td * bd_k;
          b += td * (bd_k + bd_k1);
          c -= td * bd_k1;
          if (hs1) b += z0 * L[j].g1 * theta * dt;
          d -= otd * (amO_k * skm + (apO_k - amO) * sk - apO * skp) -
               otd * (bd_k * skm - (bd_k + bd_k1) * sk + bd_k1 * skp);
        }
        const double cnk = hs2 ? dt * L[j].g2 * o0 - hflux_in<NF, PRE>(A, fc, nfs, k, L[j])
                               : -hflux_in<NF, PRE>(A, fc, nfs, k, L[j]);
        d += cnk;
        A.Cn[base + (int64_t)k * 64] = cnk;
        b -= a * cprev / bprev;
        d -= a * dprev / bprev;
        put(j, base + (int64_t)k * 64, b, c, d);
        bprev = b; cprev = c; dprev = d;
        s0 = s1; s1 = s2; s2 = s3; s3 = L[j].sn;
        z0 = z1; z1 = L[j].zn;
        o0 = o1; o1 = L[j].on;
        q0 = q1; q1 = L[j].qn;
      }
    };
    if constexpr (KR > 0) {
      // rows 1..KR: unrolled, so that every head index is a constant
#pragma unroll
      for (int g0 = 0; g0 < KR; g0 += GRP) {
        if (1 + g0 > nb) break;
        group(1 + g0, [&](int j, int64_t, double b, double c, double d) {
          hb[g0 + j] = b;
          hc[g0 + j] = c;
          hd[g0 + j] = d;
        });
      }
    }
    for (int m0 = 1 + KR; m0 <= nb; m0 += GRP)
      group(m0, [&](int, int64_t idx, double b, double c, double d) {
        A.tb[idx] = b;
        A.tc[idx] = c;
        A.td[idx] = d;
      });
  }
  // Cn: zero above and at ktop, the folded top value on [ctop, ktop]
  for (int k = 0; k <= ktop && k < nk; k++) A.Cn[base + (int64_t)k * 64] = 0;
  for (int k = ct; k <= ktop && k < nk; k++) A.Cn[base + (int64_t)k * 64] = cn0 / (1 + abs(ct - ktop));
  // back substitution into scal[ktop..Nk-1]: the rows in memory (levels
  // klo..Nk-1), then those of the register head
  if (nb > 0) {
    const int klo = KR > 0 ? ktop + 1 + KR : ktop;
    double un = 0;
    if (nk - 1 >= klo) {
      un = A.td[base + (int64_t)(nk - 1) * 64] / A.tb[base + (int64_t)(nk - 1) * 64];
      A.s[base + (int64_t)(nk - 1) * 64] = un;
    }
    if (GRP > 1) {
      for (int k0 = nk - 2; k0 >= klo; k0 -= GRP) {
        double rb[GRP], rc[GRP], rd[GRP];
#pragma unroll
        for (int j = 0; j < GRP; j++) {
          const int k = k0 - j >= klo ? k0 - j : klo;
          const int64_t idx = base + (int64_t)k * 64;
          rb[j] = A.tb[idx];
          rc[j] = A.tc[idx];
          rd[j] = A.td[idx];
        }
#pragma unroll
        for (int j = 0; j < GRP; j++) {
          const int k = k0 - j;
          if (k < klo) break;
          un = rd[j] / rb[j] - rc[j] * un / rb[j];
          A.s[base + (int64_t)k * 64] = un;
        }
      }
    } else {
      for (int k = nk - 2; k >= klo; k--) {
        const int64_t idx = base + (int64_t)k * 64;
        const double bk = A.tb[idx];
        un = A.td[idx] / bk - A.tc[idx] * un / bk;
        A.s[idx] = un;
      }
    }
    if constexpr (KR > 0) {
#pragma unroll
      for (int j = KR - 1; j >= 0; j--) {
        const int k = ktop + 1 + j;
        if (k <= nk - 1) {
          un = k == nk - 1 ? hd[j] / hb[j] : hd[j] / hb[j] - hc[j] * un / hb[j];
          A.s[base + (int64_t)k * 64] = un;
        }
      }
      un = h0d / h0b - h0c * un / h0b;
      A.s[base + (int64_t)ktop * 64] = un;
    }
  } else if (A.step > 1) {
    const double b0 = KR > 0 ? h0b : A.tb[base + (int64_t)ktop * 64];
    const double d0 = KR > 0 ? h0d : A.td[base + (int64_t)ktop * 64];
    A.s[base + (int64_t)ktop * 64] = (b0 > 0 && A.active[s] != 0) ? d0 / b0 : 0;
  }
  for (int k = 0; k < ct; k++) A.s[base + (int64_t)k * 64] = 0;
  const double stop = A.s[base + (int64_t)ktop * 64];
  for (int k = ct; k < cto; k++) A.s[base + (int64_t)k * 64] = stop;
}

// Both passes of HorizontalFaceScalars (sums, then limited face values) for
// the arguments in A, on the grid's stream.
int face_scalars_launch(sunhip_grid *g, const TvdArgs &A);

}  // namespace sunhip

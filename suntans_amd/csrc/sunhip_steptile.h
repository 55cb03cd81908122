// sunhip_steptile.h -- per-tile bodies of the stencils either side of the
// pressure solve: the Poisson source (ComputeQSource, main/phys.c:2345-2395),
// the velocity and pressure correction (Corrector, main/phys.c:2301-2332) and
// the vertical velocity from continuity (Continuity, main/phys.c:3814-3844).
//
// One lane owns one column (cell tiles) or one edge (edge tiles); a launch
// covers max(Nt, Net) tiles and each lane does the cell part and/or the edge
// part its tile index has.  Every expression keeps the reference's operation
// order (the build has no FMA contraction), so the results are the
// reference's bit for bit.
#pragma once
#include "sunhip_qtile.h"

namespace sunhip {

struct StepArgs {
  double *u;              // edge field: corrected in place (Corrector)
  double *w;              // w-type cell field: read (source), corrected (Corrector), written (Continuity)
  double *q;              // cell field: nonhydrostatic pressure, accumulates qc (Corrector)
  const double *qc;       // cell field: pressure correction
  const double *utmp2;    // edge field: old horizontal velocity
  const double *wtmp2;    // w-type cell field: old vertical velocity
  const double *dzf;      // edge field
  const double *dzz;      // cell field
  double *src;            // cell field: Poisson source
  double *D;              // edge 2-D field (by edge slot): df/dg from the source step, read by Corrector
  double theta, dt;
};

// Level-invariant data of a column's faces.
template <int NF>
struct ColFaces {
  uint32_t eb[NF];
  int nke[NF];
  double nrm[NF], df[NF];
  __device__ __forceinline__ void load(const Geo &G, int s, int nf) {
    const int NS = G.NS;
#pragma unroll
    for (int f = 0; f < NF; f++) {
      if (f < nf) {
        const int es = G.es[f * NS + s];
        eb[f] = (uint32_t)G.eb[f * NS + s];
        nke[f] = G.e_nke[es];
        nrm[f] = (double)G.nrm[f * NS + s];
        df[f] = G.df[es];
      } else {
        eb[f] = 0; nke[f] = 0; nrm[f] = 0; df[f] = 0;
      }
    }
  }
};

// ComputeQSource: src over [ctop, Nk) of every column (zero outside the
// computational cells); D = df/dg on every edge.
template <int NF>
__device__ __forceinline__ void qsource_tile(const Geo &G, const StepArgs &A, int t, int l) {
  if (t < G.Net) {
    const int s = t * 64 + l;
    if (G.e_nkc[s] > 0) A.D[s] = G.df[s] / G.dg[s];
  }
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0) return;
  const int kt = G.ctop[s];
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  if (G.comp[s] != 1) {
    for (int k = kt; k < n; k++) at32(A.src, base + (uint32_t)k * 64u) = 0;
    return;
  }
  const int nf = G.nf[s];
  ColFaces<NF> F;
  F.load(G, s, nf);
  const double tf = (1 - A.theta) / A.theta, Ac = G.Ac[s], dt = A.dt;
  double wk = at32(A.w, base + (uint32_t)kt * 64u), w2k = at32(A.wtmp2, base + (uint32_t)kt * 64u);
  // four levels' loads and face terms first, then each level's sum in the
  // reference's order
  constexpr int GL = 4;
  for (int kb = kt; kb < n; kb += GL) {
    double w1[GL], w21[GL], tm[GL][NF];
#pragma unroll
    for (int j = 0; j < GL; j++) {
      const int k = kb + j;
      if (k >= n) break;
      const uint32_t k64 = (uint32_t)k * 64u, idx = base + k64;
      w1[j] = at32(A.w, idx + 64u);
      w21[j] = at32(A.wtmp2, idx + 64u);
#pragma unroll
      for (int f = 0; f < NF; f++) {
        tm[j][f] = 0;
        if (f < nf && k < F.nke[f]) {
          const uint32_t ei = F.eb[f] + k64;
          tm[j][f] = (at32(A.u, ei) + tf * at32(A.utmp2, ei)) * at32(A.dzf, ei) * F.nrm[f] * F.df[f];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < GL; j++) {
      const int k = kb + j;
      if (k >= n) break;
      double v = Ac * (wk - w1[j]) + tf * Ac * (w2k - w21[j]);
#pragma unroll
      for (int f = 0; f < NF; f++)
        if (f < nf && k < F.nke[f]) v += tm[j][f];
      at32(A.src, base + (uint32_t)k * 64u) = v / dt;
      wk = w1[j];
      w2k = w21[j];
    }
  }
}

// Corrector: u on computational interior edges with D != 0 and more than one
// layer; w of computational columns (surface ghost -qc[ctop]); q += qc on
// columns of more than one layer.
__device__ __forceinline__ void corrector_tile(const Geo &G, const StepArgs &A, int t, int l) {
  const double dt = A.dt;
  if (t < G.Net) {
    const int s = t * 64 + l;
    const int et = G.e_etop[s], nke = G.e_nke[s];
    if (G.e_comp[s] == 1 && A.D[s] != 0 && et < nke - 1) {
      const uint32_t base = (uint32_t)(G.eoff[t] + l), c1 = (uint32_t)G.e_c1[s], c2 = (uint32_t)G.e_c2[s];
      const double f = dt / G.dg[s];
      for (int k = et; k < nke; k++) {
        const uint32_t k64 = (uint32_t)k * 64u;
        at32(A.u, base + k64) -= f * (at32(A.qc, c1 + k64) - at32(A.qc, c2 + k64));
      }
    }
  }
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return;
  const int kt = G.ctop[s];
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  double qprev = at32(A.qc, base + (uint32_t)kt * 64u), dprev = at32(A.dzz, base + (uint32_t)kt * 64u);
  const double qtop = qprev, dtop = dprev;
  for (int k = kt + 1; k < n; k++) {
    const uint32_t idx = base + (uint32_t)k * 64u;
    const double qk = at32(A.qc, idx), dk = at32(A.dzz, idx);
    at32(A.w, idx) -= 2.0 * dt / (dprev + dk) * (qprev - qk);
    qprev = qk;
    dprev = dk;
  }
  at32(A.w, base + (uint32_t)kt * 64u) += 2.0 * dt / dtop * qtop;
  if (kt < n - 1)
    for (int k = kt; k < n; k++) {
      const uint32_t idx = base + (uint32_t)k * 64u;
      at32(A.q, idx) += at32(A.qc, idx);
    }
}

// Continuity: w of computational columns, bottom up from w[Nk] = 0; zero
// above ctop.
template <int NF>
__device__ __forceinline__ void continuity_tile(const Geo &G, const StepArgs &A, int t, int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return;
  const int kt = G.ctop[s], nf = G.nf[s];
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  ColFaces<NF> F;
  F.load(G, s, nf);
  const double th = A.theta, Ac = G.Ac[s], tf = (1 - th) / th;
  for (int k = 0; k < kt; k++) at32(A.w, base + (uint32_t)k * 64u) = 0;
  at32(A.w, base + (uint32_t)n * 64u) = 0;
  double wk1 = 0, w2k1 = at32(A.wtmp2, base + (uint32_t)n * 64u);
  // four levels' gathers and face terms first (independent of the upward
  // chain), then the chain subtracts them in the reference's order
  constexpr int GL = 4;
  for (int kb = n - 1; kb >= kt; kb -= GL) {
    double w2[GL], tm[GL][NF];
#pragma unroll
    for (int j = 0; j < GL; j++) {
      const int k = kb - j;
      if (k < kt) break;
      const uint32_t k64 = (uint32_t)k * 64u;
      w2[j] = at32(A.wtmp2, base + k64);
#pragma unroll
      for (int f = 0; f < NF; f++) {
        tm[j][f] = 0;
        if (f < nf && k < F.nke[f]) {
          const uint32_t ei = F.eb[f] + k64;
          tm[j][f] = (th * at32(A.u, ei) + (1 - th) * at32(A.utmp2, ei)) * F.df[f] * F.nrm[f] / Ac / th *
                     at32(A.dzf, ei);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < GL; j++) {
      const int k = kb - j;
      if (k < kt) break;
      double v = wk1 - tf * (w2[j] - w2k1);
#pragma unroll
      for (int f = 0; f < NF; f++)
        if (f < nf && k < F.nke[f]) v -= tm[j][f];
      at32(A.w, base + (uint32_t)k * 64u) = v;
      wk1 = v;
      w2k1 = w2[j];
    }
  }
}

}  // namespace sunhip

// sunhip_qtile.h -- the column arithmetic of the pressure solve, one tile
// (64 columns, one wavefront) at a time.
//
// Reference: QCoefficients main/phys.c:3420-3471, OperatorQ :3486-3536,
// Preconditioner :3577-3619 with TriSolve main/util.c:185-203, the CG
// scalar updates of CGSolveQ main/phys.c:2452-2524.
//
// These device functions are shared by the single-grid kernels
// (sunhip_q.hip) and the batched kernels that walk a list of (member, tile)
// work items over many grids (sunhip_batch.hip), so both run the same
// operations in the same order.  Element-wise arithmetic keeps the
// reference's operation order (built without FMA contraction), so the
// operator and the preconditioner agree with the reference bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "sunhip_common.h"

namespace sunhip {

enum { QM_PLAIN = 0, QM_RESID = 1, QM_FUSED = 2, QM_RESID_DOT = 3 };

// The levels of a z-level grid whose operator coefficients follow from the
// reference layer thicknesses dz[k] (grid->dz).  Where the layers have one
// thickness (dz[k] + dz[k-1] the same for every k >= 1), QCoefficients'
// coef[k] = 2 Ac/(dz[k] + dz[k-1]) is one value per column below its top
// cell, and a face height dzf[k] is dz[k].  A solve's set-up finds, per
// column and per face, the window of levels where the stored value equals
// that expression bit for bit, with at most two levels above and two below
// it (the surface and bottom cells) for a column and one above and one below
// for a face; those exceptions are copied per column.  A tile whose columns
// all fit (tflag) reads no coef and no dzf: its kernels form them from the
// window, the exceptions and dz.  Other tiles read the stored arrays.  The
// results are the same bits either way.
struct ZReg {
  const int *cwin;     // [NS] c0 | c1 << 16: coef window of the column
  const int *fwin;     // [NF*NS] f0 | f1 << 16: dzf window of face f (row f*NS)
  const double *cex;   // [4*NS] coef at ctop, ctop+1, Nk-2, Nk-1 (row i*NS)
  const double *fex;   // [2*NF*NS] dzf at the face's first and last level (rows f*NS, (NF+f)*NS)
  const int *tflag;    // [Nt] 1: every computational column of the tile fits its windows
  const double *dzt;   // dz[k], zero past Nkmax
  double den;          // the common dz[k] + dz[k-1] (0: the layers differ)
};

// coef of a column of a regular tile: the interior value inside [c0, c1),
// the copied exceptions outside it.
template <typename T>
__device__ __forceinline__ T zsel(int k, int kt, int n, int c0, int c1, T ci, T e0, T e1, T e2, T e3) {
  const T lo = k == kt ? e0 : e1, hi = k == n - 1 ? e3 : e2;
  return k < c0 ? lo : (k < c1 ? ci : hi);
}
template <typename T>
struct ZCol {
  int kt, n, c0, c1;
  T ci, e0, e1, e2, e3;
  __device__ __forceinline__ T at(int k) const { return zsel<T>(k, kt, n, c0, c1, ci, e0, e1, e2, e3); }
};
template <typename T>
__device__ __forceinline__ ZCol<T> zcol(const ZReg &Z, const Geo &G, int s, int kt, int n) {
  ZCol<T> c;
  const int w = Z.cwin[s], NS = G.NS;
  c.kt = kt; c.n = n; c.c0 = w & 0xffff; c.c1 = w >> 16;
  c.ci = (T)(2 * G.Ac[s] / Z.den);  // QCoefficients' 2*Ac/(dzz[k]+dzz[k-1]) on the window
  c.e0 = (T)Z.cex[s]; c.e1 = (T)Z.cex[NS + s]; c.e2 = (T)Z.cex[2 * NS + s]; c.e3 = (T)Z.cex[3 * NS + s];
  return c;
}
// dzf of a face of a regular tile at level k of the face's range
template <typename T>
struct ZFace {
  int f0, f1;
  T top, bot;
  __device__ __forceinline__ T at(int k, const double *dzt) const { return k < f0 ? top : (k < f1 ? (T)dzt[k] : bot); }
};
template <typename T, int NF>
__device__ __forceinline__ ZFace<T> zface(const ZReg &Z, const Geo &G, int s, int f) {
  ZFace<T> z;
  const int NS = G.NS, w = Z.fwin[f * NS + s];
  z.f0 = w & 0xffff; z.f1 = w >> 16;
  z.top = (T)Z.fex[f * NS + s]; z.bot = (T)Z.fex[(NF + f) * NS + s];
  return z;
}

// finish operations on the CG scalars
enum { OP_NONE = 0, OP_APPLY = 1, OP_PREC_SETUP = 2, OP_PREC_ITER = 3, OP_NOPREC_SETUP = 4, OP_NOPREC_ITER = 5,
       OP_STORE = 6 };

template <typename T, typename TO = T>
struct QVec {
  const T *coef, *dzf, *D;
  const T *vin;            // PLAIN / RESID: operand
  const T *src;            // RESID: right-hand side
  const T *rtmp, *pold;    // FUSED: p = rtmp + mu*pold
  T *pnew, *x;             // FUSED
  TO *out;                 // L p, or src - L x
  const T *fcoef;          // qprecond=1: face coefficients, maxfaces rows of the cell layout
  ZReg Z;                  // z-level windows (the ZR kernels)
};

__device__ inline void q_op(CGScal *sc, int op, const double *tot) {
  switch (op) {
    case OP_APPLY:  // mu = 1/alpha ; nu = alpha/(p.z) (main/phys.c:2484-2485)
      sc->pz = tot[0];
      sc->muinv = 1 / sc->alpha;
      sc->nu = sc->alpha / tot[0];
      break;
    case OP_PREC_SETUP:  // main/phys.c:2452-2475 with qprecond=2
    case OP_NOPREC_SETUP:
      sc->alpha = sc->alpha0 = op == OP_PREC_SETUP ? tot[0] : tot[1];
      if (!sc->resnorm) sc->alpha0 = 1;
      sc->eps = sc->eps0 = op == OP_PREC_SETUP ? tot[1] : sc->alpha0;
      sc->mu = 0; sc->nu = 0; sc->iter = 0; sc->nloop = 0; sc->converged = 0;
      sc->done = (sc->eps == 0 || sc->maxit <= 0);
      sc->resid = sqrt(sc->eps / sc->eps0);
      break;
    case OP_PREC_ITER:  // main/phys.c:2496-2524
    case OP_NOPREC_ITER: {
      const double alpha = op == OP_PREC_ITER ? tot[0] : tot[1];
      sc->mu = sc->muinv * alpha;
      sc->alpha = alpha;
      sc->eps = op == OP_PREC_ITER ? tot[1] : alpha;
      sc->iter += 1;
      sc->resid = sqrt(sc->eps / sc->eps0);
      if (sc->resid < sc->tol) { sc->done = 1; sc->converged = 1; sc->nloop = sc->iter - 1; }
      else if (sc->iter >= sc->maxit || sc->eps == 0) { sc->done = 1; sc->nloop = sc->iter; }
      break;
    }
    case OP_STORE:  // keep the sum (outer residual norm of the mixed form)
      sc->pz = tot[0];
      break;
  }
}

// QCoefficients, unpreconditioned branch (main/phys.c:3456-3469), every cell.
__device__ __forceinline__ void qcoef_tile(const Geo &G, const double *__restrict__ dzz, double *__restrict__ coef,
                                           int t, int l) {
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

// The modified diagonal b' of the Thomas elimination of every column
// (main/util.c:193-196 applied to the tridiagonal of main/phys.c:3584-3601).
// It depends only on coef, so it is built once per solve; the back
// substitution of every application reads it.
template <typename T>
__device__ __forceinline__ void qfactor_tile(const Geo &G, const T *__restrict__ coef, T *__restrict__ bq, int t,
                                             int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s], kt = G.ctop[s];
  if (n == 0 || G.comp[s] != 1 || kt >= n - 1) return;
  const int64_t base = G.toff[t] + l;
  T cprev = 0, bprev = 1;
  T ck = coef[base + (int64_t)kt * 64];
  for (int k = kt; k < n; k++) {
    const int64_t idx = base + (int64_t)k * 64;
    const T ck1 = k < n - 1 ? coef[idx + 64] : T(0);
    T a = 0, b, c = 0;
    if (k == kt) { b = -2 * ck - ck1; c = ck1; }
    else if (k == n - 1) { a = ck; b = -ck; }
    else { a = ck; b = -ck - ck1; c = ck1; }
    if (k > kt) b -= a * cprev / bprev;
    bq[idx] = b;
    cprev = c; bprev = b; ck = ck1;
  }
}

// Element i of a field through a 32-bit byte offset (every field is smaller
// than 4 GiB, which the grid context checks), so an access is a base
// address in scalar registers plus one 32-bit vector register.
template <typename T>
__device__ __forceinline__ const T &at32(const T *p, uint32_t i) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T &at32(T *p, uint32_t i) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + i * (uint32_t)sizeof(T));
}

// Store of a streamed result (NTS: with the non-temporal hint).
template <bool NTS, typename T>
__device__ __forceinline__ void st_stream(T *p, T v) {
  if constexpr (NTS) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// y = L v over the computational columns of tile t (main/phys.c:3486-3536),
// in one of four modes:
//   PLAIN     out = L vin
//   RESID     out = src - L vin
//   FUSED     p = rtmp + mu*pold -> pnew ; x += nu*pold ; out = L p ; returns p.(L p)
//   RESID_DOT out = (TO)(src - L vin) ; returns the residual's r.r
// NF faces per cell at most; U levels per load group; NTS stores the FUSED
// outputs with the non-temporal hint.  Returns the lane's share of the dot
// product.

template <int MODE, int NF, int U, typename T, typename TO, bool NTS = false, bool ZR = false>
__device__ __forceinline__ double qapply_tile(const Geo &G, const QVec<T, TO> &V, T mu, T nu, int t, int l,
                                              int k0 = 0, int k1 = 0x7fffffff) {
  double dot = 0;
  if (t >= G.Nt) return 0;
  const T *__restrict__ coef = V.coef;
  const T *__restrict__ dzf = V.dzf;
  const T *__restrict__ rt = V.rtmp;
  const T *__restrict__ po = V.pold;
  const T *__restrict__ vin = V.vin;
  const int s = t * 64 + l;
  const int n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return 0;
  const int kt = G.ctop[s];
  const int NS = G.NS;
  const int nf = G.nf[s];
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  uint32_t nbb[NF], ebb[NF];
  int kmn[NF], kmx[NF], lim[NF];
  T Df[NF];
  ZFace<T> zf[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    const int nbf = f < nf ? G.nbb[f * NS + s] : -1;
    if (nbf >= 0) {
      nbb[f] = (uint32_t)nbf;
      ebb[f] = (uint32_t)G.eb[f * NS + s];
      const int kr = G.kr[f * NS + s];
      kmn[f] = kr & 0xffff;
      kmx[f] = kr >> 16;
      Df[f] = V.D[G.es[f * NS + s]];
    } else {
      nbb[f] = base; ebb[f] = 0; kmn[f] = 0; kmx[f] = 0; Df[f] = 0;
    }
    if (ZR) zf[f] = zface<T, NF>(V.Z, G, s, f);
    lim[f] = kmx[f] > 0 ? kmx[f] - 1 : 0;
  }
  const ZCol<T> zc = ZR ? zcol<T>(V.Z, G, s, kt, n) : ZCol<T>{};
  // levels [kbeg, kend) of the column: all of it, or one band [k0, k1)
  const int kbeg = max(kt, k0), kend = min(n, k1);
  T vprev = 0;
  if (kbeg > kt && kbeg < kend) {  // a band below the top: the level above it
    const uint32_t idx = base + (uint32_t)(kbeg - 1) * 64u;
    vprev = MODE == QM_FUSED ? at32(rt, idx) + mu * at32(po, idx) : at32(vin, idx);
  }
  for (int kb = kbeg; kb < kend; kb += U) {
    T v[U + 1], c[U + 1], pv[U + 1], xv[U], nv[U][NF], dz[U][NF], dzg[U];
    if (ZR) {
#pragma unroll
      for (int j = 0; j < U; j++) dzg[j] = (T)V.Z.dzt[kb + j];  // the group's layer thicknesses (uniform)
    }
#pragma unroll
    for (int j = 0; j <= U; j++) {
      const uint32_t idx = base + (uint32_t)min(kb + j, n) * 64u;
      if (MODE == QM_FUSED) { pv[j] = at32(po, idx); v[j] = at32(rt, idx); }
      else { pv[j] = 0; v[j] = at32(vin, idx); }
      if (!ZR) c[j] = at32(coef, idx);  // ZR: formed where used
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int k = kb + j;
      xv[j] = MODE == QM_FUSED ? at32((const T *)V.x, base + (uint32_t)min(k, n) * 64u) : T(0);
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const uint32_t kk = (uint32_t)min(k, lim[f]) * 64u;
        const uint32_t ni = nbb[f] + kk;
        nv[j][f] = MODE == QM_FUSED ? at32(rt, ni) + mu * at32(po, ni) : at32(vin, ni);
        if (!ZR) dz[j][f] = at32(dzf, ebb[f] + kk);
      }
    }
    if (MODE == QM_FUSED) {
#pragma unroll
      for (int j = 0; j <= U; j++) v[j] = v[j] + mu * pv[j];
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int k = kb + j;
      if (k < kend) {
        const T v0 = v[j], vm1 = j ? v[j - 1] : vprev, vp1 = v[j + 1];
        const T c0 = ZR ? zc.at(k) : c[j], cp1 = ZR ? zc.at(k + 1) : c[j + 1];
        T y = 0;
#pragma unroll
        for (int f = 0; f < NF; f++) {
          const T dzk = ZR ? (k < zf[f].f0 ? zf[f].top : (k < zf[f].f1 ? dzg[j] : zf[f].bot)) : dz[j][f];
          const T term = (nv[j][f] - v0) * dzk * Df[f];
          y = (k >= kmn[f] && k < kmx[f]) ? y + term : y;
        }
        if (kt == n - 1) y -= T(2.0) * c0 * v0;
        else if (k == kt) y += (-2 * c0 - cp1) * v0 + cp1 * vp1;
        else if (k == n - 1) y += c0 * vm1 - c0 * v0;
        else y += c0 * vm1 - (c0 + cp1) * v0 + cp1 * vp1;
        const uint32_t idx = base + (uint32_t)k * 64u;
        if (MODE == QM_FUSED) {
          st_stream<NTS>(&at32(V.pnew, idx), v0);
          st_stream<NTS>(&at32(V.x, idx), xv[j] + nu * pv[j]);
          dot += (double)v0 * (double)y;
          st_stream<NTS, TO>(&at32(V.out, idx), (TO)y);
        } else if (MODE == QM_RESID) {
          at32(V.out, idx) = at32(V.src, idx) - y;
        } else if (MODE == QM_RESID_DOT) {
          const T rk = at32(V.src, idx) - y;
          at32(V.out, idx) = (TO)rk;
          dot += (double)rk * (double)rk;
        } else {
          at32(V.out, idx) = y;
        }
      }
    }
    vprev = v[U - 1];
  }
  return dot;
}

// ---- qprecond=1: the symmetrically scaled operator ------------------------
// The face levels [kmn, kmx) and the face's D of every face of slot s with a
// neighbour (nbb >= 0); kmx = 0 for the other faces.
template <int NF>
__device__ __forceinline__ void qc_faces(const Geo &G, const double *__restrict__ D, int s, uint32_t (&nbb)[NF],
                                         int (&kmn)[NF], int (&kmx)[NF], double (&Df)[NF]) {
  const int NS = G.NS, nf = G.nf[s];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    const int nb = f < nf ? G.nbb[f * NS + s] : -1;
    if (nb >= 0) {
      const int kr = G.kr[f * NS + s];
      nbb[f] = (uint32_t)nb; kmn[f] = kr & 0xffff; kmx[f] = kr >> 16;
      Df[f] = D[G.es[f * NS + s]];
    } else {
      nbb[f] = 0; kmn[f] = 0; kmx[f] = 0; Df[f] = 0;
    }
  }
}

// ConditionQ (main/phys.c:3629-3681) of tile t's computational columns:
// c = sqrt of the diagonal of L (faces dzz*D in face order, then the
// vertical coefficients as QCoefficients forms them), 1 where the diagonal
// is not positive.
template <int NF>
__device__ __forceinline__ void qcond_tile(const Geo &G, const double *__restrict__ dzz, const double *__restrict__ D,
                                           double *__restrict__ c, int t, int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return;
  const int kt = G.ctop[s];
  const double Ac = G.Ac[s];
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  uint32_t nbb[NF];
  int kmn[NF], kmx[NF];
  double Df[NF];
  qc_faces<NF>(G, D, s, nbb, kmn, kmx, Df);
  auto Z = [&](int k) { return at32(dzz, base + (uint32_t)k * 64u); };
  // a[k] of main/phys.c:3667-3669
  auto A = [&](int k) { return k == kt ? Ac / Z(kt) : 2 * Ac / (Z(k) + Z(k - 1)); };
  for (int k = kt; k < n; k++) {
    const double zk = Z(k);
    double x = 0;
#pragma unroll
    for (int f = 0; f < NF; f++)
      if (k >= kmn[f] && k < kmx[f]) x += zk * Df[f];
    if (kt < n - 1) {
      if (k == kt) x += 2 * A(kt) + A(kt + 1);
      else if (k == n - 1) x += A(n - 1);
      else x += (A(k) + A(k + 1));
    }
    if (x <= 0) x = 1;
    at32(c, base + (uint32_t)k * 64u) = sqrt(x);
  }
}

// QCoefficients with qprecond=1 (main/phys.c:3426-3454): coef scaled by the
// conditioning c of the two levels it couples, the face coefficients
// dzz*D/(c c_neighbour) into fc (NF rows of the cell layout, row f at
// f*NT).  Every face coefficient of a level outside the face's range is 0.
template <int NF>
__device__ __forceinline__ void qcoef_pc1_tile(const Geo &G, const double *__restrict__ dzz,
                                               const double *__restrict__ D, const double *__restrict__ c,
                                               double *__restrict__ coef, double *__restrict__ fc, int64_t NT,
                                               int t, int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return;
  const int kt = G.ctop[s];
  const double Ac = G.Ac[s];
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  uint32_t nbb[NF];
  int kmn[NF], kmx[NF];
  double Df[NF];
  qc_faces<NF>(G, D, s, nbb, kmn, kmx, Df);
  auto Z = [&](int k) { return at32(dzz, base + (uint32_t)k * 64u); };
  auto Cc = [&](int k) { return at32(c, base + (uint32_t)k * 64u); };
  at32(coef, base + (uint32_t)kt * 64u) = Ac / Z(kt) / Cc(kt);
  for (int k = kt + 1; k < n; k++) at32(coef, base + (uint32_t)k * 64u) = 2 * Ac / (Z(k) + Z(k - 1)) / (Cc(k) * Cc(k - 1));
#pragma unroll
  for (int f = 0; f < NF; f++) {
    double *row = fc + (int64_t)f * NT;
    for (int k = 0; k < n; k++) {
      const uint32_t idx = base + (uint32_t)k * 64u;
      at32(row, idx) = (k >= kmn[f] && k < kmx[f]) ? Z(k) * Df[f] / (Cc(k) * at32(c, nbb[f] + (uint32_t)k * 64u))
                                                  : 0.0;
    }
  }
}

// OperatorQC (main/phys.c:3364-3409) over tile t in the modes of
// qapply_tile (PLAIN, RESID, FUSED): y = -v, plus v_neighbour * fcoef over
// the faces in face order, plus the scaled vertical neighbours (top: the one
// below; bottom: the one above).  In FUSED mode v = rtmp + mu*p_old is
// formed on the fly for the column and its neighbours, p and x are written
// as in qapply_tile and the lane's p.(L p) is returned.
template <int MODE, int NF>
__device__ __forceinline__ double qcapply_tile(const Geo &G, const QVec<double, double> &V, double mu, double nu,
                                               int64_t NT, int t, int l) {
  if (t >= G.Nt) return 0;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return 0;
  const int kt = G.ctop[s];
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  uint32_t nbb[NF];
  int kmn[NF], kmx[NF];
  double Df[NF];
  qc_faces<NF>(G, V.D, s, nbb, kmn, kmx, Df);
  auto val = [&](uint32_t idx) {
    return MODE == QM_FUSED ? at32(V.rtmp, idx) + mu * at32(V.pold, idx) : at32(V.vin, idx);
  };
  double dot = 0;
  for (int k = kt; k < n; k++) {
    const uint32_t idx = base + (uint32_t)k * 64u;
    const double v0 = val(idx);
    double y = -v0;
#pragma unroll
    for (int f = 0; f < NF; f++)
      if (k >= kmn[f] && k < kmx[f]) y += val(nbb[f] + (uint32_t)k * 64u) * at32(V.fcoef + (int64_t)f * NT, idx);
    if (kt < n - 1) {
      if (k == kt) y += at32(V.coef, idx + 64u) * val(idx + 64u);
      else if (k == n - 1) y += at32(V.coef, idx) * val(idx - 64u);
      else y += at32(V.coef, idx) * val(idx - 64u) + at32(V.coef, idx + 64u) * val(idx + 64u);
    }
    if (MODE == QM_FUSED) {
      const double po = at32(V.pold, idx);
      at32(V.pnew, idx) = v0;
      at32(V.x, idx) = at32((const double *)V.x, idx) + nu * po;
      at32(V.out, idx) = y;
      dot += v0 * y;
    } else if (MODE == QM_RESID) {
      at32(V.out, idx) = at32(V.src, idx) - y;
    } else {
      at32(V.out, idx) = y;
    }
  }
  return dot;
}

// Element-wise a = a op b on tile t's computational levels: op 0 a /= b,
// op 1 a *= b (the scalings of main/phys.c:2425-2433, 2539-2546).
__device__ __forceinline__ void qscale_tile(const Geo &G, double *__restrict__ a, const double *__restrict__ b, int op,
                                            int t, int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return;
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  for (int k = G.ctop[s]; k < n; k++) {
    const uint32_t idx = base + (uint32_t)k * 64u;
    at32(a, idx) = op == 0 ? at32(a, idx) / at32(b, idx) : at32(a, idx) * at32(b, idx);
  }
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

// The FUSED mode of qapply_tile for the four consecutive tiles t0..t0+3 of
// one 256-thread block, walking the levels together.  Each column forms
// p = rtmp + mu*p_old once per level and stages it in LDS (double-buffered,
// one barrier per group of U levels); a face whose neighbour column lies in
// the block reads its p there, and only neighbours outside the block are
// gathered from memory.  Every value is formed by the same operations in the
// same order as in qapply_tile, so the results are bit-identical.  Must be
// called by all 256 threads of the block.
template <int NF, int U, typename T, bool NTS>
__device__ __forceinline__ double qapply_block(const Geo &G, const QVec<T, T> &V, T mu, T nu, int t0) {
  constexpr int BT = 4;
  __shared__ T pl[2][U][BT * 64];
  __shared__ int bnd[2][BT];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int t = t0 + w;
  const int s = t * 64 + l;
  int n = 0, kt = 0;
  bool act = false;
  if (t < G.Nt) {
    n = G.nk[s];
    act = n > 0 && G.comp[s] == 1;
    kt = act ? G.ctop[s] : 0;
  }
  {
    const int lo = wave_min_i(act ? kt : 0x7fffffff), hi = wave_max_i(act ? n : 0);
    if (l == 0) { bnd[0][w] = lo; bnd[1][w] = hi; }
  }
  __syncthreads();
  const int kb0 = min(min(bnd[0][0], bnd[0][1]), min(bnd[0][2], bnd[0][3]));
  const int kb1 = max(max(bnd[1][0], bnd[1][1]), max(bnd[1][2], bnd[1][3]));
  if (kb1 <= kb0) return 0;  // no computational column in the block (uniform)
  const T *__restrict__ coef = V.coef;
  const T *__restrict__ dzf = V.dzf;
  const T *__restrict__ rt = V.rtmp;
  const T *__restrict__ po = V.pold;
  const int NS = G.NS;
  const uint32_t base = act ? (uint32_t)(G.toff[t] + l) : 0u;
  uint32_t nbb[NF], ebb[NF];
  int kmn[NF], kmx[NF], lim[NF], loc[NF];
  T Df[NF];
  const int nf = act ? G.nf[s] : 0;
#pragma unroll
  for (int f = 0; f < NF; f++) {
    const int nbf = f < nf ? G.nbb[f * NS + s] : -1;
    if (nbf >= 0) {
      nbb[f] = (uint32_t)nbf;
      ebb[f] = (uint32_t)G.eb[f * NS + s];
      const int kr = G.kr[f * NS + s];
      kmn[f] = kr & 0xffff;
      kmx[f] = kr >> 16;
      Df[f] = V.D[G.es[f * NS + s]];
      const int rel = G.nbs[f * NS + s] - t0 * 64;
      loc[f] = (rel >= 0 && rel < BT * 64) ? rel : -1;
    } else {
      nbb[f] = base; ebb[f] = 0; kmn[f] = 0; kmx[f] = 0; Df[f] = 0; loc[f] = -1;
    }
    lim[f] = kmx[f] > 0 ? kmx[f] - 1 : 0;
  }
  double dot = 0;
  T vprev = 0;
  int buf = 0;
  for (int kb = kb0; kb < kb1; kb += U, buf ^= 1) {
    T v[U + 1], c[U + 1], pv[U + 1], xv[U], nv[U][NF], dz[U][NF];
    if (act) {
#pragma unroll
      for (int j = 0; j <= U; j++) {
        const uint32_t idx = base + (uint32_t)min(kb + j, n) * 64u;
        pv[j] = at32(po, idx);
        v[j] = at32(rt, idx);
        c[j] = at32(coef, idx);
      }
#pragma unroll
      for (int j = 0; j < U; j++) xv[j] = at32((const T *)V.x, base + (uint32_t)min(kb + j, n) * 64u);
#pragma unroll
      for (int j = 0; j <= U; j++) v[j] = v[j] + mu * pv[j];
    } else {
#pragma unroll
      for (int j = 0; j <= U; j++) { pv[j] = 0; v[j] = 0; c[j] = 0; }
#pragma unroll
      for (int j = 0; j < U; j++) xv[j] = 0;
    }
#pragma unroll
    for (int j = 0; j < U; j++) pl[buf][j][threadIdx.x] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int k = kb + j;
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const uint32_t kk = (uint32_t)min(k, lim[f]) * 64u;
        if (loc[f] >= 0) {
          nv[j][f] = pl[buf][j][loc[f]];
        } else if (f < nf) {
          const uint32_t ni = nbb[f] + kk;
          nv[j][f] = at32(rt, ni) + mu * at32(po, ni);
        } else {
          nv[j][f] = 0;
        }
        dz[j][f] = f < nf ? at32(dzf, ebb[f] + kk) : T(0);
      }
    }
    if (act) {
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = kb + j;
        if (k >= kt && k < n) {
          const T v0 = v[j], vm1 = j ? v[j - 1] : vprev, vp1 = v[j + 1];
          const T c0 = c[j], cp1 = c[j + 1];
          T y = 0;
#pragma unroll
          for (int f = 0; f < NF; f++) {
            const T term = (nv[j][f] - v0) * dz[j][f] * Df[f];
            y = (k >= kmn[f] && k < kmx[f]) ? y + term : y;
          }
          if (kt == n - 1) y -= T(2.0) * c0 * v0;
          else if (k == kt) y += (-2 * c0 - cp1) * v0 + cp1 * vp1;
          else if (k == n - 1) y += c0 * vm1 - c0 * v0;
          else y += c0 * vm1 - (c0 + cp1) * v0 + cp1 * vp1;
          const uint32_t idx = base + (uint32_t)k * 64u;
          st_stream<NTS>(&at32(V.pnew, idx), v0);
          st_stream<NTS>(&at32(V.x, idx), xv[j] + nu * pv[j]);
          dot += (double)v0 * (double)y;
          st_stream<NTS>(&at32(V.out, idx), y);
        }
      }
    }
    vprev = v[U - 1];
  }
  return dot;
}

// The FUSED mode of qapply_tile with the neighbour values exchanged inside
// the wave.  The lanes of tile t walk the levels together (the tile's range of
// computational levels); per group of U levels each lane writes its own
// p = rtmp + mu*p_old to the wave's exchange row in LDS, the first lanes
// gather the tile's cross-tile neighbours (the compact pair list, one load per
// 64 pairs and level instead of one per face), and every face reads its
// neighbour's p from the row.  The row is private to the wave and rewritten
// only after the wave's own reads (LDS operations of a wave complete in
// order), so no barrier is needed.  Columns that are not computational but
// have layers still publish their p.  Values are formed by the same
// operations in the same order as in qapply_tile: bit-identical results.
template <int NF, int U, typename T, bool NTS, bool ZR = false>
__device__ __forceinline__ double qapply_xchg(const Geo &G, const QVec<T, T> &V, T mu, T nu, int t, int l,
                                              T (*row)[SUNHIP_TILE + SUNHIP_XPAIRS]) {
  if (t >= G.Nt) return 0;
  const T *__restrict__ coef = V.coef;
  const T *__restrict__ dzf = V.dzf;
  const T *__restrict__ rt = V.rtmp;
  const T *__restrict__ po = V.pold;
  const int s = t * 64 + l, NS = G.NS;
  const int n = G.nk[s];
  const bool own = n > 0 && G.comp[s] == 1;
  const int kt = n > 0 ? G.ctop[s] : 0;
  const int kb0 = __builtin_amdgcn_readfirstlane(wave_min_i(own ? kt : 0x7fffffff));
  const int kb1 = __builtin_amdgcn_readfirstlane(wave_max_i(own ? n : 0));
  if (kb1 <= kb0) return 0;  // no computational column in the tile (uniform)
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  const int nf = n > 0 ? G.nf[s] : 0;
  // per face: the exchange code, the face's level range [kmn, kmx) and (ZR)
  // its window [f0, f1) packed as 16-bit pairs, unpacked where compared, so
  // the per-face state fits the register budget
  int code[NF], krp[NF], lim[NF], zwp[NF];
  uint32_t ebb[NF];
  T Df[NF], ztop[NF], zbot[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    code[f] = f < nf ? G.xcode[f * NS + s] : -1;
    // the window kernel leaves faces without a neighbour (and face slots past
    // nf, and columns that are not computational) at an empty window
    if (ZR) {
      const ZFace<T> z = zface<T, NF>(V.Z, G, s, f);
      zwp[f] = z.f0 | (z.f1 << 16);
      ztop[f] = z.top;
      zbot[f] = z.bot;
    }
    if (f < nf) {
      krp[f] = G.kr[f * NS + s];
      ebb[f] = (uint32_t)G.eb[f * NS + s];
      Df[f] = G.nbb[f * NS + s] >= 0 ? V.D[G.es[f * NS + s]] : T(0);
    } else {
      krp[f] = 0; ebb[f] = 0; Df[f] = 0;
    }
    if (code[f] < 0) krp[f] = 0;
    const int kmx = krp[f] >> 16;
    lim[f] = kmx > 0 ? kmx - 1 : 0;
  }
  // every slot has its (possibly empty) window; plain locals, so they stay in registers
  int zc0 = 0, zc1 = 0;
  T zci = 0, ze0 = 0, ze1 = 0, ze2 = 0, ze3 = 0;
  if (ZR) {
    const ZCol<T> zc = zcol<T>(V.Z, G, s, kt, n);
    zc0 = zc.c0; zc1 = zc.c1; zci = zc.ci; ze0 = zc.e0; ze1 = zc.e1; ze2 = zc.e2; ze3 = zc.e3;
  }
  // cross-tile pairs this lane gathers: pair l and pair 64 + l
  const int np = G.npair[t];
  uint32_t pnb[2];
  int plim[2];
  bool pact[2];
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int pi = r * 64 + l;
    pact[r] = pi < np;
    if (pact[r]) {
      const int fs = G.pair[(size_t)t * SUNHIP_XPAIRS + pi];
      pnb[r] = (uint32_t)G.nbb[fs];
      const int mx = G.kr[fs] >> 16;
      plim[r] = mx > 0 ? mx - 1 : 0;
    } else {
      pnb[r] = base; plim[r] = 0;
    }
  }
  const bool two = np > 64;
  const int nn = n > 0 ? n : 0;
  double dot = 0;
  T vprev = 0;
  for (int kb = kb0; kb < kb1; kb += U) {
    T v[U + 1], c[U + 1], pv[U + 1], xv[U], nv[U][NF], dz[U][NF], dzg[U];
    if (ZR) {
#pragma unroll
      for (int j = 0; j < U; j++) dzg[j] = (T)V.Z.dzt[kb + j];  // the group's layer thicknesses (uniform)
    }
#pragma unroll
    for (int j = 0; j <= U; j++) {
      const uint32_t idx = base + (uint32_t)min(kb + j, nn) * 64u;
      pv[j] = at32(po, idx);
      v[j] = at32(rt, idx);
      if (!ZR) c[j] = at32(coef, idx);  // ZR: formed where used
    }
#pragma unroll
    for (int j = 0; j < U; j++) xv[j] = at32((const T *)V.x, base + (uint32_t)min(kb + j, nn) * 64u);
    // cross-tile values loaded raw and combined after the group's loads,
    // so no load of the group waits for another
    T gr0[U], gp0[U], gr1[U], gp1[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int k = kb + j;
      gr0[j] = 0; gp0[j] = 0; gr1[j] = 0; gp1[j] = 0;
      if (pact[0]) {
        const uint32_t ni = pnb[0] + (uint32_t)min(k, plim[0]) * 64u;
        gr0[j] = at32(rt, ni); gp0[j] = at32(po, ni);
      }
      if (two && pact[1]) {
        const uint32_t ni = pnb[1] + (uint32_t)min(k, plim[1]) * 64u;
        gr1[j] = at32(rt, ni); gp1[j] = at32(po, ni);
      }
#pragma unroll
      for (int f = 0; f < NF; f++)
        if (!ZR) dz[j][f] = f < nf ? at32(dzf, ebb[f] + (uint32_t)min(k, lim[f]) * 64u) : T(0);
    }
#pragma unroll
    for (int j = 0; j <= U; j++) v[j] = v[j] + mu * pv[j];
#pragma unroll
    for (int j = 0; j < U; j++) {
      row[j][l] = v[j];
      row[j][64 + l] = pact[0] ? gr0[j] + mu * gp0[j] : T(0);
      if (two) row[j][128 + l] = pact[1] ? gr1[j] + mu * gp1[j] : T(0);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int j = 0; j < U; j++)
#pragma unroll
      for (int f = 0; f < NF; f++) nv[j][f] = code[f] >= 0 ? row[j][code[f]] : T(0);
    __builtin_amdgcn_wave_barrier();
    if (own) {
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = kb + j;
        if (k >= kt && k < n) {
          const T v0 = v[j], vm1 = j ? v[j - 1] : vprev, vp1 = v[j + 1];
          const T c0 = ZR ? zsel<T>(k, kt, n, zc0, zc1, zci, ze0, ze1, ze2, ze3) : c[j];
          const T cp1 = ZR ? zsel<T>(k + 1, kt, n, zc0, zc1, zci, ze0, ze1, ze2, ze3) : c[j + 1];
          T y = 0;
#pragma unroll
          for (int f = 0; f < NF; f++) {
            const T dzk = ZR ? (k < (zwp[f] & 0xffff) ? ztop[f] : (k < (zwp[f] >> 16) ? dzg[j] : zbot[f])) : dz[j][f];
            const T term = (nv[j][f] - v0) * dzk * Df[f];
            y = (k >= (krp[f] & 0xffff) && k < (krp[f] >> 16)) ? y + term : y;
          }
          if (kt == n - 1) y -= T(2.0) * c0 * v0;
          else if (k == kt) y += (-2 * c0 - cp1) * v0 + cp1 * vp1;
          else if (k == n - 1) y += c0 * vm1 - c0 * v0;
          else y += c0 * vm1 - (c0 + cp1) * v0 + cp1 * vp1;
          const uint32_t idx = base + (uint32_t)k * 64u;
          st_stream<NTS>(&at32(V.pnew, idx), v0);
          st_stream<NTS>(&at32(V.x, idx), xv[j] + nu * pv[j]);
          dot += (double)v0 * (double)y;
          st_stream<NTS>(&at32(V.out, idx), y);
        }
      }
    }
    vprev = v[U - 1];
  }
  return dot;
}

// qapply_xchg for a tile whose columns fit their z-level windows, software
// pipelined: the loads of the next group of U levels are issued before the
// current group is exchanged, computed and stored, so its stores never stand
// between a load and its use.  The level loop is unrolled twice over two
// register buffers (A, B).  Same operations in the same order as
// qapply_xchg: bit-identical results.
template <int NF, int U, typename T, bool NTS>
__device__ __forceinline__ double qapply_xchg_pipe(const Geo &G, const QVec<T, T> &V, T mu, T nu, int t, int l,
                                                   T (*row)[SUNHIP_TILE + SUNHIP_XPAIRS]) {
  if (t >= G.Nt) return 0;
  const T *__restrict__ rt = V.rtmp;
  const T *__restrict__ po = V.pold;
  const int s = t * 64 + l, NS = G.NS;
  const int n = G.nk[s];
  const bool own = n > 0 && G.comp[s] == 1;
  const int kt = n > 0 ? G.ctop[s] : 0;
  const int kb0 = __builtin_amdgcn_readfirstlane(wave_min_i(own ? kt : 0x7fffffff));
  const int kb1 = __builtin_amdgcn_readfirstlane(wave_max_i(own ? n : 0));
  if (kb1 <= kb0) return 0;  // no computational column in the tile (uniform)
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  const int nf = n > 0 ? G.nf[s] : 0;
  int code[NF], kmn[NF], kmx[NF];
  T Df[NF];
  ZFace<T> zf[NF];
  // the per-face tables have SUNHIP_MAXF rows, so every row of the class
  // exists: load them all, then select (no branch between the loads)
#pragma unroll
  for (int f = 0; f < NF; f++) {
    const int xc = G.xcode[f * NS + s], kr = G.kr[f * NS + s], nb = G.nbb[f * NS + s], es = G.es[f * NS + s];
    zf[f] = zface<T, NF>(V.Z, G, s, f);
    const T d = V.D[max(es, 0)];
    const bool on = f < nf;
    code[f] = on ? xc : -1;
    kmn[f] = on && xc >= 0 ? (kr & 0xffff) : 0;
    kmx[f] = on && xc >= 0 ? (kr >> 16) : 0;
    Df[f] = on && nb >= 0 ? d : T(0);
  }
  int zc0, zc1;
  T zci, ze0, ze1, ze2, ze3;
  {
    const ZCol<T> zc = zcol<T>(V.Z, G, s, kt, n);
    zc0 = zc.c0; zc1 = zc.c1; zci = zc.ci; ze0 = zc.e0; ze1 = zc.e1; ze2 = zc.e2; ze3 = zc.e3;
  }
  // at most 64 cross-tile pairs (the caller sends other tiles to qapply_xchg): pair l
  const int np = G.npair[t];
  const bool pact = l < np;
  uint32_t pnb = base;
  int plim = 0;
  if (pact) {
    const int fs = G.pair[(size_t)t * SUNHIP_XPAIRS + l];
    pnb = (uint32_t)G.nbb[fs];
    const int mx = G.kr[fs] >> 16;
    plim = mx > 0 ? mx - 1 : 0;
  }
  const int nn = n > 0 ? n : 0;
  double dot = 0;
  T vprev = 0;
  // the loads of the group at kb (raw values; combined where used)
  auto load = [&](int kb, T (&pv)[U + 1], T (&v)[U + 1], T (&xv)[U], T (&r0)[U], T (&p0)[U], T (&dz)[U]) {
#pragma unroll
    for (int j = 0; j < U; j++) dz[j] = (T)V.Z.dzt[kb + j];  // the layer thicknesses of the group (uniform)
#pragma unroll
    for (int j = 0; j <= U; j++) {
      const uint32_t idx = base + (uint32_t)min(kb + j, nn) * 64u;
      pv[j] = at32(po, idx);
      v[j] = at32(rt, idx);
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int k = kb + j;
      xv[j] = at32((const T *)V.x, base + (uint32_t)min(k, nn) * 64u);
      r0[j] = 0; p0[j] = 0;
      if (pact) {
        const uint32_t ni = pnb + (uint32_t)min(k, plim) * 64u;
        r0[j] = at32(rt, ni); p0[j] = at32(po, ni);
      }
    }
  };
  // exchange, compute and store the group at kb
  auto process = [&](int kb, T (&pv)[U + 1], T (&v)[U + 1], T (&xv)[U], T (&r0)[U], T (&p0)[U], T (&dz)[U]) {
    T nv[U][NF];
#pragma unroll
    for (int j = 0; j <= U; j++) v[j] = v[j] + mu * pv[j];
#pragma unroll
    for (int j = 0; j < U; j++) {
      row[j][l] = v[j];
      row[j][64 + l] = pact ? r0[j] + mu * p0[j] : T(0);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int j = 0; j < U; j++)
#pragma unroll
      for (int f = 0; f < NF; f++) nv[j][f] = code[f] >= 0 ? row[j][code[f]] : T(0);
    __builtin_amdgcn_wave_barrier();
    if (own) {
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = kb + j;
        if (k >= kt && k < n) {
          const T v0 = v[j], vm1 = j ? v[j - 1] : vprev, vp1 = v[j + 1];
          const T c0 = zsel<T>(k, kt, n, zc0, zc1, zci, ze0, ze1, ze2, ze3);
          const T cp1 = zsel<T>(k + 1, kt, n, zc0, zc1, zci, ze0, ze1, ze2, ze3);
          T y = 0;
#pragma unroll
          for (int f = 0; f < NF; f++) {
            const T dzk = k < zf[f].f0 ? zf[f].top : (k < zf[f].f1 ? dz[j] : zf[f].bot);
            const T term = (nv[j][f] - v0) * dzk * Df[f];
            y = (k >= kmn[f] && k < kmx[f]) ? y + term : y;
          }
          if (kt == n - 1) y -= T(2.0) * c0 * v0;
          else if (k == kt) y += (-2 * c0 - cp1) * v0 + cp1 * vp1;
          else if (k == n - 1) y += c0 * vm1 - c0 * v0;
          else y += c0 * vm1 - (c0 + cp1) * v0 + cp1 * vp1;
          const uint32_t idx = base + (uint32_t)k * 64u;
          st_stream<NTS>(&at32(V.pnew, idx), v0);
          st_stream<NTS>(&at32(V.x, idx), xv[j] + nu * pv[j]);
          dot += (double)v0 * (double)y;
          st_stream<NTS>(&at32(V.out, idx), y);
        }
      }
    }
    vprev = v[U - 1];
  };
  T pvA[U + 1], vA[U + 1], xvA[U], r0A[U], p0A[U], dzA[U];
  T pvB[U + 1], vB[U + 1], xvB[U], r0B[U], p0B[U], dzB[U];
  load(kb0, pvA, vA, xvA, r0A, p0A, dzA);
  for (int kb = kb0; kb < kb1; kb += 2 * U) {
    if (kb + U < kb1) load(kb + U, pvB, vB, xvB, r0B, p0B, dzB);
    process(kb, pvA, vA, xvA, r0A, p0A, dzA);
    if (kb + U >= kb1) break;
    if (kb + 2 * U < kb1) load(kb + 2 * U, pvA, vA, xvA, r0A, p0A, dzA);
    process(kb + U, pvB, vB, xvB, r0B, p0B, dzB);
  }
  return dot;
}

// Vertical-line preconditioner of tile t: (UPDATE) r -= nu*z, then
// out = M^-1 r by Thomas down the column (b' recomputed, d' stored in out)
// and back up it (b' from bq); rz and rr collect the lane's r.out and r.r.
// RZID: rz from the forward sweep alone, r.M^-1 r = sum d'^2/b' (M = L D L^T
// with D = diag(b') and d' = L^-1 r), so the backward sweep does not read r
// again; out is unchanged, only the rounding of rz differs.
// KR > 0: the top KR levels of the column keep b', d' and c in registers
// between the sweeps (their backward step reads nothing and their d' never
// goes to memory); the levels below go through out and bq as above.  The
// arithmetic and its order are the same for every KR.
// ZR (a tile whose columns fit their z-level windows): coef formed from
// the window and the exceptions (Z), as in qapply_tile.
// KL > 0 (with KR > 0): the KL levels below the head keep b', c and d' in
// LDS between the sweeps (lrow: this lane's first word, the block's lanes
// interleaved, 3*KL words of stride 256), so they neither write d' to out
// nor read out, bq and coef back; the arithmetic is the same again.
template <bool UPDATE, int U, typename T, bool RZID = false, int KR = 0, bool ZR = false, int KL = 0>
__device__ __forceinline__ void qprec_tile(const Geo &G, const T *__restrict__ coef, const T *__restrict__ z,
                                           T *__restrict__ r, T *__restrict__ out, const T *__restrict__ bq, T nu,
                                           int t, int l, double &rz, double &rr, const ZReg *Zp = nullptr,
                                           T *lrow = nullptr) {
  static_assert(KL == 0 || (KR > 0 && KL % U == 0), "an LDS stage follows a register head, in whole load groups");
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s], kt = G.ctop[s];
  if (n == 0 || G.comp[s] != 1) return;
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  const ZCol<T> zc = ZR ? zcol<T>(*Zp, G, s, kt, n) : ZCol<T>{};
  // coef at level k (k <= n; level n is never used but stays in the tile)
  auto C = [&](int k) -> T { return ZR ? zc.at(k) : at32(coef, base + (uint32_t)k * 64u); };
  if (kt == n - 1) {
    const uint32_t idx = base + (uint32_t)kt * 64u;
    T rk = at32(r, idx);
    if (UPDATE) { rk = rk - nu * at32(z, idx); at32(r, idx) = rk; }
    const T u = T(-0.5) * rk / C(kt);
    at32(out, idx) = u;
    rz += (double)rk * (double)u;
    rr += (double)rk * (double)rk;
    return;
  }
  constexpr int KH = KR > 0 ? KR : 1;
  T hb[KH], hd[KH], hc[KH], hr[KH];
  T cprev = 0, bprev = 1, dprev = 0;
  if constexpr (KR > 0) {
    // the head in groups of HG levels: a group's loads (clamped into the
    // tile, which stores one level below its deepest column) are all issued
    // before its arithmetic, so the forward sweep waits once per group
    constexpr int HG = KR % 8 == 0 ? 8 : 4;
#pragma unroll
    for (int g0 = 0; g0 < KR; g0 += HG) {
      if (kt + g0 >= n) break;
      T rg[HG], zg[HG], cg[HG + 1];
#pragma unroll
      for (int jj = 0; jj <= HG; jj++) {
        const int k = min(kt + g0 + jj, n);
        cg[jj] = C(k);
        if (jj < HG) {
          const uint32_t idx = base + (uint32_t)k * 64u;
          rg[jj] = at32((const T *)r, idx);
          zg[jj] = UPDATE ? at32(z, idx) : T(0);
        }
      }
#pragma unroll
      for (int jj = 0; jj < HG; jj++) {
        const int j = g0 + jj, k = kt + j;
        if (k < n) {
          const uint32_t idx = base + (uint32_t)k * 64u;
          const T ck = cg[jj], ck1 = cg[jj + 1];
          T rk = rg[jj];
          if (UPDATE) { rk = rk - nu * zg[jj]; st_stream<true>(&at32(r, idx), rk); }
          rr += (double)rk * (double)rk;
          T a = 0, b, c = 0;
          if (k == kt) { b = -2 * ck - ck1; c = ck1; }
          else if (k == n - 1) { a = ck; b = -ck; }
          else { a = ck; b = -ck - ck1; c = ck1; }
          T d = rk;
          if (k > kt) {
            b -= a * cprev / bprev;
            d -= a * dprev / bprev;
          }
          if constexpr (RZID) rz += (double)d * (double)d / (double)b;
          hb[j] = b; hd[j] = d; hc[j] = c;
          if constexpr (!RZID) hr[j] = rk;
          cprev = c; bprev = b; dprev = d;
        }
      }
    }
  }
  // the LDS stage: levels kt+KR .. kt+KR+KL-1, as the memory loop below but
  // with the row kept in lrow
  const int kls = kt + KR, kle = min(n, kt + KR + KL);
  if constexpr (KL > 0) {
    for (int kb = kls; kb < kle; kb += U) {
      T rv[U], zv[U], cv[U + 1];
#pragma unroll
      for (int j = 0; j <= U; j++) {
        const uint32_t idx = base + (uint32_t)min(kb + j, n) * 64u;
        cv[j] = C(min(kb + j, n));
        if (j < U) {
          rv[j] = at32((const T *)r, idx);
          zv[j] = UPDATE ? at32(z, idx) : T(0);
        }
      }
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int k = kb + j;
        if (k < kle) {
          const uint32_t idx = base + (uint32_t)k * 64u;
          T rk = rv[j];
          if (UPDATE) { rk = rk - nu * zv[j]; st_stream<true>(&at32(r, idx), rk); }
          rr += (double)rk * (double)rk;
          const T ck = cv[j], ck1 = cv[j + 1];
          T a = 0, b, c = 0;
          if (k == kt) { b = -2 * ck - ck1; c = ck1; }
          else if (k == n - 1) { a = ck; b = -ck; }
          else { a = ck; b = -ck - ck1; c = ck1; }
          T d = rk;
          if (k > kt) {
            b -= a * cprev / bprev;
            d -= a * dprev / bprev;
          }
          if constexpr (RZID) rz += (double)d * (double)d / (double)b;
          T *w = lrow + (3 * (k - kls)) * 256;
          w[0] = b;
          w[256] = c;
          w[512] = d;
          cprev = c; bprev = b; dprev = d;
        }
      }
    }
  }
  const int klo = kt + KR + KL;  // first level that goes through memory
  for (int kb = klo; kb < n; kb += U) {
    T rv[U], zv[U], cv[U + 1];
#pragma unroll
    for (int j = 0; j <= U; j++) {
      const uint32_t idx = base + (uint32_t)min(kb + j, n) * 64u;
      cv[j] = C(min(kb + j, n));
      if (j < U) {
        rv[j] = at32((const T *)r, idx);
        zv[j] = UPDATE ? at32(z, idx) : T(0);
      }
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int k = kb + j;
      if (k < n) {
        const uint32_t idx = base + (uint32_t)k * 64u;
        T rk = rv[j];
        if (UPDATE) { rk = rk - nu * zv[j]; st_stream<true>(&at32(r, idx), rk); }
        rr += (double)rk * (double)rk;
        const T ck = cv[j], ck1 = cv[j + 1];
        T a = 0, b, c = 0;
        if (k == kt) { b = -2 * ck - ck1; c = ck1; }
        else if (k == n - 1) { a = ck; b = -ck; }
        else { a = ck; b = -ck - ck1; c = ck1; }
        T d = rk;
        if (k > kt) {
          b -= a * cprev / bprev;
          d -= a * dprev / bprev;
        }
        at32(out, idx) = d;
        if constexpr (RZID) rz += (double)d * (double)d / (double)b;
        cprev = c; bprev = b; dprev = d;
      }
    }
  }
  T unext = 0, cnext = 0;
  for (int ke = n - 1; ke >= klo; ke -= U) {
    T bv[U], dv[U], rv[U], cv[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const uint32_t idx = base + (uint32_t)max(ke - j, klo) * 64u;
      bv[j] = at32(bq, idx);
      dv[j] = at32((const T *)out, idx);
      rv[j] = RZID ? T(0) : at32((const T *)r, idx);
      cv[j] = C(max(ke - j, klo));
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int k = ke - j;
      if (k >= klo) {
        const T b = bv[j], d = dv[j];
        const T u = (k == n - 1) ? d / b : d / b - cnext * unext / b;
        st_stream<true>(&at32(out, base + (uint32_t)k * 64u), u);
        if constexpr (!RZID) rz += (double)rv[j] * (double)u;
        unext = u;
        cnext = cv[j];  // c at level k-1 is coef[k]
      }
    }
  }
  if constexpr (KL > 0) {
    for (int k = kle - 1; k >= kls; k--) {
      const T *w = lrow + (3 * (k - kls)) * 256;
      const T b = w[0], c = w[256], d = w[512];
      const T u = (k == n - 1) ? d / b : d / b - c * unext / b;
      const uint32_t idx = base + (uint32_t)k * 64u;
      st_stream<true>(&at32(out, idx), u);
      if constexpr (!RZID) rz += (double)at32((const T *)r, idx) * (double)u;
      unext = u;
    }
  }
  if constexpr (KR > 0) {
#pragma unroll
    for (int j = KR - 1; j >= 0; j--) {
      const int k = kt + j;
      if (k < n) {
        const T b = hb[j], d = hd[j];
        const T u = (k == n - 1) ? d / b : d / b - hc[j] * unext / b;
        st_stream<true>(&at32(out, base + (uint32_t)k * 64u), u);
        if constexpr (!RZID) rz += (double)hr[j] * (double)u;
        unext = u;
      }
    }
  }
}

// qprecond=0: r -= nu*z, out = r; rr collects r.r (main/phys.c:2505-2519).
template <typename T>
__device__ __forceinline__ void qnoprec_tile(const Geo &G, const T *__restrict__ z, T *__restrict__ r,
                                             T *__restrict__ out, T nu, bool update, int t, int l, double &rr) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return;
  const int64_t base = G.toff[t] + l;
  for (int k = G.ctop[s]; k < n; k++) {
    const int64_t idx = base + (int64_t)k * 64;
    T rk = r[idx];
    if (update) { rk = rk - nu * z[idx]; r[idx] = rk; }
    out[idx] = rk;
    rr += (double)rk * (double)rk;
  }
}

// x += nu*p over tile t (the last iteration's x update).
template <typename T>
__device__ __forceinline__ void qfinish_tile(const Geo &G, T *__restrict__ x, const T *__restrict__ p, T nu, int t,
                                             int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return;
  const int64_t base = G.toff[t] + l;
  for (int k = G.ctop[s]; k < n; k++) {
    const int64_t idx = base + (int64_t)k * 64;
    x[idx] = x[idx] + nu * p[idx];
  }
}

// x64 += e32 over tile t (mixed form, FP64 accumulation).
__device__ __forceinline__ void mix_update_tile(const Geo &G, double *__restrict__ x, const float *__restrict__ e,
                                                int t, int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return;
  const int64_t base = G.toff[t] + l;
  for (int k = G.ctop[s]; k < n; k++) {
    const int64_t idx = base + (int64_t)k * 64;
    x[idx] = x[idx] + (double)e[idx];
  }
}

// The inner FP32 solve's last x update (qfinish_tile, when upd) and then
// x64 += e32 (mix_update_tile) in one pass over tile t, with the same
// roundings as the two passes.
__device__ __forceinline__ void mix_finish_tile(const Geo &G, double *__restrict__ x, const float *__restrict__ e,
                                                const float *__restrict__ p, float nu, bool upd, int t, int l) {
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] != 1) return;
  const int64_t base = G.toff[t] + l;
  for (int k = G.ctop[s]; k < n; k++) {
    const int64_t idx = base + (int64_t)k * 64;
    float ek = e[idx];
    if (upd) ek = ek + nu * p[idx];
    x[idx] = x[idx] + (double)ek;
  }
}

// Fill / cast over tile t's stored levels (padding included).
template <typename T>
__device__ __forceinline__ void fill_tile(const Geo &G, T *__restrict__ p, T v, int t, int l) {
  if (t >= G.Nt) return;
  for (int64_t i = G.toff[t] + l; i < G.toff[t + 1]; i += 64) p[i] = v;
}

template <typename TI, typename TO>
__device__ __forceinline__ void cast_tile(const Geo &G, const TI *__restrict__ a, TO *__restrict__ b, int t, int l) {
  if (t >= G.Nt) return;
  for (int64_t i = G.toff[t] + l; i < G.toff[t + 1]; i += 64) b[i] = (TO)a[i];
}

}  // namespace sunhip

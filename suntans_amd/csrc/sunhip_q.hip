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
// The scalars live on the device.  Each kernel's block partial sums are
// finished either by the last block to arrive or by a one-block finisher
// kernel (a tuning choice); both add the partials in a fixed order, so a
// solve is reproducible run to run.  Element-wise arithmetic keeps the
// reference's operation order and is built without FMA contraction, so
// OperatorQ and the preconditioner agree with the reference bit for bit;
// only the summation order of the dot products differs.
//
// The kernels are templated on the storage type T of vectors and
// coefficients.  T=double is the drop-in; T=float is the inner solve of the
// mixed-precision form (sunhip_q_solve_mixed), whose dot products still
// accumulate in double and whose outer residual, convergence test and
// solution update stay in FP64.
//
// Both kernels walk a column U levels at a time (a tuning choice, 1, 2 or 4)
// and issue every load of a group before its arithmetic, with addresses
// clamped into the column (a tile stores one level more than its deepest
// column), so no load waits behind a branch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "sunhip_solvers.h"

using namespace sunhip;


// Where a launch's block sums go: the last block finishes them (cnt set) or
// they are left at part[pbase + block] for a finisher kernel.
struct Red {
  double *part;
  int pbase;
  unsigned *cnt;
  int op;
};

// Block sums of NV values, then either finish (last block) or publish.
template <int NV, int NT>
__device__ void reduce_out(const double (&v)[NV], const Red &R, CGScal *sc) {
  if (R.op == OP_NONE) return;
  if (R.cnt) {
    double tot[NV];
    if (grid_reduce<NV, NT>(v, R.part, R.cnt, tot) && threadIdx.x == 0) q_op(sc, R.op, tot);
    return;
  }
  constexpr int NW = NT / 64;
  __shared__ double sh[NV][NW];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    const double s = wave_sum(v[i]);
    if (l == 0) sh[i][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double s = 0;
      for (int j = 0; j < NW; j++) s += sh[i][j];
      R.part[(size_t)(R.pbase + blockIdx.x) * NV + i] = s;
    }
  }
}

// One block adds nparts partials in a fixed order and applies op.
template <int NV>
__global__ __launch_bounds__(1024) void k_qfinalize(const double *__restrict__ part, int nparts, CGScal *sc, int op) {
  if ((op == OP_APPLY || op == OP_PREC_ITER || op == OP_NOPREC_ITER) && sc->done) return;
  __shared__ double sh[NV][16];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  double tot[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double s = 0;
    for (int b = threadIdx.x; b < nparts; b += 1024) s += part[(size_t)b * NV + i];
    s = wave_sum(s);
    if (l == 0) sh[i][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double s = 0;
      for (int j = 0; j < 16; j++) s += sh[i][j];
      tot[i] = s;
    }
    q_op(sc, op, tot);
  }
}

__global__ __launch_bounds__(256) void k_qcoef(Geo G, const double *__restrict__ dzz, double *__restrict__ coef) {
  qcoef_tile(G, dzz, coef, blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

template <typename T>
__global__ __launch_bounds__(256) void k_qfactor(Geo G, const T *__restrict__ coef, T *__restrict__ bq) {
  qfactor_tile<T>(G, coef, bq, blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

// y = L v (qapply_tile's modes) over all tiles, TPB tiles per block.
// MINW: minimum waves per SIMD the register allocation must allow (0: the
// compiler's choice).
template <int MODE, int NF, int U, typename T, typename TO, int TPB = 4, bool NTS = false, int MINW = 0,
          bool ZR = false>
__global__ __launch_bounds__(TPB * 64, MINW > 0 ? MINW : 1) void k_qapply(Geo G, QVec<T, TO> V, CGScal *sc, Red R) {
  if (MODE == QM_FUSED && sc->done) return;
  const int t = (TPB == 4 ? tile_group(G, blockIdx.x, gridDim.x) : xcd_block(blockIdx.x, gridDim.x)) * TPB +
                (threadIdx.x >> 6);
  const T mu = MODE == QM_FUSED ? (T)sc->mu : T(0);
  const T nu = MODE == QM_FUSED ? (T)sc->nu : T(0);
  const double dot = (ZR && t < G.Nt && V.Z.tflag[t])
                         ? qapply_tile<MODE, NF, U, T, TO, NTS, true>(G, V, mu, nu, t, threadIdx.x & 63)
                         : qapply_tile<MODE, NF, U, T, TO, NTS, false>(G, V, mu, nu, t, threadIdx.x & 63);
  if (MODE == QM_FUSED || MODE == QM_RESID_DOT) {
    double dv[1] = {dot};
    reduce_out<1, TPB * 64>(dv, R, sc);
  }
}

// The fused operator over level bands of H levels: block b covers the tiles
// of group b % ngroups in band b / ngroups (band-major, so blocks running
// together on an XCD are neighbouring tiles at the same depths and share
// their rows in L2).
template <int NF, int U, typename T, bool NTS>
__global__ __launch_bounds__(256) void k_qapply_band(Geo G, QVec<T, T> V, CGScal *sc, Red R, int H, int ngroups) {
  if (sc->done) return;
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  const int band = vb / ngroups, grp = vb - band * ngroups;
  const int t = grp * 4 + (threadIdx.x >> 6);
  const double dot = qapply_tile<QM_FUSED, NF, U, T, T, NTS>(G, V, (T)sc->mu, (T)sc->nu, t, threadIdx.x & 63,
                                                            band * H, band * H + H);
  double dv[1] = {dot};
  reduce_out<1, 256>(dv, R, sc);
}

// The fused operator with block-staged neighbour values (qapply_block).
template <int NF, int U, typename T, bool NTS>
__global__ __launch_bounds__(256) void k_qapply_blk(Geo G, QVec<T, T> V, CGScal *sc, Red R) {
  if (sc->done) return;
  const int t0 = tile_group(G, blockIdx.x, gridDim.x) * 4;
  const double dot = qapply_block<NF, U, T, NTS>(G, V, (T)sc->mu, (T)sc->nu, t0);
  double dv[1] = {dot};
  reduce_out<1, 256>(dv, R, sc);
}

// The iteration's operator kernel with the neighbour exchange inside each
// wave (qapply_xchg); 4 tiles per block, one exchange row per wave.
template <int NF, int U, typename T, bool NTS, int MINW = 0, bool ZR = false>
__global__ __launch_bounds__(256, MINW > 0 ? MINW : 1) void k_qapply_x(Geo G, QVec<T, T> V, CGScal *sc, Red R) {
  if (sc->done) return;
  __shared__ T rows[4][U][SUNHIP_TILE + SUNHIP_XPAIRS];
  const int w = threadIdx.x >> 6;
  const int t = tile_group(G, blockIdx.x, gridDim.x) * 4 + w;
  const double dot =
      (ZR && t < G.Nt && V.Z.tflag[t])
          ? qapply_xchg<NF, U, T, NTS, true>(G, V, (T)sc->mu, (T)sc->nu, t, threadIdx.x & 63, rows[w])
          : qapply_xchg<NF, U, T, NTS, false>(G, V, (T)sc->mu, (T)sc->nu, t, threadIdx.x & 63, rows[w]);
  double v[1] = {dot};
  reduce_out<1, 256>(v, R, sc);
}

// The pipelined exchange kernel (qapply_xchg_pipe) for tiles that fit
// their z-level windows and have at most 64 cross-tile pairs, qapply_xchg
// for the others.
template <int NF, int U, typename T, bool NTS, int MINW = 0>
__global__ __launch_bounds__(256, MINW > 0 ? MINW : 1) void k_qapply_xp(Geo G, QVec<T, T> V, CGScal *sc, Red R) {
  if (sc->done) return;
  __shared__ T rows[4][U][SUNHIP_TILE + SUNHIP_XPAIRS];
  const int w = threadIdx.x >> 6;
  const int t = tile_group(G, blockIdx.x, gridDim.x) * 4 + w;
  const double dot =
      (t < G.Nt && V.Z.tflag[t] && G.npair[t] <= 64)
          ? qapply_xchg_pipe<NF, U, T, NTS>(G, V, (T)sc->mu, (T)sc->nu, t, threadIdx.x & 63, rows[w])
          : qapply_xchg<NF, U, T, NTS, false>(G, V, (T)sc->mu, (T)sc->nu, t, threadIdx.x & 63, rows[w]);
  double v[1] = {dot};
  reduce_out<1, 256>(v, R, sc);
}

// Vertical-line preconditioner (qprec_tile) over all tiles, 4 per block.
template <bool UPDATE, int U, typename T, bool RZID = false, int KR = 0, bool ZR = false, int KL = 0>
__global__ __launch_bounds__(256) void k_qprec(Geo G, const T *__restrict__ coef, const T *__restrict__ z,
                                               T *__restrict__ r, T *__restrict__ out, const T *__restrict__ bq,
                                               CGScal *sc, Red R, ZReg Z) {
  if (UPDATE && sc->done) return;
  // rows of the LDS stage: [level][b', c, d'][lane of the block]
  __shared__ T lrows[KL > 0 ? 3 * KL * 256 : 1];
  const int t = tile_group(G, blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  double rz = 0, rr = 0;
  if (ZR && t < G.Nt && Z.tflag[t])
    qprec_tile<UPDATE, U, T, RZID, KR, true, KL>(G, coef, z, r, out, bq, UPDATE ? (T)sc->nu : T(0), t,
                                                 threadIdx.x & 63, rz, rr, &Z, lrows + threadIdx.x);
  else
    qprec_tile<UPDATE, U, T, RZID, KR, false, KL>(G, coef, z, r, out, bq, UPDATE ? (T)sc->nu : T(0), t,
                                                  threadIdx.x & 63, rz, rr, &Z, lrows + threadIdx.x);
  double v[2] = {rz, rr};
  reduce_out<2, 256>(v, R, sc);
}

// qprecond=0: r -= nu*z, rtmp = r, alpha = eps = r.r (main/phys.c:2505-2519).
template <typename T>
__global__ __launch_bounds__(256) void k_qnoprec(Geo G, const T *__restrict__ z, T *__restrict__ r,
                                                 T *__restrict__ out, CGScal *sc, Red R, int update) {
  if (update && sc->done) return;
  const int t = tile_group(G, blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  double rr = 0;
  qnoprec_tile<T>(G, z, r, out, update ? (T)sc->nu : T(0), update != 0, t, threadIdx.x & 63, rr);
  double v[2] = {rr, rr};
  reduce_out<2, 256>(v, R, sc);
}

// x += nu*p for the last iteration (its x update is otherwise folded into the
// next apply).  p is the buffer the last apply wrote.
template <typename T>
__global__ __launch_bounds__(256) void k_qfinish(Geo G, T *__restrict__ x, const T *__restrict__ p0,
                                                 const T *__restrict__ p1, const CGScal *sc) {
  if (sc->iter == 0) return;
  qfinish_tile<T>(G, x, (sc->iter & 1) ? p1 : p0, (T)sc->nu, blockIdx.x * 4 + (threadIdx.x >> 6),
                  threadIdx.x & 63);
}

// The z-level windows (ZReg) of every computational column and face, their
// exceptions, and whether every column of the tile fits them.  A column's
// window starts below its top cell at the first level whose stored coef is
// 2 Ac/den bit for bit and runs while it is; it fits when at most two levels
// lie above it and two below.  A face's window starts at its first level
// where the stored dzf is dz[k] and runs while it is; it fits with at most
// one level on either side.
template <int NF>
__global__ __launch_bounds__(256) void k_qwin(Geo G, const double *__restrict__ coef, const double *__restrict__ dzf,
                                              double den, const double *__restrict__ dzt, int *__restrict__ cwin,
                                              int *__restrict__ fwin, double *__restrict__ cex,
                                              double *__restrict__ fex, int *__restrict__ tflag) {
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s], NS = G.NS;
  int cw = 0, fw[NF];
  double ce[4] = {0, 0, 0, 0}, fe[2][NF];
#pragma unroll
  for (int f = 0; f < NF; f++) { fw[f] = 0; fe[0][f] = 0; fe[1][f] = 0; }
  bool fits = true;
  if (n > 0 && G.comp[s] == 1) {
    const int kt = G.ctop[s];
    const uint32_t base = (uint32_t)(G.toff[t] + l);
    auto cv = [&](int k) { return at32(coef, base + (uint32_t)k * 64u); };
    if (den > 0) {
      const double ci = 2 * G.Ac[s] / den;
      int c0 = kt + 1;
      while (c0 < n && !(cv(c0) == ci)) c0++;
      int c1 = c0;
      while (c1 < n && cv(c1) == ci) c1++;
      fits = c0 - kt <= 2 && n - c1 <= 2;
      cw = c0 | (c1 << 16);
      ce[0] = cv(kt);
      if (kt + 1 < n) ce[1] = cv(kt + 1);
      if (n - 2 >= kt) ce[2] = cv(n - 2);
      ce[3] = cv(n - 1);
    } else {
      fits = false;
    }
    const int nf = G.nf[s];
#pragma unroll
    for (int f = 0; f < NF; f++) {
      if (f >= nf || G.nbb[f * NS + s] < 0) continue;
      const int kr = G.kr[f * NS + s], kmn = kr & 0xffff, kmx = kr >> 16;
      if (kmx <= kmn) continue;
      const uint32_t eb = (uint32_t)G.eb[f * NS + s];
      auto dv = [&](int k) { return at32(dzf, eb + (uint32_t)k * 64u); };
      int e0 = kmn;
      while (e0 < kmx && !(dv(e0) == dzt[e0])) e0++;
      int e1 = e0;
      while (e1 < kmx && dv(e1) == dzt[e1]) e1++;
      fits = fits && e0 - kmn <= 1 && kmx - e1 <= 1;
      fw[f] = e0 | (e1 << 16);
      fe[0][f] = dv(kmn);
      fe[1][f] = dv(kmx - 1);
    }
  }
  cwin[s] = cw;
#pragma unroll
  for (int i = 0; i < 4; i++) cex[i * NS + s] = ce[i];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    fwin[f * NS + s] = fw[f];
    fex[f * NS + s] = fe[0][f];
    fex[(NF + f) * NS + s] = fe[1][f];
  }
  const int all = wave_min_i(fits ? 1 : 0);
  if (l == 0) tflag[t] = all;
}

// ---- qprecond=1 (ConditionQ, its QCoefficients branch, OperatorQC) ------
template <int NF>
__global__ __launch_bounds__(256) void k_qcond(Geo G, const double *__restrict__ dzz, const double *__restrict__ D,
                                               double *__restrict__ c) {
  qcond_tile<NF>(G, dzz, D, c, blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

template <int NF>
__global__ __launch_bounds__(256) void k_qcoef_pc1(Geo G, const double *__restrict__ dzz,
                                                   const double *__restrict__ D, const double *__restrict__ c,
                                                   double *__restrict__ coef, double *__restrict__ fc, int64_t NT) {
  qcoef_pc1_tile<NF>(G, dzz, D, c, coef, fc, NT, blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

template <int MODE, int NF>
__global__ __launch_bounds__(256) void k_qcapply(Geo G, QVec<double, double> V, CGScal *sc, Red R, int64_t NT) {
  if (MODE == QM_FUSED && sc->done) return;
  const int t = tile_group(G, blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  const double mu = MODE == QM_FUSED ? sc->mu : 0.0, nu = MODE == QM_FUSED ? sc->nu : 0.0;
  const double dot = qcapply_tile<MODE, NF>(G, V, mu, nu, NT, t, threadIdx.x & 63);
  if (MODE == QM_FUSED) {
    double dv[1] = {dot};
    reduce_out<1, 256>(dv, R, sc);
  }
}

__global__ __launch_bounds__(256) void k_qscale(Geo G, double *__restrict__ a, const double *__restrict__ b,
                                                int op) {
  qscale_tile(G, a, b, op, blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

// The iterations read a non-computational neighbour's p where the reference
// reads it: p aliases src there (main/phys.c:2419, 2496-2522 update only the
// computational cells), so p = rtmp + mu*p_old with p_old = 0 needs rtmp =
// src on those columns.
__global__ __launch_bounds__(256) void k_qbnd_src(Geo G, const double *__restrict__ src, double *__restrict__ rtmp) {
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0 || G.comp[s] == 1) return;
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  for (int k = 0; k < n; k++) at32(rtmp, base + (uint32_t)k * 64u) = at32(src, base + (uint32_t)k * 64u);
}

template <typename T>
__global__ void k_fill(T *p, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename TI, typename TO>
__global__ void k_cast(const TI *__restrict__ a, TO *__restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (TO)a[i];
}

// x64 += e32 on the computational columns (mixed form, FP64 accumulation).
__global__ __launch_bounds__(256) void k_mix_update(Geo G, double *__restrict__ x, const float *__restrict__ e) {
  mix_update_tile(G, x, e, blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

// The inner solve's k_qfinish and k_mix_update in one launch (mixed form).
__global__ __launch_bounds__(256) void k_mix_finish(Geo G, double *__restrict__ x, const float *__restrict__ e,
                                                    const float *__restrict__ p0, const float *__restrict__ p1,
                                                    const CGScal *sc) {
  const int it = sc->iter;
  mix_finish_tile(G, x, e, (it & 1) ? p1 : p0, (float)sc->nu, it != 0, blockIdx.x * 4 + (threadIdx.x >> 6),
                  threadIdx.x & 63);
}

// ---------------------------------------------------------------------------
// host side

namespace {

void read_tuning(sunhip_grid *g) {
  if (g->tune_unroll < 0) {
    const char *e = getenv("SUNHIP_QUNROLL");
    g->tune_unroll = e ? atoi(e) : 4;
    if (g->tune_unroll != 1 && g->tune_unroll != 2 && g->tune_unroll != 4 && g->tune_unroll != 8) g->tune_unroll = 4;
  }
  if (g->tune_reduce < 0) {
    const char *e = getenv("SUNHIP_REDUCE");
    g->tune_reduce = (e && !strcmp(e, "last")) ? 0 : 1;
  }
}

int nb4(const sunhip_grid *g) { return (g->Nt + 3) / 4; }

template <typename T>
int launch_factor(sunhip_grid *g, const T *coef, T *bq) {
  k_qfactor<T><<<nb4(g), 256, 0, g->stream>>>(g->geo(), coef, bq);
  HIPCHK(hipGetLastError());
  return 0;
}

// Preconditioner over all columns; op decides what happens to the sums
// (OP_NONE for the plain drop-in application).
template <bool UPDATE, typename T>
int launch_prec(sunhip_grid *g, CGScal *sc, const T *coef, const T *z, T *r, T *out, const T *bq, int op,
                unsigned *cnt, int pk = -1, const ZReg *Z = nullptr) {
  const Geo G = g->geo();
  const int nb = nb4(g);
  Red R{g->d_part, 0, (g->tune_reduce == 0 && op != OP_NONE) ? cnt : nullptr, op};
  // the timing pair brackets the column kernel alone, not the finisher
  const int ps = pk >= 0 ? prof_begin(g, pk) : -1;
  const bool rzid = g->tune_rzid && op != OP_NONE;
  const bool zr = Z && Z->cwin;
  const ZReg ZZ = zr ? *Z : ZReg{};
#define PREC3(U, KR, KL) \
  do { \
    if (zr) { \
      if (rzid) \
        k_qprec<UPDATE, U, T, true, KR, true, KL><<<nb, 256, 0, g->stream>>>(G, coef, z, r, out, bq, sc, R, ZZ); \
      else \
        k_qprec<UPDATE, U, T, false, KR, true, KL><<<nb, 256, 0, g->stream>>>(G, coef, z, r, out, bq, sc, R, ZZ); \
    } else { \
      if (rzid) \
        k_qprec<UPDATE, U, T, true, KR, false, KL><<<nb, 256, 0, g->stream>>>(G, coef, z, r, out, bq, sc, R, ZZ); \
      else \
        k_qprec<UPDATE, U, T, false, KR, false, KL><<<nb, 256, 0, g->stream>>>(G, coef, z, r, out, bq, sc, R, ZZ); \
    } \
  } while (0)
#define PREC(U, KR) PREC3(U, KR, 0)
  // FP32 columns keep 16 levels by default: 24 and 32 measured no faster on
  // the 5e7-unknown solve (the register head costs occupancy)
  const bool fp32 = std::is_same<T, float>::value;
  int head = g->tune_phead;
  if (head < 0) head = fp32 ? (g->Nt >= 2048 ? 16 : 0) : (g->Nt >= 1024 ? 24 : 0);
  if (head == 0) {
    switch (g->tune_punroll) {
      case 1: PREC(1, 0); break;
      case 2: PREC(2, 0); break;
      case 4: PREC(4, 0); break;
      default: PREC(8, 0); break;
    }
  } else {  // register head: the memory part at 4 levels per group
    switch (head) {
      case 8: PREC(4, 8); break;
      case 16:
        if (g->tune_plds == 12) PREC3(4, 16, 12);
        else PREC(4, 16);
        break;
      case 24:
        if (g->tune_plds == 12) PREC3(4, 24, 12);
        else PREC(4, 24);
        break;
      default: PREC(4, 32); break;
    }
  }
#undef PREC
#undef PREC3
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  if (op != OP_NONE && g->tune_reduce == 1) k_qfinalize<2><<<1, 1024, 0, g->stream>>>(g->d_part, nb, sc, op);
  HIPCHK(hipGetLastError());
  return 0;
}

template <int MODE, typename T, typename TO>
int launch_apply(sunhip_grid *g, CGScal *sc, const QVec<T, TO> &V, const Red &R, int pk = -1) {
  // launch-shape variants (tiles per block, streaming stores) exist for the
  // fused iteration kernel on triangles at the default unroll
  const bool variants = MODE == QM_FUSED && g->mf <= 3 && g->tune_unroll == 4;
  const int tpb = variants ? g->tune_tiles : 4;
  const int nba = (g->Nt + tpb - 1) / tpb;
  const Geo G = g->geo();
  const int ps = pk >= 0 ? prof_begin(g, pk) : -1;
  // the z-level windows, when the solve has them (the plain and exchange kernels)
  const bool zr = V.Z.cwin != nullptr;
#define APPLY(NF, U)                                                                              \
  do {                                                                                            \
    if (zr) k_qapply<MODE, NF, U, T, TO, 4, false, 0, true><<<nba, 256, 0, g->stream>>>(G, V, sc, R); \
    else k_qapply<MODE, NF, U, T, TO><<<nba, 256, 0, g->stream>>>(G, V, sc, R);                      \
  } while (0)
#define APPLYV(TPB, NTS) k_qapply<MODE, 3, 4, T, TO, TPB, NTS><<<nba, TPB * 64, 0, g->stream>>>(G, V, sc, R)
#define APPLYW(NTS, W) k_qapply<MODE, 3, 4, T, TO, 4, NTS, W><<<nba, 256, 0, g->stream>>>(G, V, sc, R)
  if constexpr (MODE == QM_FUSED && std::is_same<T, TO>::value) {
    // the exchange kernel, unless another launch shape was asked for
    const int xmode = g->tune_xchg >= 0 ? g->tune_xchg : (g->Nt >= 4096 ? 2 : 0);
    if (zr && xmode >= 4 && g->max_pairs <= SUNHIP_XPAIRS && g->tune_unroll == 4) {
      // the pipelined exchange kernel: 4 two-level groups (xchg 4), 4 four-level
      // groups under a 4-wave register budget (5), or 2 two-level groups at
      // 4 waves (6)
      const int nbx = (g->Nt + 3) / 4;
      with_nf(g->mf, [&](auto nfc) {
        constexpr int NF = decltype(nfc)::value;
        if (xmode == 5 && NF <= 4) {
          if (g->tune_nts) k_qapply_xp<NF, 4, T, true, 4><<<nbx, 256, 0, g->stream>>>(G, V, sc, R);
          else k_qapply_xp<NF, 4, T, false, 4><<<nbx, 256, 0, g->stream>>>(G, V, sc, R);
        } else if (xmode == 6) {
          if (g->tune_nts) k_qapply_xp<NF, 2, T, true, 4><<<nbx, 256, 0, g->stream>>>(G, V, sc, R);
          else k_qapply_xp<NF, 2, T, false, 4><<<nbx, 256, 0, g->stream>>>(G, V, sc, R);
        } else {
          if (g->tune_nts) k_qapply_xp<NF, 2, T, true><<<nbx, 256, 0, g->stream>>>(G, V, sc, R);
          else k_qapply_xp<NF, 2, T, false><<<nbx, 256, 0, g->stream>>>(G, V, sc, R);
        }
        return 0;
      });
      prof_end(g, ps);
      HIPCHK(hipGetLastError());
      if (R.op != OP_NONE && !R.cnt) k_qfinalize<1><<<1, 1024, 0, g->stream>>>(g->d_part, nbx, sc, R.op);
      HIPCHK(hipGetLastError());
      return 0;
    }
    const bool xchg = xmode && g->max_pairs <= SUNHIP_XPAIRS && g->tune_tiles == 4 && !g->tune_blk &&
                      !g->tune_band && !g->tune_minw && g->tune_unroll == 4;
    if (xchg) {
      const int nbx = (g->Nt + 3) / 4;
      with_nf(g->mf, [&](auto nfc) {
        // FP32 values take half the registers: 8 levels per group on triangles and quads
        constexpr int NF = decltype(nfc)::value,
                      UX = NF <= 4 ? (std::is_same<T, float>::value ? 8 : 4) : 2;
#define XCH(U, NTS, W)                                                                          \
  do {                                                                                          \
    if (zr) k_qapply_x<NF, U, T, NTS, W, true><<<nbx, 256, 0, g->stream>>>(G, V, sc, R);       \
    else k_qapply_x<NF, U, T, NTS, W><<<nbx, 256, 0, g->stream>>>(G, V, sc, R);                 \
  } while (0)
        if (xmode == 2) {  // register budget of 4 waves per SIMD
          if (g->tune_nts) XCH(UX, true, 4); else XCH(UX, false, 4);
        } else if (xmode == 3) {  // 2 levels per group
          if (g->tune_nts) XCH(2, true, 0); else XCH(2, false, 0);
        } else {
          if (g->tune_nts) XCH(UX, true, 0); else XCH(UX, false, 0);
        }
#undef XCH
        return 0;
      });
      prof_end(g, ps);
      HIPCHK(hipGetLastError());
      if (R.op != OP_NONE && !R.cnt) k_qfinalize<1><<<1, 1024, 0, g->stream>>>(g->d_part, nbx, sc, R.op);
      HIPCHK(hipGetLastError());
      return 0;
    }
    if (variants && g->tune_band > 0 && tpb == 4) {
      const int H = g->tune_band, ngroups = (g->Nt + 3) / 4, nbands = (g->maxdepth + H - 1) / H;
      const int nblk = ngroups * nbands;
      if ((int64_t)nblk > g->part_n) {
        set_error("band=%d needs %d block partials, more than the %lld held", H, nblk, (long long)g->part_n);
        return -1;
      }
      if (g->tune_nts) k_qapply_band<3, 4, T, true><<<nblk, 256, 0, g->stream>>>(G, V, sc, R, H, ngroups);
      else k_qapply_band<3, 4, T, false><<<nblk, 256, 0, g->stream>>>(G, V, sc, R, H, ngroups);
      prof_end(g, ps);
      HIPCHK(hipGetLastError());
      if (R.op != OP_NONE && !R.cnt) k_qfinalize<1><<<1, 1024, 0, g->stream>>>(g->d_part, nblk, sc, R.op);
      HIPCHK(hipGetLastError());
      return 0;
    }
    if (variants && g->tune_blk && tpb == 4) {
      if (g->tune_nts) k_qapply_blk<3, 4, T, true><<<nba, 256, 0, g->stream>>>(G, V, sc, R);
      else k_qapply_blk<3, 4, T, false><<<nba, 256, 0, g->stream>>>(G, V, sc, R);
      prof_end(g, ps);
      HIPCHK(hipGetLastError());
      if (R.op != OP_NONE && !R.cnt) k_qfinalize<1><<<1, 1024, 0, g->stream>>>(g->d_part, nba, sc, R.op);
      HIPCHK(hipGetLastError());
      return 0;
    }
  }
  if (variants && g->tune_minw == 3 && tpb == 4) {
    if (g->tune_nts) APPLYW(true, 3); else APPLYW(false, 3);
  } else if (variants && g->tune_minw == 4 && tpb == 4) {
    if (g->tune_nts) APPLYW(true, 4); else APPLYW(false, 4);
  } else if (variants && (tpb != 4 || g->tune_nts)) {
    if (tpb == 8) { if (g->tune_nts) APPLYV(8, true); else APPLYV(8, false); }
    else if (tpb == 16) { if (g->tune_nts) APPLYV(16, true); else APPLYV(16, false); }
    else APPLYV(4, true);
  } else if (g->mf <= 3) {
    if (g->tune_unroll == 1) APPLY(3, 1); else if (g->tune_unroll == 2) APPLY(3, 2);
    else if (g->tune_unroll == 8) APPLY(3, 8); else APPLY(3, 4);
  } else if (g->mf == 4) {
    if (g->tune_unroll == 1) APPLY(4, 1); else if (g->tune_unroll == 2) APPLY(4, 2); else APPLY(4, 4);
  } else if (g->mf <= 6) {  // 2 levels per group at most: 6 or 8 gathers per level
    if (g->tune_unroll == 1) APPLY(6, 1); else APPLY(6, 2);
  } else {
    if (g->tune_unroll == 1) APPLY(8, 1); else APPLY(8, 2);
  }
#undef APPLY
#undef APPLYV
#undef APPLYW
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  if (R.op != OP_NONE && !R.cnt) k_qfinalize<1><<<1, 1024, 0, g->stream>>>(g->d_part, nba, sc, R.op);

  HIPCHK(hipGetLastError());
  return 0;
}

// OperatorQC over all tiles (qprecond=1).
template <int MODE>
int launch_qcapply(sunhip_grid *g, CGScal *sc, const QVec<double, double> &V, const Red &R, int pk = -1) {
  const int nb = nb4(g);
  const Geo G = g->geo();
  const int ps = pk >= 0 ? prof_begin(g, pk) : -1;
  with_nf(g->mf, [&](auto nfc) {
    constexpr int NF = decltype(nfc)::value;
    k_qcapply<MODE, NF><<<nb, 256, 0, g->stream>>>(G, V, sc, R, g->NT);
    return 0;
  });
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  if (R.op != OP_NONE && !R.cnt) k_qfinalize<1><<<1, 1024, 0, g->stream>>>(g->d_part, nb, sc, R.op);
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T>
int launch_noprec(sunhip_grid *g, const QSolver<T> &S, int update, int op) {
  const int nb = nb4(g);
  Red R{g->d_part, 0, g->tune_reduce == 0 ? g->d_count + 2 : nullptr, op};
  k_qnoprec<T><<<nb, 256, 0, g->stream>>>(g->geo(), update ? S.z : nullptr, S.r, S.rtmp, S.sc, R, update);
  HIPCHK(hipGetLastError());
  if (g->tune_reduce == 1) k_qfinalize<2><<<1, 1024, 0, g->stream>>>(g->d_part, nb, S.sc, op);
  HIPCHK(hipGetLastError());
  return 0;
}

// Initial residual r (already in S.r), rtmp = M^-1 r, alpha0/eps0, p_old = 0.
template <typename T>
int cg_start(sunhip_grid *g, const sunhip_cg_params *prm, QSolver<T> &S) {
  int rc;
  if (prm->precond == 2)
    rc = launch_prec<false, T>(g, S.sc, S.coef, nullptr, S.r, S.rtmp, S.bq, OP_PREC_SETUP, g->d_count + 1, -1, &S.Z);
  else
    rc = launch_noprec<T>(g, S, 0, OP_NOPREC_SETUP);
  if (rc) return rc;
  // p_old = 0 so the first apply reads p = rtmp exactly
  const int64_t n = g->NT;
  k_fill<T><<<(unsigned)((n + 255) / 256), 256, 0, g->stream>>>(S.p0, n, T(0));
  k_fill<T><<<(unsigned)((n + 255) / 256), 256, 0, g->stream>>>(S.p1, n, T(0));
  HIPCHK(hipGetLastError());
  return 0;
}

// One PCG iteration; it = iteration index (selects the p buffers).
template <typename T>
int cg_iteration(sunhip_grid *g, const sunhip_cg_params *prm, QSolver<T> &S, int it, int pk_apply, int pk_prec) {
  QVec<T, T> V{};
  V.coef = S.coef; V.dzf = S.dzf; V.D = S.D;
  V.rtmp = S.rtmp;
  V.pold = (it & 1) ? S.p1 : S.p0;
  V.pnew = (it & 1) ? S.p0 : S.p1;
  V.x = S.x;
  V.out = S.z;
  V.fcoef = S.fcoef;
  V.Z = S.Z;
  const Red RA{g->d_part, 0, g->tune_reduce == 0 ? g->d_count : nullptr, OP_APPLY};
  int rc;
  if constexpr (std::is_same<T, double>::value) {
    if (prm->precond == 1) rc = launch_qcapply<QM_FUSED>(g, S.sc, V, RA, pk_apply);
    else rc = launch_apply<QM_FUSED, T, T>(g, S.sc, V, RA, pk_apply);
  } else {
    rc = launch_apply<QM_FUSED, T, T>(g, S.sc, V, RA, pk_apply);
  }
  if (rc) return rc;
  if (prm->precond == 2)
    return launch_prec<true, T>(g, S.sc, S.coef, S.z, S.r, S.rtmp, S.bq, OP_PREC_ITER, g->d_count + 1, pk_prec,
                                &S.Z);
  const int ps = prof_begin(g, pk_prec);
  rc = launch_noprec<T>(g, S, 1, OP_NOPREC_ITER);
  prof_end(g, ps);
  return rc;
}

// Iterate until the device flag says done; the host polls every few
// iterations and kernels of a batch that start after convergence return at
// once.
// finish = false leaves the last x update to the caller (k_mix_finish).
template <typename T>
int cg_loop(sunhip_grid *g, const sunhip_cg_params *prm, QSolver<T> &S, int pk_apply, int pk_prec,
            bool finish = true) {
  int every = prm->check_every > 0 ? prm->check_every : (g->NT > (1 << 22) ? 4 : 8);
  if (every & 1) every++;
  int it = 0, rc, slot = 0;
  bool pending = false, done = false;
  // iterations enqueued after convergence return at once (device flag)
  while (!done && it < prm->maxiters) {
    for (int b = 0; b < every && it < prm->maxiters; b++, it++)
      if ((rc = cg_iteration<T>(g, prm, S, it, pk_apply, pk_prec))) return rc;
    if ((rc = poll_chunk(g, S.sc, slot, pending, &done))) return rc;
  }
  if (finish) k_qfinish<T><<<nb4(g), 256, 0, g->stream>>>(g->geo(), S.x, S.p0, S.p1, S.sc);
  HIPCHK(hipGetLastError());
  return 0;
}

int check_q(sunhip_grid *g, const sunhip_cg_params *prm) {
  if (!g->have_extents) {
    set_error("sunhip_q_*: call sunhip_set_extents first");
    return -1;
  }
  if (prm->precond != 0 && prm->precond != 1 && prm->precond != 2) {
    set_error("qprecond=%d is not supported (0, 1 or 2)", prm->precond);
    return -1;
  }
  if (g->maxdepth > 128) { set_error("column depth %d exceeds 128", g->maxdepth); return -1; }
  read_tuning(g);
  return 0;
}

struct F64 {
  Field *q, *src, *dzz, *dzf, *D, *coef, *r, *z, *rtmp, *p0, *p1, *bq;
};

int f64bufs(sunhip_grid *g, F64 &B) {
  return field(g, "q", &B.q) | field(g, "src", &B.src) | field(g, "dzz", &B.dzz) | field(g, "dzf", &B.dzf) |
         field(g, "D", &B.D) | field(g, "coef", &B.coef) | field(g, "r", &B.r) | field(g, "z", &B.z) |
         field(g, "rtmp", &B.rtmp) | field(g, "p0", &B.p0) | field(g, "p1", &B.p1) | field(g, "qbp", &B.bq);
}

QSolver<double> solver64(sunhip_grid *g, F64 &B) {
  return QSolver<double>{g->d_qs, B.coef->d, B.dzf->d, B.D->d, B.q->d, B.r->d, B.z->d, B.rtmp->d,
                         B.p0->d, B.p1->d, B.bq->d};
}

// The face coefficients of qprecond=1: one row of the cell layout per face
// slot of the kernels' face-count class (>= maxfaces).
int qc_fcoef(sunhip_grid *g, double **out) {
  auto it = g->fields.find("#qfcoef");
  if (it == g->fields.end()) {
    Field f;
    f.kind = CELL;
    f.n = (size_t)nf_class(g->mf) * (size_t)g->NT;
    if (int rc = dev_alloc(g, (void **)&f.d, f.n * sizeof(double))) return rc;
    HIPCHK(hipMemsetAsync(f.d, 0, f.n * sizeof(double), g->stream));
    it = g->fields.emplace("#qfcoef", f).first;
  }
  *out = it->second.d;
  return 0;
}

// coefficients, b', r = src - L x, then cg_start (main/phys.c:2423-2475).
// qprecond=1: c = ConditionQ into "qcond" (whose other columns keep what the
// caller put there: the reference reads a boundary neighbour's c as it
// finds it), src /= c and q *= c on the computational columns, the scaled
// coefficients, and OperatorQC in place of OperatorQ.
int q_setup(sunhip_grid *g, const sunhip_cg_params *prm, F64 &B, QSolver<double> &S, CGScal *stage = nullptr,
            CGScal *sc = nullptr) {
  int rc = check_q(g, prm);
  if (rc) return rc;
  if ((rc = f64bufs(g, B))) return rc;
  S = solver64(g, B);
  if (sc) S.sc = sc;
  if ((rc = upload_scal(g, S.sc, prm, stage))) return rc;
  const Geo G = g->geo();
  QVec<double, double> V{};
  V.coef = B.coef->d; V.dzf = B.dzf->d; V.D = B.D->d; V.vin = B.q->d; V.src = B.src->d; V.out = B.r->d;
  if (prm->precond == 1) {
    Field *c;
    double *fc;
    if ((rc = field(g, "qcond", &c)) || (rc = qc_fcoef(g, &fc))) return rc;
    with_nf(g->mf, [&](auto nfc) {
      constexpr int NF = decltype(nfc)::value;
      k_qcond<NF><<<nb4(g), 256, 0, g->stream>>>(G, B.dzz->d, B.D->d, c->d);
      k_qscale<<<nb4(g), 256, 0, g->stream>>>(G, B.src->d, c->d, 0);
      k_qscale<<<nb4(g), 256, 0, g->stream>>>(G, B.q->d, c->d, 1);
      k_qcoef_pc1<NF><<<nb4(g), 256, 0, g->stream>>>(G, B.dzz->d, B.D->d, c->d, B.coef->d, fc, g->NT);
      return 0;
    });
    HIPCHK(hipGetLastError());
    S.fcoef = fc;
    V.fcoef = fc;
    if ((rc = launch_qcapply<QM_RESID>(g, S.sc, V, Red{g->d_part, 0, nullptr, OP_NONE}))) return rc;
  } else {
    k_qcoef<<<nb4(g), 256, 0, g->stream>>>(G, B.dzz->d, B.coef->d);
    if (prm->precond == 2 && (rc = launch_factor<double>(g, B.coef->d, B.bq->d))) return rc;
    if ((rc = q_windows(g, B.coef->d, B.dzf->d, &S.Z))) return rc;
    V.Z = S.Z;
    if ((rc = launch_apply<QM_RESID, double, double>(g, S.sc, V, Red{g->d_part, 0, nullptr, OP_NONE}))) return rc;
  }
  if ((rc = cg_start<double>(g, prm, S))) return rc;
  k_qbnd_src<<<nb4(g), 256, 0, g->stream>>>(G, B.src->d, B.rtmp->d);
  HIPCHK(hipGetLastError());
  return 0;
}

// After the last x update, qprecond=1 rescales x: by c (main/phys.c:2539-2546)
// or, when the caller's c aliases the residual rows as Solve's does
// (main/phys.c:1280 passes phys->stmp3, which CGSolveQ uses as r), by the
// final residual, as the reference then does.
int q_unscale(sunhip_grid *g, const sunhip_cg_params *prm, const F64 &B) {
  if (prm->precond != 1) return 0;
  Field *c;
  if (int rc = field(g, "qcond", &c)) return rc;
  k_qscale<<<nb4(g), 256, 0, g->stream>>>(g->geo(), B.q->d, prm->qc_alias ? B.r->d : c->d, 0);
  HIPCHK(hipGetLastError());
  return 0;
}

int q_stats(sunhip_grid *g, CGScal *d, sunhip_cg_stats *st) {
  HIPCHK(hipMemcpyAsync(g->h_sc, d, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
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
    st->outer = 0;
    st->inner = 0;
  }
  return 0;
}

}  // namespace

namespace sunhip {
int field32(sunhip_grid *g, const char *name, float **out) {
  auto it = g->fields32.find(name);
  if (it == g->fields32.end()) {
    float *p = nullptr;
    const size_t n = (size_t)std::max(g->NT, g->NE);
    int rc = dev_alloc(g, (void **)&p, n * sizeof(float));
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(p, 0, n * sizeof(float), g->stream));
    it = g->fields32.emplace(name, p).first;
  }
  *out = it->second;
  return 0;
}

int upload_scal(sunhip_grid *g, CGScal *d, const sunhip_cg_params *prm, CGScal *stage) {
  CGScal h{};
  h.tol = prm->epsilon;
  h.maxit = prm->maxiters;
  h.precond = prm->precond;
  h.resnorm = prm->resnorm;
  if (!stage) stage = g->h_sc;
  *stage = h;
  HIPCHK(hipMemcpyAsync(d, stage, sizeof h, hipMemcpyHostToDevice, g->stream));
  return 0;
}

int q_begin64(sunhip_grid *g, const sunhip_cg_params *prm, QSolver<double> *S, CGScal *stage, CGScal *sc) {
  HIPCHK(hipSetDevice(g->device));
  F64 B;
  return q_setup(g, prm, B, *S, stage, sc);
}

int q_end64(sunhip_grid *g, const QSolver<double> &S) {
  k_qfinish<double><<<nb4(g), 256, 0, g->stream>>>(g->geo(), S.x, S.p0, S.p1, S.sc);
  HIPCHK(hipGetLastError());
  return 0;
}

int q_end32(sunhip_grid *g, const QSolver<float> &S) {
  k_qfinish<float><<<nb4(g), 256, 0, g->stream>>>(g->geo(), S.x, S.p0, S.p1, S.sc);
  HIPCHK(hipGetLastError());
  return 0;
}

int q_mixed_prepare(sunhip_grid *g, const sunhip_cg_params *prm, QSolver<float> *S32, QVec<double, float> *R) {
  HIPCHK(hipSetDevice(g->device));
  int rc = check_q(g, prm);
  if (rc) return rc;
  if (prm->precond != 2) { set_error("the mixed form uses the vertical-line preconditioner (qprecond=2)"); return -1; }
  F64 B;
  if ((rc = f64bufs(g, B))) return rc;
  float *c32, *f32, *x32, *r32, *z32, *t32, *pa, *pb, *b32;
  rc = field32(g, "coef", &c32) | field32(g, "dzf", &f32) | field32(g, "r", &r32) | field32(g, "z", &z32) |
       field32(g, "rtmp", &t32) | field32(g, "p0", &pa) | field32(g, "p1", &pb) | field32(g, "bq", &b32) |
       field32(g, "e", &x32);
  if (rc) return rc;
  if (!g->d_D32 && (rc = dev_alloc(g, (void **)&g->d_D32, sizeof(float) * (size_t)g->NES))) return rc;
  const Geo G = g->geo();
  k_qcoef<<<nb4(g), 256, 0, g->stream>>>(G, B.dzz->d, B.coef->d);
  k_cast<double, float><<<(unsigned)((g->NT + 255) / 256), 256, 0, g->stream>>>(B.coef->d, c32, g->NT);
  k_cast<double, float><<<(unsigned)((g->NE + 255) / 256), 256, 0, g->stream>>>(B.dzf->d, f32, g->NE);
  k_cast<double, float><<<(unsigned)((g->NES + 255) / 256), 256, 0, g->stream>>>(B.D->d, g->d_D32, g->NES);
  HIPCHK(hipGetLastError());
  if ((rc = launch_factor<float>(g, c32, b32))) return rc;
  ZReg Z{};
  if ((rc = q_windows(g, B.coef->d, B.dzf->d, &Z))) return rc;
  *S32 = QSolver<float>{g->d_ms, c32, f32, g->d_D32, x32, r32, z32, t32, pa, pb, b32};
  S32->Z = Z;
  QVec<double, float> V{};
  V.coef = B.coef->d; V.dzf = B.dzf->d; V.D = B.D->d; V.vin = B.q->d; V.src = B.src->d; V.out = r32;
  V.Z = Z;
  *R = V;
  return 0;
}

int set_zdz(sunhip_grid *g, const double *dz) {
  const size_t n = (size_t)g->Nkmax + 16;
  if (!g->d_zdzt && dev_alloc(g, (void **)&g->d_zdzt, n * sizeof(double))) return -2;
  std::vector<double> t(n, 0.0);
  for (int k = 0; k < g->Nkmax; k++) t[k] = dz[k];
  // the common dz[k] + dz[k-1] (the operands and order of QCoefficients'
  // dzz[k] + dzz[k-1]), or 0 when the layers differ
  double den = g->Nkmax > 1 ? dz[1] + dz[0] : 0.0;
  for (int k = 2; k < g->Nkmax; k++)
    if (!(dz[k] + dz[k - 1] == den)) den = 0.0;
  if (!(den > 0)) den = 0.0;
  HIPCHK(hipMemcpyAsync(g->d_zdzt, t.data(), n * sizeof(double), hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));  // the host vector goes out of scope
  g->zden = den;
  g->have_zdz = true;
  return 0;
}

int q_windows(sunhip_grid *g, const double *coef, const double *dzf, ZReg *Z) {
  *Z = ZReg{};
  if (!g->have_zdz || !g->tune_zreg) return 0;
  const int NF = nf_class(g->mf);
  if (!g->d_cwin) {
    int rc = dev_alloc(g, (void **)&g->d_cwin, sizeof(int) * (size_t)g->NS);
    if (!rc) rc = dev_alloc(g, (void **)&g->d_fwin, sizeof(int) * (size_t)g->NS * NF);
    if (!rc) rc = dev_alloc(g, (void **)&g->d_cex, sizeof(double) * (size_t)g->NS * 4);
    if (!rc) rc = dev_alloc(g, (void **)&g->d_fex, sizeof(double) * (size_t)g->NS * 2 * NF);
    if (!rc) rc = dev_alloc(g, (void **)&g->d_tflag, sizeof(int) * (size_t)g->Nt);
    if (rc) return rc;
  }
  with_nf(g->mf, [&](auto nfc) {
    constexpr int NFC = decltype(nfc)::value;
    k_qwin<NFC><<<nb4(g), 256, 0, g->stream>>>(g->geo(), coef, dzf, g->zden, g->d_zdzt, g->d_cwin, g->d_fwin,
                                               g->d_cex, g->d_fex, g->d_tflag);
    return 0;
  });
  HIPCHK(hipGetLastError());
  *Z = ZReg{g->d_cwin, g->d_fwin, g->d_cex, g->d_fex, g->d_tflag, g->d_zdzt, g->zden};
  return 0;
}

int q_inner_begin(sunhip_grid *g, const sunhip_cg_params *inner, QSolver<float> &S32, CGScal *stage) {
  int rc;
  if ((rc = upload_scal(g, S32.sc, inner, stage))) return rc;
  k_fill<float><<<(unsigned)((g->NT + 255) / 256), 256, 0, g->stream>>>(S32.x, g->NT, 0.0f);
  HIPCHK(hipGetLastError());
  return cg_start<float>(g, inner, S32);
}

int q_mix_update(sunhip_grid *g, double *x, const float *e) {
  k_mix_update<<<nb4(g), 256, 0, g->stream>>>(g->geo(), x, e);
  HIPCHK(hipGetLastError());
  return 0;
}
}  // namespace sunhip

extern "C" {

int sunhip_q_solve(sunhip_grid *g, const sunhip_cg_params *prm, sunhip_cg_stats *st) {
  HIPCHK(hipSetDevice(g->device));
  F64 B;
  QSolver<double> S;
  int rc;
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  if ((rc = q_setup(g, prm, B, S))) return rc;
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  if ((rc = cg_loop<double>(g, prm, S, PK_QAPPLY, PK_QPREC))) return rc;
  if ((rc = q_unscale(g, prm, B))) return rc;
  return q_stats(g, g->d_qs, st);
}

int sunhip_q_solve_fixed(sunhip_grid *g, const sunhip_cg_params *prm, int n, sunhip_cg_stats *st) {
  HIPCHK(hipSetDevice(g->device));
  sunhip_cg_params p = *prm;
  p.epsilon = 0;       // never satisfied
  p.maxiters = n;
  F64 B;
  QSolver<double> S;
  int rc;
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  if ((rc = q_setup(g, &p, B, S))) return rc;
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  for (int it = 0; it < n; it++)
    if ((rc = cg_iteration<double>(g, &p, S, it, PK_QAPPLY, PK_QPREC))) return rc;
  k_qfinish<double><<<nb4(g), 256, 0, g->stream>>>(g->geo(), S.x, S.p0, S.p1, S.sc);
  HIPCHK(hipGetLastError());
  if ((rc = q_unscale(g, &p, B))) return rc;
  return q_stats(g, g->d_qs, st);
}

// Mixed precision (FP32 inner PCG inside FP64 iterative refinement):
//   r = src - L x in FP64 (true residual), eps = r.r in FP64; stop when
//   sqrt(eps/eps0) < epsilon (eps0 the first true residual);
//   otherwise solve L e = r in FP32 (vertical-line PCG, dots accumulated in
//   FP64) to inner_tol relative, then x += e in FP64.
int sunhip_q_solve_mixed(sunhip_grid *g, const sunhip_cg_params *prm, double inner_tol, int max_outer,
                         sunhip_cg_stats *st, int *inner_iters) {
  HIPCHK(hipSetDevice(g->device));
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  QSolver<float> S;
  QVec<double, float> R;
  int rc = q_mixed_prepare(g, prm, &S, &R);
  if (rc) return rc;
  CGScal *so = g->d_mo;
  sunhip_cg_params inner = *prm;
  inner.epsilon = inner_tol;
  inner.resnorm = 0;
  double eps0 = -1, eps = 0;
  int outer = 0, total_inner = 0, converged = 0;
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  // FP64 true residual, cast to FP32 for the inner solve; its r.r on the host
  auto residual = [&]() -> int {
    int e = launch_apply<QM_RESID_DOT, double, float>(g, so, R, Red{g->d_part, 0, nullptr, OP_STORE}, PK_RESID);
    if (e) return e;
    HIPCHK(hipMemcpyAsync(g->h_sc, so, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
    return 0;
  };
  if ((rc = residual())) return rc;
  HIPCHK(hipStreamSynchronize(g->stream));
  for (;;) {
    eps = g->h_sc->pz;
    if (eps0 < 0) eps0 = eps;
    const double ratio = eps == 0 ? 0.0 : sqrt(eps / eps0);
    if (eps == 0 || ratio < prm->epsilon) { converged = 1; break; }
    if (outer >= max_outer) break;
    // inner FP32 PCG from e = 0, asked for no more reduction than the outer
    // target still needs (with a margin for the FP32/FP64 residual gap)
    inner.epsilon = std::max(inner_tol, 0.25 * prm->epsilon / ratio);
    if ((rc = q_inner_begin(g, &inner, S, nullptr))) return rc;
    if ((rc = cg_loop<float>(g, &inner, S, PK_MAPPLY, PK_MPREC, false))) return rc;
    // x += e (with the inner solve's last update), the next residual, then
    // one wait for both the inner count and the residual
    k_mix_finish<<<nb4(g), 256, 0, g->stream>>>(g->geo(), const_cast<double *>(R.vin), S.x, S.p0, S.p1, S.sc);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&g->h_poll[0], S.sc, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
    if ((rc = residual())) return rc;
    HIPCHK(hipStreamSynchronize(g->stream));
    total_inner += g->h_poll[0].iter;
    outer++;
  }
  HIPCHK(hipEventRecord(g->ev2, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  if (st) {
    st->niters = outer;
    st->outer = outer;
    st->inner = total_inner;
    st->converged = converged;
    st->resid = eps0 > 0 ? sqrt(eps / eps0) : 0;
    st->eps0 = eps0;
    float ms = 0;
    hipEventElapsedTime(&ms, g->ev1, g->ev2);
    st->device_ms = ms;
    hipEventElapsedTime(&ms, g->ev0, g->ev1);
    st->setup_ms = ms;
  }
  if (inner_iters) *inner_iters = total_inner;
  return 0;
}

int sunhip_set_dz(sunhip_grid *g, const double *dz) {
  HIPCHK(hipSetDevice(g->device));
  if (!dz) { set_error("sunhip_set_dz: dz is required"); return -1; }
  return set_zdz(g, dz);
}

int sunhip_q_operator(sunhip_grid *g, const char *x, const char *y) {
  HIPCHK(hipSetDevice(g->device));
  if (!g->have_extents) { set_error("sunhip_q_operator: call sunhip_set_extents first"); return -1; }
  Field *fx, *fy, *coef, *dzf, *D;
  int rc = field(g, x, &fx) | field(g, y, &fy) | field(g, "coef", &coef) | field(g, "dzf", &dzf) |
           field(g, "D", &D);
  if (rc) return rc;
  QVec<double, double> V{};
  V.coef = coef->d; V.dzf = dzf->d; V.D = D->d; V.vin = fx->d; V.out = fy->d;
  read_tuning(g);
  if ((rc = launch_apply<QM_PLAIN, double, double>(g, g->d_qs, V, Red{g->d_part, 0, nullptr, OP_NONE}))) return rc;
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

// Coefficients from the device field "dzz" into "coef".
int sunhip_q_coefficients(sunhip_grid *g) {
  HIPCHK(hipSetDevice(g->device));
  Field *dzz, *coef;
  int rc = field(g, "dzz", &dzz) | field(g, "coef", &coef);
  if (rc) return rc;
  k_qcoef<<<nb4(g), 256, 0, g->stream>>>(g->geo(), dzz->d, coef->d);
  HIPCHK(hipGetLastError());
  return 0;
}

int sunhip_q_precondition(sunhip_grid *g, const char *x, const char *xc) {
  HIPCHK(hipSetDevice(g->device));
  if (!g->have_extents) { set_error("sunhip_q_precondition: call sunhip_set_extents first"); return -1; }
  read_tuning(g);
  Field *fx, *fxc, *coef, *bq;
  int rc = field(g, x, &fx) | field(g, xc, &fxc) | field(g, "coef", &coef) | field(g, "qbp", &bq);
  if (rc) return rc;
  if ((rc = launch_factor<double>(g, coef->d, bq->d))) return rc;
  if ((rc = launch_prec<false, double>(g, g->d_qs, coef->d, nullptr, fx->d, fxc->d, bq->d, OP_NONE, nullptr)))
    return rc;
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

// Tuning knobs: "unroll" levels per load group (1, 2, 4); "reduce" 0 last
// block / 1 finisher kernel.  -1 restores the default (environment or
// built-in).
int sunhip_set_tuning(sunhip_grid *g, const char *key, int value) {
  if (!strcmp(key, "unroll")) g->tune_unroll = value;
  else if (!strcmp(key, "prec_unroll")) {
    if (value != 1 && value != 2 && value != 4 && value != 8) { set_error("prec_unroll must be 1, 2, 4 or 8"); return -1; }
    g->tune_punroll = value;
  }
  else if (!strcmp(key, "prec_head")) {
    if (value != -1 && value != 0 && value != 8 && value != 16 && value != 24 && value != 32) {
      set_error("prec_head must be -1, 0, 8, 16, 24 or 32");
      return -1;
    }
    g->tune_phead = value;
  }
  else if (!strcmp(key, "xchg")) {
    if (value < -1 || value > 6) { set_error("xchg must be -1..6"); return -1; }
    g->tune_xchg = value;
  }
  else if (!strcmp(key, "zreg")) g->tune_zreg = value != 0;
  else if (!strcmp(key, "order")) g->tune_order = value != 0;
  else if (!strcmp(key, "tvd_group")) {
    if (value != 1 && value != 2 && value != 4) { set_error("tvd_group must be 1, 2 or 4"); return -1; }
    g->tune_tvdgrp = value;
  }
  else if (!strcmp(key, "prec_lds")) {
    if (value != 0 && value != 12) { set_error("prec_lds must be 0 or 12"); return -1; }
    g->tune_plds = value;
  }
  else if (!strcmp(key, "tvd_head")) {
    if (value != 0 && value != 8 && value != 16) { set_error("tvd_head must be 0, 8 or 16"); return -1; }
    g->tune_tvdhead = value;
  }
  else if (!strcmp(key, "tvd_sums_group")) {
    if (value != 1 && value != 2 && value != 4) { set_error("tvd_sums_group must be 1, 2 or 4"); return -1; }
    g->tune_sumsgrp = value;
  }
  else if (!strcmp(key, "h_blocks")) {
    if (value != 256 && value != 512 && value != 1024 && value != 2048) {
      set_error("h_blocks must be 256, 512, 1024 or 2048");
      return -1;
    }
    g->tune_hblk = value;
  }
  else if (!strcmp(key, "reduce")) g->tune_reduce = value;
  else if (!strcmp(key, "tiles")) {
    if (value != 4 && value != 8 && value != 16) { set_error("tiles must be 4, 8 or 16"); return -1; }
    g->tune_tiles = value;
  } else if (!strcmp(key, "ntstore")) g->tune_nts = value != 0;
  else if (!strcmp(key, "rzid")) g->tune_rzid = value != 0;
  else if (!strcmp(key, "minwaves")) g->tune_minw = value;
  else if (!strcmp(key, "blocklds")) g->tune_blk = value != 0;
  else if (!strcmp(key, "band")) {
    if (value < 0 || (value > 0 && value < 8)) { set_error("band must be 0 or at least 8 levels"); return -1; }
    g->tune_band = value;
  }
  else { set_error("unknown tuning key %s", key); return -1; }
  return 0;
}

}  // extern "C"

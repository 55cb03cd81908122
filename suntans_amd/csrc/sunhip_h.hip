// sunhip_h.hip -- free-surface PCG on gfx950.
//
// Reference: CGSolve main/phys.c:3086-3217, HCoefficients :3252-3269,
// OperatorH :3335-3349, HPreconditioner :3227-3235, InnerProduct :3280-3294.
// Two kernels per iteration with the scalars on the device; each kernel
// adds the previous one's block sums itself, in one fixed order, so that no
// one-block finishing kernel runs between them (at Nc = 1e6 a launch costs
// about as much as the work).  Templated on the storage type like the
// pressure solve: T=double is the drop-in, T=float the inner solve of the
// mixed-precision form (sunhip_h_solve_mixed).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "sunhip_solvers.h"

using namespace sunhip;

// Scratch of the allocation-free free-surface solve (sunhip_h_solve_noalloc):
// the cell and face coefficients and the five CG vectors of one grid,
// (6 + maxfaces) doubles per cell slot.  Each device that runs the solve
// holds its own copy for the life of the process; calls on one device share
// it.
#define SUNHIP_NOALLOC_BYTES (32u << 20)
static __device__ __align__(256) double scratch[SUNHIP_NOALLOC_BYTES / sizeof(double)];

template <int NV>
__device__ void h_partials(const double (&v)[NV], double *part) {
  __shared__ double sh[NV][4];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    const double s = wave_sum(v[i]);
    if (l == 0) sh[i][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) part[(size_t)blockIdx.x * NV + i] = ((sh[i][0] + sh[i][1]) + sh[i][2]) + sh[i][3];
  }
}

template <int NV>
__global__ __launch_bounds__(1024) void k_hfinalize(const double *__restrict__ part, int nparts, CGScal *sc, int op) {
  if ((op == HOP_APPLY || op == HOP_ITER) && sc->done) return;
  __shared__ double sh[NV][16];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double s = 0;
    for (int b = threadIdx.x; b < nparts; b += 1024) s += part[(size_t)b * NV + i];
    s = wave_sum(s);
    if (l == 0) sh[i][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double s = 0;
      for (int j = 0; j < 16; j++) s += sh[i][j];
      tot[i] = s;
    }
    h_op(sc, op, tot);
  }
}

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

template <typename T>
__global__ void k_hzero_comp(Geo G, T *__restrict__ x) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < G.NS && G.nk[s] > 0 && G.comp[s] == 1) x[s] = 0;
}

template <int MODE, typename T, typename TO>
__global__ __launch_bounds__(256) void k_happly(Geo G, HVec<T, TO> V, CGScal *sc, double *part) {
  if (MODE == HM_FUSED && sc->done) return;
  const int s = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const T mu = MODE == HM_FUSED ? (T)sc->mu : T(0), nu = MODE == HM_FUSED ? (T)sc->nu : T(0);
  const double dot = happly_cell<MODE, T, TO>(G, V, mu, nu, s);
  if (MODE == HM_FUSED || MODE == HM_RESID_DOT) {
    double v[1] = {dot};
    h_partials<1>(v, part);
  }
}

// ---- the iteration: two kernels, each finishing the other's dot products ----
//
// Each iteration kernel starts by adding the previous kernel's block sums
// (every block the same sums in the same order, so every block holds the same
// scalars) and applying the pending step of the reference's loop to the
// scalars: HOP_ITER (main/phys.c:3175-3203) before the operator, HOP_APPLY
// (alpha / p.z) before the update.  The scalars alternate between two slots:
// a kernel reads one and its block 0 writes the other, with its own step as
// the one pending.  The block sums alternate between two areas likewise.
// At most 1024 blocks (tuning key "h_blocks"; a few cells per thread), so
// that each block reads a few KB of sums.  (Issuing a block's cell loads before the sums, or blocks
// of 1024 threads with one cell each, measured slower at Nc = 1e6.)
constexpr int HBS = 256;

// h_partials for HBS-thread blocks
template <int NV>
__device__ void h_partials_it(const double (&v)[NV], double *part) {
  __shared__ double sh[NV][HBS / 64];
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
      for (int j = 0; j < HBS / 64; j++) s += sh[i][j];
      part[(size_t)blockIdx.x * NV + i] = s;
    }
  }
}

// sum over the nparts block sums of part (NV per block), every block alike
template <int NV>
__device__ void h_sum_parts(const double *__restrict__ part, int nparts, double (&tot)[NV]) {
  __shared__ double sh[NV][HBS / 64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double s = 0;
    for (int b = threadIdx.x; b < nparts; b += HBS) s += part[(size_t)b * NV + i];
    s = wave_sum(s);
    if (l == 0) sh[i][w] = s;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double s = 0;
    for (int j = 0; j < HBS / 64; j++) s += sh[i][j];
    tot[i] = s;
  }
}

// The scalars after in's pending step on the block sums of part (NV per
// block, written by the kernel before); block 0 stores them to out with
// next_op pending unless the solve is done.  out may be in (one block).
template <int NV>
__device__ void h_state(const CGScal *in, CGScal *out, const double *__restrict__ part, int nparts, int next_op,
                        CGScal &S) {
  __shared__ CGScal sS;
  const int pend = in->pend, done = in->done;
  __syncthreads();  // every thread has read pend before block 0 may overwrite it
  double tot[NV];
  if (pend && !done) h_sum_parts<NV>(part, nparts, tot);
  if (threadIdx.x == 0) {
    CGScal c = *in;
    if (pend && !done) h_op(&c, pend, tot);
    c.pend = 0;
    sS = c;
    if (blockIdx.x == 0) {
      CGScal o = c;
      o.pend = c.done ? 0 : next_op;
      *out = o;
    }
  }
  __syncthreads();
  S = sS;
}

// One cell of the fused operator (happly_cell, mode HM_FUSED) in two halves:
// the loads, unconditional at clamped indices, then the arithmetic in
// happly_cell's order.  A thread loads its first HCPT cells before it works
// on any, so their loads are in flight together (stores of one cell would
// otherwise keep the compiler from issuing the next cell's loads early).
constexpr int HCPT = 4;

template <typename T, int NF>
struct HCellIn {
  bool ok;
  int nf;
  T po0, r0, c, x0;
  int nb[NF];
  T fc[NF], rn[NF], pn[NF];
};

template <typename T, int NF>
__device__ __forceinline__ void hcell_direct(const Geo &G, const HVec<T, T> &V, int s, HCellIn<T, NF> &I) {
  const int sc = s < G.NS ? s : 0;
  I.ok = (s < G.NS) & (G.nk[sc] > 0) & (G.comp[sc] == 1);
  I.nf = G.nf[sc];
  I.po0 = V.pold[sc];
  I.r0 = V.rtmp[sc];
  I.c = V.coef[sc];
  I.x0 = V.x[sc];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    const int row = f < G.mf ? f : G.mf - 1;  // rows past maxfaces are not stored
    I.nb[f] = G.nbs[row * G.NS + sc];
    I.fc[f] = V.fcoef[row * G.NS + sc];
  }
}

template <typename T, int NF>
__device__ __forceinline__ void hcell_gather(const HVec<T, T> &V, int sc, HCellIn<T, NF> &I) {
#pragma unroll
  for (int f = 0; f < NF; f++) {
    if (!(f < I.nf && I.nb[f] >= 0)) I.nb[f] = -1;
    const int n = I.nb[f] >= 0 ? I.nb[f] : sc;
    I.rn[f] = V.rtmp[n];
    I.pn[f] = V.pold[n];
  }
}

template <typename T, int NF>
__device__ __forceinline__ double hcell_apply(const HVec<T, T> &V, const HCellIn<T, NF> &I, T mu, T nu, int s) {
  if (!I.ok) return 0;
  const T v0 = I.r0 + mu * I.po0;
  T y = I.c * v0;
#pragma unroll
  for (int f = 0; f < NF; f++)
    if (I.nb[f] >= 0) y -= I.fc[f] * (I.rn[f] + mu * I.pn[f]);
  V.pnew[s] = v0;
  V.x[s] = I.x0 + nu * I.po0;
  V.out[s] = y;
  return (double)v0 * (double)y;
}

// p = rtmp + mu p_old ; x += nu p_old ; z = L p ; block sums of p.z
template <typename T, int NF>
__global__ __launch_bounds__(HBS) void k_happly_it(Geo G, HVec<T, T> V, const CGScal *sin, CGScal *sout,
                                                   const double *pin, double *pout, int nparts) {
  CGScal S;
  h_state<2>(sin, sout, pin, nparts, HOP_APPLY, S);
  if (S.done) return;
  const T mu = (T)S.mu, nu = (T)S.nu;
  const int stride = gridDim.x * HBS;
  const int s0 = xcd_block(blockIdx.x, gridDim.x) * HBS + threadIdx.x;
  HCellIn<T, NF> I[HCPT];
#pragma unroll
  for (int j = 0; j < HCPT; j++) hcell_direct<T, NF>(G, V, s0 + j * stride, I[j]);
#pragma unroll
  for (int j = 0; j < HCPT; j++) hcell_gather<T, NF>(V, s0 + j * stride < G.NS ? s0 + j * stride : 0, I[j]);
  double dot = 0;
#pragma unroll
  for (int j = 0; j < HCPT; j++) dot += hcell_apply<T, NF>(V, I[j], mu, nu, s0 + j * stride);
  for (int s = s0 + HCPT * stride; s < G.NS; s += stride) dot += happly_cell<HM_FUSED, T, T>(G, V, mu, nu, s);
  double v[1] = {dot};
  h_partials_it<1>(v, pout);
}

// r -= nu z ; rtmp = M^-1 r ; block sums of r.rtmp and r.r (hprec_cell's
// arithmetic; the first HCPT cells' loads issued together)
template <typename T>
__global__ __launch_bounds__(HBS) void k_hprec_it(Geo G, const T *__restrict__ coef, const T *__restrict__ z,
                                                  T *__restrict__ r, T *__restrict__ rt, const CGScal *sin,
                                                  CGScal *sout, const double *pin, double *pout, int nparts) {
  CGScal S;
  h_state<1>(sin, sout, pin, nparts, HOP_ITER, S);
  if (S.done) return;
  const T nu = (T)S.nu;
  const int pc = S.precond;
  const int stride = gridDim.x * HBS;
  const int s0 = xcd_block(blockIdx.x, gridDim.x) * HBS + threadIdx.x;
  bool ok[HCPT];
  T rk[HCPT], zk[HCPT], ck[HCPT];
#pragma unroll
  for (int j = 0; j < HCPT; j++) {
    const int s = s0 + j * stride, sc = s < G.NS ? s : 0;
    ok[j] = (s < G.NS) & (G.nk[sc] > 0) & (G.comp[sc] == 1);
    rk[j] = r[sc];
    zk[j] = z[sc];
    ck[j] = coef[sc];
  }
  double rz = 0, rr = 0;
#pragma unroll
  for (int j = 0; j < HCPT; j++) {
    if (!ok[j]) continue;
    const int s = s0 + j * stride;
    const T rn = rk[j] - nu * zk[j];
    r[s] = rn;
    const T u = pc == 1 ? rn / ck[j] : rn;
    rt[s] = u;
    rz += (double)rn * (double)u;
    rr += (double)rn * (double)rn;
  }
  for (int s = s0 + HCPT * stride; s < G.NS; s += stride) hprec_cell<T>(G, coef, z, r, rt, nu, pc, true, s, rz, rr);
  double v[2] = {rz, rr};
  h_partials_it<2>(v, pout);
}

// Apply the pending step in place (one block): the scalars the host reads,
// also written to host (pinned host memory, may be null) so that polling
// needs no copy.
__global__ __launch_bounds__(HBS) void k_hsettle(CGScal *sc, const double *pin, int nparts, CGScal *host) {
  CGScal S;
  h_state<2>(sc, sc, pin, nparts, 0, S);
  if (host && threadIdx.x == 0) *host = S;
}

// r -= nu*z ; rtmp = r/hcoef (hprecond=1) or r ; dots.
template <typename T>
__global__ __launch_bounds__(256) void k_hprec(Geo G, const T *__restrict__ coef, const T *__restrict__ z,
                                               T *__restrict__ r, T *__restrict__ rt, CGScal *sc, double *part,
                                               int update) {
  if (update && sc->done) return;
  const int s = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  double rz = 0, rr = 0;
  hprec_cell<T>(G, coef, z, r, rt, update ? (T)sc->nu : T(0), sc->precond, update != 0, s, rz, rr);
  double v[2] = {rz, rr};
  h_partials<2>(v, part);
}

template <typename T>
__global__ void k_hfinish(Geo G, T *__restrict__ x, const T *__restrict__ p0, const T *__restrict__ p1,
                          const CGScal *sc) {
  if (sc->iter == 0) return;
  const T *__restrict__ p = (sc->iter & 1) ? p1 : p0;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < G.NS && G.nk[s] > 0 && G.comp[s] == 1) x[s] = x[s] + (T)sc->nu * p[s];
}

template <typename T>
__global__ void k_hfill(T *p, int n, T v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename TI, typename TO>
__global__ void k_hcast(const TI *__restrict__ a, TO *__restrict__ b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (TO)a[i];
}

__global__ void k_hmix_update(Geo G, double *__restrict__ x, const float *__restrict__ e) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < G.NS && G.nk[s] > 0 && G.comp[s] == 1) x[s] = x[s] + (double)e[s];
}

namespace {

struct HBufs {
  Field *h, *src, *D, *coef, *r, *z, *rt, *p0, *p1;
  double *fcoef;
};

int nbh(const sunhip_grid *g) { return (g->NS + 255) / 256; }

int hbufs(sunhip_grid *g, HBufs &B) {
  int rc = field(g, "h", &B.h) | field(g, "htmp", &B.src) | field(g, "D", &B.D) | field(g, "hcoef", &B.coef) |
           field(g, "hr", &B.r) | field(g, "hz", &B.z) | field(g, "hrt", &B.rt) | field(g, "hp0", &B.p0) |
           field(g, "hp1", &B.p1);
  if (rc) return rc;
  // face coefficients: maxfaces rows of NS
  auto it = g->fields.find("#hfcoef");
  if (it == g->fields.end()) {
    Field f;
    f.kind = CELL2;
    f.n = (size_t)g->mf * g->NS;
    if ((rc = dev_alloc(g, (void **)&f.d, f.n * sizeof(double)))) return rc;
    HIPCHK(hipMemsetAsync(f.d, 0, f.n * sizeof(double), g->stream));
    it = g->fields.emplace("#hfcoef", f).first;
  }
  B.fcoef = it->second.d;
  return 0;
}

int h_coefficients(sunhip_grid *g, HBufs &B, double grav, double theta, double dt) {
  const double td = theta * dt;
  const double tmp = grav * (td * td);  // grav*pow(theta*dt,2)
  k_hcoef<<<nbh(g), 256, 0, g->stream>>>(g->geo(), B.D->d, tmp, B.coef->d, B.fcoef);
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T>
int h_start(sunhip_grid *g, HSolver<T> &S) {
  const int nb = nbh(g);
  k_hprec<T><<<nb, 256, 0, g->stream>>>(g->geo(), S.coef, nullptr, S.r, S.rt, S.sc, g->d_part, 0);
  k_hfinalize<2><<<1, 1024, 0, g->stream>>>(g->d_part, nb, S.sc, HOP_SETUP);
  k_hfill<T><<<nb, 256, 0, g->stream>>>(S.p0, g->NS, T(0));
  k_hfill<T><<<nb, 256, 0, g->stream>>>(S.p1, g->NS, T(0));
  HIPCHK(hipGetLastError());
  return 0;
}

// blocks of the iteration kernels; their sums go to d_part (p.z, nb values)
// and d_part + nb (r.rtmp and r.r, 2 nb values)
int nbh_it(const sunhip_grid *g) { return std::min(g->tune_hblk, (g->NS + HBS - 1) / HBS); }

// One iteration: S.sc[0] holds the scalars before and after it (with the
// step HOP_ITER pending), S.sc[1] those between its two kernels.
template <typename T>
int h_iteration(sunhip_grid *g, HSolver<T> &S, int it, int pk_a, int pk_p) {
  const Geo G = g->geo();
  HVec<T, T> V{};
  V.coef = S.coef; V.fcoef = S.fcoef; V.rtmp = S.rt;
  V.pold = (it & 1) ? S.p1 : S.p0;
  V.pnew = (it & 1) ? S.p0 : S.p1;
  V.x = S.x; V.out = S.z;
  const int nb = nbh_it(g);
  double *pz = g->d_part, *prz = g->d_part + nb;
  int ps = prof_begin(g, pk_a);
  with_nf(g->mf, [&](auto nfc) {
    constexpr int NF = decltype(nfc)::value;
    k_happly_it<T, NF><<<nb, HBS, 0, g->stream>>>(G, V, S.sc, S.sc + 1, prz, pz, nb);
    return 0;
  });
  prof_end(g, ps);
  ps = prof_begin(g, pk_p);
  k_hprec_it<T><<<nb, HBS, 0, g->stream>>>(G, S.coef, S.z, S.r, S.rt, S.sc + 1, S.sc, pz, prz, nb);
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  return 0;
}

// The last iteration's HOP_ITER applied to S.sc[0]; with host, the result
// also lands there.
int h_settle(sunhip_grid *g, CGScal *sc, CGScal *host = nullptr) {
  const int nb = nbh_it(g);
  k_hsettle<<<1, HBS, 0, g->stream>>>(sc, g->d_part + nb, nb, host);
  HIPCHK(hipGetLastError());
  return 0;
}

// poll_chunk for the free surface: the settle kernel writes the scalars
// into the pinned poll slot itself
int h_poll(sunhip_grid *g, CGScal *sc, int &slot, bool &pending, bool *done) {
  int rc = h_settle(g, sc, &g->h_poll[slot]);
  if (rc) return rc;
  HIPCHK(hipEventRecord(g->evp[slot], g->stream));
  *done = false;
  if (pending) {
    HIPCHK(hipEventSynchronize(g->evp[slot ^ 1]));
    *done = g->h_poll[slot ^ 1].done != 0;
  }
  pending = true;
  slot ^= 1;
  return 0;
}

template <typename T>
int h_loop(sunhip_grid *g, const sunhip_cg_params *prm, HSolver<T> &S, int pk_a, int pk_p) {
  int every = prm->check_every > 0 ? prm->check_every : 8;
  if (every & 1) every++;
  int it = 0, rc, slot = 0;
  bool pending = false, done = false;
  // iterations enqueued after convergence return at once (device flag)
  while (!done && it < prm->maxiters) {
    for (int b = 0; b < every && it < prm->maxiters; b++, it++)
      if ((rc = h_iteration<T>(g, S, it, pk_a, pk_p))) return rc;
    if ((rc = h_poll(g, S.sc, slot, pending, &done))) return rc;
  }
  k_hfinish<T><<<nbh(g), 256, 0, g->stream>>>(g->geo(), S.x, S.p0, S.p1, S.sc);
  HIPCHK(hipGetLastError());
  return 0;
}

int upload_hscal(sunhip_grid *g, CGScal *d, const sunhip_cg_params *prm) { return upload_scal(g, d, prm); }

int fill_stats(sunhip_grid *g, const CGScal &c, sunhip_cg_stats *st, int outer = 0, int inner = 0) {
  if (!st) return 0;
  st->outer = outer;
  st->inner = inner;
  st->niters = c.nloop;
  st->converged = c.converged;
  st->resid = c.resid;
  st->eps0 = c.eps0;
  float ms = 0;
  hipEventElapsedTime(&ms, g->ev1, g->ev2);
  st->device_ms = ms;
  hipEventElapsedTime(&ms, g->ev0, g->ev1);
  st->setup_ms = ms;
  return 0;
}

}  // namespace

namespace sunhip {
int h_begin64(sunhip_grid *g, const sunhip_cg_params *prm, double grav, double theta, double dt,
              HSolver<double> *S, CGScal *stage) {
  HIPCHK(hipSetDevice(g->device));
  if (prm->precond != 0 && prm->precond != 1) {
    set_error("hprecond=%d not supported (0 or 1)", prm->precond);
    return -1;
  }
  HBufs B;
  int rc = hbufs(g, B);
  if (rc) return rc;
  const Geo G = g->geo();
  const int nb = nbh(g);
  if ((rc = upload_scal(g, g->d_hs, prm, stage))) return rc;
  if ((rc = h_coefficients(g, B, grav, theta, dt))) return rc;
  // x = 0 on computational cells; r = b - L x lifts type-3 boundary values
  k_hzero_comp<double><<<nb, 256, 0, g->stream>>>(G, B.h->d);
  HVec<double, double> V{};
  V.coef = B.coef->d; V.fcoef = B.fcoef; V.vin = B.h->d; V.src = B.src->d; V.out = B.r->d;
  k_happly<HM_RESID, double, double><<<nb, 256, 0, g->stream>>>(G, V, g->d_hs, g->d_part);
  HIPCHK(hipGetLastError());
  *S = HSolver<double>{g->d_hs, B.coef->d, B.fcoef, B.h->d, B.r->d, B.z->d, B.rt->d, B.p0->d, B.p1->d};
  return h_start<double>(g, *S);
}

int h_end64(sunhip_grid *g, const HSolver<double> &S) {
  k_hfinish<double><<<nbh(g), 256, 0, g->stream>>>(g->geo(), S.x, S.p0, S.p1, S.sc);
  HIPCHK(hipGetLastError());
  return 0;
}
}  // namespace sunhip

namespace {
// The device address of scratch on each device, taken once with that device
// current (the host-side name of the array is not a device address).
constexpr int kMaxDev = 64;
double *g_scratch[kMaxDev];
std::mutex g_scratch_mu;

int scratch_on(int device, double **out) {
  if (device < 0 || device >= kMaxDev) {
    set_error("sunhip_h_solve_noalloc: device %d out of range", device);
    return -1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("sunhip_h_solve_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return -3;
    }
    g_scratch[device] = (double *)p;
  }
  *out = g_scratch[device];
  return 0;
}
}  // namespace

extern "C" {

// sunhip_h_solve with every work array in the library's static scratch:
// no device memory is allocated.  "h", "htmp" and "D" must already be on
// the device (uploaded); the same kernels run in the same order, so the
// result is the one sunhip_h_solve gives.
int sunhip_h_solve_noalloc(sunhip_grid *g, const sunhip_cg_params *prm, double grav, double theta, double dt,
                           sunhip_cg_stats *st) {
  HIPCHK(hipSetDevice(g->device));
  if (prm->precond != 0 && prm->precond != 1) {
    set_error("hprecond=%d not supported (0 or 1)", prm->precond);
    return -1;
  }
  const size_t NS = (size_t)g->NS, need = (6 + (size_t)g->mf) * NS * sizeof(double);
  if (need > SUNHIP_NOALLOC_BYTES) {
    set_error("sunhip_h_solve_noalloc: the grid needs %zu bytes of scratch ((6 + maxfaces) x %zu cell slots x 8), "
              "more than its %u-byte maximum; use sunhip_h_solve", need, NS, (unsigned)SUNHIP_NOALLOC_BYTES);
    return -1;
  }
  Field *fh, *fsrc, *fD;
  {
    auto a = g->fields.find("h"), b = g->fields.find("htmp"), c = g->fields.find("D");
    if (a == g->fields.end() || b == g->fields.end() || c == g->fields.end()) {
      set_error("sunhip_h_solve_noalloc: upload \"h\", \"htmp\" and \"D\" first (it allocates nothing)");
      return -1;
    }
    fh = &a->second; fsrc = &b->second; fD = &c->second;
  }
  double *ws = nullptr;
  if (int rc = scratch_on(g->device, &ws)) return rc;
  double *coef = ws, *fcoef = coef + NS, *r = fcoef + (size_t)g->mf * NS, *z = r + NS, *rt = z + NS, *p0 = rt + NS,
         *p1 = p0 + NS;
  const Geo G = g->geo();
  const int nb = nbh(g);
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  // the last call's contents: clear all of it (slots the kernels skip are read as 0)
  const int nz = (int)((6 + (size_t)g->mf) * NS);
  k_hfill<double><<<(nz + 255) / 256, 256, 0, g->stream>>>(ws, nz, 0.0);
  HIPCHK(hipGetLastError());
  int rc = upload_scal(g, g->d_hs, prm);
  if (rc) return rc;
  const double td = theta * dt;
  k_hcoef<<<nb, 256, 0, g->stream>>>(G, fD->d, grav * (td * td), coef, fcoef);
  k_hzero_comp<double><<<nb, 256, 0, g->stream>>>(G, fh->d);
  HVec<double, double> V{};
  V.coef = coef; V.fcoef = fcoef; V.vin = fh->d; V.src = fsrc->d; V.out = r;
  k_happly<HM_RESID, double, double><<<nb, 256, 0, g->stream>>>(G, V, g->d_hs, g->d_part);
  HIPCHK(hipGetLastError());
  HSolver<double> S{g->d_hs, coef, fcoef, fh->d, r, z, rt, p0, p1};
  if ((rc = h_start<double>(g, S))) return rc;
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  if ((rc = h_loop<double>(g, prm, S, PK_HAPPLY, PK_HPREC))) return rc;
  HIPCHK(hipMemcpyAsync(g->h_sc, g->d_hs, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipEventRecord(g->ev2, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  return fill_stats(g, *g->h_sc, st);
}

}  // extern "C"

extern "C" {

int sunhip_h_solve(sunhip_grid *g, const sunhip_cg_params *prm, double grav, double theta, double dt,
                   sunhip_cg_stats *st) {
  HIPCHK(hipSetDevice(g->device));
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  HSolver<double> S;
  int rc = h_begin64(g, prm, grav, theta, dt, &S, nullptr);
  if (rc) return rc;
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  if ((rc = h_loop<double>(g, prm, S, PK_HAPPLY, PK_HPREC))) return rc;
  HIPCHK(hipMemcpyAsync(g->h_sc, g->d_hs, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipEventRecord(g->ev2, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  return fill_stats(g, *g->h_sc, st);
}

// Mixed precision free surface: FP64 true residual r = b - L h (after the
// reference's boundary lift with h = 0 inside), FP32 inner Jacobi-PCG on
// L e = r to inner_tol (adaptive toward the target), h += e in FP64, until
// sqrt(r.r / eps0) < epsilon.  eps0 follows the reference's rule
// (main/phys.c:3150-3156): r0.r0, except that without the preconditioner
// and without resnorm it is 1, an absolute test on the residual norm.  The
// inner solves always stop on a reduction relative to their own first
// residual.
int sunhip_h_solve_mixed(sunhip_grid *g, const sunhip_cg_params *prm, double grav, double theta, double dt,
                         double inner_tol, int max_outer, sunhip_cg_stats *st, int *inner_iters) {
  HIPCHK(hipSetDevice(g->device));
  HBufs B;
  int rc = hbufs(g, B);
  if (rc) return rc;
  float *c32, *f32, *x32, *r32, *z32, *t32, *pa, *pb;
  if (!g->d_h32) {
    if ((rc = dev_alloc(g, (void **)&g->d_h32, sizeof(float) * (size_t)g->NS * (7 + g->mf)))) return rc;
  }
  c32 = g->d_h32; x32 = c32 + g->NS; r32 = x32 + g->NS; z32 = r32 + g->NS; t32 = z32 + g->NS;
  pa = t32 + g->NS; pb = pa + g->NS; f32 = pb + g->NS;
  const Geo G = g->geo();
  const int nb = nbh(g);
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  if ((rc = h_coefficients(g, B, grav, theta, dt))) return rc;
  k_hcast<double, float><<<nb, 256, 0, g->stream>>>(B.coef->d, c32, g->NS);
  k_hcast<double, float><<<(g->mf * g->NS + 255) / 256, 256, 0, g->stream>>>(B.fcoef, f32, g->mf * g->NS);
  k_hzero_comp<double><<<nb, 256, 0, g->stream>>>(G, B.h->d);
  HIPCHK(hipGetLastError());
  sunhip_cg_params inner = *prm;
  inner.resnorm = 1;  // relative inner test for both preconditioners
  inner.precond = prm->precond;
  HSolver<float> S{g->d_ms, c32, f32, x32, r32, z32, t32, pa, pb};
  HVec<double, float> R{};
  R.coef = B.coef->d; R.fcoef = B.fcoef; R.vin = B.h->d; R.src = B.src->d; R.out = r32;
  double eps0 = -1, eps = 0;
  int outer = 0, total = 0, converged = 0;
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  for (;;) {
    k_happly<HM_RESID_DOT, double, float><<<nb, 256, 0, g->stream>>>(G, R, g->d_mo, g->d_part);
    k_hfinalize<1><<<1, 1024, 0, g->stream>>>(g->d_part, nb, g->d_mo, HOP_STORE);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(g->h_sc, g->d_mo, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    eps = g->h_sc->pz;
    if (eps0 < 0) eps0 = (prm->precond == 0 && !prm->resnorm) ? 1.0 : eps;
    const double ratio = eps == 0 ? 0.0 : sqrt(eps / eps0);
    if (eps == 0 || ratio < prm->epsilon) { converged = 1; break; }
    if (outer >= max_outer) break;
    inner.epsilon = std::max(inner_tol, 0.25 * prm->epsilon / ratio);
    if ((rc = upload_hscal(g, S.sc, &inner))) return rc;
    k_hfill<float><<<nb, 256, 0, g->stream>>>(x32, g->NS, 0.0f);
    if ((rc = h_start<float>(g, S))) return rc;
    if ((rc = h_loop<float>(g, &inner, S, PK_OTHER, PK_OTHER))) return rc;
    HIPCHK(hipMemcpyAsync(g->h_sc, S.sc, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    total += g->h_sc->iter;
    k_hmix_update<<<nb, 256, 0, g->stream>>>(G, B.h->d, x32);
    HIPCHK(hipGetLastError());
    outer++;
  }
  HIPCHK(hipEventRecord(g->ev2, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  CGScal c{};
  c.nloop = outer; c.converged = converged; c.resid = eps0 > 0 ? sqrt(eps / eps0) : 0; c.eps0 = eps0;
  if (inner_iters) *inner_iters = total;
  return fill_stats(g, c, st, outer, total);
}

int sunhip_h_operator(sunhip_grid *g, double grav, double theta, double dt, const char *x, const char *y) {
  HIPCHK(hipSetDevice(g->device));
  HBufs B;
  int rc = hbufs(g, B);
  if (rc) return rc;
  Field *fx, *fy;
  if ((rc = field(g, x, &fx) | field(g, y, &fy))) return rc;
  if ((rc = h_coefficients(g, B, grav, theta, dt))) return rc;
  HVec<double, double> V{};
  V.coef = B.coef->d; V.fcoef = B.fcoef; V.vin = fx->d; V.out = fy->d;
  k_happly<HM_PLAIN, double, double><<<nbh(g), 256, 0, g->stream>>>(g->geo(), V, g->d_hs, g->d_part);
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
  std::vector<double> fc((size_t)g->mf * g->NS, 0.0);
  for (int c = 0; c < g->Nc; c++) {
    const int s = g->slot_of_cell[c];
    for (int f = 0; f < g->nfaces[c]; f++) fc[(size_t)f * g->NS + s] = hfcoef_flat[(size_t)c * g->mf + f];
  }
  HIPCHK(hipMemcpyAsync(B.fcoef, fc.data(), fc.size() * sizeof(double), hipMemcpyHostToDevice, g->stream));
  HVec<double, double> V{};
  V.coef = B.coef->d; V.fcoef = B.fcoef; V.vin = fx->d; V.out = fy->d;
  k_happly<HM_PLAIN, double, double><<<nbh(g), 256, 0, g->stream>>>(g->geo(), V, g->d_hs, g->d_part);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

}  // extern "C"

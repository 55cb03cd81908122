// sunhip_batch.hip -- many independent grids solved together on one GPU.
//
// No reference counterpart: the reference solves one grid per process.  The
// batched form runs the PCG iterations of every member (an ensemble member
// or a serving request) in the same launches:
//
// * Member table + work list.  Every active solve contributes a row (its
//   grid geometry, its vectors, its CG scalars) to a table on the device and
//   one work item (row, tile) per tile of its grid.  A kernel walks the work
//   list grid-stride, one wavefront per item, and runs exactly the tile
//   arithmetic of the single-grid kernels (sunhip_qtile.h, sunhip_htile.h),
//   so grids of any size and depth share a launch and the small ones do not
//   leave the GPU idle.  With a bounded grid (a few blocks per CU) the
//   kernels are persistent: each wavefront loops over many items.
// * Per-member reductions.  Each item writes its wavefront's partial sums;
//   a finisher kernel with one block per member adds the member's items in a
//   fixed order and updates the member's scalars.  A member's result does not
//   depend on which other members share the batch, nor on the launch shape
//   or unroll depth: alone or in a mix it is bit for bit the same.
// * Members advance independently.  Each member selects its p buffers from
//   its own iteration count and skips its items once converged, so solves
//   can join (continuous batching) and leave between any two iterations.
// * hipGraphs.  A chunk of iterations over the current table is captured
//   once and replayed until the set of members changes.
// * Autotuning.  The unroll depth and the launch shape are chosen by timing
//   the first chunks of a batch with each candidate (all candidates compute
//   the same values), and the winner is kept per batch shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "sunhip_solvers.h"

using namespace sunhip;

namespace {

constexpr int BW = 4;  // wavefronts per block

struct BRange {
  CGScal *sc;
  int wlo, whi;  // the member's work items
};

template <typename T, typename TO>
struct BQMember {
  Geo G;
  const T *coef, *dzf, *D, *bq;
  T *x, *r, *z, *rtmp, *p0, *p1;
  const T *vin, *src;  // residual operands (RESID_DOT)
  TO *out;             // residual output (RESID_DOT)
  CGScal *sc;
};

struct BHMember {
  Geo G;
  const double *coef, *fcoef;
  double *x, *r, *z, *rt, *p0, *p1;
  CGScal *sc;
};

__device__ __forceinline__ int first_item() {
  return __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * BW + (int)(threadIdx.x >> 6));
}

template <int MODE, int NF, int U, typename T, typename TO>
__global__ __launch_bounds__(256) void kb_qapply(const BQMember<T, TO> *__restrict__ M, const int2 *__restrict__ W,
                                                 int nwork, double *__restrict__ part) {
  const int l = threadIdx.x & 63, stride = gridDim.x * BW;
  for (int w = first_item(); w < nwork; w += stride) {
    const int2 it = W[w];
    const int m = __builtin_amdgcn_readfirstlane(it.x), t = __builtin_amdgcn_readfirstlane(it.y);
    const BQMember<T, TO> &B = M[m];
    CGScal *sc = B.sc;
    QVec<T, TO> V{};
    V.coef = B.coef; V.dzf = B.dzf; V.D = B.D;
    T mu = 0, nu = 0;
    if constexpr (MODE == QM_FUSED) {
      if (sc->done) continue;
      const bool odd = sc->iter & 1;
      V.rtmp = B.rtmp;
      V.pold = odd ? B.p1 : B.p0;
      V.pnew = odd ? B.p0 : B.p1;
      V.x = B.x;
      V.out = B.z;
      mu = (T)sc->mu;
      nu = (T)sc->nu;
    } else {
      V.vin = B.vin; V.src = B.src; V.out = B.out;
    }
    const double d = wave_sum(qapply_tile<MODE, NF, U, T, TO, MODE == QM_FUSED>(B.G, V, mu, nu, t, l));
    if (l == 0) part[w] = d;
  }
}

// kb_qapply in the FUSED mode with the neighbour exchange inside each wave
// (qapply_xchg): the same values and the same per-item sums.
template <int NF, int U, typename T>
__global__ __launch_bounds__(256) void kb_qapply_x(const BQMember<T, T> *__restrict__ M,
                                                      const int2 *__restrict__ W, int nwork,
                                                      double *__restrict__ part) {
  __shared__ T rows[4][U][SUNHIP_TILE + SUNHIP_XPAIRS];
  const int l = threadIdx.x & 63, stride = gridDim.x * BW;
  for (int w = first_item(); w < nwork; w += stride) {
    const int2 it = W[w];
    const int m = __builtin_amdgcn_readfirstlane(it.x), t = __builtin_amdgcn_readfirstlane(it.y);
    const BQMember<T, T> &B = M[m];
    CGScal *sc = B.sc;
    if (sc->done) continue;
    QVec<T, T> V{};
    V.coef = B.coef; V.dzf = B.dzf; V.D = B.D;
    const bool odd = sc->iter & 1;
    V.rtmp = B.rtmp;
    V.pold = odd ? B.p1 : B.p0;
    V.pnew = odd ? B.p0 : B.p1;
    V.x = B.x;
    V.out = B.z;
    const double d = wave_sum(qapply_xchg<NF, U, T, true>(B.G, V, (T)sc->mu, (T)sc->nu, t, l, rows[threadIdx.x >> 6]));
    if (l == 0) part[w] = d;
  }
}

template <int U, typename T>
__global__ __launch_bounds__(256) void kb_qprec(const BQMember<T, T> *__restrict__ M, const int2 *__restrict__ W,
                                                int nwork, double *__restrict__ part) {
  const int l = threadIdx.x & 63, stride = gridDim.x * BW;
  // each lane's rows of the LDS stage (its own words only)
  constexpr int KL = 12;
  __shared__ T lrows[3 * KL * 256];
  for (int w = first_item(); w < nwork; w += stride) {
    const int2 it = W[w];
    const int m = __builtin_amdgcn_readfirstlane(it.x), t = __builtin_amdgcn_readfirstlane(it.y);
    const BQMember<T, T> &B = M[m];
    if (B.sc->done) continue;
    double rz = 0, rr = 0;
    // the top 24 (FP64) or 16 (FP32) levels of each column stay in registers
    // between the Thomas sweeps and the next 12 in LDS, as in the
    // single-grid kernel
    qprec_tile<true, U, T, true, sizeof(T) == 8 ? 24 : 16, false, KL>(B.G, B.coef, B.z, B.r, B.rtmp, B.bq,
                                                                      (T)B.sc->nu, t, l, rz, rr, nullptr,
                                                                      lrows + threadIdx.x);
    rz = wave_sum(rz);
    rr = wave_sum(rr);
    if (l == 0) { part[2 * (size_t)w] = rz; part[2 * (size_t)w + 1] = rr; }
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void kb_happly(const BHMember *__restrict__ M, const int2 *__restrict__ W, int nwork,
                                                 double *__restrict__ part) {
  const int l = threadIdx.x & 63, stride = gridDim.x * BW;
  for (int w = first_item(); w < nwork; w += stride) {
    const int2 it = W[w];
    const int m = __builtin_amdgcn_readfirstlane(it.x), t = __builtin_amdgcn_readfirstlane(it.y);
    const BHMember &B = M[m];
    CGScal *sc = B.sc;
    if (sc->done) continue;
    const bool odd = sc->iter & 1;
    HVec<double, double> V{};
    V.coef = B.coef; V.fcoef = B.fcoef; V.rtmp = B.rt;
    V.pold = odd ? B.p1 : B.p0;
    V.pnew = odd ? B.p0 : B.p1;
    V.x = B.x; V.out = B.z;
    const double d = wave_sum(happly_cell<MODE, double, double>(B.G, V, sc->mu, sc->nu, t * 64 + l));
    if (l == 0) part[w] = d;
  }
}

__global__ __launch_bounds__(256) void kb_hprec(const BHMember *__restrict__ M, const int2 *__restrict__ W, int nwork,
                                                double *__restrict__ part) {
  const int l = threadIdx.x & 63, stride = gridDim.x * BW;
  for (int w = first_item(); w < nwork; w += stride) {
    const int2 it = W[w];
    const int m = __builtin_amdgcn_readfirstlane(it.x), t = __builtin_amdgcn_readfirstlane(it.y);
    const BHMember &B = M[m];
    CGScal *sc = B.sc;
    if (sc->done) continue;
    double rz = 0, rr = 0;
    hprec_cell<double>(B.G, B.coef, B.z, B.r, B.rt, sc->nu, sc->precond, true, t * 64 + l, rz, rr);
    rz = wave_sum(rz);
    rr = wave_sum(rr);
    if (l == 0) { part[2 * (size_t)w] = rz; part[2 * (size_t)w + 1] = rr; }
  }
}

// One block per member: add the member's item partials in a fixed order and
// apply op (KIND 0: pressure scalars, 1: free-surface scalars).
template <int NV, int KIND>
__global__ __launch_bounds__(256) void kb_finalize(const BRange *__restrict__ R, const double *__restrict__ part,
                                                   int op, int skip_done) {
  const BRange r = R[blockIdx.x];
  if (skip_done && r.sc->done) return;
  __shared__ double sh[NV][4];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double s = 0;
    for (int b = r.wlo + threadIdx.x; b < r.whi; b += 256) s += part[(size_t)b * NV + i];
    s = wave_sum(s);
    if (l == 0) sh[i][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) tot[i] = ((sh[i][0] + sh[i][1]) + sh[i][2]) + sh[i][3];
    if (KIND == 0) q_op(r.sc, op, tot);
    else h_op(r.sc, op, tot);
  }
}

__global__ void kb_gather(CGScal *const *__restrict__ src, CGScal *__restrict__ dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = *src[i];
}

// Scalar steps of many grids: the three kernels of sunhip_update_scalars,
// each over a work list of (member, cell tile) or (member, edge tile).
struct BTMember {
  Geo G;
  TvdArgs T;
  ColArgs C;
};

template <int NF>
__global__ __launch_bounds__(256) void kb_tvd_sums(const BTMember *__restrict__ M, const int2 *__restrict__ W, int n) {
  const int l = threadIdx.x & 63, stride = gridDim.x * BW;
  for (int w = first_item(); w < n; w += stride) {
    const int2 it = W[w];
    const BTMember &B = M[__builtin_amdgcn_readfirstlane(it.x)];
    if (B.T.TVD) tvd_sums_tile<NF, 4>(B.G, B.T, __builtin_amdgcn_readfirstlane(it.y), l);
  }
}

__global__ __launch_bounds__(256) void kb_tvd_faces(const BTMember *__restrict__ M, const int2 *__restrict__ W, int n) {
  const int l = threadIdx.x & 63, stride = gridDim.x * BW;
  for (int w = first_item(); w < n; w += stride) {
    const int2 it = W[w];
    const BTMember &B = M[__builtin_amdgcn_readfirstlane(it.x)];
    if (B.T.TVD) tvd_faces_tile(B.G, B.T, __builtin_amdgcn_readfirstlane(it.y), l);
  }
}

template <int NF>
__global__ __launch_bounds__(256) void kb_update_column(const BTMember *__restrict__ M, const int2 *__restrict__ W,
                                                        int n) {
  const int l = threadIdx.x & 63, stride = gridDim.x * BW;
  for (int w = first_item(); w < n; w += stride) {
    const int2 it = W[w];
    const BTMember &B = M[__builtin_amdgcn_readfirstlane(it.x)];
    // members with the faces pass's flux terms read those (TVD >= 1)
    // (and keep their top 9 Thomas rows in registers)
    if (B.C.P1) update_column_tile<NF, 2, false, true, 8>(B.G, B.C, __builtin_amdgcn_readfirstlane(it.y), l);
    else update_column_tile<NF, 2>(B.G, B.C, __builtin_amdgcn_readfirstlane(it.y), l);
  }
}

// The stencils either side of the pressure solve over a work list of
// (member, tile), tiles covering max(Nt, Net) of each member.
struct BSMember {
  Geo G;
  StepArgs A;
};

template <int NF>
__global__ __launch_bounds__(256) void kb_qsource(const BSMember *__restrict__ M, const int2 *__restrict__ W, int n) {
  const int l = threadIdx.x & 63, stride = gridDim.x * BW;
  for (int w = first_item(); w < n; w += stride) {
    const int2 it = W[w];
    const BSMember &B = M[__builtin_amdgcn_readfirstlane(it.x)];
    qsource_tile<NF>(B.G, B.A, __builtin_amdgcn_readfirstlane(it.y), l);
  }
}

__global__ __launch_bounds__(256) void kb_corrector(const BSMember *__restrict__ M, const int2 *__restrict__ W,
                                                    int n) {
  const int l = threadIdx.x & 63, stride = gridDim.x * BW;
  for (int w = first_item(); w < n; w += stride) {
    const int2 it = W[w];
    const BSMember &B = M[__builtin_amdgcn_readfirstlane(it.x)];
    corrector_tile(B.G, B.A, __builtin_amdgcn_readfirstlane(it.y), l);
  }
}

template <int NF>
__global__ __launch_bounds__(256) void kb_continuity(const BSMember *__restrict__ M, const int2 *__restrict__ W,
                                                     int n) {
  const int l = threadIdx.x & 63, stride = gridDim.x * BW;
  for (int w = first_item(); w < n; w += stride) {
    const int2 it = W[w];
    const BSMember &B = M[__builtin_amdgcn_readfirstlane(it.x)];
    continuity_tile<NF>(B.G, B.A, __builtin_amdgcn_readfirstlane(it.y), l);
  }
}

// ---------------------------------------------------------------------------
// host side

// Kinds of solve; the first three are also the kinds of iteration table.  A
// mixed-precision pressure solve (T_QMIX) walks through its refinement
// steps on its own: residual to compute (M_RESID), residual computed and
// waiting to be read (M_READ), FP32 inner solve running in the FP32 table
// (M_INNER).
enum TaskKind { T_Q64 = 0, T_H64 = 1, T_Q32 = 2, T_NKIND = 3, T_QMIX = 3 };
enum MixState { M_RESID = 1, M_READ = 2, M_INNER = 3 };

struct Task {
  sunhip_grid *g;
  int kind;
  int64_t tag;
  QSolver<double> q64;
  QSolver<float> q32;
  HSolver<double> h64;
  int stage;  // pinned staging slot of its scalar upload
  std::chrono::steady_clock::time_point t0;
  // mixed precision (T_QMIX)
  int mstate = 0, outer = 0, inner_total = 0, max_outer = 0;
  double eps0 = -1, eps = 0, inner_tol = 0;
  sunhip_cg_params prm{};
  QVec<double, float> R{};
  // the iteration table a task is in (-1 none)
  int table() const { return kind == T_QMIX ? (mstate == M_INNER ? T_Q32 : -1) : kind; }
};

template <typename T>
struct DevBuf {
  T *d = nullptr;
  size_t cap = 0;
  int ensure(size_t n) {
    if (n <= cap) return 0;
    if (d) HIPCHK(hipFree(d));
    size_t c = std::max(n, cap * 2 + 16);
    HIPCHK(hipMalloc((void **)&d, c * sizeof(T)));
    cap = c;
    return 0;
  }
  void release() {
    if (d) hipFree(d);
    d = nullptr;
    cap = 0;
  }
};

template <typename T>
struct HostBuf {
  T *h = nullptr;
  size_t cap = 0;
  int ensure(size_t n) {
    if (n <= cap) return 0;
    if (h) HIPCHK(hipHostFree(h));
    size_t c = std::max(n, cap * 2 + 16);
    HIPCHK(hipHostMalloc((void **)&h, c * sizeof(T), hipHostMallocDefault));
    cap = c;
    return 0;
  }
  void release() {
    if (h) hipHostFree(h);
    h = nullptr;
    cap = 0;
  }
};

struct Config {
  int unroll = 4;   // levels per load group (1, 2, 4)
  int bpc = 0;      // blocks per CU of a persistent launch; 0 = one work item per wavefront
};

const Config kCandidates[] = {{4, 0}, {2, 0}, {1, 0}, {4, 4}, {2, 4}, {4, 8}};
constexpr int kNCand = sizeof(kCandidates) / sizeof(kCandidates[0]);

}  // namespace

struct sunhip_batch {
  int device = 0;
  int ncu = 256;
  hipStream_t stream = nullptr;
  std::vector<Task> tasks;  // active solves
  std::vector<sunhip_grid *> attached;  // grids whose work goes to this batch's stream
  std::vector<int> free_stage;
  HostBuf<CGScal> h_stage;  // one slot per task
  int nstage = 0;
  // tables of the current set
  bool dirty = true;
  int nf = 3;
  bool xok = true;  // every member's cross-tile faces fit the exchange kernel
  DevBuf<BQMember<double, double>> tq64;
  DevBuf<BQMember<float, float>> tq32;
  DevBuf<BHMember> th64;
  DevBuf<BRange> rng[T_NKIND];
  DevBuf<int2> work[T_NKIND];
  DevBuf<double> part[T_NKIND];
  int nmem[T_NKIND] = {0, 0, 0}, nwork[T_NKIND] = {0, 0, 0};
  HostBuf<char> h_tab;  // pinned staging for table uploads
  // residual table of the mixed form
  DevBuf<BQMember<double, float>> tres;
  DevBuf<BRange> rres;
  DevBuf<int2> wres;
  DevBuf<double> pres;
  int nres = 0, nwres = 0;  // members and items of the residual table
  // polling
  DevBuf<BTMember> ttvd;
  DevBuf<int2> wtc, wte;
  DevBuf<BSMember> tstep;  // stencils either side of the pressure solve
  DevBuf<int2> wstep;
  DevBuf<CGScal *> d_scp;
  DevBuf<CGScal> d_scv;
  HostBuf<CGScal> h_scv;
  HostBuf<CGScal *> h_scp;
  // graph of one chunk over the current tables
  hipGraphExec_t gexec = nullptr;
  int gchunk = 0;
  Config gcfg;
  int chunk = 0;  // iterations per graph launch (0 = automatic)
  // tuning
  bool fixed = false;
  Config cfg;
  std::map<long, Config> tuned;
  int cand = -1;           // candidate being timed, -1 none
  double cand_ms[kNCand];
  long cand_key = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timing = false;
  int timing_items = 0;
};

namespace {

int take_stage(sunhip_batch *b) {
  if (!b->free_stage.empty()) {
    const int s = b->free_stage.back();
    b->free_stage.pop_back();
    return s;
  }
  if (b->nstage + 1 > (int)b->h_stage.cap) {
    // the old slots may still feed copies in flight
    if (hipStreamSynchronize(b->stream) != hipSuccess) return -1;
    HostBuf<CGScal> nb;
    if (nb.ensure(std::max(64, 2 * b->nstage + 2))) return -1;
    if (b->h_stage.h) memcpy(nb.h, b->h_stage.h, sizeof(CGScal) * b->nstage);
    b->h_stage.release();
    b->h_stage = nb;
  }
  return b->nstage++;
}

long shape_key(const sunhip_batch *b) {
  long k = 0;
  for (int i = 0; i < T_NKIND; i++) {
    int lg = 0;
    while ((1 << lg) < b->nwork[i]) lg++;
    k = k * 64 + (b->nwork[i] ? lg + 1 : 0);
  }
  return k * 16 + nf_class(b->nf);
}

int table_upload(sunhip_batch *b, void *dst, const void *src, size_t bytes, size_t &off) {
  // tables go through one pinned staging area so the copies are asynchronous
  char *h = b->h_tab.h + off;
  memcpy(h, src, bytes);
  HIPCHK(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, b->stream));
  off += (bytes + 255) & ~(size_t)255;
  return 0;
}

// Rebuild the member tables and work lists of the active tasks.
int rebuild(sunhip_batch *b) {
  std::vector<BQMember<double, double>> q64;
  std::vector<BQMember<float, float>> q32;
  std::vector<BHMember> h64;
  std::vector<BQMember<double, float>> res;
  std::vector<BRange> rg[T_NKIND], rgr;
  std::vector<int2> wk[T_NKIND], wkr;
  int nf = 3;
  bool xok = true;
  for (const Task &t : b->tasks) {
    sunhip_grid *g = t.g;
    nf = std::max(nf, g->mf);
    xok = xok && g->max_pairs <= SUNHIP_XPAIRS;
    if (t.kind == T_QMIX && t.mstate == M_RESID) {
      // FP64 true residual of a mixed solve
      const int row = (int)rgr.size();
      BRange r{g->d_mo, (int)wkr.size(), 0};
      for (int i = 0; i < g->Nt; i++) wkr.push_back(make_int2(row, i));
      r.whi = (int)wkr.size();
      BQMember<double, float> m{};
      m.G = g->geo(); m.coef = t.R.coef; m.dzf = t.R.dzf; m.D = t.R.D; m.vin = t.R.vin; m.src = t.R.src;
      m.out = t.R.out; m.sc = g->d_mo;
      res.push_back(m);
      rgr.push_back(r);
      continue;
    }
    const int k = t.table();
    if (k < 0) continue;
    const int row = (int)rg[k].size();
    const int ntiles = k == T_H64 ? (g->NS + 63) / 64 : g->Nt;
    BRange r{nullptr, (int)wk[k].size(), 0};
    for (int i = 0; i < ntiles; i++) wk[k].push_back(make_int2(row, i));
    r.whi = (int)wk[k].size();
    if (k == T_Q64) {
      const QSolver<double> &S = t.q64;
      BQMember<double, double> m{};
      m.G = g->geo(); m.coef = S.coef; m.dzf = S.dzf; m.D = S.D; m.bq = S.bq; m.x = S.x; m.r = S.r; m.z = S.z;
      m.rtmp = S.rtmp; m.p0 = S.p0; m.p1 = S.p1; m.sc = S.sc;
      q64.push_back(m);
      r.sc = S.sc;
    } else if (k == T_Q32) {
      const QSolver<float> &S = t.q32;
      BQMember<float, float> m{};
      m.G = g->geo(); m.coef = S.coef; m.dzf = S.dzf; m.D = S.D; m.bq = S.bq; m.x = S.x; m.r = S.r; m.z = S.z;
      m.rtmp = S.rtmp; m.p0 = S.p0; m.p1 = S.p1; m.sc = S.sc;
      q32.push_back(m);
      r.sc = S.sc;
    } else {
      const HSolver<double> &S = t.h64;
      BHMember m{};
      m.G = g->geo(); m.coef = S.coef; m.fcoef = S.fcoef; m.x = S.x; m.r = S.r; m.z = S.z; m.rt = S.rt;
      m.p0 = S.p0; m.p1 = S.p1; m.sc = S.sc;
      h64.push_back(m);
      r.sc = S.sc;
    }
    rg[k].push_back(r);
  }
  int rc = 0;
  size_t bytes = q64.size() * sizeof(q64[0]) + q32.size() * sizeof(q32[0]) + h64.size() * sizeof(h64[0]) +
                 res.size() * sizeof(res[0]) + rgr.size() * sizeof(BRange) + wkr.size() * sizeof(int2) + 16384;
  for (int k = 0; k < T_NKIND; k++) bytes += rg[k].size() * sizeof(BRange) + wk[k].size() * sizeof(int2) + 512;
  // the staging area may still feed the previous upload
  HIPCHK(hipStreamSynchronize(b->stream));
  if ((rc = b->h_tab.ensure(bytes))) return rc;
  size_t off = 0;
  if ((rc = b->tq64.ensure(q64.size() + 1)) || (rc = b->tq32.ensure(q32.size() + 1)) ||
      (rc = b->th64.ensure(h64.size() + 1)) || (rc = b->tres.ensure(res.size() + 1)) ||
      (rc = b->rres.ensure(rgr.size() + 1)) || (rc = b->wres.ensure(wkr.size() + 1)) ||
      (rc = b->pres.ensure(wkr.size() + 1)))
    return rc;
  if (!q64.empty()) rc |= table_upload(b, b->tq64.d, q64.data(), q64.size() * sizeof(q64[0]), off);
  if (!q32.empty()) rc |= table_upload(b, b->tq32.d, q32.data(), q32.size() * sizeof(q32[0]), off);
  if (!h64.empty()) rc |= table_upload(b, b->th64.d, h64.data(), h64.size() * sizeof(h64[0]), off);
  if (!res.empty()) {
    rc |= table_upload(b, b->tres.d, res.data(), res.size() * sizeof(res[0]), off);
    rc |= table_upload(b, b->rres.d, rgr.data(), rgr.size() * sizeof(BRange), off);
    rc |= table_upload(b, b->wres.d, wkr.data(), wkr.size() * sizeof(int2), off);
  }
  b->nres = (int)rgr.size();
  b->nwres = (int)wkr.size();
  for (int k = 0; k < T_NKIND; k++) {
    b->nmem[k] = (int)rg[k].size();
    b->nwork[k] = (int)wk[k].size();
    if ((rc = b->rng[k].ensure(rg[k].size() + 1)) || (rc = b->work[k].ensure(wk[k].size() + 1)) ||
        (rc = b->part[k].ensure(2 * wk[k].size() + 2)))
      return rc;
    if (!rg[k].empty()) {
      rc |= table_upload(b, b->rng[k].d, rg[k].data(), rg[k].size() * sizeof(BRange), off);
      rc |= table_upload(b, b->work[k].d, wk[k].data(), wk[k].size() * sizeof(int2), off);
    }
  }
  if (rc) return rc;
  b->nf = nf;
  b->xok = xok;
  b->dirty = false;
  if (b->gexec) {
    hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  return rc;
}

// FP64 true residuals of the mixed solves waiting for one (one launch for
// all of them, then their sums into each member's outer scalars).
int launch_residuals(sunhip_batch *b) {
  if (!b->nres) return 0;
  const int nw = b->nwres, nb = (nw + BW - 1) / BW;
  with_nf(b->nf, [&](auto nf) {
    constexpr int NF = decltype(nf)::value;
    kb_qapply<QM_RESID_DOT, NF, NF <= 4 ? 4 : 2, double, float><<<nb, 256, 0, b->stream>>>(b->tres.d, b->wres.d, nw,
                                                                                          b->pres.d);
    return 0;
  });
  kb_finalize<1, 0><<<b->nres, 256, 0, b->stream>>>(b->rres.d, b->pres.d, OP_STORE, 0);
  HIPCHK(hipGetLastError());
  for (Task &t : b->tasks)
    if (t.kind == T_QMIX && t.mstate == M_RESID) t.mstate = M_READ;
  b->dirty = true;  // the residual table is spent
  return 0;
}

int grid_for(const sunhip_batch *b, int nwork, const Config &c) {
  const int full = (nwork + BW - 1) / BW;
  if (c.bpc <= 0) return std::max(full, 1);
  return std::max(1, std::min(full, c.bpc * b->ncu));
}

// One iteration of every active solve, enqueued on b->stream.
int enqueue_iteration(sunhip_batch *b, const Config &c) {
  // pressure, FP64
  if (b->nwork[T_Q64]) {
    const int n = b->nwork[T_Q64], nb = grid_for(b, n, c);
    auto *M = b->tq64.d;
    double *P = b->part[T_Q64].d;
#define QA(NF, U) kb_qapply<QM_FUSED, NF, U, double, double><<<nb, 256, 0, b->stream>>>(M, b->work[T_Q64].d, n, P)
    if (b->xok && c.unroll == 4)  // neighbour exchange inside each wave
      with_nf(b->nf, [&](auto nfc) {
        constexpr int NF = decltype(nfc)::value;
        kb_qapply_x<NF, NF <= 4 ? 4 : 2, double><<<nb, 256, 0, b->stream>>>(M, b->work[T_Q64].d, n, P);
        return 0;
      });
    else if (b->nf <= 3) { if (c.unroll == 1) QA(3, 1); else if (c.unroll == 2) QA(3, 2); else QA(3, 4); }
    else if (b->nf == 4) { if (c.unroll == 1) QA(4, 1); else if (c.unroll == 2) QA(4, 2); else QA(4, 4); }
    else if (b->nf <= 6) { if (c.unroll == 1) QA(6, 1); else QA(6, 2); }
    else { if (c.unroll == 1) QA(8, 1); else QA(8, 2); }
#undef QA
    kb_finalize<1, 0><<<b->nmem[T_Q64], 256, 0, b->stream>>>(b->rng[T_Q64].d, P, OP_APPLY, 1);
    if (c.unroll == 1) kb_qprec<1, double><<<nb, 256, 0, b->stream>>>(M, b->work[T_Q64].d, n, P);
    else if (c.unroll == 2) kb_qprec<2, double><<<nb, 256, 0, b->stream>>>(M, b->work[T_Q64].d, n, P);
    else kb_qprec<4, double><<<nb, 256, 0, b->stream>>>(M, b->work[T_Q64].d, n, P);
    kb_finalize<2, 0><<<b->nmem[T_Q64], 256, 0, b->stream>>>(b->rng[T_Q64].d, P, OP_PREC_ITER, 1);
  }
  // pressure, FP32 inner solves of the mixed form
  if (b->nwork[T_Q32]) {
    const int n = b->nwork[T_Q32], nb = grid_for(b, n, c);
    auto *M = b->tq32.d;
    double *P = b->part[T_Q32].d;
#define QA(NF, U) kb_qapply<QM_FUSED, NF, U, float, float><<<nb, 256, 0, b->stream>>>(M, b->work[T_Q32].d, n, P)
    if (b->xok && c.unroll == 4)  // neighbour exchange inside each wave
      with_nf(b->nf, [&](auto nfc) {
        constexpr int NF = decltype(nfc)::value;
        kb_qapply_x<NF, NF <= 4 ? 4 : 2, float><<<nb, 256, 0, b->stream>>>(M, b->work[T_Q32].d, n, P);
        return 0;
      });
    else if (b->nf <= 3) { if (c.unroll == 1) QA(3, 1); else if (c.unroll == 2) QA(3, 2); else QA(3, 4); }
    else if (b->nf == 4) { if (c.unroll == 1) QA(4, 1); else if (c.unroll == 2) QA(4, 2); else QA(4, 4); }
    else if (b->nf <= 6) { if (c.unroll == 1) QA(6, 1); else QA(6, 2); }
    else { if (c.unroll == 1) QA(8, 1); else QA(8, 2); }
#undef QA
    kb_finalize<1, 0><<<b->nmem[T_Q32], 256, 0, b->stream>>>(b->rng[T_Q32].d, P, OP_APPLY, 1);
    if (c.unroll == 1) kb_qprec<1, float><<<nb, 256, 0, b->stream>>>(M, b->work[T_Q32].d, n, P);
    else if (c.unroll == 2) kb_qprec<2, float><<<nb, 256, 0, b->stream>>>(M, b->work[T_Q32].d, n, P);
    else kb_qprec<4, float><<<nb, 256, 0, b->stream>>>(M, b->work[T_Q32].d, n, P);
    kb_finalize<2, 0><<<b->nmem[T_Q32], 256, 0, b->stream>>>(b->rng[T_Q32].d, P, OP_PREC_ITER, 1);
  }
  // free surface
  if (b->nwork[T_H64]) {
    const int n = b->nwork[T_H64], nb = grid_for(b, n, c);
    double *P = b->part[T_H64].d;
    kb_happly<HM_FUSED><<<nb, 256, 0, b->stream>>>(b->th64.d, b->work[T_H64].d, n, P);
    kb_finalize<1, 1><<<b->nmem[T_H64], 256, 0, b->stream>>>(b->rng[T_H64].d, P, HOP_APPLY, 1);
    kb_hprec<<<nb, 256, 0, b->stream>>>(b->th64.d, b->work[T_H64].d, n, P);
    kb_finalize<2, 1><<<b->nmem[T_H64], 256, 0, b->stream>>>(b->rng[T_H64].d, P, HOP_ITER, 1);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int chunk_len(const sunhip_batch *b) {
  if (b->chunk > 0) return b->chunk;
  // large batches: poll often (an iteration is long); small: amortise the poll
  const long items = (long)b->nwork[T_Q64] + b->nwork[T_Q32] + b->nwork[T_H64] / 8;
  return items > 60000 ? 2 : items > 8000 ? 4 : 8;
}

int capture(sunhip_batch *b, const Config &c, int chunk) {
  if (b->gexec) {
    HIPCHK(hipGraphExecDestroy(b->gexec));
    b->gexec = nullptr;
  }
  hipGraph_t graph;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int i = 0; i < chunk && !rc; i++) rc = enqueue_iteration(b, c);
  hipError_t e = hipStreamEndCapture(b->stream, &graph);
  if (rc) return rc;
  HIPCHK(e);
  HIPCHK(hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0));
  HIPCHK(hipGraphDestroy(graph));
  b->gchunk = chunk;
  b->gcfg = c;
  return 0;
}

// Settle the timed candidate's measurement (after a synchronisation).
void finish_timing(sunhip_batch *b) {
  if (!b->timing) return;
  b->timing = false;
  float ms = 0;
  if (hipEventElapsedTime(&ms, b->ev0, b->ev1) != hipSuccess) return;
  b->cand_ms[b->cand] = ms / std::max(1, b->timing_items);
  b->cand++;
  if (b->cand >= kNCand) {
    int best = 0;
    for (int i = 1; i < kNCand; i++)
      if (b->cand_ms[i] < b->cand_ms[best]) best = i;
    b->tuned[b->cand_key] = kCandidates[best];
    b->cfg = kCandidates[best];
    b->cand = -1;
  }
}

}  // namespace

extern "C" {

int sunhip_batch_create(sunhip_batch **out, int device) {
  *out = nullptr;
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    set_error("sunhip_batch_create: no HIP device %d (%d visible)", device, n);
    return -1;
  }
  HIPCHK(hipSetDevice(device));
  sunhip_batch *b = new sunhip_batch;
  b->device = device;
  hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, device);
  HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&b->ev0));
  HIPCHK(hipEventCreate(&b->ev1));
  if (b->h_tab.ensure(1 << 16)) return -3;
  *out = b;
  return 0;
}

int sunhip_batch_destroy(sunhip_batch *b) {
  if (!b) return 0;
  hipSetDevice(b->device);
  hipStreamSynchronize(b->stream);
  // grids outliving the batch get their own streams back
  for (sunhip_grid *g : b->attached) {
    g->stream = g->own_stream;
    g->batch = nullptr;
  }
  b->attached.clear();
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  b->tq64.release(); b->tq32.release(); b->th64.release();
  for (int k = 0; k < T_NKIND; k++) { b->rng[k].release(); b->work[k].release(); b->part[k].release(); }
  b->tres.release(); b->rres.release(); b->wres.release(); b->pres.release();
  b->ttvd.release(); b->wtc.release(); b->wte.release(); b->tstep.release(); b->wstep.release();
  b->d_scp.release(); b->d_scv.release(); b->h_scv.release(); b->h_stage.release(); b->h_tab.release();
  hipEventDestroy(b->ev0);
  hipEventDestroy(b->ev1);
  hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int sunhip_batch_attach(sunhip_batch *b, sunhip_grid *g) {
  if (g->batch == b) return 0;
  if (g->batch) { set_error("grid is attached to another batch"); return -1; }
  if (g->device != b->device) { set_error("grid and batch are on different devices"); return -1; }
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(g->stream));
  g->stream = b->stream;
  g->batch = b;
  b->attached.push_back(g);
  return 0;
}

}  // extern "C"

namespace sunhip {
// A grid being destroyed: drop its solves from the batch and detach it.
void batch_forget_grid(sunhip_batch *b, sunhip_grid *g) {
  hipStreamSynchronize(b->stream);
  std::vector<Task> keep;
  for (const Task &t : b->tasks) {
    if (t.g == g) {
      if (t.stage >= 0) b->free_stage.push_back(t.stage);
      b->dirty = true;
    } else {
      keep.push_back(t);
    }
  }
  b->tasks.swap(keep);
  g->stream = g->own_stream;
  g->batch = nullptr;
  b->attached.erase(std::remove(b->attached.begin(), b->attached.end(), g), b->attached.end());
}
}  // namespace sunhip

extern "C" {

int sunhip_batch_detach(sunhip_batch *b, sunhip_grid *g) {
  if (g->batch != b) return 0;
  for (const Task &t : b->tasks)
    if (t.g == g) { set_error("grid still has a solve in the batch"); return -1; }
  HIPCHK(hipStreamSynchronize(b->stream));
  g->stream = g->own_stream;
  g->batch = nullptr;
  b->attached.erase(std::remove(b->attached.begin(), b->attached.end(), g), b->attached.end());
  return 0;
}

int sunhip_batch_begin_q(sunhip_batch *b, sunhip_grid *g, const sunhip_cg_params *prm, int64_t tag) {
  int rc = sunhip_batch_attach(b, g);
  if (rc) return rc;
  if (prm->precond != 2) { set_error("batched pressure solves use qprecond=2"); return -1; }
  Task t{};
  t.g = g; t.kind = T_Q64; t.tag = tag;
  if ((t.stage = take_stage(b)) < 0) return -3;
  if ((rc = q_begin64(g, prm, &t.q64, b->h_stage.h + t.stage))) return rc;
  t.t0 = std::chrono::steady_clock::now();
  b->tasks.push_back(t);
  b->dirty = true;
  return 0;
}

int sunhip_batch_begin_h(sunhip_batch *b, sunhip_grid *g, const sunhip_cg_params *prm, const sunhip_h_phys *ph,
                         int64_t tag) {
  int rc = sunhip_batch_attach(b, g);
  if (rc) return rc;
  Task t{};
  t.g = g; t.kind = T_H64; t.tag = tag;
  if ((t.stage = take_stage(b)) < 0) return -3;
  if ((rc = h_begin64(g, prm, ph->grav, ph->theta, ph->dt, &t.h64, b->h_stage.h + t.stage))) return rc;
  t.t0 = std::chrono::steady_clock::now();
  b->tasks.push_back(t);
  b->dirty = true;
  return 0;
}

int sunhip_batch_active(const sunhip_batch *b) { return (int)b->tasks.size(); }

int sunhip_batch_advance(sunhip_batch *b, int iters) {
  HIPCHK(hipSetDevice(b->device));
  if (b->tasks.empty()) return 0;
  int rc;
  const bool fresh = b->dirty;
  if (b->dirty && (rc = rebuild(b))) return rc;
  if ((rc = launch_residuals(b))) return rc;
  const int chunk = chunk_len(b);
  if (!b->nwork[T_Q64] && !b->nwork[T_Q32] && !b->nwork[T_H64]) return 0;  // only residuals this time
  // choose the configuration: fixed, tuned for this shape, or the next
  // candidate to time
  Config c = b->cfg;
  bool time_it = false;
  if (!b->fixed) {
    // tuned shapes keep their winner; a batch is tuned once, on the first
    // shape it runs, and every other shape uses that winner (the candidates
    // are timed per work item, so members leaving during tuning do not bias
    // the comparison much)
    auto it = b->tuned.find(shape_key(b));
    if (it != b->tuned.end()) {
      c = it->second;
    } else if (b->tuned.empty()) {
      if (b->cand < 0) { b->cand = 0; b->cand_key = shape_key(b); }
      c = kCandidates[b->cand];
      time_it = !b->timing;
    }
  }
  // A set of members that has just changed is launched directly: it may
  // change again at the next poll, and capturing a graph for one replay
  // costs more than it saves.  A set that survived a chunk is captured once
  // and replayed.
  const bool graph = !fresh && !time_it;
  if (graph && (!b->gexec || b->gchunk != chunk || b->gcfg.unroll != c.unroll || b->gcfg.bpc != c.bpc))
    if ((rc = capture(b, c, chunk))) return rc;
  for (int done = 0; done < iters; done += chunk) {
    const bool t = time_it && done == 0;
    if (t) HIPCHK(hipEventRecord(b->ev0, b->stream));
    if (graph) {
      HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    } else {
      for (int i = 0; i < chunk; i++)
        if ((rc = enqueue_iteration(b, c))) return rc;
    }
    if (t) {
      HIPCHK(hipEventRecord(b->ev1, b->stream));
      b->timing = true;
      b->timing_items = (b->nwork[T_Q64] + b->nwork[T_Q32] + b->nwork[T_H64]) * chunk;
    }
  }
  return 0;
}

int sunhip_batch_poll(sunhip_batch *b, int64_t *tags, int *kinds, sunhip_cg_stats *st, int max) {
  HIPCHK(hipSetDevice(b->device));
  const int n = (int)b->tasks.size();
  if (n == 0) return 0;
  int rc;
  // the scalars each solve waits on (a mixed solve: its outer residual or
  // its inner solve's scalars)
  if ((rc = b->h_scp.ensure(n)) || (rc = b->d_scp.ensure(n)) || (rc = b->d_scv.ensure(n)) ||
      (rc = b->h_scv.ensure(n)))
    return rc;
  for (int i = 0; i < n; i++) {
    const Task &t = b->tasks[i];
    CGScal *p = t.kind == T_Q64 ? t.q64.sc : t.kind == T_H64 ? t.h64.sc : t.kind == T_Q32 ? t.q32.sc
              : (t.mstate == M_INNER ? t.q32.sc : t.g->d_mo);
    b->h_scp.h[i] = p;
  }
  HIPCHK(hipMemcpyAsync(b->d_scp.d, b->h_scp.h, n * sizeof(CGScal *), hipMemcpyHostToDevice, b->stream));
  kb_gather<<<(n + 255) / 256, 256, 0, b->stream>>>(b->d_scp.d, b->d_scv.d, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b->h_scv.h, b->d_scv.d, n * sizeof(CGScal), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  finish_timing(b);
  int out = 0;
  std::vector<Task> keep;
  const auto now = std::chrono::steady_clock::now();
  for (int i = 0; i < n; i++) {
    const CGScal &c = b->h_scv.h[i];
    Task &t = b->tasks[i];
    if (t.kind == T_QMIX) {
      if (t.mstate == M_READ) {
        // one refinement step decided (the rule of sunhip_q_solve_mixed)
        t.eps = c.pz;
        if (t.eps0 < 0) t.eps0 = t.eps;
        const double ratio = t.eps == 0 ? 0.0 : sqrt(t.eps / t.eps0);
        const bool conv = t.eps == 0 || ratio < t.prm.epsilon;
        if (conv || t.outer >= t.max_outer) {
          if (out >= max) { keep.push_back(t); continue; }
          if (tags) tags[out] = t.tag;
          if (kinds) kinds[out] = T_QMIX;
          if (st) {
            st[out].niters = t.inner_total;
            st[out].converged = conv;
            st[out].resid = ratio;
            st[out].eps0 = t.eps0;
            st[out].device_ms = std::chrono::duration<double, std::milli>(now - t.t0).count();
            st[out].setup_ms = 0;
            st[out].outer = t.outer;
            st[out].inner = t.inner_total;
          }
          if (t.stage >= 0) b->free_stage.push_back(t.stage);
          out++;
          b->dirty = true;
          continue;
        }
        sunhip_cg_params inner = t.prm;
        inner.resnorm = 0;
        inner.epsilon = std::max(t.inner_tol, 0.25 * t.prm.epsilon / ratio);
        if ((rc = q_inner_begin(t.g, &inner, t.q32, b->h_stage.h + t.stage))) return rc;
        t.mstate = M_INNER;
        b->dirty = true;
      } else if (t.mstate == M_INNER && c.done) {
        t.inner_total += c.iter;
        if ((rc = q_end32(t.g, t.q32))) return rc;
        if ((rc = q_mix_update(t.g, const_cast<double *>(t.R.vin), t.q32.x))) return rc;
        t.outer++;
        t.mstate = M_RESID;
        b->dirty = true;
      }
      keep.push_back(t);
      continue;
    }
    if (!c.done || out >= max) {
      keep.push_back(t);
      continue;
    }
    if (t.kind == T_Q64) rc = q_end64(t.g, t.q64);
    else if (t.kind == T_Q32) rc = q_end32(t.g, t.q32);
    else rc = h_end64(t.g, t.h64);
    if (rc) return rc;
    if (tags) tags[out] = t.tag;
    if (kinds) kinds[out] = t.kind;
    if (st) {
      st[out].niters = c.nloop;
      st[out].converged = c.converged;
      st[out].resid = c.resid;
      st[out].eps0 = c.eps0;
      st[out].device_ms = std::chrono::duration<double, std::milli>(now - t.t0).count();
      st[out].setup_ms = 0;
      st[out].outer = 0;
      st[out].inner = 0;
    }
    if (t.stage >= 0) b->free_stage.push_back(t.stage);
    out++;
    b->dirty = true;
  }
  b->tasks.swap(keep);
  return out;
}

int sunhip_batch_begin_q_mixed(sunhip_batch *b, sunhip_grid *g, const sunhip_cg_params *prm, double inner_tol,
                               int max_outer, int64_t tag) {
  int rc = sunhip_batch_attach(b, g);
  if (rc) return rc;
  Task t{};
  t.g = g; t.kind = T_QMIX; t.tag = tag;
  if ((t.stage = take_stage(b)) < 0) return -3;
  if ((rc = q_mixed_prepare(g, prm, &t.q32, &t.R))) return rc;
  t.mstate = M_RESID;
  t.prm = *prm;
  t.inner_tol = inner_tol;
  t.max_outer = max_outer;
  t.t0 = std::chrono::steady_clock::now();
  b->tasks.push_back(t);
  b->dirty = true;
  return 0;
}

int sunhip_batch_sync(sunhip_batch *b) {
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int sunhip_batch_set_tuning(sunhip_batch *b, int unroll, int blocks_per_cu, int chunk) {
  if (unroll == 0) {
    b->fixed = false;
  } else {
    if (unroll != 1 && unroll != 2 && unroll != 4) { set_error("unroll must be 1, 2 or 4"); return -1; }
    b->fixed = true;
    b->cfg.unroll = unroll;
    b->cfg.bpc = std::max(0, blocks_per_cu);
  }
  b->chunk = std::max(0, chunk);
  return 0;
}

int sunhip_batch_tuning(const sunhip_batch *b, int *unroll, int *blocks_per_cu, int *chunk) {
  Config c = b->cfg;
  if (!b->fixed) {
    auto it = b->tuned.find(shape_key(b));
    if (it != b->tuned.end()) c = it->second;
  }
  if (unroll) *unroll = c.unroll;
  if (blocks_per_cu) *blocks_per_cu = c.bpc;
  if (chunk) *chunk = chunk_len(b);
  return (int)b->tuned.size();
}

// Run every active solve to completion; stats in completion order.
static int drain(sunhip_batch *b, std::vector<int64_t> &tags, std::vector<sunhip_cg_stats> &sts) {
  std::vector<int64_t> tg(b->tasks.size() + 1);
  std::vector<sunhip_cg_stats> s(b->tasks.size() + 1);
  int rc;
  while (!b->tasks.empty()) {
    if ((rc = sunhip_batch_advance(b, chunk_len(b)))) return rc;
    const int k = sunhip_batch_poll(b, tg.data(), nullptr, s.data(), (int)tg.size());
    if (k < 0) return k;
    for (int i = 0; i < k; i++) { tags.push_back(tg[i]); sts.push_back(s[i]); }
  }
  return 0;
}

int sunhip_batch_q_solve(sunhip_batch *b, sunhip_grid **gs, int n, const sunhip_cg_params *prm,
                         sunhip_cg_stats *st) {
  if (!b->tasks.empty()) { set_error("sunhip_batch_q_solve needs an idle batch"); return -1; }
  int rc;
  for (int i = 0; i < n; i++)
    if ((rc = sunhip_batch_begin_q(b, gs[i], prm, i))) return rc;
  std::vector<int64_t> tags;
  std::vector<sunhip_cg_stats> sts;
  if ((rc = drain(b, tags, sts))) return rc;
  for (size_t i = 0; i < tags.size(); i++) if (st) st[tags[i]] = sts[i];
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int sunhip_batch_h_solve(sunhip_batch *b, sunhip_grid **gs, int n, const sunhip_cg_params *prm,
                         const sunhip_h_phys *ph, sunhip_cg_stats *st) {
  if (!b->tasks.empty()) { set_error("sunhip_batch_h_solve needs an idle batch"); return -1; }
  int rc;
  for (int i = 0; i < n; i++)
    if ((rc = sunhip_batch_begin_h(b, gs[i], prm, ph + i, i))) return rc;
  std::vector<int64_t> tags;
  std::vector<sunhip_cg_stats> sts;
  if ((rc = drain(b, tags, sts))) return rc;
  for (size_t i = 0; i < tags.size(); i++) if (st) st[tags[i]] = sts[i];
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

// Mixed precision over many grids: each member refines on its own (its FP64
// residuals of a round computed in one launch with the others', its FP32
// inner solves iterating in the same launches as the others'), so no member
// waits for another's inner solve.
int sunhip_batch_q_solve_mixed(sunhip_batch *b, sunhip_grid **gs, int n, const sunhip_cg_params *prm,
                               double inner_tol, int max_outer, sunhip_cg_stats *st, int *inner_iters) {
  if (!b->tasks.empty()) { set_error("sunhip_batch_q_solve_mixed needs an idle batch"); return -1; }
  HIPCHK(hipSetDevice(b->device));
  int rc;
  for (int i = 0; i < n; i++)
    if ((rc = sunhip_batch_begin_q_mixed(b, gs[i], prm, inner_tol, max_outer, i))) return rc;
  std::vector<int64_t> tags;
  std::vector<sunhip_cg_stats> sts;
  if ((rc = drain(b, tags, sts))) return rc;
  for (size_t k = 0; k < tags.size(); k++) {
    if (st) { st[tags[k]] = sts[k]; st[tags[k]].niters = sts[k].outer; }
    if (inner_iters) inner_iters[tags[k]] = sts[k].inner;
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int sunhip_batch_update_scalars(sunhip_batch *b, sunhip_grid **gs, int n, const sunhip_scalar_params *sp,
                                const char *scal) {
  HIPCHK(hipSetDevice(b->device));
  int rc;
  std::vector<BTMember> tab(n);
  std::vector<int2> wc, we;
  for (int i = 0; i < n; i++) {
    sunhip_grid *g = gs[i];
    if ((rc = sunhip_batch_attach(b, g))) return rc;
    tab[i].G = g->geo();
    if ((rc = tvd_args(g, sp + i, scal, &tab[i].T, &tab[i].C))) return rc;
    if (sp[i].TVD) {
      // as sunhip_update_scalars: the faces pass forms the column's flux
      // terms, and the column update runs in place
      Field *p1, *p2;
      if ((rc = field(g, "tvdP1", &p1) | field(g, "tvdP2", &p2))) return rc;
      tab[i].T.P1 = p1->d; tab[i].T.P2 = p2->d;
      tab[i].C.P1 = p1->d; tab[i].C.P2 = p2->d;
      tab[i].C.s_old = tab[i].C.s;
    }
    for (int t = 0; t < g->Nt; t++) wc.push_back(make_int2(i, t));
    for (int t = 0; t < g->Net; t++) we.push_back(make_int2(i, t));
  }
  if (n == 0) return 0;
  // table staging is reused: wait for its previous uploads
  HIPCHK(hipStreamSynchronize(b->stream));
  const size_t bytes = tab.size() * sizeof(BTMember) + (wc.size() + we.size()) * sizeof(int2) + 1024;
  if ((rc = b->h_tab.ensure(bytes)) || (rc = b->ttvd.ensure(tab.size())) || (rc = b->wtc.ensure(wc.size() + 1)) ||
      (rc = b->wte.ensure(we.size() + 1)))
    return rc;
  size_t off = 0;
  rc = table_upload(b, b->ttvd.d, tab.data(), tab.size() * sizeof(BTMember), off) |
       table_upload(b, b->wtc.d, wc.data(), wc.size() * sizeof(int2), off) |
       table_upload(b, b->wte.d, we.data(), we.size() * sizeof(int2), off);
  if (rc) return rc;
  const int nc = (int)wc.size(), ne = (int)we.size();
  // face values from the current scalar, then the column update from its
  // copy (the order of sunhip_update_scalars)
  int mf = 3;
  for (int i = 0; i < n; i++) mf = std::max(mf, gs[i]->mf);
  with_nf(mf, [&](auto nf) {
    kb_tvd_sums<decltype(nf)::value><<<(nc + BW - 1) / BW, 256, 0, b->stream>>>(b->ttvd.d, b->wtc.d, nc);
    return 0;
  });
  if (ne) kb_tvd_faces<<<(ne + BW - 1) / BW, 256, 0, b->stream>>>(b->ttvd.d, b->wte.d, ne);
  HIPCHK(hipGetLastError());
  for (int i = 0; i < n; i++)
    if (!sp[i].TVD)  // first-order upwind reads the neighbours' old values: keep them in stmp
      HIPCHK(hipMemcpyAsync(const_cast<double *>(tab[i].C.s_old), tab[i].C.s,
                            gs[i]->fields.at(scal).n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
  with_nf(mf, [&](auto nf) {
    kb_update_column<decltype(nf)::value><<<(nc + BW - 1) / BW, 256, 0, b->stream>>>(b->ttvd.d, b->wtc.d, nc);
    return 0;
  });
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

// Member table and work list of a stencil launch; returns the number of work
// items (0: nothing to launch) or a negative error.
static int step_tables(sunhip_batch *b, sunhip_grid **gs, int n, const double *theta, const double *dt,
                       const char *qc, const char *q, const char *w, const char *src, int *mf) {
  int rc;
  std::vector<BSMember> tab(n);
  std::vector<int2> wk;
  *mf = 3;
  for (int i = 0; i < n; i++) {
    sunhip_grid *g = gs[i];
    if ((rc = sunhip_batch_attach(b, g))) return rc;
    tab[i].G = g->geo();
    if ((rc = step_args(g, theta[i], dt[i], qc, q, w, src, &tab[i].A))) return rc;
    for (int t = 0; t < std::max(g->Nt, g->Net); t++) wk.push_back(make_int2(i, t));
    *mf = std::max(*mf, g->mf);
  }
  if (n == 0) return 0;
  HIPCHK(hipStreamSynchronize(b->stream));  // the staging area is reused
  const size_t bytes = tab.size() * sizeof(BSMember) + wk.size() * sizeof(int2) + 1024;
  if ((rc = b->h_tab.ensure(bytes)) || (rc = b->tstep.ensure(tab.size())) || (rc = b->wstep.ensure(wk.size() + 1)))
    return rc;
  size_t off = 0;
  rc = table_upload(b, b->tstep.d, tab.data(), tab.size() * sizeof(BSMember), off) |
       table_upload(b, b->wstep.d, wk.data(), wk.size() * sizeof(int2), off);
  if (rc) return rc;
  return (int)wk.size();
}

int sunhip_batch_compute_qsource(sunhip_batch *b, sunhip_grid **gs, int n, const double *theta, const double *dt,
                                 const char *src) {
  HIPCHK(hipSetDevice(b->device));
  int mf;
  const int nw = step_tables(b, gs, n, theta, dt, "q", "qp", "w", src, &mf);
  if (nw <= 0) return nw;
  with_nf(mf, [&](auto nf) {
    kb_qsource<decltype(nf)::value><<<(nw + BW - 1) / BW, 256, 0, b->stream>>>(b->tstep.d, b->wstep.d, nw);
    return 0;
  });
  HIPCHK(hipGetLastError());
  return 0;
}

int sunhip_batch_correct(sunhip_batch *b, sunhip_grid **gs, int n, const double *theta, const double *dt,
                         const char *qc, const char *q, const char *w) {
  HIPCHK(hipSetDevice(b->device));
  int mf;
  const int nw = step_tables(b, gs, n, theta, dt, qc, q, w, "src", &mf);
  if (nw <= 0) return nw;
  kb_corrector<<<(nw + BW - 1) / BW, 256, 0, b->stream>>>(b->tstep.d, b->wstep.d, nw);
  with_nf(mf, [&](auto nf) {
    kb_continuity<decltype(nf)::value><<<(nw + BW - 1) / BW, 256, 0, b->stream>>>(b->tstep.d, b->wstep.d, nw);
    return 0;
  });
  HIPCHK(hipGetLastError());
  return 0;
}

}  // extern "C"

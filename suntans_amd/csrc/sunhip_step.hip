// sunhip_step.hip -- the stencils either side of the pressure solve on
// gfx950: the Poisson source (ComputeQSource, main/phys.c:2345-2395), the
// projection of the velocity and pressure (Corrector, main/phys.c:2301-2332)
// and the vertical velocity from continuity (Continuity,
// main/phys.c:3814-3844).  One launch each over max(Nt, Net) tiles: lanes
// of cell tiles own a column, lanes of edge tiles an edge.  The bodies live
// in sunhip_steptile.h, shared with the batched engine.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sunhip_steptile.h"
#include "sunhip_solvers.h"

using namespace sunhip;

template <int NF>
__global__ __launch_bounds__(256) void k_qsource(Geo G, StepArgs A) {
  qsource_tile<NF>(G, A, xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

__global__ __launch_bounds__(256) void k_corrector(Geo G, StepArgs A) {
  corrector_tile(G, A, xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

template <int NF>
__global__ __launch_bounds__(256) void k_continuity(Geo G, StepArgs A) {
  continuity_tile<NF>(G, A, xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

namespace sunhip {

// Device pointers of a step's fields (created on demand).  qc, q and w name
// the correction, the pressure and the vertical velocity fields; the others
// are fixed: u, utmp2, wtmp2, dzf, dzz, src, D.
int step_args(sunhip_grid *g, double theta, double dt, const char *qc, const char *q, const char *w,
              const char *src, StepArgs *out) {
  if (!g->have_extents) { set_error("time-step stencils: call sunhip_set_extents first"); return -1; }
  if (!(theta > 0) || !(dt > 0)) { set_error("time-step stencils: theta=%g and dt=%g must be positive", theta, dt); return -1; }
  Field *fu, *fw, *fq, *fqc, *fu2, *fw2, *fdzf, *fdzz, *fsrc, *fD;
  int rc = field(g, "u", &fu) | field(g, w, &fw) | field(g, q, &fq) | field(g, qc, &fqc) | field(g, "utmp2", &fu2) |
           field(g, "wtmp2", &fw2) | field(g, "dzf", &fdzf) | field(g, "dzz", &fdzz) | field(g, src, &fsrc) |
           field(g, "D", &fD);
  if (rc) return rc;
  if (fw->kind != CELLW || fw2->kind != CELLW) { set_error("time-step stencils: %s must be a w-type field", w); return -1; }
  if (fq->kind != CELL || fqc->kind != CELL || fsrc->kind != CELL) {
    set_error("time-step stencils: %s, %s and %s must be cell fields", q, qc, src);
    return -1;
  }
  StepArgs A{};
  A.u = fu->d; A.w = fw->d; A.q = fq->d; A.qc = fqc->d; A.utmp2 = fu2->d; A.wtmp2 = fw2->d; A.dzf = fdzf->d;
  A.dzz = fdzz->d; A.src = fsrc->d; A.D = fD->d; A.theta = theta; A.dt = dt;
  *out = A;
  return 0;
}

}  // namespace sunhip

extern "C" {

int sunhip_compute_qsource(sunhip_grid *g, double theta, double dt, const char *src) {
  HIPCHK(hipSetDevice(g->device));
  StepArgs A;
  int rc = step_args(g, theta, dt, "q", "qp", "w", src, &A);
  if (rc) return rc;
  const Geo G = g->geo();
  const int nb = (std::max(g->Nt, g->Net) + 3) / 4;
  const int ps = prof_begin(g, PK_QSOURCE);
  with_nf(g->mf, [&](auto nf) {
    k_qsource<decltype(nf)::value><<<nb, 256, 0, g->stream>>>(G, A);
    return 0;
  });
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  return 0;
}

int sunhip_corrector(sunhip_grid *g, double dt, const char *qc, const char *q) {
  HIPCHK(hipSetDevice(g->device));
  StepArgs A;
  int rc = step_args(g, 1.0, dt, qc, q, "w", "src", &A);
  if (rc) return rc;
  const int nb = (std::max(g->Nt, g->Net) + 3) / 4;
  const int ps = prof_begin(g, PK_CORRECT);
  k_corrector<<<nb, 256, 0, g->stream>>>(g->geo(), A);
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  return 0;
}

int sunhip_continuity(sunhip_grid *g, double theta, const char *w) {
  HIPCHK(hipSetDevice(g->device));
  StepArgs A;
  int rc = step_args(g, theta, 1.0, "q", "qp", w, "src", &A);
  if (rc) return rc;
  const Geo G = g->geo();
  const int nb = (g->Nt + 3) / 4;
  const int ps = prof_begin(g, PK_CONTINUITY);
  with_nf(g->mf, [&](auto nf) {
    k_continuity<decltype(nf)::value><<<nb, 256, 0, g->stream>>>(G, A);
    return 0;
  });
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // extern "C"

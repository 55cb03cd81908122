// sunhip_upred.hip -- the horizontal-momentum predictor (UPredictor,
// main/phys.c:2579-3018) on gfx950, in the two parts either side of the
// reference's open-boundary hooks (BoundaryVelocities, OpenBoundaryFluxes,
// which are user code in each case's boundaries.c):
//   sunhip_u_predict  main/phys.c:2596-2942: wtmp2 = w, utmp2 = u, and per
//                     edge U*** (utmp), E = A^-1 1 and D = dzf^T E
//   sunhip_u_solve    main/phys.c:2951-3018: htmp, the free-surface PCG,
//                     drying, u = U*** - g theta dt E grad h
// The bodies are in sunhip_utile.h.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sunhip_solvers.h"
#include "sunhip_utile.h"

using namespace sunhip;

__global__ __launch_bounds__(256) void k_upred_cells(Geo G, UPArgs A) {
  upred_cells_tile(G, A, xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}
__global__ __launch_bounds__(256) void k_upred_edges(Geo G, UPArgs A) {
  upred_edge_tile(G, A, xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}
__global__ __launch_bounds__(256) void k_upred_hterm(Geo G, UPArgs A) {
  upred_hterm_tile(G, A, xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}
template <int NF>
__global__ __launch_bounds__(256) void k_upred_htmp(Geo G, UPArgs A) {
  upred_htmp_tile<NF>(G, A, xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}
__global__ __launch_bounds__(256) void k_upred_dry(Geo G, UPArgs A) {
  upred_dry_tile(G, A, xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}
__global__ __launch_bounds__(256) void k_upred_velocity(Geo G, UPArgs A) {
  upred_velocity_tile(G, A, xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

namespace sunhip {

static int upred_args(sunhip_grid *g, const sunhip_upred_params *p, UPArgs *out) {
  if (!g->have_extents) { set_error("UPredictor: call sunhip_set_extents first"); return -1; }
  if (!(p->dt > 0) || !(p->theta > 0)) { set_error("UPredictor: dt=%g and theta=%g must be positive", p->dt, p->theta); return -1; }
  struct { const char *name; double **ptr; } f[] = {
      {"u", nullptr}, {"utmp", nullptr}, {"utmp2", nullptr}, {"ue", nullptr}, {"dzf", nullptr},
      {"up_c", nullptr}, {"up_bp", nullptr}, {"up_ep", nullptr}, {"D", nullptr}, {"tau_T", nullptr},
      {"CdT", nullptr}, {"CdB", nullptr}, {"dzz", nullptr}, {"nu_tv", nullptr}, {"nu_lax", nullptr},
      {"w", nullptr}, {"wtmp2", nullptr}, {"s", nullptr}, {"h", nullptr}, {"htmp", nullptr},
      {"active", nullptr}, {"hcorr", nullptr}, {"dv", nullptr}, {"up_hterm", nullptr}};
  UPArgs A{};
  double **dst[] = {&A.u, &A.utmp, &A.utmp2, &A.E, (double **)&A.dzf, &A.sc, &A.sbp, &A.sep, &A.D, (double **)&A.tau_T, (double **)&A.CdT, (double **)&A.CdB,
                    (double **)&A.dzz, (double **)&A.nu_tv, (double **)&A.nu_lax, (double **)&A.w, &A.wtmp2, &A.s,
                    &A.h, &A.htmp, &A.active, &A.hcorr, (double **)&A.dv, &A.hterm};
  for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); i++) {
    Field *fp;
    int rc = field(g, f[i].name, &fp);
    if (rc) return rc;
    *dst[i] = fp->d;
  }
  if (!g->d_uerr && dev_alloc(g, (void **)&g->d_uerr, 2 * sizeof(int))) return -2;
  A.err = g->d_uerr;
  A.dt = p->dt; A.theta = p->theta; A.grav = p->grav; A.nu = p->nu; A.lax = p->lax_vertical;
  A.thetaM = p->thetaM; A.nonlinear = p->nonlinear;
  *out = A;
  return 0;
}

}  // namespace sunhip

extern "C" {

int sunhip_u_predict(sunhip_grid *g, const sunhip_upred_params *p) {
  HIPCHK(hipSetDevice(g->device));
  UPArgs A;
  int rc = upred_args(g, p, &A);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(A.err, 0, 2 * sizeof(int), g->stream));
  const Geo G = g->geo();
  k_upred_cells<<<(g->Nt + 3) / 4, 256, 0, g->stream>>>(G, A);
  k_upred_edges<<<(g->Net + 3) / 4, 256, 0, g->stream>>>(G, A);
  HIPCHK(hipGetLastError());
  int err[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(err, A.err, sizeof(err), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  if (err[0]) { set_error("UPredictor: an edge level has two cells of zero height (main/phys.c:2900-2904)"); return -5; }
  if (err[1]) { set_error("UPredictor: U*** = nan (main/phys.c:2945-2950)"); return -6; }
  return 0;
}

int sunhip_u_solve(sunhip_grid *g, const sunhip_upred_params *p, const sunhip_cg_params *prm, sunhip_cg_stats *st) {
  HIPCHK(hipSetDevice(g->device));
  UPArgs A;
  int rc = upred_args(g, p, &A);
  if (rc) return rc;
  const Geo G = g->geo();
  k_upred_hterm<<<(g->Net + 3) / 4, 256, 0, g->stream>>>(G, A);
  with_nf(g->mf, [&](auto nf) {
    k_upred_htmp<decltype(nf)::value><<<(g->Nt + 3) / 4, 256, 0, g->stream>>>(G, A);
    return 0;
  });
  HIPCHK(hipGetLastError());
  if ((rc = sunhip_h_solve(g, prm, p->grav, p->theta, p->dt, st))) return rc;
  k_upred_dry<<<(g->Nt + 3) / 4, 256, 0, g->stream>>>(G, A);
  k_upred_velocity<<<(g->Net + 3) / 4, 256, 0, g->stream>>>(G, A);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

}  // extern "C"

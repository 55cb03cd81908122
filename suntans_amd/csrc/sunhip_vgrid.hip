// sunhip_vgrid.hip -- the vertical grid that moves with the free surface,
// on gfx950: UpdateDZ (main/phys.c:816-968: layer heights dzz, top cells
// ctop/etop, the old copies) and SetFluxHeight (main/phys.c:4338-4435: the
// flux-face heights dzf).  One lane per column, then one lane per edge; the
// operations keep the reference's order, so the results are bit-identical.
// The per-face level ranges the solvers read are rebuilt from the new
// extents on the device, so no extents travel from the host afterwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "sunhip_solvers.h"

using namespace sunhip;

namespace sunhip {

struct VGArgs {
  double *dzz, *dzzold, *h, *dv, *dzbot, *dzfB, *dzf;
  const double *u, *dz;
  int *ctop, *ctopold, *etop, *etopold;
  double dzsmall, smoothbot;
  int Nkmax, stairstep, fixdzz;
};

__device__ __forceinline__ double vmax(double x1, double x2) { return x2 > x1 ? x2 : x1; }  // main/util.c:212-217
__device__ __forceinline__ double vmin(double x, double y) { return x < y ? x : y; }      // main/util.c:241-246
__device__ __forceinline__ double vupwind(double u, double a, double b) {                // main/util.c:219-232
  if (fabs(u) > 0) return u > 0 ? b : a;
  return vmax(a, b);
}

}  // namespace sunhip

// UpdateDZ, the column part (main/phys.c:824-848, 850-920, 934-951).
template <int OPT>
__global__ __launch_bounds__(256) void k_udz_cells(Geo G, VGArgs A) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0) return;
  const uint32_t base = (uint32_t)(G.toff[t] + l);
  auto Z = [&](double *p, int k) -> double & { return at32(p, base + (uint32_t)k * 64u); };
  if (OPT == 0) {
    const int ct = A.ctop[s];
    A.ctopold[s] = ct;
    for (int k = 0; k < ct; k++) Z(A.dzzold, k) = 0;
    for (int k = ct; k < n; k++) Z(A.dzzold, k) = Z(A.dzz, k);
  }
  if (OPT == -1) A.h[s] = 0.0;
  const double h = A.h[s], dv = A.dv[s];
  {
    double z = 0;
    for (int k = 0; k < n; k++) z -= A.dz[k];
    Z(A.dzz, n - 1) = A.dz[n - 1] + dv + z;
  }
  int ct = A.ctop[s];
  if (A.Nkmax > 1) {
    double z = 0;
    bool found = false;
    for (int k = 0; k < n; k++) {
      z -= A.dz[k];
      if (h >= z) {
        if (!found) {
          if (k == n - 1) { Z(A.dzz, k) = h + dv; ct = k; }
          else if (h == z) { Z(A.dzz, k) = 0; ct = k + 1; }
          else { Z(A.dzz, k) = h - z; ct = k; }
          found = true;
        } else if (k == n - 1) {
          Z(A.dzz, k) = A.dz[k] + dv + z;
        } else {
          Z(A.dzz, k) = z < -dv ? 0 : A.dz[k];
        }
      } else {
        Z(A.dzz, k) = 0;
      }
    }
  } else {
    Z(A.dzz, 0) = dv + h;
  }
  A.ctop[s] = ct;
  if (OPT == -1) {
    const int k = n - 1;
    A.dzbot[s] = Z(A.dzz, k);
    if (!A.stairstep && A.fixdzz && Z(A.dzz, k) < A.dz[k] * A.dzsmall) A.dv[s] += (A.dz[k] * A.dzsmall - Z(A.dzz, k));
  }
  if (OPT == 1) {
    A.ctopold[s] = ct;
    for (int k = 0; k < n; k++) Z(A.dzzold, k) = Z(A.dzz, k);
  }
}

// UpdateDZ, the edge part (main/phys.c:922-932, 952-966).
template <int OPT>
__global__ __launch_bounds__(256) void k_udz_edges(Geo G, VGArgs A) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (t >= G.Net) return;
  const int s = t * 64 + l;
  if (G.e_nkc[s] == 0) return;  // padding
  if (OPT == 0) A.etopold[s] = A.etop[s];
  int s1 = G.e_s1[s], s2 = G.e_s2[s];
  int et;
  if (s1 < 0) et = A.ctop[s2];
  else if (s2 < 0) et = A.ctop[s1];
  else et = A.ctop[s1] < A.ctop[s2] ? A.ctop[s1] : A.ctop[s2];
  A.etop[s] = et;
  if (OPT == 1) {
    A.etopold[s] = et;
    int b1 = G.e_c1[s], b2 = G.e_c2[s];
    if (s1 < 0) { s1 = s2; b1 = b2; }
    if (s2 < 0) { s2 = s1; b2 = b1; }
    double z1 = at32(A.dzz, (uint32_t)b1 + (uint32_t)(G.nk[s1] - 1) * 64u);
    double z2 = at32(A.dzz, (uint32_t)b2 + (uint32_t)(G.nk[s2] - 1) * 64u);
    double z = 0;
    for (int k = 0; k < G.e_nke[s] - 1; k++) z -= A.dz[k];
    if (A.h[s1] < z) z1 = z1 - z + A.h[s1];
    if (A.h[s2] < z) z2 = z2 - z + A.h[s2];
    A.dzfB[s] = vmin(z1, z2);
  }
}

// SetFluxHeight's smoothbot bounds on the bottom cells (main/phys.c:4345-4349
// before, :4429-4432 after).
template <bool BEFORE>
__global__ __launch_bounds__(256) void k_sfh_bottom(Geo G, VGArgs A) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (t >= G.Nt) return;
  const int s = t * 64 + l, n = G.nk[s];
  if (n == 0) return;
  double &zb = at32(A.dzz, (uint32_t)(G.toff[t] + l) + (uint32_t)(n - 1) * 64u);
  if (BEFORE) zb = vmax(A.dzbot[s], A.smoothbot * A.dz[n - 1]);
  else zb = vmax(zb, A.dzsmall * A.dz[n - 1]);
}

// SetFluxHeight, one lane per edge (main/phys.c:4342-4344, 4351-4427).
__global__ __launch_bounds__(256) void k_sfh_edges(Geo G, VGArgs A) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (t >= G.Net) return;
  const int s = t * 64 + l, nkc = G.e_nkc[s];
  if (nkc == 0) return;
  const uint32_t base = (uint32_t)(G.eoff[t] + l);
  auto F = [&](int k) -> double & { return at32(A.dzf, base + (uint32_t)k * 64u); };
  for (int k = 0; k < nkc; k++) F(k) = 0;
  int s1 = G.e_s1[s], s2 = G.e_s2[s], b1 = G.e_c1[s], b2 = G.e_c2[s];
  if (s1 < 0) { s1 = s2; b1 = b2; }
  if (s2 < 0) { s2 = s1; b2 = b1; }
  const int et = A.etop[s], ne = G.e_nke[s];
  auto Z1 = [&](int k) { return at32(A.dzz, (uint32_t)b1 + (uint32_t)k * 64u); };
  auto Z2 = [&](int k) { return at32(A.dzz, (uint32_t)b2 + (uint32_t)k * 64u); };
  for (int k = et; k < ne; k++) F(k) = vupwind(at32(A.u, base + (uint32_t)k * 64u), Z1(k), Z2(k));
  const int k = ne - 1;
  if (et == k)
    F(k) = vmax(0, vupwind(at32(A.u, base + (uint32_t)k * 64u), A.h[s1], A.h[s2]) + vmin(A.dv[s1], A.dv[s2]));
  else
    F(k) = vmin(Z1(k), Z2(k));
  for (int kk = et; kk < ne; kk++)
    if (F(kk) <= 1e-10) F(kk) = 0;
}

namespace sunhip {

static int vg_args(sunhip_grid *g, const sunhip_vgrid_params *vp, VGArgs *out) {
  if (!g->have_extents) { set_error("vertical grid: call sunhip_set_extents first"); return -1; }
  if (!vp || !vp->dz) { set_error("vertical grid: dz is required"); return -1; }
  const char *names[] = {"dzz", "dzzold", "h", "dv", "dzbot", "dzfB", "dzf", "u"};
  double *ptr[8];
  for (int i = 0; i < 8; i++) {
    Field *f;
    int rc = field(g, names[i], &f);
    if (rc) return rc;
    ptr[i] = f->d;
  }
  if (!g->d_dz && dev_alloc(g, (void **)&g->d_dz, sizeof(double) * (size_t)std::max(g->Nkmax, 1))) return -2;
  HIPCHK(hipMemcpyAsync(g->d_dz, vp->dz, sizeof(double) * (size_t)g->Nkmax, hipMemcpyHostToDevice, g->stream));
  if (int rc = set_zdz(g, vp->dz)) return rc;
  VGArgs A{};
  A.dzz = ptr[0]; A.dzzold = ptr[1]; A.h = ptr[2]; A.dv = ptr[3]; A.dzbot = ptr[4]; A.dzfB = ptr[5];
  A.dzf = ptr[6]; A.u = ptr[7]; A.dz = g->d_dz;
  A.ctop = g->d_ctop; A.ctopold = g->d_ctopold; A.etop = g->d_e_etop; A.etopold = g->d_e_etopold;
  A.dzsmall = vp->dzsmall; A.smoothbot = vp->smoothbot;
  A.Nkmax = g->Nkmax; A.stairstep = vp->stairstep; A.fixdzz = vp->fixdzz;
  *out = A;
  return 0;
}

}  // namespace sunhip

extern "C" {

int sunhip_update_dz(sunhip_grid *g, const sunhip_vgrid_params *vp, int option) {
  HIPCHK(hipSetDevice(g->device));
  if (option != 0 && option != 1 && option != -1) { set_error("sunhip_update_dz: option must be 0, 1 or -1"); return -1; }
  if (vp->linearFS) return 0;  // main/phys.c:820-822
  VGArgs A;
  int rc = vg_args(g, vp, &A);
  if (rc) return rc;
  const Geo G = g->geo();
  const int nc = (g->Nt + 3) / 4, ne = (g->Net + 3) / 4;
  if (option == 0) { k_udz_cells<0><<<nc, 256, 0, g->stream>>>(G, A); k_udz_edges<0><<<ne, 256, 0, g->stream>>>(G, A); }
  else if (option == 1) { k_udz_cells<1><<<nc, 256, 0, g->stream>>>(G, A); k_udz_edges<1><<<ne, 256, 0, g->stream>>>(G, A); }
  else { k_udz_cells<-1><<<nc, 256, 0, g->stream>>>(G, A); k_udz_edges<-1><<<ne, 256, 0, g->stream>>>(G, A); }
  HIPCHK(hipGetLastError());
  return refresh_face_ranges(g);
}

int sunhip_set_flux_height(sunhip_grid *g, const sunhip_vgrid_params *vp) {
  HIPCHK(hipSetDevice(g->device));
  VGArgs A;
  int rc = vg_args(g, vp, &A);
  if (rc) return rc;
  const Geo G = g->geo();
  const int nc = (g->Nt + 3) / 4, ne = (g->Net + 3) / 4;
  if (vp->smoothbot) k_sfh_bottom<true><<<nc, 256, 0, g->stream>>>(G, A);
  k_sfh_edges<<<ne, 256, 0, g->stream>>>(G, A);
  if (vp->smoothbot) k_sfh_bottom<false><<<nc, 256, 0, g->stream>>>(G, A);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // extern "C"

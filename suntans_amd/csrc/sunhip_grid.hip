// sunhip_grid.hip -- grid context: spatial tiling, device metadata, field
// staging between the flat host layout and the tiled device layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>

#include "sunhip_internal.h"

namespace sunhip {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

int dev_alloc(sunhip_grid *g, void **p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return -2;
  }
  g->bytes += (int64_t)bytes;
  return 0;
}

Kind kind_of(const char *name) {
  static const char *edge[] = {"dzf", "u", "utmp", "utmp2", "u0", "ue", "SfHp", "SfHm", "ef0", "ef1", "bscal",
                               "bw", "tvdP1", "tvdP2"};
  static const char *edge2[] = {"D", "df", "dg", "Dq", "Dh", "tau_T", "CdT", "CdB", "dzfB", "def1"};
  static const char *cellw[] = {"Cn_W", "Cn_W2"};
  static const char *cell2[] = {"h", "htmp", "hold", "htmp2", "htmp3", "hcoef", "hr", "hz", "hrt", "hp0", "hp1",
                                "hx", "hy", "active", "htmp0", "dv", "hcorr", "dzbot", "Ftop", "Fbot"};
  if (!strncmp(name, "up_", 3)) return EDGE;  // UPredictor's per-edge scratch
  if (!strcmp(name, "wc")) return CELL;       // cell-centred w: Nk values per column
  for (auto n : cellw) if (!strcmp(name, n)) return CELLW;
  for (auto n : edge) if (!strcmp(name, n)) return EDGE;
  for (auto n : edge2) if (!strcmp(name, n)) return EDGE2;
  for (auto n : cell2) if (!strcmp(name, n)) return CELL2;
  if (name[0] == 'w') return CELLW;
  return CELL;
}

int field(sunhip_grid *g, const char *name, Field **out) {
  auto it = g->fields.find(name);
  if (it == g->fields.end()) {
    Field f;
    f.kind = kind_of(name);
    switch (f.kind) {
      case CELL: case CELLW: f.n = (size_t)g->NT; break;
      case EDGE: f.n = (size_t)g->NE; break;
      case CELL2: f.n = (size_t)g->NS; break;
      case EDGE2: f.n = (size_t)g->NES; break;
    }
    int rc = dev_alloc(g, (void **)&f.d, f.n * sizeof(double));
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(f.d, 0, f.n * sizeof(double), g->stream));
    it = g->fields.emplace(name, f).first;
  }
  *out = &it->second;
  return 0;
}

}  // namespace sunhip

using namespace sunhip;

sunhip::Geo sunhip_grid::geo() const {
  Geo G;
  G.Nt = Nt; G.NS = NS; G.Net = Net; G.NES = NES; G.mf = mf;
  G.toff = d_toff; G.tdep = d_tdep; G.eoff = d_eoff; G.edep = d_edep;
  G.nk = d_nk; G.ctop = d_ctop; G.ctopold = d_ctopold; G.comp = d_comp; G.nf = d_nf;
  G.nbb = d_nbb; G.nbs = d_nbs; G.eb = d_eb; G.es = d_es; G.nrm = d_nrm; G.kr = d_kr; G.Ac = d_Ac;
  G.e_c1 = d_e_c1; G.e_c2 = d_e_c2; G.e_s1 = d_e_s1; G.e_s2 = d_e_s2;
  G.e_nke = d_e_nke; G.e_nkc = d_e_nkc; G.e_etop = d_e_etop; G.e_mark = d_e_mark; G.e_comp = d_e_comp;
  G.df = d_df; G.dg = d_dg;
  G.xcode = d_xcode; G.pair = d_pair; G.npair = d_npair;
  G.b4perm = tune_order ? d_b4perm : nullptr;
  return G;
}

// ---------------------------------------------------------------------------
// staging kernels: one block per tile, one thread per lane

__global__ void k_scatter_cols(const double *__restrict__ flat, double *__restrict__ dst,
                               const int64_t *__restrict__ fstart, const int *__restrict__ ncol,
                               int extra, const int *__restrict__ toff, const int *__restrict__ tdep) {
  const int t = blockIdx.x, l = threadIdx.x, s = t * SUNHIP_TILE + l;
  const int64_t f0 = fstart[s];
  const int n = f0 < 0 ? 0 : ncol[s] + extra;
  double *col = dst + toff[t] + l;
  for (int k = 0; k < tdep[t]; k++) col[(int64_t)k * SUNHIP_TILE] = k < n ? flat[f0 + k] : 0.0;
}

__global__ void k_gather_cols(double *__restrict__ flat, const double *__restrict__ src,
                              const int64_t *__restrict__ fstart, const int *__restrict__ ncol, int extra,
                              const int *__restrict__ toff) {
  const int t = blockIdx.x, l = threadIdx.x, s = t * SUNHIP_TILE + l;
  const int64_t f0 = fstart[s];
  if (f0 < 0) return;
  const int n = ncol[s] + extra;
  const double *col = src + toff[t] + l;
  for (int k = 0; k < n; k++) flat[f0 + k] = col[(int64_t)k * SUNHIP_TILE];
}

__global__ void k_scatter_2d(const double *__restrict__ flat, double *__restrict__ dst,
                             const int *__restrict__ slot, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[slot[i]] = flat[i];
}

__global__ void k_gather_2d(double *__restrict__ flat, const double *__restrict__ src,
                            const int *__restrict__ slot, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flat[i] = src[slot[i]];
}

// Per-call horizontal coupling range of every face: [max(ctop_i, ctop_nc), Nke).
__global__ void k_face_ranges(int NS, int mf, const int *__restrict__ nbs, const int *__restrict__ es,
                              const int *__restrict__ ctop, const int *__restrict__ nke_e,
                              int *__restrict__ kr) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= NS) return;
  for (int f = 0; f < mf; f++) {
    const int nb = nbs[f * NS + s], e = es[f * NS + s];
    int v = 0;
    if (nb >= 0 && e >= 0) {
      const int kmin = ctop[nb] > ctop[s] ? ctop[nb] : ctop[s];
      v = kmin | (nke_e[e] << 16);
    } else if (e >= 0) {
      v = ctop[s] | (nke_e[e] << 16);
    }
    kr[f * NS + s] = v;
  }
}

// ---------------------------------------------------------------------------

static uint32_t spread16(uint32_t v) {
  v &= 0xffff;
  v = (v | (v << 8)) & 0x00ff00ff;
  v = (v | (v << 4)) & 0x0f0f0f0f;
  v = (v | (v << 2)) & 0x33333333;
  v = (v | (v << 1)) & 0x55555555;
  return v;
}

namespace sunhip {
int refresh_face_ranges(sunhip_grid *g) {
  k_face_ranges<<<(g->NS + 255) / 256, 256, 0, g->stream>>>(g->NS, g->mf, g->d_nbs, g->d_es, g->d_ctop, g->d_e_nke,
                                                            g->d_kr);
  HIPCHK(hipGetLastError());
  return 0;
}
}  // namespace sunhip

template <class T>
static int upload_vec(sunhip_grid *g, T **d, const std::vector<T> &h) {
  int rc = dev_alloc(g, (void **)d, h.size() * sizeof(T));
  if (rc) return rc;
  if (!h.empty()) HIPCHK(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

extern "C" {

const char *sunhip_error(void) { return g_err; }

int sunhip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// Host-side checks of everything the kernels index with: a value outside its
// range would make a kernel read or write outside a device array.
static int validate_topology(const sunhip_topology *T) {
  if (!T || T->Nc <= 0 || T->Ne <= 0 || T->maxfaces <= 0) {
    set_error("sunhip_grid_create: empty topology");
    return -1;
  }
  if (T->maxfaces > SUNHIP_MAXF) {
    set_error("sunhip_grid_create: maxfaces=%d exceeds %d", T->maxfaces, SUNHIP_MAXF);
    return -1;
  }
  const int Nc = T->Nc, Ne = T->Ne, mf = T->maxfaces;
  for (int i = 0; i < Nc; i++) {
    if (T->Nk[i] < 1 || T->Nk[i] > T->Nkmax || T->nfaces[i] < 1 || T->nfaces[i] > mf) {
      set_error("sunhip_grid_create: cell %d has Nk=%d (Nkmax %d) nfaces=%d (maxfaces %d)", i, T->Nk[i], T->Nkmax,
                T->nfaces[i], mf);
      return -1;
    }
    for (int f = 0; f < T->nfaces[i]; f++) {
      const int nb = T->neigh[(size_t)i * mf + f], e = T->face[(size_t)i * mf + f];
      if (nb < -1 || nb >= Nc || e < 0 || e >= Ne) {
        set_error("sunhip_grid_create: cell %d face %d has neigh=%d face=%d outside [-1,%d) and [0,%d)", i, f, nb, e,
                  Nc, Ne);
        return -1;
      }
    }
  }
  for (int e = 0; e < Ne; e++) {
    const int a = T->grad[2 * (size_t)e], b = T->grad[2 * (size_t)e + 1];
    if (a < -1 || a >= Nc || b < -1 || b >= Nc || (a < 0 && b < 0) || T->Nkc[e] < 0 || T->Nkc[e] > T->Nkmax) {
      set_error("sunhip_grid_create: edge %d has grad=(%d,%d) Nkc=%d (Nc %d, Nkmax %d)", e, a, b, T->Nkc[e], Nc,
                T->Nkmax);
      return -1;
    }
  }
  std::vector<char> seen((size_t)std::max(Nc, Ne), 0);
  for (int i = 0; i < Nc; i++) {
    const int c = T->cellp[i];
    if (c < 0 || c >= Nc || seen[c]) {
      set_error("sunhip_grid_create: cellp is not a permutation of the cells (entry %d = %d)", i, c);
      return -1;
    }
    seen[c] = 1;
  }
  std::fill(seen.begin(), seen.end(), 0);
  for (int j = 0; j < Ne; j++) {
    const int e = T->edgep[j];
    if (e < 0 || e >= Ne || seen[e]) {
      set_error("sunhip_grid_create: edgep is not a permutation of the edges (entry %d = %d)", j, e);
      return -1;
    }
    seen[e] = 1;
  }
  for (int k = 0; k < 3; k++)
    if (T->celldist[k] < 0 || T->celldist[k] > Nc || (k && T->celldist[k] < T->celldist[k - 1])) {
      set_error("sunhip_grid_create: celldist is not a nondecreasing split of [0,%d]", Nc);
      return -1;
    }
  for (int k = 0; k < 6; k++)
    if (T->edgedist[k] < 0 || T->edgedist[k] > Ne || (k && T->edgedist[k] < T->edgedist[k - 1])) {
      set_error("sunhip_grid_create: edgedist is not a nondecreasing split of [0,%d]", Ne);
      return -1;
    }
  return 0;
}

// The moving extents against the topology's column sizes.
static int validate_extents(int Nc, int Ne, const int *Nk, const int *Nkc, const sunhip_extents *X) {
  for (int c = 0; c < Nc; c++) {
    const int ct = X->ctop[c], cto = X->ctopold ? X->ctopold[c] : ct;
    if (ct < 0 || ct > Nk[c] || cto < 0 || cto > Nk[c]) {
      set_error("sunhip_set_extents: cell %d has ctop=%d ctopold=%d outside [0,Nk=%d]", c, ct, cto, Nk[c]);
      return -1;
    }
  }
  for (int e = 0; e < Ne; e++) {
    const int nke = X->Nke[e], et = X->etop ? X->etop[e] : 0;
    if (nke < 0 || nke > Nkc[e] || et < 0 || et > Nkc[e]) {
      set_error("sunhip_set_extents: edge %d has Nke=%d etop=%d outside [0,Nkc=%d]", e, nke, et, Nkc[e]);
      return -1;
    }
  }
  return 0;
}

int sunhip_validate(const sunhip_topology *T, const sunhip_extents *X) {
  if (int rc = validate_topology(T)) return rc;
  return X ? validate_extents(T->Nc, T->Ne, T->Nk, T->Nkc, X) : 0;
}

int sunhip_grid_create(sunhip_grid **out, const sunhip_topology *T, int device) {
  *out = nullptr;
  if (int rc = validate_topology(T)) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("sunhip_grid_create: no HIP device");
    return -4;
  }
  if (device < 0 || device >= ndev) {
    set_error("sunhip_grid_create: device %d out of range (%d devices)", device, ndev);
    return -1;
  }
  HIPCHK(hipSetDevice(device));
  sunhip_grid *g = new sunhip_grid();
  g->device = device;
  const int Nc = T->Nc, Ne = T->Ne, mf = T->maxfaces;
  g->Nc = Nc; g->Ne = Ne; g->mf = mf; g->Nkmax = T->Nkmax;
  g->Nk.assign(T->Nk, T->Nk + Nc);
  g->Nkc.assign(T->Nkc, T->Nkc + Ne);
  g->face.assign(T->face, T->face + (size_t)Nc * mf);
  g->neigh.assign(T->neigh, T->neigh + (size_t)Nc * mf);
  g->normal.assign(T->normal, T->normal + (size_t)Nc * mf);
  g->nfaces.assign(T->nfaces, T->nfaces + Nc);
  g->grad.assign(T->grad, T->grad + 2 * (size_t)Ne);
  g->mark.assign(T->mark, T->mark + Ne);
  g->cellp.assign(T->cellp, T->cellp + Nc);
  g->edgep.assign(T->edgep, T->edgep + Ne);
  g->celldist.assign(T->celldist, T->celldist + 3);
  g->edgedist.assign(T->edgedist, T->edgedist + 6);

  // --- cell order: Morton curve of the Voronoi points
  double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
  for (int i = 0; i < Nc; i++) {
    xmin = std::min(xmin, T->xv[i]); xmax = std::max(xmax, T->xv[i]);
    ymin = std::min(ymin, T->yv[i]); ymax = std::max(ymax, T->yv[i]);
  }
  const double span = std::max(std::max(xmax - xmin, ymax - ymin), 1e-30);
  std::vector<uint32_t> key(Nc);
  for (int i = 0; i < Nc; i++) {
    uint32_t ix = (uint32_t)std::min(65535.0, std::floor((T->xv[i] - xmin) / span * 65535.0));
    uint32_t iy = (uint32_t)std::min(65535.0, std::floor((T->yv[i] - ymin) / span * 65535.0));
    key[i] = spread16(ix) | (spread16(iy) << 1);
  }
  std::vector<int> order(Nc);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  g->Nt = (Nc + SUNHIP_TILE - 1) / SUNHIP_TILE;
  g->NS = g->Nt * SUNHIP_TILE;
  g->slot_of_cell.assign(Nc, -1);
  g->cell_of_slot.assign(g->NS, -1);
  for (int s = 0; s < Nc; s++) { g->slot_of_cell[order[s]] = s; g->cell_of_slot[s] = order[s]; }
  g->tdep.assign(g->Nt, 1);
  g->toff.assign(g->Nt + 1, 0);
  for (int t = 0; t < g->Nt; t++) {
    int d = 0;
    for (int l = 0; l < SUNHIP_TILE; l++) {
      int c = g->cell_of_slot[t * SUNHIP_TILE + l];
      if (c >= 0) d = std::max(d, T->Nk[c]);
    }
    g->tdep[t] = d + 1;
    g->maxdepth = std::max(g->maxdepth, d);
  }
  int64_t acc = 0;
  for (int t = 0; t < g->Nt; t++) { g->toff[t] = (int)acc; acc += (int64_t)g->tdep[t] * SUNHIP_TILE; }
  if (acc >= (int64_t)1 << 29) {  // fields are addressed by 32-bit byte offsets
    set_error("sunhip_grid_create: %lld cell-layer slots exceed the 32-bit layout", (long long)acc);
    delete g;
    return -1;
  }
  g->toff[g->Nt] = (int)acc;
  g->NT = acc;

  // --- edge order: follow the first adjacent cell's slot
  std::vector<int64_t> ekey(Ne);
  for (int j = 0; j < Ne; j++) {
    int a = T->grad[2 * j], b = T->grad[2 * j + 1];
    int sa = a >= 0 ? g->slot_of_cell[a] : INT32_MAX, sb = b >= 0 ? g->slot_of_cell[b] : INT32_MAX;
    ekey[j] = (int64_t)std::min(sa, sb) * 4 + (sa < sb ? 0 : 1);
  }
  std::vector<int> eorder(Ne);
  std::iota(eorder.begin(), eorder.end(), 0);
  std::stable_sort(eorder.begin(), eorder.end(), [&](int a, int b) { return ekey[a] < ekey[b]; });
  g->Net = (Ne + SUNHIP_TILE - 1) / SUNHIP_TILE;
  g->NES = g->Net * SUNHIP_TILE;
  g->slot_of_edge.assign(Ne, -1);
  g->edge_of_slot.assign(g->NES, -1);
  for (int s = 0; s < Ne; s++) { g->slot_of_edge[eorder[s]] = s; g->edge_of_slot[s] = eorder[s]; }
  g->edep.assign(g->Net, 1);
  g->eoff.assign(g->Net + 1, 0);
  acc = 0;
  for (int t = 0; t < g->Net; t++) {
    int d = 1;
    for (int l = 0; l < SUNHIP_TILE; l++) {
      int e = g->edge_of_slot[t * SUNHIP_TILE + l];
      if (e >= 0) d = std::max(d, T->Nkc[e]);
    }
    g->edep[t] = d;
    g->eoff[t] = (int)acc;
    acc += (int64_t)d * SUNHIP_TILE;
  }
  if (acc >= (int64_t)1 << 29) {  // fields are addressed by 32-bit byte offsets
    set_error("sunhip_grid_create: edge layout exceeds 2^29 values (32-bit byte offsets)");
    delete g;
    return -1;
  }
  g->eoff[g->Net] = (int)acc;
  g->NE = acc;

  // --- flat offsets
  g->cstart.assign(Nc + 1, 0);
  for (int i = 0; i < Nc; i++) g->cstart[i + 1] = g->cstart[i] + T->Nk[i];
  g->Nflat = g->cstart[Nc];
  g->NflatW = g->Nflat + Nc;
  g->estart.assign(Ne + 1, 0);
  for (int j = 0; j < Ne; j++) g->estart[j + 1] = g->estart[j] + T->Nkc[j];
  g->NflatE = g->estart[Ne];

  // --- per-slot metadata
  const int NS = g->NS, NES = g->NES;
  auto cell_elem = [&](int c) { int s = g->slot_of_cell[c]; return g->toff[s / 64] + (s % 64); };
  auto edge_elem = [&](int e) { int s = g->slot_of_edge[e]; return g->eoff[s / 64] + (s % 64); };
  // per-face tables: one row of NS per face slot of the kernels' face-count
  // class (>= maxfaces), so a kernel may load every row of its class
  const int mfc = nf_class(mf);
  std::vector<int> nk(NS, 0), nf(NS, 0), comp(NS, 0), nbb(mfc * NS, -1), nbs(mfc * NS, -1), eb(mfc * NS, -1),
      es(mfc * NS, -1), nrm(mfc * NS, 0);
  std::vector<double> Ac(NS, 0.0);
  std::vector<int64_t> cst(NS, -1), cstw(NS, -1);
  for (int s = 0; s < NS; s++) {
    int c = g->cell_of_slot[s];
    if (c < 0) continue;
    nk[s] = T->Nk[c];
    nf[s] = T->nfaces[c];
    Ac[s] = T->Ac[c];
    cst[s] = g->cstart[c];
    cstw[s] = g->cstart[c] + c;
    for (int f = 0; f < T->nfaces[c]; f++) {
      int nb = T->neigh[c * mf + f], e = T->face[c * mf + f];
      if (nb >= 0) { nbb[f * NS + s] = cell_elem(nb); nbs[f * NS + s] = g->slot_of_cell[nb]; }
      if (e >= 0) { eb[f * NS + s] = edge_elem(e); es[f * NS + s] = g->slot_of_edge[e]; }
      nrm[f * NS + s] = T->normal[c * mf + f];
    }
  }
  for (int ip = T->celldist[0]; ip < T->celldist[1]; ip++) comp[g->slot_of_cell[T->cellp[ip]]] = 1;
  for (int ip = T->celldist[1]; ip < T->celldist[2]; ip++) comp[g->slot_of_cell[T->cellp[ip]]] = 2;
  std::vector<int> ec1(NES, -1), ec2(NES, -1), es1(NES, -1), es2(NES, -1), enkc(NES, 0), emark(NES, 0),
      ecomp(NES, 0);
  std::vector<double> df(NES, 0.0), dg(NES, 0.0);
  std::vector<int64_t> est(NES, -1);
  for (int s = 0; s < NES; s++) {
    int e = g->edge_of_slot[s];
    if (e < 0) continue;
    int a = T->grad[2 * e], b = T->grad[2 * e + 1];
    if (a >= 0) { ec1[s] = cell_elem(a); es1[s] = g->slot_of_cell[a]; }
    if (b >= 0) { ec2[s] = cell_elem(b); es2[s] = g->slot_of_cell[b]; }
    enkc[s] = T->Nkc[e];
    emark[s] = T->mark[e];
    df[s] = T->df[e];
    dg[s] = T->dg[e];
    est[s] = g->estart[e];
  }
  for (int jp = T->edgedist[0]; jp < T->edgedist[1]; jp++) ecomp[g->slot_of_edge[T->edgep[jp]]] = 1;
  for (int jp = T->edgedist[2]; jp < T->edgedist[3]; jp++) ecomp[g->slot_of_edge[T->edgep[jp]]] = 2;
  // each column's faces in the order their edges appear in edgep (the order
  // of the reference's edge loops that scatter onto cells)
  std::vector<int> ford(mfc * NS, -1);
  {
    std::vector<int> pos(Ne, 0);
    for (int jp = 0; jp < Ne; jp++) pos[T->edgep[jp]] = jp;
    for (int s = 0; s < NS; s++) {
      const int c = g->cell_of_slot[s];
      if (c < 0) continue;
      std::vector<int> fs(T->nfaces[c]);
      for (int f = 0; f < T->nfaces[c]; f++) fs[f] = f;
      std::stable_sort(fs.begin(), fs.end(),
                       [&](int x, int y) { return pos[T->face[c * mf + x]] < pos[T->face[c * mf + y]]; });
      for (int r = 0; r < T->nfaces[c]; r++) ford[r * NS + s] = fs[r];
    }
  }
  g->n_type4 = T->edgedist[5] - T->edgedist[4];
  g->type2_ok = 1;
  for (int jp = T->edgedist[2]; jp < T->edgedist[3]; jp++)
    if (T->grad[2 * T->edgep[jp]] < 0) g->type2_ok = 0;

  int rc = 0;
  HIPCHK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  g->own_stream = g->stream;
  rc |= upload_vec(g, &g->d_toff, g->toff);
  rc |= upload_vec(g, &g->d_tdep, g->tdep);
  rc |= upload_vec(g, &g->d_eoff, g->eoff);
  rc |= upload_vec(g, &g->d_edep, g->edep);
  rc |= upload_vec(g, &g->d_nk, nk);
  rc |= upload_vec(g, &g->d_nf, nf);
  rc |= upload_vec(g, &g->d_comp, comp);
  rc |= upload_vec(g, &g->d_nbb, nbb);
  rc |= upload_vec(g, &g->d_nbs, nbs);
  {
    // neighbour exchange: same-tile neighbours by lane, the others as a
    // compact list of cross-tile (face, slot) pairs per tile
    std::vector<int> xcode(mfc * NS, -1), pair((size_t)g->Nt * SUNHIP_XPAIRS, -1), npair(g->Nt, 0);
    for (int t = 0; t < g->Nt; t++) {
      int m = 0;
      for (int l = 0; l < SUNHIP_TILE; l++) {
        const int s = t * SUNHIP_TILE + l;
        for (int f = 0; f < nf[s]; f++) {
          const int nb = nbs[f * NS + s];
          if (nb < 0) continue;
          if (nb / SUNHIP_TILE == t) {
            xcode[f * NS + s] = nb % SUNHIP_TILE;
          } else {
            if (m < SUNHIP_XPAIRS) { xcode[f * NS + s] = SUNHIP_TILE + m; pair[(size_t)t * SUNHIP_XPAIRS + m] = f * NS + s; }
            m++;
          }
        }
      }
      npair[t] = std::min(m, SUNHIP_XPAIRS);
      g->max_pairs = std::max(g->max_pairs, m);
    }
    rc |= upload_vec(g, &g->d_xcode, xcode);
    rc |= upload_vec(g, &g->d_pair, pair);
    rc |= upload_vec(g, &g->d_npair, npair);
  }
  rc |= upload_vec(g, &g->d_eb, eb);
  rc |= upload_vec(g, &g->d_es, es);
  rc |= upload_vec(g, &g->d_nrm, nrm);
  rc |= upload_vec(g, &g->d_Ac, Ac);
  rc |= upload_vec(g, &g->d_cstart, cst);
  rc |= upload_vec(g, &g->d_cstartw, cstw);
  rc |= upload_vec(g, &g->d_e_c1, ec1);
  rc |= upload_vec(g, &g->d_e_c2, ec2);
  rc |= upload_vec(g, &g->d_e_s1, es1);
  rc |= upload_vec(g, &g->d_e_s2, es2);
  rc |= upload_vec(g, &g->d_e_nkc, enkc);
  rc |= upload_vec(g, &g->d_e_mark, emark);
  rc |= upload_vec(g, &g->d_e_comp, ecomp);
  rc |= upload_vec(g, &g->d_df, df);
  rc |= upload_vec(g, &g->d_dg, dg);
  rc |= upload_vec(g, &g->d_estart, est);
  rc |= upload_vec(g, &g->d_ford, ford);
  {
    std::vector<int> cs(g->slot_of_cell), esv(g->slot_of_edge);
    rc |= upload_vec(g, &g->d_cellslot, cs);
    rc |= upload_vec(g, &g->d_edgeslot, esv);
  }
  std::vector<int> zeros(NS, 0), ezeros(NES, 0), kr0(mfc * NS, 0);
  rc |= upload_vec(g, &g->d_ctop, zeros);
  rc |= upload_vec(g, &g->d_ctopold, zeros);
  rc |= upload_vec(g, &g->d_kr, kr0);
  {
    // depth-first dispatch order of the 4-tile blocks: the groups of each
    // XCD's contiguous run (xcd_block's split of nb4 blocks over 8 XCDs)
    // sorted by their deepest tile, deepest first, Morton order among equals
    const int nb4 = (g->Nt + 3) / 4, q = nb4 >> 3, r8 = nb4 & 7;
    std::vector<int> perm(nb4), gdep(nb4, 0);
    for (int b = 0; b < nb4; b++) {
      perm[b] = b;
      for (int t = 4 * b; t < std::min(4 * b + 4, g->Nt); t++) gdep[b] = std::max(gdep[b], g->tdep[t]);
    }
    for (int x = 0, start = 0; x < 8; x++) {
      const int len = q + (x < r8 ? 1 : 0);
      std::stable_sort(perm.begin() + start, perm.begin() + start + len,
                       [&](int a, int b) { return gdep[a] > gdep[b]; });
      start += len;
    }
    rc |= upload_vec(g, &g->d_b4perm, perm);
  }
  rc |= upload_vec(g, &g->d_e_nke, ezeros);
  rc |= upload_vec(g, &g->d_e_etop, ezeros);
  rc |= upload_vec(g, &g->d_e_etopold, ezeros);
  // reductions: one partial pair per block of the widest launch
  const int nblk = std::max(g->Nt, g->Net) + 64;
  g->part_n = 4 * (int64_t)nblk;
  rc |= dev_alloc(g, (void **)&g->d_part, sizeof(double) * (size_t)g->part_n);
  rc |= dev_alloc(g, (void **)&g->d_count, sizeof(unsigned) * 64);
  rc |= dev_alloc(g, (void **)&g->d_qs, sizeof(CGScal));
  rc |= dev_alloc(g, (void **)&g->d_hs, 2 * sizeof(CGScal));  // two: the free-surface iteration alternates them
  rc |= dev_alloc(g, (void **)&g->d_ms, 2 * sizeof(CGScal));
  rc |= dev_alloc(g, (void **)&g->d_mo, sizeof(CGScal));
  if (rc) { sunhip_grid_destroy(g); return rc; }
  HIPCHK(hipMemset(g->d_count, 0, sizeof(unsigned) * 64));
  HIPCHK(hipHostMalloc((void **)&g->h_sc, sizeof(CGScal), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&g->h_poll, 2 * sizeof(CGScal), hipHostMallocDefault));
  HIPCHK(hipEventCreateWithFlags(&g->evp[0], hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&g->evp[1], hipEventDisableTiming));
  HIPCHK(hipEventCreate(&g->ev0));
  HIPCHK(hipEventCreate(&g->ev1));
  HIPCHK(hipEventCreate(&g->ev2));
  *out = g;
  return 0;
}

int sunhip_grid_destroy(sunhip_grid *g) {
  if (!g) return 0;
  hipSetDevice(g->device);
  if (g->stream) hipStreamSynchronize(g->stream);
  if (g->xstream) hipStreamSynchronize(g->xstream);
  if (g->xdown) hipStreamSynchronize(g->xdown);
  if (g->batch) batch_forget_grid(g->batch, g);
  for (auto &kv : g->fields) hipFree(kv.second.d);
  for (auto &kv : g->fields32) hipFree(kv.second);
  void *ptrs[] = {g->d_toff, g->d_tdep, g->d_eoff, g->d_edep, g->d_nk, g->d_ctop, g->d_ctopold, g->d_comp,
                  g->d_nf, g->d_nbb, g->d_nbs, g->d_eb, g->d_es, g->d_nrm, g->d_kr, g->d_Ac, g->d_df,
                  g->d_xcode, g->d_pair, g->d_npair,
                  g->d_dg, g->d_e_c1, g->d_e_c2, g->d_e_s1, g->d_e_s2, g->d_e_nke, g->d_e_nkc,
                  g->d_e_etop, g->d_e_mark, g->d_e_comp, g->d_cstart, g->d_cstartw, g->d_estart,
                  g->d_cellslot, g->d_edgeslot, g->d_stage[0], g->d_stage[1], g->d_stage[2], g->d_stage[3], g->d_part, g->d_count, g->d_qs, g->d_hs,
                  g->d_class_tiles, g->d_ms, g->d_mo, g->d_D32, g->d_h32, g->d_uerr, g->d_dz, g->d_e_etopold,
                  g->d_ford, g->d_zdzt, g->d_cwin, g->d_fwin, g->d_cex, g->d_fex, g->d_tflag, g->d_b4perm};
  for (void *p : ptrs) if (p) hipFree(p);
  for (int i = 0; i < SUNHIP_PIN_SLOTS; i++) {
    if (g->h_pin[i]) hipHostFree(g->h_pin[i]);
    if (g->ev_pin[i]) hipEventDestroy(g->ev_pin[i]);
  }
  for (int i = 0; i < 4; i++)
    if (g->ev_stage[i]) hipEventDestroy(g->ev_stage[i]);
  for (int i = 0; i < 2; i++)
    for (int p = 0; p < SUNHIP_MAXPARTS; p++)
      if (g->dl[i].ev_part[p]) hipEventDestroy(g->dl[i].ev_part[p]);
  if (g->ev_hand) hipEventDestroy(g->ev_hand);
  if (g->xstream) hipStreamDestroy(g->xstream);
  if (g->xdown) hipStreamDestroy(g->xdown);
  if (g->h_sc) hipHostFree(g->h_sc);
  if (g->h_poll) hipHostFree(g->h_poll);
  for (hipEvent_t e : g->evp) if (e) hipEventDestroy(e);
  if (g->ev0) hipEventDestroy(g->ev0);
  if (g->ev1) hipEventDestroy(g->ev1);
  if (g->ev2) hipEventDestroy(g->ev2);
  if (g->own_stream) hipStreamDestroy(g->own_stream);
  delete g;
  return 0;
}

int64_t sunhip_grid_bytes(const sunhip_grid *g) { return g ? g->bytes : 0; }

int sunhip_grid_sizes(const sunhip_grid *g, int64_t *o) {
  o[0] = g->Nflat; o[1] = g->NflatW; o[2] = g->NflatE; o[3] = g->NT; o[4] = g->Nt; o[5] = g->NE;
  return 0;
}

int sunhip_get_extents(sunhip_grid *g, int *ctop, int *ctopold, int *etop, int *etopold) {
  HIPCHK(hipSetDevice(g->device));
  const int NS = g->NS, NES = g->NES;
  std::vector<int> ct(NS), cto(NS), et(NES), eto(NES);
  HIPCHK(hipMemcpyAsync(ct.data(), g->d_ctop, NS * sizeof(int), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipMemcpyAsync(cto.data(), g->d_ctopold, NS * sizeof(int), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipMemcpyAsync(et.data(), g->d_e_etop, NES * sizeof(int), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipMemcpyAsync(eto.data(), g->d_e_etopold, NES * sizeof(int), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  for (int c = 0; c < g->Nc; c++) {
    const int s = g->slot_of_cell[c];
    if (ctop) ctop[c] = ct[s];
    if (ctopold) ctopold[c] = cto[s];
  }
  for (int e = 0; e < g->Ne; e++) {
    const int s = g->slot_of_edge[e];
    if (etop) etop[e] = et[s];
    if (etopold) etopold[e] = eto[s];
  }
  return 0;
}

int sunhip_set_extents(sunhip_grid *g, const sunhip_extents *X) {
  if (int rc = validate_extents(g->Nc, g->Ne, g->Nk.data(), g->Nkc.data(), X)) return rc;
  HIPCHK(hipSetDevice(g->device));
  const int NS = g->NS, NES = g->NES;
  std::vector<int> ct(NS, 0), cto(NS, 0), nke(NES, 0), et(NES, 0);
  for (int s = 0; s < g->Nc; s++) {
    int c = g->cell_of_slot[s];
    ct[s] = X->ctop[c];
    cto[s] = X->ctopold ? X->ctopold[c] : X->ctop[c];
  }
  for (int s = 0; s < g->Ne; s++) {
    int e = g->edge_of_slot[s];
    nke[s] = X->Nke[e];
    et[s] = X->etop ? X->etop[e] : 0;
  }
  HIPCHK(hipMemcpyAsync(g->d_ctop, ct.data(), NS * sizeof(int), hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipMemcpyAsync(g->d_ctopold, cto.data(), NS * sizeof(int), hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipMemcpyAsync(g->d_e_nke, nke.data(), NES * sizeof(int), hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipMemcpyAsync(g->d_e_etop, et.data(), NES * sizeof(int), hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipMemcpyAsync(g->d_e_etopold, et.data(), NES * sizeof(int), hipMemcpyHostToDevice, g->stream));
  if (int rc = refresh_face_ranges(g)) return rc;
  HIPCHK(hipStreamSynchronize(g->stream));  // the host vectors go out of scope
  g->have_extents = true;
  return 0;
}

// Take the next staging buffer of direction dir (0 upload, 1 download; at
// least n values) for a transfer whose first step runs on stream s: s waits
// until the buffer's last work is over.  The copy streams and their events
// are made on first use.
static int take_stage(sunhip_grid *g, size_t n, int dir, hipStream_t s, int *out) {
  const int b = g->stage_next[dir];
  if (g->stage_n[b] < n) {
    if (g->stage_busy[b]) HIPCHK(hipEventSynchronize(g->ev_stage[b]));
    g->stage_busy[b] = false;
    if (g->d_stage[b]) { hipFree(g->d_stage[b]); g->bytes -= (int64_t)(g->stage_n[b] * sizeof(double)); }
    g->d_stage[b] = nullptr;
    g->stage_n[b] = 0;
    int rc = dev_alloc(g, (void **)&g->d_stage[b], n * sizeof(double));
    if (rc) return rc;
    g->stage_n[b] = n;
  }
  if (g->stage_busy[b]) HIPCHK(hipStreamWaitEvent(s, g->ev_stage[b], 0));
  g->stage_next[dir] = b ^ 1;
  *out = b;
  return 0;
}

static int copy_streams(sunhip_grid *g) {
  if (g->xstream) return 0;
  HIPCHK(hipStreamCreateWithFlags(&g->xstream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&g->xdown, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&g->ev_hand, hipEventDisableTiming));
  for (int i = 0; i < 4; i++) HIPCHK(hipEventCreateWithFlags(&g->ev_stage[i], hipEventDisableTiming));
  return 0;
}

// Work on staging buffer b, last enqueued on stream s, is over when s gets here.
static int release_stage(sunhip_grid *g, int b, hipStream_t s) {
  HIPCHK(hipEventRecord(g->ev_stage[b], s));
  g->stage_busy[b] = true;
  g->stage_uses[b]++;
  return 0;
}

// Stream b continues only after what is enqueued on stream a so far.
static int hand_over(sunhip_grid *g, hipStream_t a, hipStream_t b) {
  HIPCHK(hipEventRecord(g->ev_hand, a));
  HIPCHK(hipStreamWaitEvent(b, g->ev_hand, 0));
  return 0;
}

static size_t flat_len(const sunhip_grid *g, Kind k) {
  switch (k) {
    case CELL: return (size_t)g->Nflat;
    case CELLW: return (size_t)g->NflatW;
    case EDGE: return (size_t)g->NflatE;
    case CELL2: return (size_t)g->Nc;
    case EDGE2: return (size_t)g->Ne;
  }
  return 0;
}

int sunhip_pinned(sunhip_grid *g, int slot, int64_t n, double **out) {
  if (slot < 0 || slot >= SUNHIP_PIN_SLOTS || n < 0) {
    set_error("sunhip_pinned: slot %d / size %lld", slot, (long long)n);
    return -1;
  }
  HIPCHK(hipSetDevice(g->device));
  if (g->pin_pending[slot]) {
    HIPCHK(hipEventSynchronize(g->ev_pin[slot]));
    g->pin_pending[slot] = false;
  }
  if (!g->ev_pin[slot]) HIPCHK(hipEventCreateWithFlags(&g->ev_pin[slot], hipEventDisableTiming));
  if (g->h_pin_n[slot] < (size_t)n) {
    if (g->h_pin[slot]) HIPCHK(hipHostFree(g->h_pin[slot]));
    g->h_pin[slot] = nullptr;
    g->h_pin_n[slot] = 0;
    HIPCHK(hipHostMalloc((void **)&g->h_pin[slot], (size_t)std::max<int64_t>(n, 1) * sizeof(double),
                         hipHostMallocDefault));
    g->h_pin_n[slot] = (size_t)std::max<int64_t>(n, 1);
  }
  *out = g->h_pin[slot];
  return 0;
}

// The pinned slot flat lies in, or -1.
static int pinned_slot(const sunhip_grid *g, const double *flat, size_t n) {
  for (int i = 0; i < SUNHIP_PIN_SLOTS; i++)
    if (g->h_pin[i] && flat >= g->h_pin[i] && flat + n <= g->h_pin[i] + g->h_pin_n[i]) return i;
  return -1;
}

}  // extern "C"

static void scatter_field(sunhip_grid *g, const Field *f, const double *st) {
  switch (f->kind) {
    case CELL: case CELLW:
      k_scatter_cols<<<g->Nt, SUNHIP_TILE, 0, g->stream>>>(st, f->d, f->kind == CELL ? g->d_cstart : g->d_cstartw,
                                                           g->d_nk, f->kind == CELLW, g->d_toff, g->d_tdep);
      break;
    case EDGE:
      k_scatter_cols<<<g->Net, SUNHIP_TILE, 0, g->stream>>>(st, f->d, g->d_estart, g->d_e_nkc, 0, g->d_eoff,
                                                            g->d_edep);
      break;
    case CELL2:
      k_scatter_2d<<<(g->Nc + 255) / 256, 256, 0, g->stream>>>(st, f->d, g->d_cellslot, g->Nc);
      break;
    case EDGE2:
      k_scatter_2d<<<(g->Ne + 255) / 256, 256, 0, g->stream>>>(st, f->d, g->d_edgeslot, g->Ne);
      break;
  }
}

static void gather_field(sunhip_grid *g, const Field *f, double *st) {
  switch (f->kind) {
    case CELL: case CELLW:
      k_gather_cols<<<g->Nt, SUNHIP_TILE, 0, g->stream>>>(st, f->d, f->kind == CELL ? g->d_cstart : g->d_cstartw,
                                                          g->d_nk, f->kind == CELLW, g->d_toff);
      break;
    case EDGE:
      k_gather_cols<<<g->Net, SUNHIP_TILE, 0, g->stream>>>(st, f->d, g->d_estart, g->d_e_nkc, 0, g->d_eoff);
      break;
    case CELL2:
      k_gather_2d<<<(g->Nc + 255) / 256, 256, 0, g->stream>>>(st, f->d, g->d_cellslot, g->Nc);
      break;
    case EDGE2:
      k_gather_2d<<<(g->Ne + 255) / 256, 256, 0, g->stream>>>(st, f->d, g->d_edgeslot, g->Ne);
      break;
  }
}

// A download in parts not yet waited for in full.
static int parts_outstanding(sunhip_grid *g) {
  if (g->dl_count) {
    set_error("a download in parts has not been waited for in full");
    return -1;
  }
  return 0;
}

// The last copy of an upload from staging buffer b is enqueued on the copy
// stream: scatter b into f on the compute stream after it, and mark what the
// caller may reuse when.
static int finish_upload(sunhip_grid *g, const Field *f, int b, const double *flat, size_t n) {
  const int pin = pinned_slot(g, flat, n);
  if (pin >= 0) {  // the slot is reused only through sunhip_pinned, which waits for the copy
    HIPCHK(hipEventRecord(g->ev_pin[pin], g->xstream));
    g->pin_pending[pin] = true;
  }
  int rc = hand_over(g, g->xstream, g->stream);
  if (rc) return rc;
  scatter_field(g, f, g->d_stage[b]);
  HIPCHK(hipGetLastError());
  if ((rc = release_stage(g, b, g->stream))) return rc;
  if (pin < 0) HIPCHK(hipStreamSynchronize(g->xstream));  // the caller may reuse its buffer
  return 0;
}

extern "C" {

int sunhip_upload(sunhip_grid *g, const char *name, const double *flat) {
  HIPCHK(hipSetDevice(g->device));
  Field *f;
  int rc = field(g, name, &f);
  if (rc) return rc;
  if (g->up_filled) {
    set_error("sunhip_upload: an upload in parts is in progress");
    return -1;
  }
  const size_t n = flat_len(g, f->kind);
  int b;
  if ((rc = copy_streams(g)) || (rc = take_stage(g, n, 0, g->xstream, &b))) return rc;
  HIPCHK(hipMemcpyAsync(g->d_stage[b], flat, n * sizeof(double), hipMemcpyHostToDevice, g->xstream));
  return finish_upload(g, f, b, flat, n);
}

int sunhip_upload_part(sunhip_grid *g, const char *name, const double *flat, int64_t off, int64_t n, int last) {
  HIPCHK(hipSetDevice(g->device));
  Field *f;
  int rc = field(g, name, &f);
  if (rc) return rc;
  const int64_t len = (int64_t)flat_len(g, f->kind);
  if (n < 0 || off != g->up_filled || off + n > len || (last && off + n != len)) {
    set_error("sunhip_upload_part %s: part [%lld, %lld) after %lld of %lld values%s", name, (long long)off,
              (long long)(off + n), (long long)g->up_filled, (long long)len, last ? " (last)" : "");
    g->up_filled = 0;
    return -1;
  }
  if (off == 0 && ((rc = copy_streams(g)) || (rc = take_stage(g, (size_t)len, 0, g->xstream, &g->up_stage))))
    return rc;
  const int b = g->up_stage;
  if (n > 0)
    HIPCHK(hipMemcpyAsync(g->d_stage[b] + off, flat + off, (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                          g->xstream));
  g->up_filled = off + n;
  if (!last) return 0;
  g->up_filled = 0;
  return finish_upload(g, f, b, flat, (size_t)len);
}

int sunhip_download(sunhip_grid *g, const char *name, double *flat) {
  HIPCHK(hipSetDevice(g->device));
  Field *f;
  int rc = field(g, name, &f);
  if (rc) return rc;
  if ((rc = parts_outstanding(g))) return rc;
  const size_t n = flat_len(g, f->kind);
  int b;
  if ((rc = copy_streams(g)) || (rc = take_stage(g, n, 1, g->stream, &b))) return rc;
  gather_field(g, f, g->d_stage[b]);
  HIPCHK(hipGetLastError());
  if ((rc = hand_over(g, g->stream, g->xdown))) return rc;
  HIPCHK(hipMemcpyAsync(flat, g->d_stage[b], n * sizeof(double), hipMemcpyDeviceToHost, g->xdown));
  if ((rc = release_stage(g, b, g->xdown))) return rc;
  HIPCHK(hipStreamSynchronize(g->xdown));
  return 0;
}

int sunhip_download_parts(sunhip_grid *g, const char *name, double *flat, int nparts) {
  HIPCHK(hipSetDevice(g->device));
  Field *f;
  int rc = field(g, name, &f);
  if (rc) return rc;
  if (g->dl_count >= 2) {
    set_error("sunhip_download_parts: two downloads in parts are outstanding");
    return -1;
  }
  if (g->up_filled) {
    set_error("sunhip_download_parts: an upload in parts is in progress");
    return -1;
  }
  if (nparts < 1 || nparts > SUNHIP_MAXPARTS) {
    set_error("sunhip_download_parts: %d parts (1..%d)", nparts, SUNHIP_MAXPARTS);
    return -1;
  }
  const int64_t n = (int64_t)flat_len(g, f->kind);
  int b;
  if ((rc = copy_streams(g)) || (rc = take_stage(g, (size_t)n, 1, g->stream, &b))) return rc;
  gather_field(g, f, g->d_stage[b]);
  HIPCHK(hipGetLastError());
  if ((rc = hand_over(g, g->stream, g->xdown))) return rc;
  sunhip_grid::Download &D = g->dl[(g->dl_head + g->dl_count) & 1];
  for (int p = 0; p <= nparts; p++) D.part_bound[p] = n * p / nparts;
  for (int p = 0; p < nparts; p++) {
    if (!D.ev_part[p]) HIPCHK(hipEventCreateWithFlags(&D.ev_part[p], hipEventDisableTiming));
    const int64_t a = D.part_bound[p], e = D.part_bound[p + 1];
    if (e > a)
      HIPCHK(hipMemcpyAsync(flat + a, g->d_stage[b] + a, (size_t)(e - a) * sizeof(double), hipMemcpyDeviceToHost,
                            g->xdown));
    HIPCHK(hipEventRecord(D.ev_part[p], g->xdown));
  }
  if ((rc = release_stage(g, b, g->xdown))) return rc;
  D.stage = b;
  D.use = g->stage_uses[b];
  D.nparts = nparts;
  D.parts_done = 0;
  g->dl_count++;
  return 0;
}

int sunhip_download_wait(sunhip_grid *g, int64_t upto) {
  HIPCHK(hipSetDevice(g->device));
  if (!g->dl_count) return 0;
  sunhip_grid::Download &D = g->dl[g->dl_head];
  while (D.parts_done < D.nparts && D.part_bound[D.parts_done] < upto) {
    HIPCHK(hipEventSynchronize(D.ev_part[D.parts_done]));
    D.parts_done++;
  }
  if (D.parts_done == D.nparts) {
    // every copy out of its staging buffer is over: the next transfer through
    // that buffer need not wait on the device for the copy stream
    if (g->stage_uses[D.stage] == D.use) g->stage_busy[D.stage] = false;
    g->dl_head ^= 1;
    g->dl_count--;
  }
  return 0;
}

}  // extern "C"

__global__ void k_scale(double *__restrict__ d, const double *__restrict__ x, double a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = a * x[i];
}

extern "C" {

int sunhip_copy_field(sunhip_grid *g, const char *dst, const char *src) {
  HIPCHK(hipSetDevice(g->device));
  Field *a, *b;
  int rc = field(g, dst, &a);
  if (rc) return rc;
  if ((rc = field(g, src, &b))) return rc;
  if (a->n != b->n) {
    set_error("sunhip_copy_field: %s and %s differ in size", dst, src);
    return -1;
  }
  HIPCHK(hipMemcpyAsync(a->d, b->d, a->n * sizeof(double), hipMemcpyDeviceToDevice, g->stream));
  return 0;
}

int sunhip_field_scale(sunhip_grid *g, const char *dst, const char *src, double a) {
  HIPCHK(hipSetDevice(g->device));
  Field *d, *x;
  int rc = field(g, dst, &d);
  if (rc) return rc;
  if ((rc = field(g, src, &x))) return rc;
  if (d->n != x->n) {
    set_error("sunhip_field_scale: %s and %s differ in size", dst, src);
    return -1;
  }
  k_scale<<<(unsigned)((d->n + 255) / 256), 256, 0, g->stream>>>(d->d, x->d, a, (int64_t)d->n);
  HIPCHK(hipGetLastError());
  return 0;
}

int sunhip_sync(sunhip_grid *g) {
  HIPCHK(hipSetDevice(g->device));
  HIPCHK(hipStreamSynchronize(g->stream));
  if (g->xstream) HIPCHK(hipStreamSynchronize(g->xstream));
  if (g->xdown) HIPCHK(hipStreamSynchronize(g->xdown));
  return 0;
}

// Resolve recorded event pairs into per-kernel totals.
static void prof_flush(sunhip_grid *g) {
  if (g->prof_used == 0) return;
  hipStreamSynchronize(g->stream);
  for (int i = 0; i < g->prof_used; i++) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, g->prof_pool[2 * i], g->prof_pool[2 * i + 1]) == hipSuccess) {
      g->prof_ms[g->prof_kind[i]] += ms;
      g->prof_n[g->prof_kind[i]] += 1;
    }
  }
  g->prof_used = 0;
}

int sunhip_profile(sunhip_grid *g, int on) {
  HIPCHK(hipSetDevice(g->device));
  prof_flush(g);
  g->prof = on;
  for (int i = 0; i < SUNHIP_NPROF; i++) { g->prof_ms[i] = 0; g->prof_n[i] = 0; }
  if (on && g->prof_pool.empty()) {
    g->prof_pool.resize(2 * 4096);
    g->prof_kind.resize(4096);
    for (auto &e : g->prof_pool) HIPCHK(hipEventCreate(&e));
  }
  return 0;
}

// Totals per kernel id since sunhip_profile(g, 1); returns the number of ids.
int sunhip_kernel_times(sunhip_grid *g, double *ms, int *n) {
  hipSetDevice(g->device);
  prof_flush(g);
  for (int i = 0; i < PK_COUNT; i++) { ms[i] = g->prof_ms[i]; n[i] = g->prof_n[i]; }
  return PK_COUNT;
}

}  // extern "C"

namespace sunhip {
int poll_chunk(sunhip_grid *g, const CGScal *d, int &slot, bool &pending, bool *done) {
  HIPCHK(hipMemcpyAsync(&g->h_poll[slot], d, sizeof(CGScal), hipMemcpyDeviceToHost, g->stream));
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

int prof_begin(sunhip_grid *g, int kind) {
  if (!g->prof) return -1;
  if (g->prof_used * 2 + 2 > (int)g->prof_pool.size()) prof_flush(g);
  const int slot = g->prof_used++;
  g->prof_kind[slot] = kind;
  hipEventRecord(g->prof_pool[2 * slot], g->stream);
  return slot;
}
void prof_end(sunhip_grid *g, int slot) {
  if (slot >= 0) hipEventRecord(g->prof_pool[2 * slot + 1], g->stream);
}
}  // namespace sunhip

// Plain streaming copy, one 16-byte element per thread: the calibration
// kernel for HBM byte counters and the measure of attainable copy bandwidth.
__global__ __launch_bounds__(256) void k_copy_probe(const double2 *__restrict__ a, double2 *__restrict__ b,
                                                    int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

extern "C" int sunhip_pcie_probe(sunhip_grid *g, int64_t nbytes, int reps, double *h2d_ms, double *d2h_ms) {
  HIPCHK(hipSetDevice(g->device));
  if (nbytes <= 0 || reps <= 0 || nbytes > ((int64_t)1 << 36)) { set_error("sunhip_pcie_probe: bad size"); return -1; }
  void *h = nullptr, *d = nullptr;
  HIPCHK(hipHostMalloc(&h, (size_t)nbytes, hipHostMallocDefault));
  HIPCHK(hipMalloc(&d, (size_t)nbytes));
  memset(h, 0, (size_t)nbytes);
  double *out[2] = {h2d_ms, d2h_ms};
  for (int dir = 0; dir < 2; dir++) {
    void *dst = dir == 0 ? d : h;
    const void *src = dir == 0 ? h : d;
    const hipMemcpyKind kind = dir == 0 ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    HIPCHK(hipMemcpyAsync(dst, src, (size_t)nbytes, kind, g->stream));  // warm
    HIPCHK(hipEventRecord(g->ev0, g->stream));
    for (int r = 0; r < reps; r++) HIPCHK(hipMemcpyAsync(dst, src, (size_t)nbytes, kind, g->stream));
    HIPCHK(hipEventRecord(g->ev1, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    *out[dir] = ms / reps;
  }
  HIPCHK(hipFree(d));
  HIPCHK(hipHostFree(h));
  return 0;
}

extern "C" int sunhip_copy_probe(sunhip_grid *g, int64_t nbytes, int reps, double *avg_ms) {
  HIPCHK(hipSetDevice(g->device));
  const int64_t n = nbytes / (int64_t)sizeof(double2);
  if (n <= 0 || reps <= 0 || n > ((int64_t)1 << 40)) { set_error("sunhip_copy_probe: bad size"); return -1; }
  double2 *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, n * sizeof(double2)));
  HIPCHK(hipMalloc(&b, n * sizeof(double2)));
  HIPCHK(hipMemsetAsync(a, 0, n * sizeof(double2), g->stream));
  const unsigned blocks = (unsigned)((n + 255) / 256);
  k_copy_probe<<<blocks, 256, 0, g->stream>>>(a, b, n);  // warm
  HIPCHK(hipEventRecord(g->ev0, g->stream));
  for (int r = 0; r < reps; r++) k_copy_probe<<<blocks, 256, 0, g->stream>>>(a, b, n);
  HIPCHK(hipEventRecord(g->ev1, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  *avg_ms = ms / reps;
  HIPCHK(hipFree(a));
  HIPCHK(hipFree(b));
  return 0;
}

// sunhip_tvd.hip -- TVD scalar transport on gfx950.
//
// Reference: HorizontalFaceScalars main/tvd.c:36-120, Psi :130-151,
// GetApAm :161-204, UpdateScalars main/scalars.c:35-325, TriSolve
// main/util.c:185-203, UpWind/Max/Min main/util.c:212-246.
//
// Three kernels per scalar step:
//   sums  (one lane per column): inflow flux sumQ and inflow-weighted jump
//         sumQC per cell-layer (pass 1 of HorizontalFaceScalars)
//   faces (one lane per edge column): limited face values SfHp/SfHm
//   column(one lane per column): the theta-implicit vertical tridiagonal with
//         the explicit horizontal TVD flux divergence, assembled level by
//         level and eliminated on the fly (GetApAm is evaluated per level
//         from the column's own values), then back-substituted.
// Operation order follows the reference so results agree bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>

#include "sunhip_tvdtile.h"
#include "sunhip_solvers.h"

using namespace sunhip;


template <int NF, int GRP>
__global__ __launch_bounds__(256) void k_tvd_sums(Geo G, TvdArgs A) {
  tvd_sums_tile<NF, GRP>(G, A, xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

__global__ __launch_bounds__(256) void k_tvd_faces(Geo G, TvdArgs A) {
  tvd_faces_tile(G, A, xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

template <int NF, int GRP, bool SRC = false, bool PRE = false, int KR = 0>
__global__ __launch_bounds__(256) void k_update_column(Geo G, ColArgs A) {
  update_column_tile<NF, GRP, SRC, PRE, KR>(G, A, xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6),
                                            threadIdx.x & 63);
}

namespace sunhip {
int tvd_args(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal, TvdArgs *T, ColArgs *Cc) {
  if (!g->have_extents) { set_error("scalar step: call sunhip_set_extents first"); return -1; }
  Field *fs, *so, *wn, *w2, *dzz, *dzo, *ktv, *u, *ut, *dzf, *hp, *hm, *act, *cn, *tb, *tc, *td, *sq, *sqc;
  int rc = field(g, scal, &fs) | field(g, "stmp", &so) | field(g, "wnew", &wn) | field(g, "wtmp2", &w2) |
           field(g, "dzz", &dzz) | field(g, "dzzold", &dzo) | field(g, "kappa_tv", &ktv) | field(g, "u", &u) |
           field(g, "utmp2", &ut) | field(g, "dzf", &dzf) | field(g, "SfHp", &hp) | field(g, "SfHm", &hm) |
           field(g, "active", &act) | field(g, "Cn", &cn) | field(g, "us_b", &tb) | field(g, "us_c", &tc) |
           field(g, "us_d", &td) | field(g, "gradSx", &sq) | field(g, "gradSy", &sqc);
  if (rc) return rc;
  Field *bs = nullptr;
  auto it = g->fields.find("bscal");
  if (it != g->fields.end()) bs = &it->second;
  TvdArgs A{};
  A.s = fs->d; A.u = u->d; A.utmp2 = ut->d; A.dzf = dzf->d; A.sumQ = sq->d; A.sumQC = sqc->d;
  A.oQ = sq->d; A.oQC = sqc->d; A.SfHp = hp->d; A.SfHm = hm->d; A.bscal = bs ? bs->d : nullptr;
  A.theta = sp->theta; A.dt = sp->dt; A.TVD = sp->TVD;
  *T = A;
  ColArgs C{};
  C.s_old = so->d; C.s = fs->d; C.wnew = wn->d; C.wtmp2 = w2->d; C.dzz = dzz->d; C.dzzold = dzo->d;
  C.kappa_tv = ktv->d; C.u = u->d; C.utmp2 = ut->d; C.dzf = dzf->d; C.SfHp = hp->d; C.SfHm = hm->d;
  C.active = act->d; C.Cn = cn->d; C.tb = tb->d; C.tc = tc->d; C.td = td->d;
  C.theta = sp->theta; C.dt = sp->dt; C.kappa = sp->kappa; C.TVD = sp->TVD; C.step = sp->step;
  C.alpha_top = sp->alpha_top; C.alpha_bot = sp->alpha_bot;
  *Cc = C;
  return 0;
}
}  // namespace sunhip

extern "C" {

int sunhip_face_scalars(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal) {
  HIPCHK(hipSetDevice(g->device));
  TvdArgs A;
  ColArgs C;
  int rc = tvd_args(g, sp, scal, &A, &C);
  if (rc) return rc;
  return face_scalars_launch(g, A);
}

}  // extern "C"

namespace sunhip {

int face_scalars_launch(sunhip_grid *g, const TvdArgs &A) {
  const Geo G = g->geo();
  int ps = prof_begin(g, PK_TSUMS);
  with_nf(g->mf, [&](auto nf) {
    constexpr int NF = decltype(nf)::value;
    const unsigned nb = (unsigned)((g->Nt + 3) / 4);
    if (g->tune_sumsgrp == 4) k_tvd_sums<NF, 4><<<nb, 256, 0, g->stream>>>(G, A);
    else if (g->tune_sumsgrp == 2) k_tvd_sums<NF, 2><<<nb, 256, 0, g->stream>>>(G, A);
    else k_tvd_sums<NF, 1><<<nb, 256, 0, g->stream>>>(G, A);
    return 0;
  });
  prof_end(g, ps);
  ps = prof_begin(g, PK_TFACES);
  k_tvd_faces<<<(g->Net + 3) / 4, 256, 0, g->stream>>>(G, A);
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace sunhip

extern "C" {

int sunhip_update_scalars(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal) {
  return sunhip_update_scalars_src(g, sp, scal, nullptr, nullptr, nullptr, nullptr, nullptr);
}

}  // extern "C"

// The scalar step in phases: 1 the faces pass (TVD >= 1; with the column's
// flux terms when there are no source terms), 2 the column update (after
// phase 1 when TVD >= 1 and there are no source terms), 3 both.
static int update_scalars(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal, const char *src1,
                          const char *src2, const char *ftop, const char *fbot, const char *bcell, int phases) {
  HIPCHK(hipSetDevice(g->device));
  TvdArgs T;
  ColArgs A;
  int rc = tvd_args(g, sp, scal, &T, &A);
  if (rc) return rc;
  // the source fields: cell fields for src1/src2, 2-D cell fields for the fluxes
  const char *names[5] = {src1, src2, ftop, fbot, bcell};
  const double **dst[5] = {&A.src1, &A.src2, &A.Ftop, &A.Fbot, &A.bcell};
  for (int i = 0; i < 5; i++) {
    if (!names[i]) continue;
    const Kind want = (i < 2 || i == 4) ? CELL : CELL2;
    if (kind_of(names[i]) != want) {
      set_error("sunhip_update_scalars_src: %s must name a %s field", names[i], want == CELL ? "cell" : "2-D cell");
      return -1;
    }
    Field *f;
    if ((rc = field(g, names[i], &f))) return rc;
    *dst[i] = f->d;
  }
  // without source terms the faces pass also forms the column's flux terms
  const bool srcs = A.src1 || A.src2 || A.Ftop || A.Fbot;
  const bool pre = sp->TVD && !srcs;
  if (pre) {
    Field *p1, *p2;
    if ((rc = field(g, "tvdP1", &p1) | field(g, "tvdP2", &p2))) return rc;
    T.P1 = p1->d; T.P2 = p2->d;
    A.P1 = p1->d; A.P2 = p2->d;
  }
  if ((phases & 1) && sp->TVD && (rc = face_scalars_launch(g, T))) return rc;
  if (!(phases & 2)) return 0;
  if (sp->TVD) {
    // a column reads old values of its own cells only, all before it writes
    // them: the update runs in place
    A.s_old = A.s;
  } else {
    // first-order upwind reads the neighbours' old values: keep them in stmp
    HIPCHK(hipMemcpyAsync(const_cast<double *>(A.s_old), A.s, g->fields.at(scal).n * sizeof(double),
                          hipMemcpyDeviceToDevice, g->stream));
  }
  const int ps = prof_begin(g, PK_TCOLUMN);
  with_nf(g->mf, [&](auto nf) {
    constexpr int NF = decltype(nf)::value;
    const unsigned nb = (unsigned)((g->Nt + 3) / 4);
    if (srcs)
      k_update_column<NF, 1, true><<<nb, 256, 0, g->stream>>>(g->geo(), A);  // sources: a variant of its own
    else if (pre && g->tune_tvdgrp == 4) k_update_column<NF, 4, false, true><<<nb, 256, 0, g->stream>>>(g->geo(), A);
    else if (pre && g->tune_tvdgrp == 2 && g->tune_tvdhead == 16)
      k_update_column<NF, 2, false, true, 16><<<nb, 256, 0, g->stream>>>(g->geo(), A);
    else if (pre && g->tune_tvdgrp == 2 && g->tune_tvdhead == 8)
      k_update_column<NF, 2, false, true, 8><<<nb, 256, 0, g->stream>>>(g->geo(), A);
    else if (pre && g->tune_tvdgrp == 2) k_update_column<NF, 2, false, true><<<nb, 256, 0, g->stream>>>(g->geo(), A);
    else if (pre) k_update_column<NF, 1, false, true><<<nb, 256, 0, g->stream>>>(g->geo(), A);
    else if (g->tune_tvdgrp == 4) k_update_column<NF, 4><<<nb, 256, 0, g->stream>>>(g->geo(), A);
    else if (g->tune_tvdgrp == 2) k_update_column<NF, 2><<<nb, 256, 0, g->stream>>>(g->geo(), A);
    else k_update_column<NF, 1><<<nb, 256, 0, g->stream>>>(g->geo(), A);
    return 0;
  });
  prof_end(g, ps);
  HIPCHK(hipGetLastError());
  if (phases == 3) HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

extern "C" {

int sunhip_update_scalars_src(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal, const char *src1,
                              const char *src2, const char *ftop, const char *fbot, const char *bcell) {
  return update_scalars(g, sp, scal, src1, src2, ftop, fbot, bcell, 3);
}

int sunhip_update_scalars_faces(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal) {
  if (!sp || !sp->TVD) {
    set_error("sunhip_update_scalars_faces: TVD schemes only (TVD >= 1)");
    return -1;
  }
  return update_scalars(g, sp, scal, nullptr, nullptr, nullptr, nullptr, nullptr, 1);
}

int sunhip_update_scalars_column(sunhip_grid *g, const sunhip_scalar_params *sp, const char *scal) {
  if (!sp || !sp->TVD) {
    set_error("sunhip_update_scalars_column: TVD schemes only (TVD >= 1)");
    return -1;
  }
  return update_scalars(g, sp, scal, nullptr, nullptr, nullptr, nullptr, nullptr, 2);
}

}  // extern "C"

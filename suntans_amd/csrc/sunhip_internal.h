// sunhip_internal.h -- device context and layout shared by the sunhip kernels.
//
// Layout.  Cells are ordered along a Morton curve of their Voronoi points and
// cut into tiles of 64 consecutive cells, one wavefront each, one lane per
// column.  A tile stores level k of its 64 columns contiguously; its depth is
// the deepest column (+1 so w-type fields with Nk+1 values fit), so the value
// of cell slot s = 64*t + lane at level k lives at toff[t] + 64*k + lane.
// Edges get the same treatment with their own tiles.  Every horizontal gather
// (neighbour column, face edge) is then a 64-lane access at one level, and
// the vertical coupling of a column stays in the lane's registers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "sunhip.h"

#define SUNHIP_MAXF 8   // faces per cell supported (triangles to octagons)
#define SUNHIP_TILE 64
#define SUNHIP_XPAIRS 128  // cross-tile faces per tile the exchange kernel handles

namespace sunhip {

void set_error(const char *fmt, ...);

// Face-count classes the column kernels are instantiated for: a grid of
// maxfaces mf runs the smallest class >= mf (face slots past a cell's own
// count are skipped).  with_nf calls fn(std::integral_constant<int, NF>).
inline int nf_class(int mf) { return mf <= 3 ? 3 : mf <= 4 ? 4 : mf <= 6 ? 6 : 8; }
template <typename F>
inline auto with_nf(int mf, F &&fn) {
  switch (nf_class(mf)) {
    case 3: return fn(std::integral_constant<int, 3>{});
    case 4: return fn(std::integral_constant<int, 4>{});
    case 6: return fn(std::integral_constant<int, 6>{});
    default: return fn(std::integral_constant<int, 8>{});
  }
}

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::sunhip::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                          __FILE__, __LINE__);                                    \
      return -3;                                                                  \
    }                                                                             \
  } while (0)

// Kinds of named fields.
enum Kind { CELL = 1, CELLW = 2, EDGE = 3, CELL2 = 4, EDGE2 = 5 };

struct Field {
  Kind kind;
  double *d = nullptr;
  size_t n = 0;  // device doubles
};

// Device-resident CG state (one per solver kind).
struct CGScal {
  double alpha, alpha0, eps, eps0, nu, mu, muinv, pz, tol, resid;
  int iter;       // loop iterations completed
  int nloop;      // the reference's loop counter n at exit
  int maxit, done, converged, precond, resnorm;
  int pend;       // op still to apply to the last kernel's block sums (free-surface iteration), 0 none
};

// Per-call geometry the kernels read, gathered into one argument block.
struct Geo {
  int Nt, NS, Net, NES, mf;
  const int *toff;     // [Nt+1] element offset of each cell tile
  const int *tdep;     // [Nt] tile depth (levels stored)
  const int *eoff;     // [Net+1]
  const int *edep;     // [Net]
  const int *nk;       // [NS] Nk per slot (0 = padding)
  const int *ctop;     // [NS]
  const int *ctopold;  // [NS]
  const int *comp;     // [NS] 1 computational (celldist[0..1)), 2 type-3 boundary, 0 none
  const int *nf;       // [NS]
  const int *nbb;      // [MAXF*NS] neighbour element base, -1 none
  const int *nbs;      // [MAXF*NS] neighbour slot, -1 none
  const int *eb;       // [MAXF*NS] face edge element base
  const int *es;       // [MAXF*NS] face edge slot
  const int *nrm;      // [MAXF*NS] normal sign
  const int *kr;       // [MAXF*NS] kmin | kmax << 16 for the horizontal coupling
  const double *Ac;    // [NS]
  // edges, by edge slot
  const int *e_c1, *e_c2;  // element base of grad[0], grad[1] (-1 none)
  const int *e_s1, *e_s2;  // slots of grad[0], grad[1] (-1 none)
  const int *e_nke, *e_nkc, *e_etop, *e_mark, *e_comp;  // e_comp: 1 for edgedist[0..1)
  const double *df, *dg;
  // neighbour exchange of the operator kernel: per (face, slot) where the
  // neighbour's value is found in the tile's exchange row (0..63: the lane of
  // a neighbour in the same tile; 64 + i: cross-tile pair i; -1 none), and
  // per tile the pairs' (face, slot) indices and count
  const int *xcode, *pair, *npair;
  // dispatch order of the 4-tile blocks of the column kernels (b4perm[v] =
  // the block group of virtual block v): within each XCD's run of groups the
  // deepest go first, so the launch does not end on a few deep columns;
  // null: identity
  const int *b4perm;
};

}  // namespace sunhip

struct sunhip_batch;

struct sunhip_grid {
  int device = 0;
  hipStream_t stream = nullptr;      // where this grid's work is enqueued
  hipStream_t own_stream = nullptr;  // the grid's own stream (stream differs while attached to a batch)
  sunhip_batch *batch = nullptr;     // batch this grid is attached to, if any
  int Nc = 0, Ne = 0, mf = 0, Nkmax = 0;
  int Nt = 0, Net = 0, NS = 0, NES = 0;
  int64_t NT = 0, NE = 0;           // padded element counts (cells, edges)
  int64_t Nflat = 0, NflatW = 0, NflatE = 0;
  int maxdepth = 0;                 // deepest tile
  std::vector<int> slot_of_cell, cell_of_slot, slot_of_edge, edge_of_slot;
  std::vector<int> toff, tdep, eoff, edep;
  std::vector<int> Nk, Nkc;         // host copies
  std::vector<int64_t> cstart, estart;  // flat offsets per cell / edge
  std::vector<int> face, neigh, normal, nfaces, grad, mark, cellp, edgep, celldist, edgedist;
  // device metadata
  int *d_toff = nullptr, *d_tdep = nullptr, *d_eoff = nullptr, *d_edep = nullptr;
  int *d_nk = nullptr, *d_ctop = nullptr, *d_ctopold = nullptr, *d_comp = nullptr, *d_nf = nullptr;
  int *d_nbb = nullptr, *d_nbs = nullptr, *d_eb = nullptr, *d_es = nullptr, *d_nrm = nullptr, *d_kr = nullptr;
  int *d_xcode = nullptr, *d_pair = nullptr, *d_npair = nullptr;  // neighbour exchange (Geo)
  int *d_b4perm = nullptr;          // depth-first dispatch order of 4-tile blocks (Geo)
  int tune_order = 1;               // 1: use it
  int tune_tvdgrp = 2;              // levels per load group of the scalar column update (1, 2, 4)
  int tune_plds = 12;               // levels below the preconditioner's register head kept in LDS (0, 12)
  int tune_tvdhead = 8;             // rows of the scalar column update kept in registers (0, 8, 16)
  int tune_hblk = 1024;             // most blocks of the free-surface iteration kernels
  int tune_sumsgrp = 4;             // levels per load group of the TVD inflow sums (1, 2, 4)
  int *d_uerr = nullptr;             // UPredictor's exit conditions
  double *d_dz = nullptr;            // [Nkmax] reference layer thicknesses (vertical grid updates)
  // z-level windows of the pressure solve (ZReg): dz[k] (Nkmax + 16
  // entries, zero past Nkmax), the windows, exceptions and tile flags of the
  // current solve
  double *d_zdzt = nullptr;
  double zden = 0;                   // the common dz[k] + dz[k-1], 0 when the layers differ
  bool have_zdz = false;
  int *d_cwin = nullptr, *d_fwin = nullptr, *d_tflag = nullptr;
  double *d_cex = nullptr, *d_fex = nullptr;
  int tune_zreg = 1;
  int *d_e_etopold = nullptr;        // [NES] etop before the last UpdateDZ
  int *d_ford = nullptr;             // [MAXF*NS] each column's faces in edgep order of their edges
  int n_type4 = 0;                   // no-slip (type-4) edges
  int type2_ok = 1;                  // every type-2 edge has its cell on the grad[2j] side
  bool have_def1 = false;            // "def1" holds the face distances (sunhip_set_face_distances)
  int max_pairs = 0;                 // most cross-tile faces of any tile
  double *d_Ac = nullptr, *d_df = nullptr, *d_dg = nullptr;
  int *d_e_c1 = nullptr, *d_e_c2 = nullptr, *d_e_s1 = nullptr, *d_e_s2 = nullptr;
  int *d_e_nke = nullptr, *d_e_nkc = nullptr, *d_e_etop = nullptr, *d_e_mark = nullptr, *d_e_comp = nullptr;
  int64_t *d_cstart = nullptr, *d_cstartw = nullptr, *d_estart = nullptr;  // per slot flat start, -1 padding
  int *d_cellslot = nullptr, *d_edgeslot = nullptr;  // host index -> slot
  // Transfers: device staging buffers taken in turn, two for uploads (0, 1)
  // and two for downloads (2, 3), the copies of each direction on a stream of
  // their own, so that one field's copy overlaps the scatter or gather kernel
  // of the one before it and uploads overlap downloads.
  double *d_stage[4] = {};
  size_t stage_n[4] = {};
  int stage_next[2] = {0, 2};                    // the buffer the next upload / download takes
  hipEvent_t ev_stage[4] = {};                   // the last work on buffer b is over
  bool stage_busy[4] = {};                       // ... recorded and maybe not reached yet
  unsigned stage_uses[4] = {};                   // transfers through buffer b so far
  hipStream_t xstream = nullptr;                 // host-to-device copies (default priority)
  hipStream_t xdown = nullptr;                   // device-to-host copies (default priority)
  hipEvent_t ev_hand = nullptr;                  // hand-over between a copy stream and stream
  double *h_pin[SUNHIP_PIN_SLOTS] = {};      // pinned host slots (sunhip_pinned)
  size_t h_pin_n[SUNHIP_PIN_SLOTS] = {};
  hipEvent_t ev_pin[SUNHIP_PIN_SLOTS] = {};  // last upload read from the slot
  bool pin_pending[SUNHIP_PIN_SLOTS] = {};
  // transfers in parts (sunhip_upload_part, sunhip_download_parts)
  int64_t up_filled = 0;                     // values of the upload in progress copied so far
  int up_stage = 0;                          // its staging buffer
  // downloads in parts not yet waited for in full, oldest first (at most 2)
  struct Download {
    int nparts = 0, parts_done = 0;            // parts issued, parts waited for
    int64_t part_bound[SUNHIP_MAXPARTS + 1] = {};
    hipEvent_t ev_part[SUNHIP_MAXPARTS] = {};  // part p is on the host
    int stage = 0;                             // the staging buffer its copies read
    unsigned use = 0;                          // ... and which use of it
  } dl[2];
  int dl_head = 0, dl_count = 0;
  bool have_extents = false;
  // reductions
  double *d_part = nullptr;         // block partials
  int64_t part_n = 0;               // its capacity in doubles
  unsigned *d_count = nullptr;      // last-block tickets (one per reduction site)
  sunhip::CGScal *d_qs = nullptr, *d_hs = nullptr;
  sunhip::CGScal *d_ms = nullptr, *d_mo = nullptr;  // mixed form: inner FP32 CG, outer residual
  sunhip::CGScal *h_sc = nullptr;   // pinned mirror
  sunhip::CGScal *h_poll = nullptr; // two pinned slots of pipelined convergence polling
  hipEvent_t evp[2] = {nullptr, nullptr};
  std::map<std::string, sunhip::Field> fields;
  std::map<std::string, float *> fields32;  // FP32 copies for the mixed form (cell layout)
  float *d_D32 = nullptr;
  float *d_h32 = nullptr;  // mixed free surface: FP32 coefficients and CG vectors
  int64_t bytes = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;

  int *d_class_tiles = nullptr;  // reserved
  // tuning: levels per load group of the column kernels (1, 2, 4) and
  // reduction finish (0 last block, 1 finisher kernel); -1 = default
  int tune_unroll = -1, tune_reduce = -1;
  // tiles per apply block (4, 8, 16), non-temporal apply stores, rz from the
  // forward sweep (see qprec_tile)
  int tune_tiles = 4, tune_nts = 1, tune_rzid = 1;
  int tune_minw = 0;  // register budget of the operator kernel: waves per SIMD (0, 3, 4)
  int tune_blk = 0;   // operator kernel with block-staged neighbour values
  int tune_punroll = 8;  // levels per load group of the preconditioner (1, 2, 4, 8)
  int tune_xchg = -1;    // operator kernel with the neighbour exchange inside each wave: 0 off, 1 on,
                         // 2 on with a 4-waves-per-SIMD register budget, 3 on with 2 levels per group,
                         // -1 (default) 2 from 4096 tiles up, off below (short walks favour the plain kernel)
  int tune_phead = -1;   // top levels per column the preconditioner keeps in registers (0, 8, 16, 24, 32;
                         // FP32 solves use at most 16); -1 (default) 24 from 1024 tiles up in FP64, 16
                         // from 2048 tiles up in FP32, 0 below
  int tune_band = 0;     // operator kernel over level bands of this many levels (0: whole columns)

  // per-kernel timing (sunhip_profile): an event pair around every launch of
  // the solver kernels, resolved by sunhip_kernel_times
  int prof = 0;
  std::vector<hipEvent_t> prof_pool;
  std::vector<int> prof_kind;   // kernel id per recorded pair
  int prof_used = 0;
  double prof_ms[SUNHIP_NPROF] = {0};
  int prof_n[SUNHIP_NPROF] = {0};

  sunhip::Geo geo() const;
};

namespace sunhip {
// Bracket one launch with events when profiling is on.  Returns the pair
// index or -1.
int prof_begin(sunhip_grid *g, int kind);
// Pipelined convergence polling of a CG loop: after each chunk of iterations
// the scalars at d are copied to one of two pinned slots; the host then
// waits for the previous chunk's copy while the chunk just enqueued runs, so
// the GPU does not idle during the poll.  Returns 0 with *done set, or an
// error code.
int poll_chunk(sunhip_grid *g, const CGScal *d, int &slot, bool &pending, bool *done);
void prof_end(sunhip_grid *g, int slot);
enum { PK_QAPPLY = 0, PK_QPREC = 1, PK_HAPPLY = 2, PK_HPREC = 3, PK_MAPPLY = 4, PK_MPREC = 5, PK_RESID = 6,
       PK_OTHER = 7, PK_TSUMS = 8, PK_TFACES = 9, PK_TCOLUMN = 10, PK_QSOURCE = 11, PK_CORRECT = 12,
       PK_CONTINUITY = 13, PK_WSOURCE = 14, PK_WCOLUMN = 15, PK_COUNT = 16 };
}  // namespace sunhip

namespace sunhip {
int field(sunhip_grid *g, const char *name, Field **out);  // create on demand
int field32(sunhip_grid *g, const char *name, float **out);
Kind kind_of(const char *name);
int dev_alloc(sunhip_grid *g, void **p, size_t bytes);
// Per-face level ranges (kmin | kmax << 16) from the device extents.
int refresh_face_ranges(sunhip_grid *g);
void batch_forget_grid(sunhip_batch *b, sunhip_grid *g);
}  // namespace sunhip

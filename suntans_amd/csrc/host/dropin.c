/*
 * dropin.c -- the reference-signature entry points (include/suntans_hip.h)
 * over the thin device ABI (include/sunhip.h).  Plain C.
 *
 * Each call packs the jagged REAL** rows it needs into one flat buffer, hands
 * them to the device, runs the kernels and writes back the rows the
 * reference writes.  The device copy of a grid's topology is built on the
 * first call with that gridT and reused; the moving vertical extents
 * (ctop, ctopold, Nke, etop) and all field data are refreshed on every call,
 * since the caller may have changed them between calls.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <pthread.h>
#include <unistd.h>
#ifdef __SSE2__
#include <emmintrin.h>
#endif

#include "sunhip.h"
#include "suntans_hip.h"

#define MAXCTX 16

/* What identifies a grid between calls without reading it whole: the
 * addresses of its topology arrays, its sizes and a sampled hash of the
 * arrays' contents (taken when the context is built and compared on every
 * call).  A grid edited in place in a way the samples miss needs
 * sunhip_dropin_reset. */
#define NIDENT 17
typedef struct {
  const void *p[NIDENT];
  int Nc, Ne, maxfaces, Nkmax;
} ident_t;

typedef struct {
  const gridT *key;
  ident_t id;
  uint64_t fp;
  int Nc, Ne;
  sunhip_grid *dev;
  double *flat;      /* the pinned slot in use (sunhip_pinned) */
  size_t flat_n;
  int slot, dslot;   /* the slot the next upload (0, 1) and download (2, 3) takes */
  size_t *cst, *wst, *est;  /* flat offsets of cell (Nk), w-type (Nk+1) and edge (Nkc) rows */
  int *cell_rank, *edge_rank;  /* position of cell i in cellp, of edge j in edgep */
  struct { double *flat; int parts; } dl[2];  /* downloads in flight, oldest first */
  int dl_head, dl_n;
  int *x_last;       /* the extents last sent (ctop, ctopold | Nke, etop), NULL: none */
} ctx_t;

static ctx_t g_ctx[MAXCTX];
static int g_nctx = 0;
static int g_device = 0;
static int g_mixed = 0;           /* 1: CGSolveQ/CGSolve use the mixed-precision form */
static double g_inner_tol = 1e-4;  /* its inner FP32 tolerance */
static struct { int valid; sunhip_cg_stats st; } g_last[2];

static void fatal(const char *where, const char *msg) {
  fprintf(stderr, "suntans_hip: %s: %s\n", where, msg);
  exit(EXIT_FAILURE);
}

static void check(int rc, const char *where) {
  if (rc) fatal(where, sunhip_error());
}

static void ident(const gridT *g, ident_t *id) {
  const void *p[NIDENT] = {g->Nk, g->Nkc, g->nfaces, g->neigh, g->face, g->normal, g->grad, g->mark, g->cellp,
                           g->celldist, g->edgep, g->edgedist, g->Ac, g->df, g->dg, g->xv, g->yv};
  memset(id, 0, sizeof *id);
  memcpy(id->p, p, sizeof p);
  id->Nc = g->Nc; id->Ne = g->Ne; id->maxfaces = g->maxfaces; id->Nkmax = g->Nkmax;
}

/* FNV-1a over at most 64 evenly spaced elements of an array */
static uint64_t fnv_sample(uint64_t h, const void *p, size_t n, size_t size) {
  const unsigned char *b = (const unsigned char *)p;
  const size_t m = n < 64 ? n : 64;
  for (size_t j = 0; j < m; j++) {
    const size_t i = m > 1 ? j * (n - 1) / (m - 1) : 0;
    for (size_t k = 0; k < size; k++) h = (h ^ b[i * size + k]) * 1099511628211ull;
  }
  return h;
}
static uint64_t fingerprint(const gridT *g) {
  uint64_t h = 1469598103934665603ull;
  const size_t nc = (size_t)g->Nc, ne = (size_t)g->Ne, nf = (size_t)g->maxfaces * nc;
  h = fnv_sample(h, g->Nk, nc, sizeof(int));
  h = fnv_sample(h, g->Nkc, ne, sizeof(int));
  h = fnv_sample(h, g->nfaces, nc, sizeof(int));
  h = fnv_sample(h, g->neigh, nf, sizeof(int));
  h = fnv_sample(h, g->face, nf, sizeof(int));
  h = fnv_sample(h, g->normal, nf, sizeof(int));
  h = fnv_sample(h, g->grad, 2 * ne, sizeof(int));
  h = fnv_sample(h, g->mark, ne, sizeof(int));
  h = fnv_sample(h, g->cellp, nc, sizeof(int));
  h = fnv_sample(h, g->edgep, ne, sizeof(int));
  h = fnv_sample(h, g->celldist, 3, sizeof(int));
  h = fnv_sample(h, g->edgedist, 6, sizeof(int));
  h = fnv_sample(h, g->Ac, nc, sizeof(double));
  h = fnv_sample(h, g->df, ne, sizeof(double));
  h = fnv_sample(h, g->dg, ne, sizeof(double));
  h = fnv_sample(h, g->xv, nc, sizeof(double));
  h = fnv_sample(h, g->yv, nc, sizeof(double));
  return h;
}

static void ctx_free(ctx_t *c) {
  sunhip_grid_destroy(c->dev);  /* frees the pinned slots too */
  free(c->cst);
  free(c->wst);
  free(c->est);
  free(c->cell_rank);
  free(c->edge_rank);
  free(c->x_last);
}

static int *ranks(const int *order, int n) {
  int *r = (int *)malloc(sizeof(int) * ((size_t)n + 1));
  if (!r) return NULL;
  for (int i = 0; i < n; i++) r[i] = -1;
  for (int ip = 0; ip < n; ip++)
    if (order[ip] >= 0 && order[ip] < n) r[order[ip]] = ip;
  return r;
}

static size_t *offsets(const int *len, int n, int extra) {
  size_t *o = (size_t *)malloc(sizeof(size_t) * ((size_t)n + 1));
  if (!o) return NULL;
  o[0] = 0;
  for (int i = 0; i < n; i++) o[i + 1] = o[i] + (size_t)(len[i] + extra);
  return o;
}

static ctx_t *get_ctx(gridT *g, const char *where) {
  ident_t id;
  ident(g, &id);
  for (int i = 0; i < g_nctx; i++)
    if (g_ctx[i].key == g && !memcmp(&g_ctx[i].id, &id, sizeof id) && g_ctx[i].fp == fingerprint(g))
      return &g_ctx[i];
  /* same address, different grid: drop the stale copy */
  for (int i = 0; i < g_nctx; i++)
    if (g_ctx[i].key == g) {
      ctx_free(&g_ctx[i]);
      g_ctx[i] = g_ctx[g_nctx - 1];
      g_nctx--;
      break;
    }
  if (g_nctx == MAXCTX) {
    /* evict the oldest */
    ctx_free(&g_ctx[0]);
    memmove(g_ctx, g_ctx + 1, sizeof(ctx_t) * (MAXCTX - 1));
    g_nctx--;
  }
  ctx_t *c = &g_ctx[g_nctx];
  memset(c, 0, sizeof *c);
  c->fp = fingerprint(g);
  c->id = id;
  sunhip_topology T;
  T.Nc = g->Nc; T.Ne = g->Ne; T.maxfaces = g->maxfaces; T.Nkmax = g->Nkmax;
  T.nfaces = g->nfaces; T.neigh = g->neigh; T.face = g->face; T.normal = g->normal;
  T.grad = g->grad; T.mark = g->mark; T.Nk = g->Nk; T.Nkc = g->Nkc;
  T.cellp = g->cellp; T.celldist = g->celldist; T.edgep = g->edgep; T.edgedist = g->edgedist;
  T.Ac = g->Ac; T.df = g->df; T.dg = g->dg; T.xv = g->xv; T.yv = g->yv;
  check(sunhip_grid_create(&c->dev, &T, g_device), where);
  if (g->dz) check(sunhip_set_dz(c->dev, g->dz), where);  /* the z-level windows of the pressure solve */
  int64_t sz[6];
  sunhip_grid_sizes(c->dev, sz);
  size_t n = (size_t)sz[1];
  if ((size_t)sz[2] > n) n = (size_t)sz[2];
  if ((size_t)g->maxfaces * g->Nc > n) n = (size_t)g->maxfaces * g->Nc;
  c->flat_n = n;
  c->cst = offsets(g->Nk, g->Nc, 0);
  c->wst = offsets(g->Nk, g->Nc, 1);
  c->est = offsets(g->Nkc, g->Ne, 0);
  c->cell_rank = ranks(g->cellp, g->Nc);
  c->edge_rank = ranks(g->edgep, g->Ne);
  if (!c->cst || !c->wst || !c->est || !c->cell_rank || !c->edge_rank) fatal(where, "out of host memory");
  c->key = g; c->Nc = g->Nc; c->Ne = g->Ne;
  g_nctx++;
  return c;
}

/* The next pinned slot for an upload, 0 or 1 (it waits until the device has
 * read what was last uploaded from it); it becomes c->flat. */
static double *next_buf(ctx_t *c, const char *where) {
  check(sunhip_pinned(c->dev, c->slot, (int64_t)c->flat_n, &c->flat), where);
  c->slot ^= 1;
  return c->flat;
}
/* The next pinned slot for a download, 2 or 3, so that downloads in flight
 * leave the upload slots free. */
static double *next_dbuf(ctx_t *c, const char *where) {
  double *p;
  check(sunhip_pinned(c->dev, 2 + c->dslot, (int64_t)c->flat_n, &p), where);
  c->dslot ^= 1;
  return p;
}

/* Send the moving vertical extents unless they equal the ones last sent. */
static void extents(ctx_t *c, gridT *g, const char *where) {
  const size_t nc = (size_t)g->Nc * sizeof(int), ne = (size_t)g->Ne * sizeof(int);
  if (c->x_last && !memcmp(c->x_last, g->ctop, nc) && !memcmp(c->x_last + g->Nc, g->ctopold, nc) &&
      !memcmp(c->x_last + 2 * (size_t)g->Nc, g->Nke, ne) && !memcmp(c->x_last + 2 * (size_t)g->Nc + g->Ne, g->etop, ne))
    return;
  sunhip_extents X = {g->ctop, g->ctopold, g->Nke, g->etop};
  check(sunhip_set_extents(c->dev, &X), where);
  if (!c->x_last) c->x_last = (int *)malloc(2 * (nc + ne));
  if (!c->x_last) return;
  memcpy(c->x_last, g->ctop, nc);
  memcpy(c->x_last + g->Nc, g->ctopold, nc);
  memcpy(c->x_last + 2 * (size_t)g->Nc, g->Nke, ne);
  memcpy(c->x_last + 2 * (size_t)g->Nc + g->Ne, g->etop, ne);
}
/* The device's extents were changed on the device: send them next time. */
static void extents_changed(ctx_t *c) {
  free(c->x_last);
  c->x_last = NULL;
}

/* ---- jagged rows <-> flat, split over threads -------------------------- */
enum { ROWS_PACK, ROWS_UNPACK, ROWS_UNPACK_FROM, ROWS_UNPACK_UPTO, ROWS_UNPACK_ZERO };
typedef struct {
  int mode;
  double *flat;
  REAL **rows;
  const size_t *start;     /* row i at flat + start[i], start[i+1] - start[i] values */
  const int *rank;         /* unpack: row i's position in cellp / edgep (NULL: every row) */
  int rlo, rhi, rlo2, rhi2;  /* ... only rows whose position is in [rlo, rhi) or [rlo2, rhi2) */
  const int *lvl;          /* FROM: first level written; UPTO: levels written */
  int lo, hi;              /* rows lo..hi-1 */
} rows_job;

/* Copy n doubles with streaming stores: the destination is written whole and
 * not read again soon, so its lines need not be fetched first (on x86 a plain
 * copy reads every destination line before writing it).  The caller ends a
 * run of these with copy_fence. */
static void copy_stream(double *d, const double *s, size_t n) {
#ifdef __SSE2__
  size_t i = 0;
  if (((uintptr_t)d & 15) && n) { d[0] = s[0]; i = 1; }
  for (; i + 2 <= n; i += 2) _mm_stream_pd(d + i, _mm_loadu_pd(s + i));
  if (i < n) d[i] = s[i];
#else
  memcpy(d, s, n * sizeof(double));
#endif
}
static void copy_fence(void) {
#ifdef __SSE2__
  _mm_sfence();
#endif
}

static void rows_run(const rows_job *J, int lo, int hi) {
  for (int i = lo; i < hi; i++) {
    const size_t o = J->start[i], n = J->start[i + 1] - o;
    if (J->rank) {
      const int r = J->rank[i];
      if ((r < J->rlo || r >= J->rhi) && (r < J->rlo2 || r >= J->rhi2)) {
        if (J->mode == ROWS_UNPACK_ZERO) memset(J->rows[i], 0, n * sizeof(double));
        continue;
      }
    }
    switch (J->mode) {
      case ROWS_PACK: copy_stream(J->flat + o, J->rows[i], n); break;
      case ROWS_UNPACK: case ROWS_UNPACK_ZERO: copy_stream(J->rows[i], J->flat + o, n); break;
      case ROWS_UNPACK_FROM: {
        const size_t k0 = (size_t)J->lvl[i];
        if (k0 < n) copy_stream(J->rows[i] + k0, J->flat + o + k0, n - k0);
        break;
      }
      case ROWS_UNPACK_UPTO: copy_stream(J->rows[i], J->flat + o, (size_t)J->lvl[i]); break;
    }
  }
  copy_fence();  /* this thread's stores are visible before its share is reported done */
}

/* Threads for packing: SUNHIP_DROPIN_THREADS, else the online cores, at most
 * 16.  They are started on first use and wait for work between jobs. */
static int n_threads(void) {
  static int n = 0;
  if (!n) {
    const char *e = getenv("SUNHIP_DROPIN_THREADS");
    long v = e ? atol(e) : sysconf(_SC_NPROCESSORS_ONLN);
    if (v < 1) v = 1;
    if (!e && v > 16) v = 16;
    if (v > 64) v = 64;
    n = (int)v;
  }
  return n;
}

static struct {
  pthread_mutex_t mu;
  pthread_cond_t go, done;
  int started, T;          /* threads running (besides the caller), shares per job */
  unsigned gen;            /* job counter */
  int pending;             /* shares not yet finished */
  const rows_job *job;
} g_pool = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER, PTHREAD_COND_INITIALIZER, 0, 1, 0, 0, NULL};

static void share(const rows_job *J, int t, int T, int *lo, int *hi) {
  const int n = J->hi - J->lo;
  *lo = J->lo + (int)((int64_t)n * t / T);
  *hi = J->lo + (int)((int64_t)n * (t + 1) / T);
}

static void *pool_worker(void *arg) {
  const int t = (int)(intptr_t)arg;
  unsigned seen = 0;
  for (;;) {
    pthread_mutex_lock(&g_pool.mu);
    while (g_pool.gen == seen) pthread_cond_wait(&g_pool.go, &g_pool.mu);
    seen = g_pool.gen;
    const rows_job *J = g_pool.job;
    const int T = g_pool.T;
    pthread_mutex_unlock(&g_pool.mu);
    if (t < T) {
      int lo, hi;
      share(J, t, T, &lo, &hi);
      rows_run(J, lo, hi);
    }
    pthread_mutex_lock(&g_pool.mu);
    if (--g_pool.pending == 0) pthread_cond_signal(&g_pool.done);
    pthread_mutex_unlock(&g_pool.mu);
  }
  return NULL;
}

/* Run J over the pool: share 0 on the calling thread, share t on worker t. */
static void rows_par(const rows_job *J, size_t nvalues) {
  int T = n_threads();
  if (nvalues < ((size_t)1 << 16) || J->hi - J->lo < 2 * T) T = 1;
  if (T > 1 && g_pool.started < n_threads() - 1) {
    while (g_pool.started < n_threads() - 1) {
      pthread_t th;
      if (pthread_create(&th, NULL, pool_worker, (void *)(intptr_t)(g_pool.started + 1))) break;
      pthread_detach(th);
      g_pool.started++;
    }
  }
  if (T > g_pool.started + 1) T = g_pool.started + 1;
  if (T <= 1) {
    rows_run(J, J->lo, J->hi);
    return;
  }
  pthread_mutex_lock(&g_pool.mu);
  g_pool.job = J;
  g_pool.T = T;
  g_pool.pending = g_pool.started;  /* every worker answers each job */
  g_pool.gen++;
  pthread_cond_broadcast(&g_pool.go);
  pthread_mutex_unlock(&g_pool.mu);
  int lo, hi;
  share(J, 0, T, &lo, &hi);
  rows_run(J, lo, hi);
  pthread_mutex_lock(&g_pool.mu);
  while (g_pool.pending) pthread_cond_wait(&g_pool.done, &g_pool.mu);
  pthread_mutex_unlock(&g_pool.mu);
}

/* Parts a transfer of nvalues is split into: about 16 MB each, 1..8. */
static int n_parts(size_t nvalues) {
  size_t k = nvalues * sizeof(double) / ((size_t)16 << 20);
  return k < 1 ? 1 : k > 8 ? 8 : (int)k;
}

/* Pack rows 0..n-1 into the next pinned slot part by part, each part's copy
 * starting while the next is packed, and upload them into field name. */
static void upload_rows(ctx_t *c, const char *name, REAL **rows, const size_t *start, int n, const char *where) {
  double *flat = next_buf(c, where);
  const int K = n_parts(start[n]);
  for (int p = 0; p < K; p++) {
    rows_job J = {ROWS_PACK, flat, rows, start, NULL, 0, 0, 0, 0, NULL, (int)((int64_t)n * p / K),
                  (int)((int64_t)n * (p + 1) / K)};
    rows_par(&J, start[J.hi] - start[J.lo]);
    check(sunhip_upload_part(c->dev, name, flat, (int64_t)start[J.lo], (int64_t)(start[J.hi] - start[J.lo]),
                             p == K - 1),
          where);
  }
}
static void upload_cells(ctx_t *c, gridT *g, const char *name, REAL **rows, int extra, const char *where) {
  upload_rows(c, name, rows, extra ? c->wst : c->cst, g->Nc, where);
}
static void upload_edges(ctx_t *c, gridT *g, const char *name, REAL **rows, const char *where) {
  upload_rows(c, name, rows, c->est, g->Ne, where);
}

/* Start the download of field name into the next pinned slot in parts.
 * Two may be in flight: the device gathers and copies the second while the
 * host unpacks the first.  unpack_parts takes them in order. */
static void download(ctx_t *c, const char *name, const char *where) {
  if (c->dl_n == 2) fatal(where, "a third download in flight");
  double *flat = next_dbuf(c, where);
  if (c->dl_n && c->dl[c->dl_head].flat == flat) fatal(where, "a download into a slot not yet unpacked");
  const int parts = n_parts(c->flat_n);
  check(sunhip_download_parts(c->dev, name, flat, parts), where);
  const int i = (c->dl_head + c->dl_n) & 1;
  c->dl[i].flat = flat;
  c->dl[i].parts = parts;
  c->dl_n++;
}
/* The whole of field name, in a pinned slot (valid until the slot's next use). */
static double *download_all(ctx_t *c, const char *name, const char *where) {
  if (c->dl_n) fatal(where, "a whole download with others in flight");
  download(c, name, where);
  check(sunhip_download_wait(c->dev, INT64_MAX), where);
  c->dl_n = 0;
  return c->dl[c->dl_head].flat;
}

/* Write J's rows back from the oldest download in flight, part by part. */
static void unpack_parts(ctx_t *c, rows_job J) {
  if (!c->dl_n) fatal("unpack", "no download in flight");
  J.flat = c->dl[c->dl_head].flat;
  const int lo = J.lo, n = J.hi - J.lo, K = c->dl[c->dl_head].parts;
  for (int p = 0; p < K; p++) {
    J.lo = lo + (int)((int64_t)n * p / K);
    J.hi = lo + (int)((int64_t)n * (p + 1) / K);
    check(sunhip_download_wait(c->dev, (int64_t)J.start[J.hi]), "unpack");
    rows_par(&J, J.start[J.hi] - J.start[J.lo]);
  }
  check(sunhip_download_wait(c->dev, INT64_MAX), "unpack");
  c->dl_head ^= 1;
  c->dl_n--;
}

/* What a download writes back: the rows of J from field name. */
typedef struct {
  const char *name;
  rows_job J;
} fetch_t;

/* Download and write back F[0..n-1] with two in flight. */
static void fetch(ctx_t *c, const fetch_t *F, int n, const char *where) {
  int issued = 0;
  for (int i = 0; i < n; i++) {
    for (; issued < n && issued < i + 2; issued++) download(c, F[issued].name, where);
    unpack_parts(c, F[i].J);
  }
}

/* cells of cellp[celldist[lo]..celldist[hi]) */
static rows_job cells_job(ctx_t *c, const gridT *g, REAL **rows, int lo, int hi) {
  rows_job J = {ROWS_UNPACK, NULL, rows, c->cst, c->cell_rank, g->celldist[lo], g->celldist[hi], 0, 0, NULL, 0,
                g->Nc};
  return J;
}
/* ... and zero the other rows */
static rows_job cells_zero_job(ctx_t *c, const gridT *g, REAL **rows, int lo, int hi) {
  rows_job J = cells_job(c, g, rows, lo, hi);
  J.mode = ROWS_UNPACK_ZERO;
  return J;
}
/* rows of w-type cells (Nk+1 values) of cellp[celldist[lo]..celldist[hi]) */
static rows_job cells_w_job(ctx_t *c, const gridT *g, REAL **rows, int lo, int hi) {
  rows_job J = cells_job(c, g, rows, lo, hi);
  J.start = c->wst;
  return J;
}
/* levels [ctop, Nk) of every cell */
static rows_job cells_wet_job(ctx_t *c, const gridT *g, REAL **rows) {
  rows_job J = {ROWS_UNPACK_FROM, NULL, rows, c->cst, NULL, 0, 0, 0, 0, g->ctop, 0, g->Nc};
  return J;
}
/* every edge row */
static rows_job edges_job(ctx_t *c, const gridT *g, REAL **rows) {
  rows_job J = {ROWS_UNPACK, NULL, rows, c->est, NULL, 0, 0, 0, 0, NULL, 0, g->Ne};
  return J;
}
/* levels [0, Nke) of every edge row */
static rows_job edges_wet_job(ctx_t *c, const gridT *g, REAL **rows) {
  rows_job J = {ROWS_UNPACK_UPTO, NULL, rows, c->est, NULL, 0, 0, 0, 0, g->Nke, 0, g->Ne};
  return J;
}
/* levels [0, Nke) of the edges of edgep[edgedist[0..1)) and edgep[edgedist[2..3)) */
static rows_job face_values_job(ctx_t *c, const gridT *g, REAL **rows) {
  rows_job J = {ROWS_UNPACK_UPTO, NULL, rows, c->est, c->edge_rank, g->edgedist[0], g->edgedist[1],
                g->edgedist[2], g->edgedist[3], g->Nke, 0, g->Ne};
  return J;
}
static void fetch1(ctx_t *c, const char *name, rows_job J, const char *where) {
  fetch_t F = {name, J};
  fetch(c, &F, 1, where);
}

/* qprecond=1: 1 when c's rows are the residual rows phys->stmp3 CGSolveQ
 * uses as r (main/phys.c:2419), as in Solve's call (main/phys.c:1280). */
static int c_aliases_r(REAL **c, const gridT *g, const physT *p, const char *W) {
  int same = 0, n = 0;
  for (int ip = g->celldist[0]; ip < g->celldist[1]; ip++, n++) {
    const int i = g->cellp[ip];
    same += c[i] == p->stmp3[i];
  }
  if (same != 0 && same != n) fatal(W, "c shares some rows, but not all, with the residual rows phys->stmp3");
  return same != 0;
}

void CGSolveQ(REAL **q, REAL **src, REAL **c, gridT *grid, physT *phys, propT *prop, int myproc, int numprocs,
              MPI_Comm comm) {
  static const char *W = "CGSolveQ";
  (void)myproc; (void)comm;
  if (numprocs != 1) fatal(W, "only numprocs=1 (the whole grid on one GPU) is supported");
  if (prop->qprecond != 0 && prop->qprecond != 1 && prop->qprecond != 2) fatal(W, "qprecond must be 0, 1 or 2");
  if (prop->qprecond == 1 && !c) fatal(W, "qprecond=1 needs the conditioning rows c");
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  upload_cells(C, grid, "q", q, 0, W);
  upload_cells(C, grid, "src", src, 0, W);
  upload_cells(C, grid, "dzz", grid->dzz, 0, W);
  upload_edges(C, grid, "dzf", grid->dzf, W);
  check(sunhip_upload(C->dev, "D", phys->D), W);
  sunhip_cg_params P = {prop->qmaxiters, prop->qprecond, (int)prop->resnorm, prop->qepsilon, 0, 0};
  if (prop->qprecond == 1) {
    /* a boundary neighbour's c is read as the caller left it */
    upload_cells(C, grid, "qcond", c, 0, W);
    P.qc_alias = c_aliases_r(c, grid, phys, W);
  }
  if (g_mixed) {
    if (prop->qprecond != 2) fatal(W, "the mixed-precision form needs qprecond=2");
    int inner = 0;
    check(sunhip_q_solve_mixed(C->dev, &P, g_inner_tol, 50, &g_last[0].st, &inner), W);
    g_last[0].st.niters = inner;
  } else {
    check(sunhip_q_solve(C->dev, &P, &g_last[0].st), W);
  }
  g_last[0].valid = 1;
  fetch1(C, "q", cells_job(C, grid, q, 0, 1), W);
}

void CGSolve(gridT *grid, physT *phys, propT *prop, int myproc, int numprocs, MPI_Comm comm) {
  static const char *W = "CGSolve";
  (void)myproc; (void)comm;
  if (numprocs != 1) fatal(W, "only numprocs=1 (the whole grid on one GPU) is supported");
  if (prop->hprecond != 0 && prop->hprecond != 1) fatal(W, "hprecond must be 0 or 1");
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  check(sunhip_upload(C->dev, "h", phys->h), W);
  check(sunhip_upload(C->dev, "htmp", phys->htmp), W);
  check(sunhip_upload(C->dev, "D", phys->D), W);
  sunhip_cg_params P = {prop->maxiters, prop->hprecond, (int)prop->resnorm, prop->epsilon, 0, 0};
  if (g_mixed) {
    int inner = 0;
    check(sunhip_h_solve_mixed(C->dev, &P, prop->grav, prop->theta, prop->dt, g_inner_tol, 50, &g_last[1].st,
                               &inner), W);
    g_last[1].st.niters = inner;
  } else {
    check(sunhip_h_solve(C->dev, &P, prop->grav, prop->theta, prop->dt, &g_last[1].st), W);
  }
  g_last[1].valid = 1;
  const double *h = download_all(C, "h", W);
  for (int ip = grid->celldist[0]; ip < grid->celldist[1]; ip++) {
    const int i = grid->cellp[ip];
    phys->h[i] = h[i];
  }
}

int sunhip_q_coefficients(sunhip_grid *g);

void OperatorQ(REAL **coef, REAL **x, REAL **y, REAL **c, gridT *grid, physT *phys, propT *prop) {
  static const char *W = "OperatorQ";
  (void)c; (void)prop;
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  upload_cells(C, grid, "coef", coef, 0, W);
  upload_edges(C, grid, "dzf", grid->dzf, W);
  check(sunhip_upload(C->dev, "D", phys->D), W);
  upload_cells(C, grid, "opx", x, 0, W);
  check(sunhip_q_operator(C->dev, "opx", "opy"), W);
  fetch1(C, "opy", cells_job(C, grid, y, 0, 1), W);
}

void Preconditioner(REAL **x, REAL **xc, REAL **coef, gridT *grid, physT *phys, propT *prop) {
  static const char *W = "Preconditioner";
  (void)phys; (void)prop;
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  upload_cells(C, grid, "coef", coef, 0, W);
  upload_cells(C, grid, "opx", x, 0, W);
  check(sunhip_q_precondition(C->dev, "opx", "opy"), W);
  fetch1(C, "opy", cells_job(C, grid, xc, 0, 1), W);
}

int sunhip_h_operator_coef(sunhip_grid *g, const double *hfcoef_flat, const char *x, const char *y);

void OperatorH(REAL *x, REAL *y, REAL *coef, REAL *fcoef, gridT *grid, physT *phys, propT *prop) {
  static const char *W = "OperatorH";
  (void)phys; (void)prop;
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  check(sunhip_upload(C->dev, "hcoef", coef), W);
  check(sunhip_upload(C->dev, "hx", x), W);
  check(sunhip_h_operator_coef(C->dev, fcoef, "hx", "hy"), W);
  const double *hy = download_all(C, "hy", W);
  for (int ip = grid->celldist[0]; ip < grid->celldist[1]; ip++) {
    const int i = grid->cellp[ip];
    y[i] = hy[i];
  }
}

/* open-boundary rows (edges of edgedist[2..5)) into edge field name */
static void upload_brows(ctx_t *C, gridT *g, REAL **bscal, const char *name, const char *W) {
  const int lo = g->edgedist[2], hi = g->edgedist[5];
  if (!bscal || hi <= lo) return;
  double *b = next_buf(C, W);
  memset(b, 0, C->est[g->Ne] * sizeof(double));
  for (int jp = lo; jp < hi; jp++) {
    const int j = g->edgep[jp];
    memcpy(b + C->est[j], bscal[jp - lo], (size_t)g->Nke[j] * sizeof(double));
  }
  check(sunhip_upload(C->dev, name, b), W);
}
static void upload_bscal(ctx_t *C, gridT *g, REAL **bscal, const char *W) { upload_brows(C, g, bscal, "bscal", W); }

static void scalar_inputs(ctx_t *C, gridT *grid, physT *phys, REAL **scal, REAL **bscal, const char *W) {
  extents(C, grid, W);
  upload_cells(C, grid, "s", scal, 0, W);
  upload_edges(C, grid, "u", phys->u, W);
  upload_edges(C, grid, "utmp2", phys->utmp2, W);
  upload_edges(C, grid, "dzf", grid->dzf, W);
  upload_bscal(C, grid, bscal, W);
}

void HorizontalFaceScalars(gridT *grid, physT *phys, propT *prop, REAL **scal, REAL **boundary_scal, int TVD,
                           MPI_Comm comm, int myproc) {
  static const char *W = "HorizontalFaceScalars";
  (void)comm; (void)myproc;
  ctx_t *C = get_ctx(grid, W);
  scalar_inputs(C, grid, phys, scal, boundary_scal, W);
  sunhip_scalar_params S = {prop->theta, prop->dt, 0.0, TVD, prop->n, 0, 0};
  check(sunhip_face_scalars(C->dev, &S, "s"), W);
  const fetch_t F[] = {{"SfHp", face_values_job(C, grid, phys->SfHp)},
                       {"SfHm", face_values_job(C, grid, phys->SfHm)},
                       {"gradSx", cells_job(C, grid, phys->gradSx, 0, 2)},
                       {"gradSy", cells_job(C, grid, phys->gradSy, 0, 2)}};
  fetch(C, F, 4, W);
}

void UpdateScalars(gridT *grid, physT *phys, propT *prop, REAL **wnew, REAL **scal, REAL **boundary_scal,
                   REAL **Cn, REAL kappa, REAL kappaH, REAL **kappa_tv, REAL theta, REAL **src1, REAL **src2,
                   REAL *Ftop, REAL *Fbot, int alpha_top, int alpha_bot, MPI_Comm comm, int myproc,
                   int checkflag, int TVDscheme) {
  static const char *W = "UpdateScalars";
  (void)kappaH; (void)comm; (void)myproc; (void)checkflag;
  prop->TVD = TVDscheme;
  prop->horiTVD = 1;
  prop->vertTVD = 1;
  ctx_t *C = get_ctx(grid, W);
  scalar_inputs(C, grid, phys, scal, boundary_scal, W);
  sunhip_scalar_params S = {theta, prop->dt, kappa, TVDscheme, prop->n, alpha_top, alpha_bot};
  /* TVD without source terms: the faces pass runs as soon as its inputs are
   * on the device, and its face values come back while the column update's
   * inputs go over (the link carries both directions at once) */
  const int split = TVDscheme && !src1 && !src2 && !Ftop && !Fbot;
  if (split) {
    check(sunhip_update_scalars_faces(C->dev, &S, "s"), W);
    download(C, "SfHp", W);
    download(C, "SfHm", W);
  }
  upload_cells(C, grid, "wnew", wnew, 1, W);
  upload_cells(C, grid, "wtmp2", phys->wtmp2, 1, W);
  upload_cells(C, grid, "dzz", grid->dzz, 0, W);
  upload_cells(C, grid, "dzzold", grid->dzzold, 0, W);
  upload_cells(C, grid, "kappa_tv", kappa_tv, 0, W);
  double *act = next_buf(C, W);
  for (int i = 0; i < grid->Nc; i++) act[i] = phys->active[i] ? 1.0 : 0.0;
  check(sunhip_upload(C->dev, "active", act), W);
  /* source terms: cell rows (Nk values each; src2 may be w-type rows, of
   * which the step reads Nk) and the two per-column fluxes */
  if (src1) upload_cells(C, grid, "src1", src1, 0, W);
  if (src2) upload_cells(C, grid, "src2", src2, 0, W);
  if (Ftop) check(sunhip_upload(C->dev, "Ftop", Ftop), W);
  if (Fbot) check(sunhip_upload(C->dev, "Fbot", Fbot), W);
  /* first-order upwind with open boundaries: the rows the reference builds
   * in stmp2 (main/scalars.c:76-96), zero in the computational columns, then
   * each open-boundary edge's values in its cell's rows, in edge order */
  const int nob = grid->edgedist[5] - grid->edgedist[2];
  const int bcell = !TVDscheme && boundary_scal && nob > 0;
  if (bcell) {
    double *b = next_buf(C, W);
    memset(b, 0, C->cst[grid->Nc] * sizeof(double));
    for (int jp = grid->edgedist[2]; jp < grid->edgedist[5]; jp++) {
      const int ib = grid->grad[2 * grid->edgep[jp]];
      for (int k = grid->ctop[ib]; k < grid->Nk[ib]; k++) b[C->cst[ib] + k] = boundary_scal[jp - grid->edgedist[2]][k];
    }
    check(sunhip_upload(C->dev, "bcell", b), W);
  }
  /* Cn: the computational columns, zero in the others */
  const fetch_t F[] = {{"s", cells_job(C, grid, scal, 0, 1)},
                       {"Cn", cells_zero_job(C, grid, Cn, 0, 1)},
                       {"SfHp", face_values_job(C, grid, phys->SfHp)},
                       {"SfHm", face_values_job(C, grid, phys->SfHm)}};
  if (split) {
    /* the column update is enqueued before the face values are unpacked */
    check(sunhip_update_scalars_column(C->dev, &S, "s"), W);
    unpack_parts(C, F[2].J);
    unpack_parts(C, F[3].J);
    fetch(C, F, 2, W);
    return;
  }
  check(sunhip_update_scalars_src(C->dev, &S, "s", src1 ? "src1" : NULL, src2 ? "src2" : NULL,
                                  Ftop ? "Ftop" : NULL, Fbot ? "Fbot" : NULL, bcell ? "bcell" : NULL), W);
  fetch(C, F, TVDscheme ? 4 : 2, W);
}


void ComputeQSource(REAL **src, gridT *grid, physT *phys, propT *prop, int myproc, int numprocs) {
  static const char *W = "ComputeQSource";
  (void)myproc;
  if (numprocs != 1) fatal(W, "only numprocs=1 (the whole grid on one GPU) is supported");
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  upload_edges(C, grid, "u", phys->u, W);
  upload_edges(C, grid, "utmp2", phys->utmp2, W);
  upload_edges(C, grid, "dzf", grid->dzf, W);
  upload_cells(C, grid, "w", phys->w, 1, W);
  upload_cells(C, grid, "wtmp2", phys->wtmp2, 1, W);
  check(sunhip_compute_qsource(C->dev, prop->theta, prop->dt, "src"), W);
  fetch1(C, "src", cells_wet_job(C, grid, src), W);
  memcpy(phys->D, download_all(C, "D", W), (size_t)grid->Ne * sizeof(double));
}

void Corrector(REAL **qc, gridT *grid, physT *phys, propT *prop, int myproc, int numprocs, MPI_Comm comm) {
  static const char *W = "Corrector";
  (void)myproc; (void)comm;
  if (numprocs != 1) fatal(W, "only numprocs=1 (the whole grid on one GPU) is supported");
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  upload_cells(C, grid, "q", qc, 0, W);
  upload_cells(C, grid, "qp", phys->q, 0, W);
  upload_cells(C, grid, "dzz", grid->dzz, 0, W);
  upload_cells(C, grid, "w", phys->w, 1, W);
  upload_edges(C, grid, "u", phys->u, W);
  check(sunhip_upload(C->dev, "D", phys->D), W);
  check(sunhip_corrector(C->dev, prop->dt, "q", "qp"), W);
  const fetch_t F[] = {{"u", edges_job(C, grid, phys->u)},
                       {"w", cells_w_job(C, grid, phys->w, 0, 1)},
                       {"qp", cells_job(C, grid, phys->q, 0, 1)}};
  fetch(C, F, 3, W);
}

void Continuity(REAL **w, gridT *grid, physT *phys, propT *prop) {
  static const char *W = "Continuity";
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  upload_edges(C, grid, "u", phys->u, W);
  upload_edges(C, grid, "utmp2", phys->utmp2, W);
  upload_edges(C, grid, "dzf", grid->dzf, W);
  upload_cells(C, grid, "wtmp2", phys->wtmp2, 1, W);
  check(sunhip_continuity(C->dev, prop->theta, "w"), W);
  fetch1(C, "w", cells_w_job(C, grid, w, 0, 1), W);
}


static sunhip_upred_params upred_params(const propT *prop) {
  sunhip_upred_params P = {prop->dt, prop->theta, prop->grav, prop->nu, prop->laxWendroff_Vertical, prop->thetaM,
                           prop->nonlinear};
  return P;
}

void sunhip_UPredictor_begin(gridT *grid, physT *phys, propT *prop, int myproc, int numprocs, MPI_Comm comm) {
  static const char *W = "UPredictor";
  (void)myproc; (void)comm;
  if (numprocs != 1) fatal(W, "only numprocs=1 (the whole grid on one GPU) is supported");
  for (int jp = grid->edgedist[0]; jp < grid->edgedist[1]; jp++) {
    const int j = grid->edgep[jp];
    if (grid->Nke[j] - grid->etop[j] == 1 && (phys->CdB[j] == -1 || phys->CdT[j] == -1))
      fatal(W, "no-slip drag (CdB or CdT = -1) on a one-layer edge reads a coefficient left from the previous "
               "edge in the reference (main/phys.c:2803-2815, 2880-2889); not built for the device");
  }
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  upload_edges(C, grid, "u", phys->u, W);
  upload_edges(C, grid, "utmp", phys->utmp, W);
  upload_edges(C, grid, "ue", phys->ut, W);
  upload_edges(C, grid, "dzf", grid->dzf, W);
  upload_cells(C, grid, "w", phys->w, 1, W);
  upload_cells(C, grid, "dzz", grid->dzz, 0, W);
  upload_cells(C, grid, "nu_tv", phys->nu_tv, 0, W);
  upload_cells(C, grid, "nu_lax", phys->nu_lax, 0, W);
  check(sunhip_upload(C->dev, "h", phys->h), W);
  check(sunhip_upload(C->dev, "tau_T", phys->tau_T), W);
  check(sunhip_upload(C->dev, "CdT", phys->CdT), W);
  check(sunhip_upload(C->dev, "CdB", phys->CdB), W);
  sunhip_upred_params P = upred_params(prop);
  const int rc = sunhip_u_predict(C->dev, &P);
  if (rc == -5) { fprintf(stderr, "%s: %s\n", W, sunhip_error()); exit(0); }  /* main/phys.c:2902 exits with 0 */
  if (rc == -6) { fprintf(stderr, "%s: %s\n", W, sunhip_error()); exit(EXIT_FAILURE); }
  check(rc, W);
  const fetch_t F[] = {{"utmp", edges_job(C, grid, phys->utmp)},
                       {"ue", edges_job(C, grid, phys->ut)},
                       {"utmp2", edges_wet_job(C, grid, phys->utmp2)},
                       {"wtmp2", cells_w_job(C, grid, phys->wtmp2, 0, 2)}};
  fetch(C, F, 4, W);
  memcpy(phys->D, download_all(C, "D", W), (size_t)grid->Ne * sizeof(double));
}

void sunhip_UPredictor_solve(gridT *grid, physT *phys, propT *prop, int myproc, int numprocs, MPI_Comm comm) {
  static const char *W = "UPredictor";
  (void)myproc; (void)comm;
  if (numprocs != 1) fatal(W, "only numprocs=1 (the whole grid on one GPU) is supported");
  if (prop->cgsolver != 1) fatal(W, "cgsolver must be 1 (GSSolve is not built for the device)");
  if (prop->hprecond != 0 && prop->hprecond != 1) fatal(W, "hprecond must be 0 or 1");
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  upload_edges(C, grid, "u", phys->u, W);
  upload_edges(C, grid, "utmp", phys->utmp, W);
  upload_edges(C, grid, "ue", phys->ut, W);
  upload_edges(C, grid, "dzf", grid->dzf, W);
  upload_cells(C, grid, "dzz", grid->dzz, 0, W);
  upload_cells(C, grid, "s", phys->s, 0, W);
  check(sunhip_upload(C->dev, "h", phys->h), W);
  check(sunhip_upload(C->dev, "dv", grid->dv), W);
  check(sunhip_upload(C->dev, "hcorr", phys->hcorr), W);
  check(sunhip_upload(C->dev, "D", phys->D), W);
  sunhip_upred_params P = upred_params(prop);
  sunhip_cg_params G = {prop->maxiters, prop->hprecond, (int)prop->resnorm, prop->epsilon, 0, 0};
  check(sunhip_u_solve(C->dev, &P, &G, &g_last[1].st), W);
  g_last[1].valid = 1;
  memcpy(phys->h, download_all(C, "h", W), (size_t)grid->Nc * sizeof(double));
  memcpy(phys->hcorr, download_all(C, "hcorr", W), (size_t)grid->Nc * sizeof(double));
  const double *act = download_all(C, "active", W);
  for (int i = 0; i < grid->Nc; i++) phys->active[i] = act[i] != 0;
  const fetch_t F[] = {{"u", edges_job(C, grid, phys->u)}, {"s", cells_job(C, grid, phys->s, 0, 2)}};
  fetch(C, F, 2, W);
}

static sunhip_vgrid_params vgrid_params(const gridT *grid, const propT *prop) {
  sunhip_vgrid_params V = {grid->dz, grid->dzsmall, grid->smoothbot, grid->stairstep, grid->fixdzz, prop->linearFS};
  return V;
}

void UpdateDZ(gridT *grid, physT *phys, propT *prop, int option) {
  static const char *W = "UpdateDZ";
  if (prop->linearFS) return;  /* main/phys.c:820-822 */
  if (option != 0 && option != 1 && option != -1) fatal(W, "option must be 0, 1 or -1");
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  upload_cells(C, grid, "dzz", grid->dzz, 0, W);
  check(sunhip_upload(C->dev, "h", phys->h), W);
  check(sunhip_upload(C->dev, "dv", grid->dv), W);
  sunhip_vgrid_params V = vgrid_params(grid, prop);
  check(sunhip_update_dz(C->dev, &V, option), W);
  extents_changed(C);
  check(sunhip_get_extents(C->dev, grid->ctop, option != -1 ? grid->ctopold : NULL, grid->etop,
                           option != -1 ? grid->etopold : NULL), W);
  const fetch_t F[] = {{"dzz", cells_job(C, grid, grid->dzz, 0, 2)},
                       {"dzzold", cells_job(C, grid, grid->dzzold, 0, 2)}};
  fetch(C, F, option != -1 ? 2 : 1, W);
  if (option == -1) {
    memcpy(phys->h, download_all(C, "h", W), (size_t)grid->Nc * sizeof(double));
    memcpy(grid->dv, download_all(C, "dv", W), (size_t)grid->Nc * sizeof(double));
    memcpy(grid->dzbot, download_all(C, "dzbot", W), (size_t)grid->Nc * sizeof(double));
  }
  if (option == 1) memcpy(grid->dzfB, download_all(C, "dzfB", W), (size_t)grid->Ne * sizeof(double));
}

void SetFluxHeight(gridT *grid, physT *phys, propT *prop) {
  static const char *W = "SetFluxHeight";
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  upload_edges(C, grid, "u", phys->u, W);
  upload_cells(C, grid, "dzz", grid->dzz, 0, W);
  check(sunhip_upload(C->dev, "h", phys->h), W);
  check(sunhip_upload(C->dev, "dv", grid->dv), W);
  if (grid->smoothbot) check(sunhip_upload(C->dev, "dzbot", grid->dzbot), W);
  sunhip_vgrid_params V = vgrid_params(grid, prop);
  check(sunhip_set_flux_height(C->dev, &V), W);
  const fetch_t F[] = {{"dzf", edges_job(C, grid, grid->dzf)}, {"dzz", cells_job(C, grid, grid->dzz, 0, 2)}};
  fetch(C, F, grid->smoothbot ? 2 : 1, W);
}

/* WPredictor (main/phys.c:2025-2290), file-static in the reference: the
 * same signature under this name, for the call at main/phys.c:1271. */
void sunhip_WPredictor(gridT *grid, physT *phys, propT *prop, int myproc, int numprocs, MPI_Comm comm) {
  static const char *W = "WPredictor";
  (void)myproc; (void)comm;
  if (numprocs != 1) fatal(W, "only numprocs=1 (the whole grid on one GPU) is supported");
  const int scheme = prop->nonlinear;
  if (scheme != 0 && scheme != 1 && scheme != 2 && scheme != 5)
    fatal(W, "nonlinear must be 0, 1, 2 or 5 (scheme 4's vertical Lax-Wendroff term reads dzz one past the "
             "column bottom, main/phys.c:2230-2233; not built for the device)");
  if (prop->conserveMomentum)
    fatal(W, "conserveMomentum reads scratch left from the previous column in the reference "
             "(main/phys.c:2136-2138); not built for the device");
  if (grid->edgedist[5] > grid->edgedist[4])
    fatal(W, "type-4 (no-slip) edges are not built for the device (the reference's type-4 term runs to Nke of "
             "a cell index, main/phys.c:2199)");
  ctx_t *C = get_ctx(grid, W);
  extents(C, grid, W);
  upload_cells(C, grid, "w", phys->w, 1, W);
  upload_cells(C, grid, "wtmp", phys->wtmp, 1, W);
  upload_cells(C, grid, "Cn_W", phys->Cn_W, 1, W);
  upload_cells(C, grid, "Cn_W2", phys->Cn_W2, 1, W);
  upload_cells(C, grid, "qp", phys->q, 0, W);
  upload_cells(C, grid, "dzz", grid->dzz, 0, W);
  upload_cells(C, grid, "nu_tv", phys->nu_tv, 0, W);
  upload_cells(C, grid, "nu_lax", phys->nu_lax, 0, W);
  upload_cells(C, grid, "wc", phys->wc, 0, W);
  upload_cells(C, grid, "stmp", phys->stmp, 0, W);
  upload_edges(C, grid, "ue", phys->ut, W);
  upload_edges(C, grid, "utmp2", phys->utmp2, W);
  upload_brows(C, grid, phys->boundary_w, "bw", W);
  if (scheme == 2) {
    /* main/phys.c:4272-4273 on the host, with the reference's own libm */
    double *def1 = next_buf(C, W);
    for (int j = 0; j < grid->Ne; j++) {
      int nc1 = grid->grad[2 * j];
      if (nc1 == -1) nc1 = grid->grad[2 * j + 1];
      def1[j] = sqrt(pow(grid->xv[nc1] - grid->xe[j], 2) + pow(grid->yv[nc1] - grid->ye[j], 2));
    }
    check(sunhip_set_face_distances(C->dev, def1), W);
  }
  if (scheme == 5) {
    upload_edges(C, grid, "u", phys->u, W);
    upload_edges(C, grid, "dzf", grid->dzf, W);
    upload_edges(C, grid, "SfHp", phys->SfHp, W);
    upload_edges(C, grid, "SfHm", phys->SfHm, W);
    upload_cells(C, grid, "gradSx", phys->gradSx, 0, W);
    upload_cells(C, grid, "gradSy", phys->gradSy, 0, W);
  }
  sunhip_wpred_params P = {prop->dt, prop->theta, prop->nu, prop->nu_H, prop->CdW, prop->laxWendroff_Vertical,
                           scheme, prop->n, prop->AB, prop->TVDmomentum, prop->conserveMomentum};
  check(sunhip_w_predict(C->dev, &P), W);
  const fetch_t F[] = {{"w", cells_w_job(C, grid, phys->w, 0, 2)},
                       {"wtmp", cells_w_job(C, grid, phys->wtmp, 0, 2)},
                       {"Cn_W", cells_w_job(C, grid, phys->Cn_W, 0, 2)},
                       {"Cn_W2", cells_w_job(C, grid, phys->Cn_W2, 0, 2)},
                       {"stmp", cells_job(C, grid, phys->stmp, 0, 2)},
                       {"wc", cells_job(C, grid, phys->wc, 0, 2)},
                       {"ue", edges_job(C, grid, phys->ut)},
                       {"SfHp", edges_job(C, grid, phys->SfHp)},
                       {"SfHm", edges_job(C, grid, phys->SfHm)},
                       {"gradSx", cells_job(C, grid, phys->gradSx, 0, 2)},
                       {"gradSy", cells_job(C, grid, phys->gradSy, 0, 2)}};
  fetch(C, F, scheme == 5 ? 11 : scheme ? 7 : 5, W);
}

int sunhip_last_solve(int which, int *niters, double *resid, double *device_ms) {
  if (which < 0 || which > 1 || !g_last[which].valid) return -1;
  if (niters) *niters = g_last[which].st.niters;
  if (resid) *resid = g_last[which].st.resid;
  if (device_ms) *device_ms = g_last[which].st.device_ms;
  return 0;
}

void sunhip_dropin_reset(void) {
  for (int i = 0; i < g_nctx; i++) ctx_free(&g_ctx[i]);
  memset(g_ctx, 0, sizeof g_ctx);
  g_nctx = 0;
}

void sunhip_dropin_set_precision(int mixed, double inner_tol) {
  g_mixed = mixed ? 1 : 0;
  if (inner_tol > 0) g_inner_tol = inner_tol;
}

void sunhip_dropin_set_mode(int device) {
  if (device != g_device) sunhip_dropin_reset();
  g_device = device;
}

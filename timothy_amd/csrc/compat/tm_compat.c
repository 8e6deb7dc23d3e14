/*
 * tm_compat.c -- the reference's octree / potential / patches / fieldgradient
 * C API on top of the native C ABI (timothy_hip.h).
 *
 * Reference functions served (mcytu/timothy):
 *   octbuild, octfree                    src/octree.c:145-207, 321-327
 *   octcomputebox, octcomputeboxr        src/octree.c:214-266
 *   octlocateregion                      src/octree.c:271-316
 *   octheapinit/push/pop/free            src/octree.c:332-371
 *   computepotential, computegradient    src/potential.c:87-138, 209-259
 *   computeremotepotential/gradient      src/potential.c:144-188, 265-310 (one rank: no halo)
 *   patches_*                            src/patches.c:47-657
 *   fieldgradient                        src/fieldgradient.c:256-277
 *
 * One implicit device context serves the calls; it is created on first use
 * and is not thread-safe (the reference calls these from one host thread).
 * Failures go through the host program's terminate() when it defines one
 * (the reference's utils.c), else print and abort.
 */
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "timothy_compat.h"
#include "timothy_hip.h"

double3dv_t affShift;
double affScale;

/* the host program's terminate (src/utils.c:103-109), when it has one */
extern void terminate(systeminfo_t systeminfo, char *msg, char *file, int line) __attribute__((weak));

#define MAXVALUE 536870912.0 /* powf(2,29), src/global.h:67 */
#define ROOT_LEVEL 29

static struct {
  tm_ctx *ctx;
  int device;
  int precision;
  int host_octree;
  int configured;
  /* snapshot of the last octbuild */
  const cellsinfo_t *built_for;
  int64_t n;
  int have_forces;
  double *h_x, *h_y, *h_z, *h_size, *h_out;
  float *h_young;
  int32_t *h_ctype, *h_phase;
  int64_t h_cap;
  void *h_block; /* page-locked block holding the arrays above */
  /* the emitted octree: page-locked, kept between steps (handed to the
   * caller as cellsinfo->octree by octbuild, taken back by octfree) */
  octnode_t *oct_buf;
  int64_t oct_cap;
  int oct_out;
  /* two page-locked chunks for copies from and to the caller's own
   * (pageable) arrays */
  void *stage[2];
  /* grid transfer buffers */
  double *d_prod, *d_field, *d_grad, *d_interp;
  int64_t grid_cap, interp_cap;
  int deposited;
  /* per-field copies for the interpolation */
  double **d_envf, **d_envg;
  int env_nf, env_ready;
  int64_t env_nodes;
  /* updatepositions staging (host and device, one block each) */
  void *mv_host, *mv_dev;
  int64_t mv_cap;
  /* environment solve: one device solution per field, the source */
  double **d_envsol, *d_envsrc;
  int envsol_nf;
  int64_t envsol_nodes;
  int env_iters;
} S;

static void free_env(void);

static void fail(systeminfo_t si, const char *msg, const char *file, int line) {
  if (terminate) terminate(si, (char *)msg, (char *)file, line);
  fprintf(stderr, "error: %s (%s, line %d)\n", msg, file, line);
  fflush(stderr);
  abort();
}

#define FAIL(si, msg) fail((si), (msg), __FILE__, __LINE__)

static void check(systeminfo_t si, tm_status s, const char *what) {
  if (s == TM_OK) return;
  char buf[512];
  snprintf(buf, sizeof(buf), "%s: %s %s", what, tm_status_name(s), S.ctx ? tm_last_error(S.ctx) : "");
  FAIL(si, buf);
}

static void configure(void) {
  if (S.configured) return;
  const char *p = getenv("TM_PRECISION");
  S.precision = (p && !strcmp(p, "mixed")) ? 1 : 0;
  const char *d = getenv("TM_DEVICE");
  S.device = d ? atoi(d) : 0;
  const char *o = getenv("TM_COMPAT_HOST_OCTREE");
  S.host_octree = o ? atoi(o) != 0 : 1;
  S.configured = 1;
}

static tm_ctx *ctx_for(systeminfo_t si) {
  configure();
  if (si.size > 1) FAIL(si, "the MI355X cell-mechanics path serves one MPI rank (systeminfo.size == 1)");
  if (!S.ctx) {
    /* The GPU runtime draws from rand() while it starts up.  Timothy's own
     * random numbers (cell placement, moves, division, the initial fields)
     * come from the same rand() stream, so the context is created -- and the
     * runtime exercised once -- on a private generator state, and the
     * program's state is put back untouched. */
    static char scratch[256];
    char *prog = initstate(1u, scratch, sizeof(scratch));
    tm_status s = tm_create(&S.ctx, S.device);
    if (s == TM_OK) {
      void *p = NULL;
      s = tm_malloc(S.ctx, &p, 64);
      if (s == TM_OK) s = tm_memset(S.ctx, p, 0, 64);
      if (s == TM_OK) s = tm_sync(S.ctx);
      if (p) tm_free(S.ctx, p);
    }
    setstate(prog);
    check(si, s, "tm_create");
  }
  return S.ctx;
}

void tm_compat_set_precision(int precision) {
  configure();
  S.precision = precision ? 1 : 0;
}
void tm_compat_set_device(int device) {
  configure();
  S.device = device;
}
void tm_compat_set_host_octree(int enable) {
  configure();
  S.host_octree = enable ? 1 : 0;
}
void *tm_compat_context(void) { return S.ctx; }

static void free_host(void) {
  if (S.h_block && S.ctx) tm_host_free(S.ctx, S.h_block);
  S.h_block = NULL;
  S.h_x = S.h_y = S.h_z = S.h_size = S.h_out = NULL;
  S.h_young = NULL;
  S.h_ctype = S.h_phase = NULL;
  S.h_cap = 0;
  for (int b = 0; b < 2; b++) {
    if (S.stage[b] && S.ctx) tm_host_free(S.ctx, S.stage[b]);
    S.stage[b] = NULL;
  }
  /* the octree buffer, unless a caller still holds it (octfree then takes it
   * back; it lives on until the process ends) */
  if (S.oct_buf && S.ctx && !S.oct_out) {
    tm_host_free(S.ctx, S.oct_buf);
    S.oct_buf = NULL;
    S.oct_cap = 0;
  }
}

void tm_compat_shutdown(void) {
  free_host();
  if (S.ctx) {
    free_env();
    if (S.d_prod) tm_free(S.ctx, S.d_prod);
    if (S.d_field) tm_free(S.ctx, S.d_field);
    if (S.d_grad) tm_free(S.ctx, S.d_grad);
    if (S.d_interp) tm_free(S.ctx, S.d_interp);
    if (S.mv_dev) tm_free(S.ctx, S.mv_dev);
    if (S.mv_host) tm_host_free(S.ctx, S.mv_host);
    for (int f = 0; f < S.envsol_nf; f++)
      if (S.d_envsol[f]) tm_free(S.ctx, S.d_envsol[f]);
    if (S.d_envsrc) tm_free(S.ctx, S.d_envsrc);
    tm_destroy(S.ctx);
  }
  free(S.d_envsol);
  S.d_envsol = NULL;
  S.d_envsrc = NULL;
  S.envsol_nf = 0;
  S.envsol_nodes = 0;
  S.mv_host = S.mv_dev = NULL;
  S.mv_cap = 0;
  S.d_prod = S.d_field = S.d_grad = S.d_interp = NULL;
  S.grid_cap = S.interp_cap = 0;
  S.ctx = NULL;
  S.built_for = NULL;
}

/* Copies between the caller's pageable arrays and the device, pipelined
 * through two page-locked chunks: threads copy one chunk on the host while
 * the previous one crosses the link (a pageable copy stages through the
 * runtime's buffers with one host thread). */
#define STAGE_BYTES ((size_t)32 << 20)

static void ensure_stage(systeminfo_t si) {
  for (int b = 0; b < 2; b++)
    if (!S.stage[b]) check(si, tm_host_alloc(S.ctx, &S.stage[b], STAGE_BYTES), "copy staging");
}

static void par_memcpy(void *dst, const void *src, size_t bytes) {
  const size_t blk = (size_t)1 << 20;
  const int64_t nb = (int64_t)((bytes + blk - 1) / blk);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nb; i++) {
    const size_t o = (size_t)i * blk, m = bytes - o < blk ? bytes - o : blk;
    memcpy((char *)dst + o, (const char *)src + o, m);
  }
}

static void copy_to_device(systeminfo_t si, void *dst, const void *src, size_t bytes, const char *what) {
  ensure_stage(si);
  for (size_t o = 0, i = 0; o < bytes; o += STAGE_BYTES, i++) {
    const size_t m = bytes - o < STAGE_BYTES ? bytes - o : STAGE_BYTES;
    void *buf = S.stage[i & 1];
    par_memcpy(buf, (const char *)src + o, m); /* while chunk i - 1 crosses from the other buffer */
    check(si, tm_sync(S.ctx), what);         /* chunk i - 1 is across: its buffer is free again */
    check(si, tm_memcpy_htod_async(S.ctx, (char *)dst + o, buf, m), what);
  }
  check(si, tm_sync(S.ctx), what);
}

static void copy_to_host(systeminfo_t si, void *dst, const void *src, size_t bytes, const char *what) {
  ensure_stage(si);
  if (bytes == 0) return;
  size_t m0 = bytes < STAGE_BYTES ? bytes : STAGE_BYTES;
  check(si, tm_memcpy_dtoh_async(S.ctx, S.stage[0], src, m0), what);
  for (size_t o = 0, i = 0; o < bytes; o += STAGE_BYTES, i++) {
    const size_t m = bytes - o < STAGE_BYTES ? bytes - o : STAGE_BYTES;
    check(si, tm_sync(S.ctx), what); /* chunk i is in its buffer */
    const size_t o1 = o + STAGE_BYTES;
    if (o1 < bytes) {
      const size_t m1 = bytes - o1 < STAGE_BYTES ? bytes - o1 : STAGE_BYTES;
      check(si, tm_memcpy_dtoh_async(S.ctx, S.stage[(i + 1) & 1], (const char *)src + o1, m1), what);
    }
    par_memcpy((char *)dst + o, S.stage[i & 1], m); /* while chunk i + 1 crosses */
  }
}

/* the per-cell staging arrays, in one page-locked block (the uploads and
 * downloads through it run at the link rate) */
static void ensure_host(systeminfo_t si, int64_t n) {
  if (n <= S.h_cap) return;
  if (S.h_block) tm_host_free(S.ctx, S.h_block);
  S.h_block = NULL;
  S.h_cap = 0;
  const size_t a = ((size_t)n * sizeof(double) + 255) & ~(size_t)255; /* one array of doubles, aligned */
  const size_t b = ((size_t)n * sizeof(int32_t) + 255) & ~(size_t)255;
  void *p = NULL;
  check(si, tm_host_alloc(S.ctx, &p, 10 * a + 3 * b), "host staging");
  char *q = (char *)p;
  S.h_x = (double *)q;
  S.h_y = (double *)(q + a);
  S.h_z = (double *)(q + 2 * a);
  S.h_size = (double *)(q + 3 * a);
  S.h_out = (double *)(q + 4 * a); /* 6 n doubles */
  S.h_young = (float *)(q + 10 * a);
  S.h_ctype = (int32_t *)(q + 10 * a + b);
  S.h_phase = (int32_t *)(q + 10 * a + 2 * b);
  S.h_block = p;
  S.h_cap = n;
}

/* ------------------------------------------------------------------ */
/* octree                                                              */
/* ------------------------------------------------------------------ */

void octbuild(systeminfo_t si, cellsinfo_t *ci, celltype_t *celltype) {
  const int64_t n = (int64_t)ci->localcount.n;
  if (n == 0) {
    /* src/octree.c:152-153: no tree; the step's CIC calls still run (they
     * deposit nothing and interpolate onto no cells) */
    S.built_for = ci;
    S.n = 0;
    S.have_forces = 0;
    S.deposited = 0;
    return;
  }
  tm_ctx *c = ctx_for(si);
  ensure_host(si, n);
  int maxt = 0, neg = 0;
#pragma omp parallel for schedule(static) reduction(max : maxt) reduction(| : neg)
  for (int64_t i = 0; i < n; i++) {
    const celldata_t *cd = &ci->cells[i];
    S.h_x[i] = cd->x;
    S.h_y[i] = cd->y;
    S.h_z[i] = cd->z;
    S.h_size[i] = cd->size;
    S.h_young[i] = cd->young;
    S.h_ctype[i] = cd->ctype;
    S.h_phase[i] = cd->phase;
    if (cd->ctype > maxt) maxt = cd->ctype;
    neg |= cd->ctype < 0;
  }
  if (neg) FAIL(si, "negative cell type");
  int ntypes = maxt + 1;
  if (ntypes > 16) FAIL(si, "at most 16 cell types");
  tm_celltype t[16];
  for (int k = 0; k < ntypes; k++) {
    t[k].size = celltype[k].size;
    t[k].h = celltype[k].h;
    t[k].h2 = celltype[k].h2;
    t[k].h3 = celltype[k].h3;
    t[k].h4 = celltype[k].h4;
  }
  int dim = ci->dimension == 2 ? 2 : 3;
  check(si, tm_set_celltypes(c, t, ntypes, dim), "tm_set_celltypes");
  check(si, tm_upload(c, n, 1, NULL, S.h_x, S.h_y, S.h_z, S.h_size, S.h_young, S.h_ctype, S.h_phase), "tm_upload");
  check(si, tm_build(c), "octbuild");
  double sh[3], sc;
  check(si, tm_get_affine(c, 0, sh, &sc), "tm_get_affine");
  affShift.x = sh[0];
  affShift.y = sh[1];
  affShift.z = sh[2];
  affScale = sc;
  S.built_for = ci;
  S.n = n;
  S.have_forces = 0;
  S.deposited = 0;
  const octnode_t *prev = ci->octree;
  ci->octree = NULL;
  ci->octsize = 0;
  ci->octmaxsize = 0;
  if (S.host_octree) {
    int64_t cnt = 0;
    check(si, tm_get_octree(c, 0, NULL, &cnt), "tm_get_octree");
    /* the kept page-locked buffer when it is free (or still this caller's,
     * from an octbuild without octfree), else a plain allocation */
    if (S.oct_out && prev == S.oct_buf) S.oct_out = 0;
    octnode_t *o = NULL;
    if (!S.oct_out) {
      if (cnt > S.oct_cap) {
        if (S.oct_buf) tm_host_free(c, S.oct_buf);
        S.oct_buf = NULL;
        S.oct_cap = 0;
        void *p = NULL;
        const int64_t want = cnt + cnt / 8;
        check(si, tm_host_alloc(c, &p, sizeof(octnode_t) * (size_t)want), "cellsinfo->octree");
        S.oct_buf = (octnode_t *)p;
        S.oct_cap = want;
      }
      o = S.oct_buf;
      S.oct_out = 1;
    } else {
      o = (octnode_t *)malloc(sizeof(octnode_t) * (size_t)cnt);
      if (!o) FAIL(si, "cannot allocate cellsinfo->octree");
    }
    check(si, tm_get_octree(c, 0, o, &cnt), "tm_get_octree");
    ci->octree = o;
    ci->octsize = cnt;
    ci->octmaxsize = cnt;
  }
}

void octfree(cellsinfo_t *ci) {
  if (ci->localcount.n == 0) return; /* src/octree.c:323 */
  if (ci->octree && ci->octree == S.oct_buf) S.oct_out = 0; /* kept for the next octbuild */
  else free(ci->octree);
  ci->octree = NULL;
  if (S.built_for == ci) S.built_for = NULL;
}

/* src/octree.c:246-266 */
void octcomputebox(int64_t c, uint3dv_t *mn, uint3dv_t *mx, cellsinfo_t ci, celltype_t *celltype) {
  const celldata_t *cd = &ci.cells[c];
  double h = celltype[cd->ctype].h;
  mn->x = (unsigned int)(((cd->x - h - affShift.x) / affScale) * MAXVALUE);
  mn->y = (unsigned int)(((cd->y - h - affShift.y) / affScale) * MAXVALUE);
  mn->z = (unsigned int)(((cd->z - h - affShift.z) / affScale) * MAXVALUE);
  mx->x = (unsigned int)(((cd->x + h - affShift.x) / affScale) * MAXVALUE);
  mx->y = (unsigned int)(((cd->y + h - affShift.y) / affScale) * MAXVALUE);
  mx->z = (unsigned int)(((cd->z + h - affShift.z) / affScale) * MAXVALUE);
}

/* src/octree.c:214-241: halo cells, box cropped to the unit cube */
void octcomputeboxr(int64_t c, uint3dv_t *mn, uint3dv_t *mx, cellcommdata_t cc, celltype_t *celltype) {
  const cellindata_t *cd = &cc.recvcellindata[c];
  double h = celltype[cd->ctype].h;
  double lo[3] = {(cd->x - h - affShift.x) / affScale, (cd->y - h - affShift.y) / affScale,
                  (cd->z - h - affShift.z) / affScale};
  double hi[3] = {(cd->x + h - affShift.x) / affScale, (cd->y + h - affShift.y) / affScale,
                  (cd->z + h - affShift.z) / affScale};
  for (int a = 0; a < 3; a++) {
    lo[a] = lo[a] < 0.0 ? 0.0 : lo[a];
    hi[a] = hi[a] > 1.0 ? 1.0 : hi[a];
  }
  mn->x = (unsigned int)(lo[0] * MAXVALUE);
  mn->y = (unsigned int)(lo[1] * MAXVALUE);
  mn->z = (unsigned int)(lo[2] * MAXVALUE);
  mx->x = (unsigned int)(hi[0] * MAXVALUE);
  mx->y = (unsigned int)(hi[1] * MAXVALUE);
  mx->z = (unsigned int)(hi[2] * MAXVALUE);
}

static inline unsigned child_index(unsigned x, unsigned y, unsigned z, unsigned bit) {
  return ((x >> bit) & 1u) + 2u * ((y >> bit) & 1u) + 4u * ((z >> bit) & 1u);
}

/* src/octree.c:271-316: smallest node containing the box.  A missing child
 * ends the descent at its parent (the reference reads octree[-1] there). */
int octlocateregion(uint3dv_t mn, uint3dv_t mx, cellsinfo_t ci) {
  unsigned dx = mn.x ^ mx.x, dy = mn.y ^ mx.y, dz = mn.z ^ mx.z;
  unsigned l1 = ROOT_LEVEL, l2 = ROOT_LEVEL, lmin = ROOT_LEVEL;
  while (!(dx & (1u << l1)) && l1) l1--;
  while (!(dy & (1u << l2)) && (l2 > l1)) l2--;
  while (!(dz & (1u << lmin)) && (lmin > l2)) lmin--;
  lmin++;
  unsigned level = ROOT_LEVEL - 1;
  unsigned steps = level - lmin + 1;
  int64_t idx = 0, parent = 0;
  while (steps--) {
    unsigned ci_ = child_index(mn.x, mn.y, mn.z, level);
    level--;
    parent = idx;
    idx = ci.octree[idx].child[ci_];
    if (idx == -1) break;
    if (ci.octree[idx].data != -2) break;
  }
  if (idx == -1) idx = parent;
  return (int)idx;
}

/* src/octree.c:332-371 */
void octheapinit(systeminfo_t si, octheap_t *h) {
  h->size = 64;
  h->count = 0;
  if (!(h->data = malloc(sizeof(int) * h->size))) FAIL(si, "cannot allocate ttheap->data");
}
void octheappush(systeminfo_t si, octheap_t *h, int idx) {
  if (h->count == h->size) {
    h->size += 64;
    int *d = realloc(h->data, sizeof(int) * h->size);
    if (!d) FAIL(si, "cannot reallocate ttheap->data");
    h->data = d;
  }
  h->data[h->count++] = idx;
}
int octheappop(octheap_t *h) { return h->data[--h->count]; }
void octheapfree(octheap_t *h) { free(h->data); }

/* ------------------------------------------------------------------ */
/* potential                                                           */
/* ------------------------------------------------------------------ */

static void require_build(systeminfo_t si, const cellsinfo_t *ci) {
  if (S.built_for != ci || S.n != (int64_t)ci->localcount.n)
    FAIL(si, "computepotential/computegradient need octbuild on the same cells first");
}

static void run_pairs(systeminfo_t si, cellsinfo_t *ci) {
  tm_ctx *c = S.ctx;
  const int64_t n = S.n;
  check(si, tm_neighbors(c), "neighbour lists");
  check(si, tm_sph(c, S.precision ? TM_MIXED : TM_FP64), "pair sums");
  double *o = S.h_out;
  check(si, tm_get_results(c, o, o + n, o + 2 * n, o + 3 * n, o + 4 * n, o + 5 * n), "results");
  S.have_forces = 1;
  (void)ci;
}

/* src/potential.c:87-138 */
void computepotential(systeminfo_t si, cellsinfo_t *ci, celltype_t *celltype, cellcommdata_t cc) {
  if (ci->localcount.n <= 1) return; /* src/potential.c:89 */
  ctx_for(si);
  require_build(si, ci);
  run_pairs(si, ci);
  const int64_t n = S.n;
  const double *o = S.h_out;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    ci->cells[i].density = o[i];
    ci->cells[i].v = o[n + i];
    ci->cells[i].mindist = o[2 * n + i];
  }
  (void)celltype;
  (void)cc;
}

/* src/potential.c:209-259 */
void computegradient(systeminfo_t si, cellsinfo_t *ci, celltype_t *celltype, cellcommdata_t cc) {
  if (ci->localcount.n <= 1) return; /* src/potential.c:212 */
  ctx_for(si);
  require_build(si, ci);
  if (!S.have_forces) run_pairs(si, ci);
  const int64_t n = S.n;
  const double *o = S.h_out;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    ci->forces[i].x = o[3 * n + i];
    ci->forces[i].y = o[4 * n + i];
    ci->forces[i].z = o[5 * n + i];
  }
  (void)celltype;
  (void)cc;
}

/* src/potential.c:144-188 and 265-310: contributions of halo cells, which
 * one rank does not have */
void computeremotepotential(systeminfo_t si, cellsinfo_t *ci, celltype_t *celltype, cellcommdata_t cc) {
  if (cc.numimp <= 0) return;
  FAIL(si, "halo cells (numimp > 0) need more than one rank; not served on one GPU");
  (void)ci;
  (void)celltype;
}
void computeremotegradient(systeminfo_t si, cellsinfo_t *ci, celltype_t *celltype, cellcommdata_t cc) {
  if (cc.numimp <= 0) return;
  FAIL(si, "halo cells (numimp > 0) need more than one rank; not served on one GPU");
  (void)ci;
  (void)celltype;
}

/* ------------------------------------------------------------------ */
/* moving the cells                                                    */
/* ------------------------------------------------------------------ */

/* bytes per cell of the staging block: x, y, z, size, fx, fy, fz (56),
 * ctype (4), three draws (12), removecell (1) */
#define MV_BYTES 73

static void ensure_mv(systeminfo_t si, tm_ctx *c, int64_t n) {
  if (n <= S.mv_cap) return;
  if (S.mv_host) tm_host_free(c, S.mv_host);
  if (S.mv_dev) tm_free(c, S.mv_dev);
  S.mv_host = S.mv_dev = NULL;
  S.mv_cap = 0;
  check(si, tm_host_alloc(c, &S.mv_host, (size_t)MV_BYTES * (size_t)n), "host staging");
  check(si, tm_malloc(c, &S.mv_dev, (size_t)MV_BYTES * (size_t)n), "cell staging");
  S.mv_cap = n;
}

/* src/cells.c:342-365.  The random steps are drawn from the program's rand()
 * exactly as the reference draws them (cell after cell; x, y, then z), so the
 * rest of the program sees the same rand() stream; the update itself runs on
 * the GPU over the current cells and forces. */
void updatepositions(settings_t st, celltype_t *celltype, cellsinfo_t *ci, unsigned char *removecell,
                     int64_t *removecount) {
  systeminfo_t si;
  memset(&si, 0, sizeof(si));
  si.size = 1;
  const int64_t n = (int64_t)ci->localcount.n;
  if (n <= 0) return;
  const int dim = st.dimension == 2 ? 2 : 3;
  tm_ctx *c = ctx_for(si);
  if (st.numberofcelltypes < 1 || st.numberofcelltypes > 16) FAIL(si, "at most 16 cell types");
  tm_move m;
  memset(&m, 0, sizeof(m));
  for (int t = 0; t < st.numberofcelltypes; t++) m.randommove[t] = celltype[t].randommove;
  m.box[0] = BOXSIZEX;
  m.box[1] = BOXSIZEY;
  m.box[2] = BOXSIZEZ;
  ensure_mv(si, c, n);
  /* the same layout on host and device */
  double *hx = (double *)S.mv_host, *hy = hx + n, *hz = hy + n, *hs = hz + n, *hf = hs + n;
  int32_t *ht = (int32_t *)(hf + 3 * n);
  float *hd = (float *)(ht + n);
  uint8_t *hr = (uint8_t *)(hd + 3 * n);
  double *dx = (double *)S.mv_dev, *dy = dx + n, *dz = dy + n, *ds = dz + n, *df = ds + n;
  int32_t *dt = (int32_t *)(df + 3 * n);
  float *dd = (float *)(dt + n);
  uint8_t *dr = (uint8_t *)(dd + 3 * n);
  int64_t k = 0;
  for (int64_t i = 0; i < n; i++) {
    const celldata_t *cd = &ci->cells[i];
    hx[i] = cd->x;
    hy[i] = cd->y;
    hz[i] = cd->z;
    hs[i] = cd->size;
    hf[i] = ci->forces[i].x;
    hf[n + i] = ci->forces[i].y;
    hf[2 * n + i] = ci->forces[i].z;
    if (cd->ctype < 0 || cd->ctype >= st.numberofcelltypes) FAIL(si, "ctype out of range");
    ht[i] = cd->ctype;
    for (int a = 0; a < dim; a++) hd[k++] = (float)rand() / RAND_MAX;
    hr[i] = removecell[i];
  }
  check(si, tm_memcpy_htod(c, S.mv_dev, S.mv_host, (size_t)MV_BYTES * (size_t)n), "move upload");
  int64_t added = 0;
  check(si,
        tm_move_cells(c, n, dim, dx, dy, dz, ds, dt, df, df + n, df + 2 * n, &m, dd, dr, &added),
        "updatepositions");
  check(si, tm_memcpy_dtoh(c, hx, dx, sizeof(double) * 3 * (size_t)n), "move download");
  check(si, tm_memcpy_dtoh(c, hr, dr, (size_t)n), "move download");
  for (int64_t i = 0; i < n; i++) {
    ci->cells[i].x = hx[i];
    ci->cells[i].y = hy[i];
    if (dim == 3) ci->cells[i].z = hz[i];
    removecell[i] = hr[i];
  }
  *removecount += added;
}

/* src/io.c:372-516, one rank: vtk/step%08d.vtk from the host cells and
 * forces, packed on the GPU (positions and forces to big-endian float) */
void writevtk(systeminfo_t si, settings_t st, cellsinfo_t ci) {
  if (si.size > 1) FAIL(si, "the MI355X cell-mechanics path serves one MPI rank");
  const int64_t n = (int64_t)ci.localcount.n;
  tm_ctx *c = ctx_for(si);
  char path[64];
  snprintf(path, sizeof(path), "vtk/step%08d.vtk", st.step);
  if (n == 0) {
    check(si, tm_write_vtk_arrays(c, path, 0, NULL, NULL, NULL, NULL, NULL, NULL, NULL, si.rank), "writevtk");
    return;
  }
  ensure_mv(si, c, n);
  /* the staging block's x, y, z and force columns, and its types */
  double *hx = (double *)S.mv_host, *hy = hx + n, *hz = hy + n, *hs = hz + n, *hf = hs + n;
  int32_t *ht = (int32_t *)(hf + 3 * n);
  double *dx = (double *)S.mv_dev, *dy = dx + n, *dz = dy + n, *ds = dz + n, *df = ds + n;
  int32_t *dt = (int32_t *)(df + 3 * n);
  for (int64_t i = 0; i < n; i++) {
    hx[i] = ci.cells[i].x;
    hy[i] = ci.cells[i].y;
    hz[i] = ci.cells[i].z;
    hf[i] = ci.forces[i].x;
    hf[n + i] = ci.forces[i].y;
    hf[2 * n + i] = ci.forces[i].z;
    ht[i] = ci.cells[i].ctype;
  }
  check(si, tm_memcpy_htod(c, S.mv_dev, S.mv_host, (size_t)(60 * n)), "writevtk upload");
  check(si, tm_write_vtk_arrays(c, path, n, dx, dy, dz, dt, df, df + n, df + 2 * n, si.rank), "writevtk");
}

/* ------------------------------------------------------------------ */
/* environment                                                         */
/* ------------------------------------------------------------------ */

static tm_grid grid_of(const grid_t *g, int rank);

/* src/environment.c:33-52 */
void environment_allocate(systeminfo_t si, settings_t st, grid_t grid, environment_t **environment,
                          solverdata_t *sd, solversettings_t *ss) {
  (void)sd;
  if (st.numberoffields == 0) return;
  const int nf = st.numberoffields;
  const size_t nodes = (size_t)(grid.localsize.x * grid.localsize.y * grid.localsize.z);
  ss->dt = calloc(nf, sizeof(double));
  ss->z = calloc(nf, sizeof(double));
  ss->vartypes = calloc(nf, sizeof(int));
  ss->stencil = calloc(nf, sizeof(void *));
  if (!ss->dt || !ss->z || !ss->vartypes || !ss->stencil) FAIL(si, "cannot allocate solver settings");
  for (int i = 0; i < nf; i++)
    if (!((*environment)[i].data = calloc(nodes, sizeof(double)))) FAIL(si, "cannot allocate environment->data");
  for (int i = 0; i < nf; i++)
    if (!((*environment)[i].production = calloc(nodes, sizeof(double))))
      FAIL(si, "cannot allocate environment->production");
  for (int i = 0; i < nf; i++)
    if (!((*environment)[i].gradient = calloc(nodes * 3, sizeof(double))))
      FAIL(si, "cannot allocate environment->gradient");
}

/* src/environment.c:54-70: mean + rand()/RAND_MAX * variance per node, in the
 * reference's order (fields, then i, j, k), so the program's rand() stream
 * advances as it would */
void environment_init(systeminfo_t si, settings_t st, grid_t grid, environment_t **environment) {
  (void)si;
  if (st.numberoffields == 0) return;
  const int64_t lx = grid.localsize.x, ly = grid.localsize.y, lz = grid.localsize.z;
  for (int f = 0; f < st.numberoffields; f++)
    for (int64_t i = 0; i < lx; i++)
      for (int64_t j = 0; j < ly; j++)
        for (int64_t k = 0; k < lz; k++)
          (*environment)[f].data[ly * lz * i + lz * j + k] =
              (*environment)[f].initialconditionmean +
              ((double)(rand()) / RAND_MAX) * (*environment)[f].initialconditionvariance;
}

/* src/environment.c:131-243: the number of steps per call and dt (the
 * matrix itself is formed on the GPU at each solve) */
void environment_initsystem(systeminfo_t si, settings_t st, grid_t *grid, environment_t **environment,
                            solverdata_t *sd, solversettings_t *ss) {
  (void)sd;
  if (st.numberoffields == 0) return;
  if ((st.secondsperstep / st.gfdt) > 1) {
    const int intdiv = (int)(st.secondsperstep / st.gfdt) + 1;
    const float a1 = st.secondsperstep / intdiv;
    const float a2 = st.secondsperstep / (intdiv - 1);
    if ((st.gfdt - a1) < (a2 - st.gfdt)) {
      ss->numberofiters = intdiv;
      for (int v = 0; v < st.numberoffields; v++) ss->dt[v] = a1;
    } else {
      ss->numberofiters = intdiv - 1;
      for (int v = 0; v < st.numberoffields; v++) ss->dt[v] = a2;
    }
  } else {
    ss->numberofiters = 1;
    for (int v = 0; v < st.numberoffields; v++) ss->dt[v] = st.secondsperstep;
  }
  ss->enviter = 0;
  ss->lower[0] = (int)grid->loweridx[si.rank].x;
  ss->lower[1] = (int)grid->loweridx[si.rank].y;
  ss->lower[2] = (int)grid->loweridx[si.rank].z;
  ss->upper[0] = (int)grid->upperidx[si.rank].x;
  ss->upper[1] = (int)grid->upperidx[si.rank].y;
  ss->upper[2] = (int)grid->upperidx[si.rank].z;
  /* src/environment.c:229-231: the matrix uses gfdt whatever the steps */
  for (int v = 0; v < st.numberoffields; v++) {
    ss->dt[v] = st.gfdt;
    ss->z[v] = (*environment)[v].diffusioncoefficient * ss->dt[v] / (grid->resolution * grid->resolution);
  }
}

/* src/environment.c:516-523 (initboundary + initsolver + solve) */
void environment_compute(systeminfo_t si, settings_t st, grid_t *grid, environment_t **environment,
                         solverdata_t *sd, solversettings_t *ss) {
  (void)sd;
  if (st.numberoffields == 0) return;
  if (si.rank == 0) {
    printf("Solving field.");
    fflush(stdout);
  }
  tm_ctx *c = ctx_for(si);
  const int64_t lx = grid->localsize.x, ly = grid->localsize.y, lz = grid->localsize.z;
  const int64_t nodes = lx * ly * lz;
  const int nf = st.numberoffields;
  if (S.envsol_nf != nf || S.envsol_nodes != nodes) {
    for (int f = 0; f < S.envsol_nf; f++)
      if (S.d_envsol[f]) tm_free(c, S.d_envsol[f]);
    if (S.d_envsrc) tm_free(c, S.d_envsrc);
    free(S.d_envsol);
    S.d_envsol = calloc(nf, sizeof(double *));
    if (!S.d_envsol) FAIL(si, "cannot allocate the solution table");
    for (int f = 0; f < nf; f++)
      check(si, tm_malloc(c, (void **)&S.d_envsol[f], sizeof(double) * (size_t)nodes), "environment solution");
    check(si, tm_malloc(c, (void **)&S.d_envsrc, sizeof(double) * (size_t)nodes), "environment source");
    S.envsol_nf = nf;
    S.envsol_nodes = nodes;
  }
  /* environment_pcfunction (src/environment.c:252-269): 100 at local node
   * (5,5,5) from step 1 on, nothing else */
  double *host = calloc((size_t)nodes, sizeof(double));
  if (!host) FAIL(si, "cannot allocate host staging");
  if (st.step != 0 && lx > 5 && ly > 5 && lz > 5) host[(5 * ly + 5) * lz + 5] = 100.0;
  check(si, tm_memcpy_htod(c, S.d_envsrc, host, sizeof(double) * (size_t)nodes), "environment source");
  tm_grid g = grid_of(grid, si.rank);
  S.env_iters = 0;
  for (int f = 0; f < nf; f++) {
    /* the right-hand side starts from initialconditionmean, not from data
     * (src/environment.c:309-323) */
    for (int64_t q = 0; q < nodes; q++) host[q] = (*environment)[f].initialconditionmean;
    check(si, tm_memcpy_htod(c, S.d_envsol[f], host, sizeof(double) * (size_t)nodes), "environment field");
    tm_diffusion d;
    d.diffusion = (*environment)[f].diffusioncoefficient;
    d.lambda = (*environment)[f].lambdadelay;
    d.boundary = (*environment)[f].boundarycondition;
    d.dt = ss->dt[f];
    d.tol = 1.0e-12;
    d.maxiter = 50;
    d.substeps = ss->numberofiters;
    int32_t it = 0;
    double rel = 0.0;
    check(si, tm_diffusion_step(c, &g, &d, S.d_envsrc, S.d_envsol[f], &it, &rel), "environment solve");
    S.env_iters += it;
  }
  ss->enviter += ss->numberofiters;
  free(host);
}

int tm_compat_environment_solution(int field, double *out) {
  if (!S.ctx || field < 0 || field >= S.envsol_nf || !out) return -1;
  return tm_memcpy_dtoh(S.ctx, out, S.d_envsol[field], sizeof(double) * (size_t)S.envsol_nodes) == TM_OK ? 0 : -1;
}

int tm_compat_environment_iterations(void) { return S.env_iters; }

/* ------------------------------------------------------------------ */
/* patches: cells <-> grid                                             */
/* ------------------------------------------------------------------ */

static tm_grid grid_of(const grid_t *g, int rank) {
  tm_grid t;
  t.lower[0] = g->lowercorner.x;
  t.lower[1] = g->lowercorner.y;
  t.lower[2] = g->lowercorner.z;
  t.resolution = g->resolution;
  t.loweridx[0] = g->loweridx[rank].x;
  t.loweridx[1] = g->loweridx[rank].y;
  t.loweridx[2] = g->loweridx[rank].z;
  t.upperidx[0] = g->upperidx[rank].x;
  t.upperidx[1] = g->upperidx[rank].y;
  t.upperidx[2] = g->upperidx[rank].z;
  t.localsize[0] = g->localsize.x;
  t.localsize[1] = g->localsize.y;
  t.localsize[2] = g->localsize.z;
  return t;
}

static int64_t grid_nodes(const grid_t *g) { return g->localsize.x * g->localsize.y * g->localsize.z; }

static void ensure_grid(systeminfo_t si, int64_t nodes) {
  if (nodes <= S.grid_cap) return;
  if (S.d_prod) tm_free(S.ctx, S.d_prod);
  if (S.d_field) tm_free(S.ctx, S.d_field);
  if (S.d_grad) tm_free(S.ctx, S.d_grad);
  S.d_prod = S.d_field = S.d_grad = NULL;
  check(si, tm_malloc(S.ctx, (void **)&S.d_prod, sizeof(double) * nodes), "grid buffer");
  check(si, tm_malloc(S.ctx, (void **)&S.d_field, sizeof(double) * nodes), "grid buffer");
  check(si, tm_malloc(S.ctx, (void **)&S.d_grad, sizeof(double) * 3 * nodes), "grid buffer");
  S.grid_cap = nodes;
}

/* src/patches.c:47-148, one rank: the patch is the bounding box of the
 * in-range corner indices of all cells */
void patches_alloc(systeminfo_t si, settings_t st, patches_t *pt, cellsinfo_t *ci, grid_t *grid) {
  if (si.size > 1) FAIL(si, "the MI355X cell-mechanics path serves one MPI rank");
  memset(pt, 0, sizeof(*pt));
  pt->data = calloc(1, sizeof(double *));
  pt->intersect = calloc(1, sizeof(int));
  pt->lowercorner = calloc(1, sizeof(int643dv_t));
  pt->uppercorner = calloc(1, sizeof(int643dv_t));
  pt->lowercornerR = calloc(1, sizeof(int643dv_t));
  pt->uppercornerR = calloc(1, sizeof(int643dv_t));
  pt->size = calloc(1, sizeof(int643dv_t));
  if (!pt->data || !pt->intersect || !pt->lowercorner || !pt->uppercorner || !pt->lowercornerR ||
      !pt->uppercornerR || !pt->size)
    FAIL(si, "cannot allocate patches");
  const int d3 = ci->dimension == 3;
  int64_t lox = INT64_MAX, loy = INT64_MAX, loz = d3 ? INT64_MAX : 0;
  int64_t hix = INT64_MIN, hiy = INT64_MIN, hiz = d3 ? INT64_MIN : 0;
  int any = 0;
  const int643dv_t L = grid->loweridx[0], U = grid->upperidx[0];
  const double gx = grid->lowercorner.x, gy = grid->lowercorner.y, gz = grid->lowercorner.z;
  const double res = grid->resolution;
  const int64_t n = (int64_t)ci->localcount.n;
  /* minima, maxima and "any": the same result in any order */
#pragma omp parallel for schedule(static) reduction(min : lox, loy, loz) reduction(max : hix, hiy, hiz) \
    reduction(| : any)
  for (int64_t c = 0; c < n; c++) {
    int643dv_t k;
    k.x = (int64_t)((ci->cells[c].x - gx) / res);
    k.y = (int64_t)((ci->cells[c].y - gy) / res);
    k.z = (int64_t)((ci->cells[c].z - gz) / res);
    for (int ax = 0; ax < 2; ax++)
      for (int ay = 0; ay < 2; ay++)
        for (int az = 0; az < 2; az++) {
          if (!(k.x + ax >= L.x && k.y + ay >= L.y && k.z + az >= L.z && k.x + ax <= U.x && k.y + ay <= U.y &&
                k.z + az <= U.z))
            continue;
          if (k.x + ax < lox) lox = k.x + ax;
          if (k.y + ay < loy) loy = k.y + ay;
          if (k.x + ax > hix) hix = k.x + ax;
          if (k.y + ay > hiy) hiy = k.y + ay;
          if (d3) {
            if (k.z + az < loz) loz = k.z + az;
            if (k.z + az > hiz) hiz = k.z + az;
          }
          any = 1;
        }
  }
  const int643dv_t lo = {lox, loy, loz}, hi = {hix, hiy, hiz};
  if (any) pt->intersect[0] = 1;
  pt->lowercorner[0] = lo;
  pt->uppercorner[0] = hi;
  if (pt->intersect[0]) {
    pt->size[0].x = hi.x - lo.x + 1;
    pt->size[0].y = hi.y - lo.y + 1;
    pt->size[0].z = ci->dimension == 3 ? hi.z - lo.z + 1 : 1;
  }
  (void)st;
}

/* src/patches.c:156-234 + 242-286: CIC deposit of field 0 on the device */
void patches_cells2envinit(systeminfo_t si, settings_t st, patches_t *pt, cellsinfo_t *ci, grid_t *grid) {
  if (st.numberoffields == 0) return;
  ctx_for(si);
  require_build(si, ci);
  const int64_t nodes = grid_nodes(grid);
  ensure_grid(si, nodes);
  tm_grid g = grid_of(grid, si.rank);
  if (S.n > 0)
    check(si, tm_cic_deposit(S.ctx, &g, 0, S.d_prod), "CIC deposit");
  else /* no cells: nothing deposited (src/patches.c:172 loops over no cells) */
    check(si, tm_memset(S.ctx, S.d_prod, 0, sizeof(double) * (size_t)nodes), "clear production");
  pt->receiver = calloc(1, sizeof(int));
  pt->sender = calloc(1, sizeof(int));
  if (!pt->receiver || !pt->sender) FAIL(si, "cannot allocate patches");
  pt->receiver[0] = pt->intersect[0];
  pt->sender[0] = pt->intersect[0];
  pt->lowercornerR[0] = pt->lowercorner[0];
  pt->uppercornerR[0] = pt->uppercorner[0];
  S.deposited = 1;
}

/* src/patches.c:292-367: production[f] = 0, then field 0 gets the deposit */
void patches_cells2envwait(systeminfo_t si, settings_t st, patches_t *pt, grid_t *grid, environment_t **env) {
  if (st.numberoffields == 0) return;
  ctx_for(si);
  if (!S.deposited) FAIL(si, "patches_cells2envwait without patches_cells2envinit");
  const int64_t nodes = grid_nodes(grid);
  for (int f = 1; f < st.numberoffields; f++) memset((*env)[f].production, 0, sizeof(double) * nodes);
  copy_to_host(si, (*env)[0].production, S.d_prod, sizeof(double) * nodes, "production");
  (void)pt;
}

static void free_env(void) {
  for (int f = 0; f < S.env_nf; f++) {
    if (S.d_envf[f]) tm_free(S.ctx, S.d_envf[f]);
    if (S.d_envg[f]) tm_free(S.ctx, S.d_envg[f]);
  }
  free(S.d_envf);
  free(S.d_envg);
  S.d_envf = S.d_envg = NULL;
  S.env_nf = 0;
  S.env_nodes = 0;
}

/* src/patches.c:380-505: copies the fields and their gradients the
 * interpolation reads (the reference packs them into commbuff here) */
void patches_env2cellsinit(systeminfo_t si, settings_t st, patches_t *pt, grid_t *grid, environment_t **env) {
  if (st.numberoffields == 0) return;
  ctx_for(si);
  const int64_t nodes = grid_nodes(grid);
  const int nf = st.numberoffields;
  if (nf != S.env_nf || nodes > S.env_nodes) {
    free_env();
    S.d_envf = calloc(nf, sizeof(double *));
    S.d_envg = calloc(nf, sizeof(double *));
    if (!S.d_envf || !S.d_envg) FAIL(si, "cannot allocate field copies");
    S.env_nf = nf;
    for (int f = 0; f < nf; f++) {
      check(si, tm_malloc(S.ctx, (void **)&S.d_envf[f], sizeof(double) * nodes), "field copy");
      check(si, tm_malloc(S.ctx, (void **)&S.d_envg[f], sizeof(double) * 3 * nodes), "gradient copy");
    }
    S.env_nodes = nodes;
  }
  for (int f = 0; f < nf; f++) {
    copy_to_device(si, S.d_envf[f], (*env)[f].data, sizeof(double) * nodes, "field upload");
    copy_to_device(si, S.d_envg[f], (*env)[f].gradient, sizeof(double) * 3 * nodes, "gradient upload");
  }
  S.env_ready = 1;
  (void)pt;
}

/* src/patches.c:512-593: cellenvdata[f][c] = CIC interpolation of field f and
 * its gradient; cellenvdata is allocated here and freed by the caller
 * (step_clean, src/step.c:102-108) */
void patches_env2cellswait(systeminfo_t si, settings_t st, patches_t *pt, cellsinfo_t *ci, grid_t *grid,
                           cellenvdata_t ***ced) {
  if (st.numberoffields == 0) return;
  ctx_for(si);
  require_build(si, ci);
  if (!S.env_ready || S.env_nf != st.numberoffields) FAIL(si, "patches_env2cellswait without patches_env2cellsinit");
  const int64_t n = (int64_t)ci->localcount.n;
  if (n > S.interp_cap) {
    if (S.d_interp) tm_free(S.ctx, S.d_interp);
    S.d_interp = NULL;
    check(si, tm_malloc(S.ctx, (void **)&S.d_interp, sizeof(double) * 4 * n), "interpolation buffer");
    S.interp_cap = n;
  }
  if (!((*ced) = (cellenvdata_t **)calloc(st.numberoffields, sizeof(cellenvdata_t *))))
    FAIL(si, "cannot allocate cellenvdata");
  tm_grid g = grid_of(grid, si.rank);
  for (int f = 0; f < st.numberoffields; f++) {
    /* (at least one element, so a cell-less rank still gets a pointer to free) */
    if (!((*ced)[f] = (cellenvdata_t *)calloc(n > 0 ? n : 1, sizeof(cellenvdata_t))))
      FAIL(si, "cannot allocate cellenvdata[f]");
    if (n == 0) continue;
    check(si, tm_cic_interp(S.ctx, &g, S.d_envf[f], S.d_envg[f], S.d_interp), "CIC interpolation");
    copy_to_host(si, (*ced)[f], S.d_interp, sizeof(cellenvdata_t) * n, "cellenvdata");
  }
  S.env_ready = 0;
  (void)pt;
}

/* src/fieldgradient.c:149-277 (no neighbouring ranks: one-sided faces) */
void fieldgradient(systeminfo_t si, settings_t st, environment_t **env, grid_t *grid) {
  if (st.numberoffields == 0) return;
  ctx_for(si);
  const int64_t nodes = grid_nodes(grid);
  ensure_grid(si, nodes);
  tm_grid g = grid_of(grid, si.rank);
  for (int f = 0; f < st.numberoffields; f++) {
    copy_to_device(si, S.d_field, (*env)[f].data, sizeof(double) * nodes, "field upload");
    check(si, tm_field_gradient(S.ctx, &g, S.d_field, S.d_grad), "field gradient");
    copy_to_host(si, (*env)[f].gradient, S.d_grad, sizeof(double) * 3 * nodes, "gradient");
  }
}

void patches_free(patches_t *pt) {
  free(pt->receiver);
  free(pt->sender);
  free(pt->intersect);
  free(pt->lowercorner);
  free(pt->uppercorner);
  free(pt->lowercornerR);
  free(pt->uppercornerR);
  free(pt->size);
  free(pt->data);
  memset(pt, 0, sizeof(*pt));
}

/*
 * oracle.c -- CPU restatement of Timothy's cell-mechanics hot path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker and the timed CPU
 * baseline, never linked into the product.
 *
 * Written from the reference's semantics (mcytu/timothy, the C sources under src/); the
 * floating-point expressions keep the reference's evaluation order so the
 * results are what its source computes under IEEE arithmetic.
 */
#include "oracle.h"

#include <float.h>
#include <stdio.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_ROOT_LEVEL 29
/* MAXVALUE is powf(2,29) in the reference (src/global.h:67); exact in float */
#define OR_MAXVALUE 536870912.0

/* ------------------------------------------------------------------ */
/* octree build: src/octree.c:40-207                                   */
/* ------------------------------------------------------------------ */

static int or_grow(or_tree *t) {
  if (t->nnodes + 1 < t->cap) return 0;
  int64_t ncap = t->cap ? t->cap * 2 : 1024;
  or_node *nn = (or_node *)realloc(t->nodes, sizeof(or_node) * (size_t)ncap);
  if (!nn) return -2;
  t->nodes = nn;
  t->cap = ncap;
  return 0;
}

/* src/octree.c:40-79 (octemptynode): appends a node; returns its index */
static int64_t or_emptynode(or_tree *t, int64_t father, unsigned level, unsigned xbit, unsigned ybit,
                            unsigned zbit) {
  if (or_grow(t)) return -1;
  int64_t k = t->nnodes;
  or_node *o = &t->nodes[k];
  if (father == -1) {
    o->xcode = o->ycode = o->zcode = 0;
    o->xlimit = o->ylimit = o->zlimit = 1u << level;
  } else {
    const or_node *f = &t->nodes[father];
    unsigned len = (f->xlimit - f->xcode) / 2;
    o->xcode = f->xcode | xbit;
    o->ycode = f->ycode | ybit;
    o->zcode = f->zcode | zbit;
    o->xlimit = o->xcode + len;
    o->ylimit = o->ycode + len;
    o->zlimit = o->zcode + len;
  }
  o->father = father;
  for (int i = 0; i < 8; i++) o->child[i] = -1;
  o->data = -1;
  o->level = level;
  t->nnodes++;
  return k;
}

/* child octant index at bit position `bit`: x + 2y + 4z (src/octree.c:101-107) */
static inline unsigned or_childindex(const uint32_t *lc, unsigned bit) {
  return ((lc[0] >> bit) & 1u) + 2u * ((lc[1] >> bit) & 1u) + 4u * ((lc[2] >> bit) & 1u);
}

/* src/octree.c:84-140 (octinsertcell).  `level` is the child bit of the
 * current node, as in the reference.  Returns -1 where the reference would
 * shift by a negative count (split below node level 3). */
static int or_insert(or_tree *t, int64_t c) {
  int64_t idx = 0;
  unsigned level = OR_ROOT_LEVEL - 1;
  const uint32_t *lc = &t->loc[3 * c];
  for (;;) {
    int32_t data = t->nodes[idx].data;
    if (data == -1) {
      t->nodes[idx].data = (int32_t)c;
      return 0;
    }
    if (level < 2) return -1;
    if (data == -2) {
      unsigned cbb = 1u << level;
      unsigned ci = or_childindex(lc, level);
      level--;
      int64_t child = t->nodes[idx].child[ci];
      if (child == -1) {
        child = or_emptynode(t, idx, level, lc[0] & cbb, lc[1] & cbb, lc[2] & cbb);
        if (child < 0) return -2;
        t->nodes[idx].child[ci] = child;
      }
      idx = child;
    } else {
      /* occupied leaf: push the resident cell one level down, retry here */
      const uint32_t *oc = &t->loc[3 * (int64_t)data];
      unsigned cbb = 1u << level;
      unsigned ci = or_childindex(oc, level);
      int64_t child = or_emptynode(t, idx, level - 1, oc[0] & cbb, oc[1] & cbb, oc[2] & cbb);
      if (child < 0) return -2;
      t->nodes[idx].child[ci] = child;
      t->nodes[child].data = data;
      t->nodes[idx].data = -2;
    }
  }
}

int or_octbuild(const or_cells *c, const or_celltype *types, or_tree *t) {
  memset(t, 0, sizeof(*t));
  if (c->n == 0) return 0; /* src/octree.c:152-153 */
  double bmin[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
  double bmax[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  const double epsilon = 0.01;
  for (int64_t i = 0; i < c->n; i++) { /* src/octree.c:161-174 */
    double x = c->x[i], y = c->y[i], z = c->z[i];
    double e = types[c->ctype[i]].h + epsilon;
    bmin[0] = (x - e < bmin[0] ? x - e : bmin[0]);
    bmax[0] = (x + e > bmax[0] ? x + e : bmax[0]);
    bmin[1] = (y - e < bmin[1] ? y - e : bmin[1]);
    bmax[1] = (y + e > bmax[1] ? y + e : bmax[1]);
    bmin[2] = (z - e < bmin[2] ? z - e : bmin[2]);
    bmax[2] = (z + e > bmax[2] ? z + e : bmax[2]);
  }
  double s = bmax[0] - bmin[0]; /* src/octree.c:175-180 */
  s = (s > bmax[1] - bmin[1] ? s : bmax[1] - bmin[1]);
  s = (s > bmax[2] - bmin[2] ? s : bmax[2] - bmin[2]);
  t->scale = s;
  t->shift[0] = bmin[0];
  t->shift[1] = bmin[1];
  t->shift[2] = bmin[2];
  t->loc = (uint32_t *)malloc(sizeof(uint32_t) * 3 * (size_t)c->n);
  if (!t->loc) return -2;
  for (int64_t i = 0; i < c->n; i++) { /* src/octree.c:187-191 */
    t->loc[3 * i + 0] = (uint32_t)(((c->x[i] - t->shift[0]) / t->scale) * OR_MAXVALUE);
    t->loc[3 * i + 1] = (uint32_t)(((c->y[i] - t->shift[1]) / t->scale) * OR_MAXVALUE);
    t->loc[3 * i + 2] = (uint32_t)(((c->z[i] - t->shift[2]) / t->scale) * OR_MAXVALUE);
  }
  if (or_emptynode(t, -1, OR_ROOT_LEVEL, 0, 0, 0) < 0) return -2;
  for (int64_t i = 0; i < c->n; i++) { /* src/octree.c:200-202 */
    int r = or_insert(t, i);
    if (r) return r;
  }
  return 0;
}

void or_treefree(or_tree *t) {
  free(t->loc);
  free(t->nodes);
  memset(t, 0, sizeof(*t));
}

void or_leaf_levels(const or_tree *t, int64_t n, int32_t *level) {
  for (int64_t i = 0; i < n; i++) level[i] = -1;
  for (int64_t k = 0; k < t->nnodes; k++)
    if (t->nodes[k].data >= 0) level[t->nodes[k].data] = (int32_t)t->nodes[k].level;
}

/* ------------------------------------------------------------------ */
/* region location: src/octree.c:246-316, src/inline.h:26-32           */
/* ------------------------------------------------------------------ */

void or_computebox(const or_cells *c, const or_celltype *types, const or_tree *t, int64_t p, uint32_t mn[3],
                   uint32_t mx[3]) {
  double h = types[c->ctype[p]].h;
  double pc[3] = {c->x[p], c->y[p], c->z[p]};
  for (int a = 0; a < 3; a++) {
    double lo = (pc[a] - h - t->shift[a]) / t->scale;
    double hi = (pc[a] + h - t->shift[a]) / t->scale;
    mn[a] = (uint32_t)(lo * OR_MAXVALUE);
    mx[a] = (uint32_t)(hi * OR_MAXVALUE);
  }
}

/* src/octree.c:271-290.  A missing child ends the descent at its parent
 * (the reference reads octree[-1] there, then takes the parent). */
static int64_t or_traversetolevel(const or_tree *t, unsigned level, const uint32_t *mn, unsigned lmin) {
  int64_t cellidx = 0, parent = 0;
  unsigned n = level - lmin + 1;
  while (n--) {
    unsigned ci = or_childindex(mn, level);
    level--;
    parent = cellidx;
    cellidx = t->nodes[cellidx].child[ci];
    if (cellidx == -1) break;
    if (t->nodes[cellidx].data != -2) break;
  }
  if (cellidx == -1) cellidx = parent;
  return cellidx;
}

int64_t or_locateregion(const or_tree *t, const uint32_t mn[3], const uint32_t mx[3]) {
  uint32_t d[3] = {mn[0] ^ mx[0], mn[1] ^ mx[1], mn[2] ^ mx[2]};
  unsigned l1 = OR_ROOT_LEVEL, l2 = OR_ROOT_LEVEL, lmin = OR_ROOT_LEVEL;
  while (!(d[0] & (1u << l1)) && l1) l1--;
  while (!(d[1] & (1u << l2)) && (l2 > l1)) l2--;
  while (!(d[2] & (1u << lmin)) && (lmin > l2)) lmin--;
  lmin++;
  return or_traversetolevel(t, OR_ROOT_LEVEL - 1, mn, lmin);
}

static inline int or_intersects(const or_node *o, const uint32_t *mn, const uint32_t *mx) {
  if (mx[0] <= o->xcode || mn[0] >= o->xlimit) return 0;
  if (mx[1] <= o->ycode || mn[1] >= o->ylimit) return 0;
  if (mx[2] <= o->zcode || mn[2] >= o->zlimit) return 0;
  return 1;
}

/* octheap_t (src/octree.c:332-371): growable int stack, one malloc per walk
 * like the reference so the timed baseline pays what the reference pays. */
typedef struct {
  int size, count;
  int64_t *data;
} or_heap;

static inline void or_heapinit(or_heap *h) {
  h->size = 64;
  h->count = 0;
  h->data = (int64_t *)malloc(sizeof(int64_t) * (size_t)h->size);
}
static inline void or_heappush(or_heap *h, int64_t v) {
  if (h->count == h->size) {
    h->size += 64;
    h->data = (int64_t *)realloc(h->data, sizeof(int64_t) * (size_t)h->size);
  }
  h->data[h->count++] = v;
}
static inline int64_t or_heappop(or_heap *h) { return h->data[--h->count]; }
static inline void or_heapfree(or_heap *h) { free(h->data); }

/* start node + its intersecting children (src/potential.c:108-116) */
static void or_walk_start(const or_tree *t, const uint32_t *mn, const uint32_t *mx, or_heap *h) {
  int64_t start = or_locateregion(t, mn, mx);
  for (int s = 0; s < 8; s++) {
    int64_t idx = t->nodes[start].child[s];
    if (idx != -1 && or_intersects(&t->nodes[idx], mn, mx)) or_heappush(h, idx);
  }
}

/* ------------------------------------------------------------------ */
/* pair formulas: src/inline.h:40-83, src/potential.c:39-78            */
/* ------------------------------------------------------------------ */

float or_sphkernel(int dimension, double dist, double h, double h2, double h3) {
  double u, c = 1.0;
  if (dist < 0.0) return 0.0f;
  u = dist / h;
  if (dimension == 2) c = 40 / (7 * M_PI * h2);
  if (dimension == 3) c = 8 / (M_PI * h3);
  if (u >= 0.0 && u <= 0.5) return (float)(c * (1 - 6 * u * u + 6 * u * u * u));
  if (u > 0.5 && u <= 1.0) return (float)(c * (2 * (1 - u) * (1 - u) * (1 - u)));
  return 0.0f;
}

void or_sphgradient(int dimension, double px, double py, double pz, double qx, double qy, double qz, double h,
                    double h3, double h4, double f[3]) {
  double u = 1.0, c = 1.0, w = 1.0, dist;
  f[0] = f[1] = f[2] = 0.0;
  dist = sqrt((px - qx) * (px - qx) + (py - qy) * (py - qy) + (pz - qz) * (pz - qz));
  if (dist >= 0.0 && dist <= h) {
    u = dist / h;
    if (dimension == 2) c = (40 * 8) / (7 * M_PI * h3);
    if (dimension == 3) c = 48 / (M_PI * h4);
    if (u >= 0.0 && u <= 0.5) w = -2.0 * u + 3.0 * u * u;
    if (u > 0.5 && u <= 1.0) w = -(1.0 - u) * (1.0 - u);
    if (u > 1.0) w = 0.0;
    f[0] = w * c * (qx - px) / dist;
    f[1] = w * c * (qy - py) / dist;
    if (dimension == 3) f[2] = w * c * (qz - pz) / dist;
  }
}

double or_potential(int dimension, double px, double py, double pz, double psize, float pyoung, double qx,
                    double qy, double qz, double qsize, float qyoung, const or_celltype *ptype, double *density,
                    double *mindist) {
  double dist, h, h2, h3, xc, D, csize;
  const double poisson = 0.33;
  dist = sqrt((px - qx) * (px - qx) + (py - qy) * (py - qy) + (pz - qz) * (pz - qz));
  *mindist = (dist < *mindist ? dist : *mindist);
  h = ptype->h;
  h2 = ptype->h2;
  h3 = ptype->h3;
  csize = ptype->size;
  if (dist <= h) {
    double r01, r02, area, sc;
    sc = (dimension == 2 ? h2 : h3);
    *density += sc * (qsize / csize) * or_sphkernel(dimension, dist, h, h2, h3);
    xc = psize + qsize - dist;
    if (xc <= 0.0) return 0.0;
    D = 0.75 * ((1.0 - poisson * poisson) / pyoung + (1.0 - poisson * poisson / qyoung));
    r01 = (psize * psize - qsize * qsize + dist * dist) / (2 * dist);
    r02 = dist - r01;
    area = M_PI * ((psize * psize * (psize - r01) - (psize * psize * psize - r01 * r01 * r01) / 3) +
                   (qsize * qsize * (qsize - r02) - (qsize * qsize * qsize - r02 * r02 * r02) / 3));
    /* pow(xc, 5 / 2) is pow(xc, 2): integer division (src/potential.c:75);
       written as the product, which is what pow(x, 2.0) returns */
    return (2.0 * (xc * xc) / (5.0 * D)) * sqrt((psize * qsize) / (psize + qsize)) + area * 0.1;
  }
  return 0.0;
}

/* ------------------------------------------------------------------ */
/* tree walks: src/potential.c:87-138 and :209-259                     */
/* ------------------------------------------------------------------ */

void or_computepotential(const or_cells *c, const or_celltype *types, const or_tree *t, int dimension,
                         double *density, double *v, double *mindist) {
  if (c->n <= 1) return;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t p = 0; p < c->n; p++) {
    or_heap h;
    uint32_t mn[3], mx[3];
    const or_celltype *pt = &types[c->ctype[p]];
    or_heapinit(&h);
    double dens = 0.0, vv = 0.0, md = DBL_MAX;
    or_computebox(c, types, t, p, mn, mx);
    or_walk_start(t, mn, mx, &h);
    while (h.count > 0) {
      int64_t idx = or_heappop(&h);
      int32_t cell = t->nodes[idx].data;
      if (cell >= 0) {
        if (cell != p)
          vv = or_potential(dimension, c->x[p], c->y[p], c->z[p], c->size[p], c->young[p], c->x[cell], c->y[cell],
                            c->z[cell], c->size[cell], c->young[cell], pt, &dens, &md);
      } else {
        for (int s = 0; s < 8; s++) {
          int64_t ni = t->nodes[idx].child[s];
          if (ni != -1 && or_intersects(&t->nodes[ni], mn, mx)) or_heappush(&h, ni);
        }
      }
    }
    or_heapfree(&h);
    density[p] = dens;
    v[p] = vv;
    mindist[p] = md;
  }
}

void or_computegradient(const or_cells *c, const or_celltype *types, const or_tree *t, int dimension, double *fx,
                        double *fy, double *fz) {
  if (c->n <= 1) return;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t p = 0; p < c->n; p++) {
    or_heap h;
    uint32_t mn[3], mx[3];
    const or_celltype *pt = &types[c->ctype[p]];
    or_heapinit(&h);
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    or_computebox(c, types, t, p, mn, mx);
    or_walk_start(t, mn, mx, &h);
    while (h.count > 0) {
      int64_t idx = or_heappop(&h);
      int32_t cell = t->nodes[idx].data;
      if (cell >= 0) {
        if (cell != p) {
          double f[3];
          or_sphgradient(dimension, c->x[p], c->y[p], c->z[p], c->x[cell], c->y[cell], c->z[cell], pt->h, pt->h3,
                         pt->h4, f);
          f0 += f[0];
          f1 += f[1];
          f2 += f[2];
        }
      } else {
        for (int s = 0; s < 8; s++) {
          int64_t ni = t->nodes[idx].child[s];
          if (ni != -1 && or_intersects(&t->nodes[ni], mn, mx)) or_heappush(&h, ni);
        }
      }
    }
    or_heapfree(&h);
    fx[p] = f0;
    fy[p] = f1;
    fz[p] = f2;
  }
}

int64_t or_walk_candidates(const or_cells *c, const or_celltype *types, const or_tree *t, int64_t *cand_off,
                           int32_t *cand, uint8_t *in_h) {
  int64_t total = 0;
  if (c->n <= 1) {
    for (int64_t p = 0; p <= c->n; p++) cand_off[p] = 0;
    return 0;
  }
  /* serial: the output order is the pop order, offsets come from a count pass */
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1 && !cand) break;
    total = 0;
    for (int64_t p = 0; p < c->n; p++) {
      or_heap h;
      uint32_t mn[3], mx[3];
      double hp = types[c->ctype[p]].h;
      if (pass == 0) cand_off[p] = total;
      int64_t k = cand_off[p];
      or_heapinit(&h);
      or_computebox(c, types, t, p, mn, mx);
      or_walk_start(t, mn, mx, &h);
      while (h.count > 0) {
        int64_t idx = or_heappop(&h);
        int32_t cell = t->nodes[idx].data;
        if (cell >= 0) {
          if (cell != p) {
            if (pass == 1) {
              double dx = c->x[p] - c->x[cell], dy = c->y[p] - c->y[cell], dz = c->z[p] - c->z[cell];
              double d = sqrt(dx * dx + dy * dy + dz * dz);
              cand[k] = cell;
              if (in_h) in_h[k] = (uint8_t)(d <= hp);
            }
            k++;
            total++;
          }
        } else {
          for (int s = 0; s < 8; s++) {
            int64_t ni = t->nodes[idx].child[s];
            if (ni != -1 && or_intersects(&t->nodes[ni], mn, mx)) or_heappush(&h, ni);
          }
        }
      }
      or_heapfree(&h);
    }
    if (pass == 0) cand_off[c->n] = total;
  }
  return total;
}

int64_t or_brute_neighbors(const or_cells *c, const or_celltype *types, int64_t *off, int32_t *nbr) {
  int64_t n = c->n;
  int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t p = 0; p < n; p++) {
    double hp = types[c->ctype[p]].h;
    int64_t k = 0;
    for (int64_t q = 0; q < n; q++) {
      if (q == p) continue;
      double dx = c->x[p] - c->x[q], dy = c->y[p] - c->y[q], dz = c->z[p] - c->z[q];
      if (sqrt(dx * dx + dy * dy + dz * dz) <= hp) k++;
    }
    cnt[p] = k;
  }
  int64_t total = 0;
  for (int64_t p = 0; p < n; p++) {
    off[p] = total;
    total += cnt[p];
  }
  off[n] = total;
  free(cnt);
  if (!nbr) return total;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t p = 0; p < n; p++) {
    double hp = types[c->ctype[p]].h;
    int64_t k = off[p];
    for (int64_t q = 0; q < n; q++) {
      if (q == p) continue;
      double dx = c->x[p] - c->x[q], dy = c->y[p] - c->y[q], dz = c->z[p] - c->z[q];
      if (sqrt(dx * dx + dy * dy + dz * dz) <= hp) nbr[k++] = (int32_t)q;
    }
  }
  return total;
}

/* ------------------------------------------------------------------ */
/* CIC and field gradient: src/patches.c:156-234,512-593,              */
/* src/fieldgradient.c:149-224                                         */
/* ------------------------------------------------------------------ */

static inline int or_inrange(const int64_t ci[3], int ax, int ay, int az, const int64_t lo[3], const int64_t hi[3]) {
  return ci[0] + ax >= lo[0] && ci[1] + ay >= lo[1] && ci[2] + az >= lo[2] && ci[0] + ax <= hi[0] &&
         ci[1] + ay <= hi[1] && ci[2] + az <= hi[2];
}

/* cicCoord = lower + cellIdx*resolution: int64 * float is evaluated in float
 * (C usual arithmetic conversions), then widened for the double sum. */
static inline void or_cicweights(double px, double py, double pz, const double lower[3], float res, int64_t ci[3],
                                 double d[3], double tt[3]) {
  ci[0] = (int64_t)((px - lower[0]) / res);
  ci[1] = (int64_t)((py - lower[1]) / res);
  ci[2] = (int64_t)((pz - lower[2]) / res);
  double pc[3] = {px, py, pz};
  for (int a = 0; a < 3; a++) {
    float off = (float)ci[a] * res;
    double cic = lower[a] + off;
    d[a] = (pc[a] - cic) / res;
    tt[a] = 1.0 - d[a];
  }
}

void or_cic_deposit(const or_cells *c, int weight_mode, const double lower[3], float resolution,
                    const int64_t loweridx[3], const int64_t upperidx[3], const int64_t localsize[3],
                    double *production) {
  int64_t G = localsize[0] * localsize[1] * localsize[2];
  memset(production, 0, sizeof(double) * (size_t)G);
  for (int64_t i = 0; i < c->n; i++) {
    int64_t ci[3];
    double d[3], t[3];
    or_cicweights(c->x[i], c->y[i], c->z[i], lower, resolution, ci, d, t);
    double wt;
    if (weight_mode == 1) {
      double r = c->size[i];
      wt = ((4.0 / 3.0) * M_PI) * (r * r * r);
    } else {
      if (c->phase[i] == 5) continue;
      wt = (c->phase[i] == 0 ? 0.75 : 1.0);
    }
    for (int ax = 0; ax < 2; ax++)
      for (int ay = 0; ay < 2; ay++)
        for (int az = 0; az < 2; az++) {
          if (!or_inrange(ci, ax, ay, az, loweridx, upperidx)) continue;
          int64_t gx = ci[0] + ax - loweridx[0], gy = ci[1] + ay - loweridx[1], gz = ci[2] + az - loweridx[2];
          production[(gx * localsize[1] + gy) * localsize[2] + gz] +=
              wt * (ax * d[0] + (1 - ax) * t[0]) * (ay * d[1] + (1 - ay) * t[1]) * (az * d[2] + (1 - az) * t[2]);
        }
  }
}

void or_cic_interp(const or_cells *c, const double lower[3], float resolution, const int64_t loweridx[3],
                   const int64_t upperidx[3], const int64_t localsize[3], const double *field,
                   const double *gradient, double *out) {
  for (int64_t i = 0; i < c->n; i++) {
    int64_t ci[3];
    double d[3], t[3];
    double val = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    or_cicweights(c->x[i], c->y[i], c->z[i], lower, resolution, ci, d, t);
    for (int ax = 0; ax < 2; ax++)
      for (int ay = 0; ay < 2; ay++)
        for (int az = 0; az < 2; az++) {
          if (!or_inrange(ci, ax, ay, az, loweridx, upperidx)) continue;
          int64_t nx = ci[0] + ax - loweridx[0], ny = ci[1] + ay - loweridx[1], nz = ci[2] + az - loweridx[2];
          int64_t node = (nx * localsize[1] + ny) * localsize[2] + nz;
          double wx = (ax * d[0] + (1 - ax) * t[0]);
          double wy = (ay * d[1] + (1 - ay) * t[1]);
          double wz = (az * d[2] + (1 - az) * t[2]);
          val += field[node] * wx * wy * wz;
          gx += gradient[3 * node + 0] * wx * wy * wz;
          gy += gradient[3 * node + 1] * wx * wy * wz;
          gz += gradient[3 * node + 2] * wx * wy * wz;
        }
    out[4 * i + 0] = val;
    out[4 * i + 1] = gx;
    out[4 * i + 2] = gy;
    out[4 * i + 3] = gz;
  }
}

void or_fieldgradient(const int64_t ls[3], float resolution, const double *f, double *g) {
  int64_t lx = ls[0], ly = ls[1], lz = ls[2];
#define NODE(i, j, k) (((i) * ly + (j)) * lz + (k))
  for (int64_t i = 0; i < lx; i++)
    for (int64_t j = 0; j < ly; j++)
      for (int64_t k = 0; k < lz; k++) {
        if (i != 0 && i != lx - 1) g[3 * NODE(i, j, k) + 0] = (f[NODE(i + 1, j, k)] - f[NODE(i - 1, j, k)]) / resolution;
        if (j != 0 && j != ly - 1) g[3 * NODE(i, j, k) + 1] = (f[NODE(i, j + 1, k)] - f[NODE(i, j - 1, k)]) / resolution;
        if (k != 0 && k != lz - 1) g[3 * NODE(i, j, k) + 2] = (f[NODE(i, j, k + 1)] - f[NODE(i, j, k - 1)]) / resolution;
      }
  for (int64_t j = 0; j < ly; j++)
    for (int64_t k = 0; k < lz; k++) {
      double x0 = f[NODE(0, j, k)], x1 = f[NODE(lx - 1, j, k)];
      g[3 * NODE(0, j, k) + 0] = (f[NODE(1, j, k)] - x0) / resolution;
      g[3 * NODE(lx - 1, j, k) + 0] = (x1 - f[NODE(lx - 2, j, k)]) / resolution;
    }
  for (int64_t i = 0; i < lx; i++)
    for (int64_t k = 0; k < lz; k++) {
      double y0 = f[NODE(i, 0, k)], y1 = f[NODE(i, ly - 1, k)];
      g[3 * NODE(i, 0, k) + 1] = (f[NODE(i, 1, k)] - y0) / resolution;
      g[3 * NODE(i, ly - 1, k) + 1] = (y1 - f[NODE(i, ly - 2, k)]) / resolution;
    }
  for (int64_t i = 0; i < lx; i++)
    for (int64_t j = 0; j < ly; j++) {
      double z0 = f[NODE(i, j, 0)], z1 = f[NODE(i, j, lz - 1)];
      g[3 * NODE(i, j, 0) + 2] = (f[NODE(i, j, 1)] - z0) / resolution;
      g[3 * NODE(i, j, lz - 1) + 2] = (z1 - f[NODE(i, j, lz - 2)]) / resolution;
    }
#undef NODE
}

/* Philox4x32-10: one round multiplies counter words 0 and 2 by the two
 * Philox multipliers and mixes the 64-bit products' halves with words 1, 3
 * and the key; the key advances by the Weyl constants between rounds. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

void or_philox_draws(int64_t n, int dimension, uint64_t seed, uint64_t step, float *draws) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int64_t c = 0; c < n; c++) {
    const uint32_t ctr[4] = {(uint32_t)c, (uint32_t)((uint64_t)c >> 32), (uint32_t)step, (uint32_t)(step >> 32)};
    uint32_t w[4];
    or_philox4x32_10(ctr, key, w);
    for (int a = 0; a < dimension; a++) draws[dimension * c + a] = (float)(w[a] >> 1) / 2147483648.0f;
  }
}

/* src/cells.c:41-58 */
static int outsidethebox(int dimension, double x, double y, double z, double r, const double box[3]) {
  if (x - r <= -box[0] / 2.0 || x + r >= box[0] / 2.0) return 1;
  if (y - r <= -box[1] / 2.0 || y + r >= box[1] / 2.0) return 1;
  if (dimension == 3 && (z - r <= -box[2] / 2.0 || z + r >= box[2] / 2.0)) return 1;
  return 0;
}

/* src/cells.c:342-365 */
int64_t or_updatepositions(int64_t n, int dimension, double *x, double *y, double *z, const double *size,
                           const int32_t *ctype, const float *randommove, const double *fx, const double *fy,
                           const double *fz, const float *draws, const double box[3], uint8_t *removecell) {
  int64_t removed = 0, k = 0;
  for (int64_t c = 0; c < n; c++) {
    const float rm = randommove[ctype[c]];
    x[c] += fx[c];
    y[c] += fy[c];
    if (dimension == 3) z[c] += fz[c];
    const float ux = draws ? draws[k++] : (float)rand() / RAND_MAX;
    x[c] += rm * (ux * 2.0 - 1.0);
    const float uy = draws ? draws[k++] : (float)rand() / RAND_MAX;
    y[c] += rm * (uy * 2.0 - 1.0);
    if (dimension == 3) {
      const float uz = draws ? draws[k++] : (float)rand() / RAND_MAX;
      z[c] += rm * (uz * 2.0 - 1.0);
    }
    if (outsidethebox(dimension, x[c], y[c], z[c], size[c], box) && removecell[c] == 0) {
      removecell[c] = 1;
      removed++;
    }
  }
  return removed;
}

/* src/environment.c:228-243 (matrix) and :309-381 / :447-483 (right-hand
 * side): the reference's HYPRE assembly restated on the [i][j][k] layout */
static double diffusion_z(float res, double D, double dt) {
  const float r2 = res * res;
  return D * dt / r2;
}

void or_diffusion_rhs(const int64_t ls[3], float res, double D, double lambda, double bc, double dt,
                      const double *src, const double *uold, double *b) {
  const int64_t lx = ls[0], ly = ls[1], lz = ls[2];
  const double z = diffusion_z(res, D, dt), zbc = z * bc;
  (void)lambda;
  for (int64_t i = 0; i < lx; i++)
    for (int64_t j = 0; j < ly; j++)
      for (int64_t k = 0; k < lz; k++) {
        const int64_t n = (i * ly + j) * lz + k;
        double v = uold[n];
        if (i == 0) v += zbc;
        if (i == lx - 1) v += zbc;
        if (j == 0) v += zbc;
        if (j == ly - 1) v += zbc;
        if (k == 0) v += zbc;
        if (k == lz - 1) v += zbc;
        if (src) v += src[n];
        b[n] = v;
      }
}

void or_diffusion_apply(const int64_t ls[3], float res, double D, double lambda, double dt, const double *u,
                        double *Au) {
  const int64_t lx = ls[0], ly = ls[1], lz = ls[2];
  const double z = diffusion_z(res, D, dt);
  const double diag = 1 + 6.0 * z + lambda * dt;
  for (int64_t i = 0; i < lx; i++)
    for (int64_t j = 0; j < ly; j++)
      for (int64_t k = 0; k < lz; k++) {
        const int64_t n = (i * ly + j) * lz + k;
        double nb = 0.0;
        if (i > 0) nb += u[n - ly * lz];
        if (i < lx - 1) nb += u[n + ly * lz];
        if (j > 0) nb += u[n - lz];
        if (j < ly - 1) nb += u[n + lz];
        if (k > 0) nb += u[n - 1];
        if (k < lz - 1) nb += u[n + 1];
        Au[n] = diag * u[n] - z * nb;
      }
}

int or_diffusion_solve(const int64_t ls[3], float res, double D, double lambda, double bc, double dt,
                       const double *src, double *x, double tol, int maxiter, int substeps) {
  const int64_t n = ls[0] * ls[1] * ls[2];
  double *b = malloc(sizeof(double) * n), *r = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
  int total = 0;
  for (int s = 0; s < substeps; s++) {
    or_diffusion_rhs(ls, res, D, lambda, bc, dt, src, x, b);
    or_diffusion_apply(ls, res, D, lambda, dt, x, Ap);
    double rr = 0.0, bb = 0.0;
    for (int64_t q = 0; q < n; q++) {
      r[q] = b[q] - Ap[q];
      p[q] = r[q];
      rr += r[q] * r[q];
      bb += b[q] * b[q];
    }
    for (int it = 0; it < maxiter && rr > tol * tol * bb; it++) {
      or_diffusion_apply(ls, res, D, lambda, dt, p, Ap);
      double pAp = 0.0;
      for (int64_t q = 0; q < n; q++) pAp += p[q] * Ap[q];
      if (!(pAp > 0.0)) break;
      const double alpha = rr / pAp;
      double rn = 0.0;
      for (int64_t q = 0; q < n; q++) {
        x[q] += alpha * p[q];
        r[q] -= alpha * Ap[q];
        rn += r[q] * r[q];
      }
      const double beta = rn / rr;
      rr = rn;
      for (int64_t q = 0; q < n; q++) p[q] = r[q] + beta * p[q];
      total++;
    }
  }
  free(b);
  free(r);
  free(p);
  free(Ap);
  return total;
}

/* src/io.c:372-516 with swapnbyte (the host is little-endian) */
static void put_be32(FILE *f, uint32_t v) {
  const unsigned char b[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8),
                              (unsigned char)v};
  fwrite(b, 1, 4, f);
}
static void put_f32(FILE *f, float v) {
  uint32_t u;
  memcpy(&u, &v, 4);
  put_be32(f, u);
}

int or_writevtk(const char *path, int64_t n, const double *x, const double *y, const double *z,
                const int32_t *ctype, const double *fx, const double *fy, const double *fz, int rank) {
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  fprintf(f, "# vtk DataFile Version 2.0\nTimothy output\nBINARY\nDATASET UNSTRUCTURED_GRID");
  fprintf(f, "\nPOINTS %lld float\n", (long long)n);
  for (int64_t i = 0; i < n; i++) {
    put_f32(f, (float)x[i]);
    put_f32(f, (float)y[i]);
    put_f32(f, (float)z[i]);
  }
  fprintf(f, "\nCELL_TYPES %lld\n", (long long)n);
  for (int64_t i = 0; i < n; i++) put_be32(f, 1u);
  fprintf(f, "\nPOINT_DATA %lld", (long long)n);
  fprintf(f, "\nSCALARS rank integer 1\nLOOKUP_TABLE default\n");
  for (int64_t i = 0; i < n; i++) put_be32(f, (uint32_t)rank);
  fprintf(f, "\nSCALARS type integer 1\nLOOKUP_TABLE default\n");
  for (int64_t i = 0; i < n; i++) put_be32(f, (uint32_t)ctype[i]);
  fprintf(f, "\nSCALARS density float 1\nLOOKUP_TABLE default\n");
  for (int64_t i = 0; i < n; i++) put_f32(f, 1.0f);
  fprintf(f, "\nVECTORS force float\n");
  for (int64_t i = 0; i < n; i++) {
    put_f32(f, (float)fx[i]);
    put_f32(f, (float)fy[i]);
    put_f32(f, (float)fz[i]);
  }
  return fclose(f) == 0 ? 0 : -1;
}

double or_wtime(void) {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return 0.0;
#endif
}

int or_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/*
 * step_harness.c -- drives the reference's cell-mechanics C API
 * (timothy_compat.h, served by libtimothy_compat.so) in the order of the
 * reference's step_compute (src/step.c:41-100), one rank, step > 0.
 *
 * Also walks the octree that octbuild left in cellsinfo->octree with the
 * reference's own traversal (octcomputebox, octlocateregion, octheap*,
 * strict octant overlap, src/potential.c:108-133) for the first `nwalk`
 * cells, and records the candidates in pop order.
 *
 * usage: step_harness <input.bin> <output.bin>
 * input:  int64 n, ntypes, G, nfields, nwalk;  ntypes x {size,h,h2,h3,h4};
 *         grid {lower[3] double, res float(as double)};  n x {x,y,z,size} double;
 *         n x young float; n x ctype int32; n x phase int32;
 *         nfields x (G^3 data double, 3 G^3 gradient double)
 * output: affShift[3], affScale, octsize (as double);  n x {density, v, mindist,
 *         fx, fy, fz};  nfields x G^3 production;  nfields x n x {value,gx,gy,gz};
 *         nfields x 3G^3 gradient;  int64 walk offsets[nwalk+1], int32 walk cells
 */
#include <stdio.h>
#include <sys/stat.h>
#include <stdlib.h>
#include <string.h>

#include "timothy_compat.h"

static void rd(FILE *f, void *p, size_t sz, size_t n) {
  if (fread(p, sz, n, f) != n) {
    fprintf(stderr, "short read\n");
    exit(2);
  }
}

static int intersects(const octnode_t *o, uint3dv_t mn, uint3dv_t mx) {
  if (mx.x <= o->xcode || mn.x >= o->xlimit) return 0;
  if (mx.y <= o->ycode || mn.y >= o->ylimit) return 0;
  if (mx.z <= o->zcode || mn.z >= o->zlimit) return 0;
  return 1;
}

/* usage: step_harness IN OUT [SEED RANDOMMOVE MOVE_OUT]: with the last three,
 * cellsupdate's updatepositions (src/cells.c:538) follows the step after
 * srand(SEED), every cell type taking RANDOMMOVE, and MOVE_OUT receives x, y,
 * z, removecell, removecount and the next rand() value */
int main(int argc, char **argv) {
  if (argc != 3 && argc != 6) return 1;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 1;
  int64_t hdr[5];
  rd(f, hdr, 8, 5);
  const int64_t n = hdr[0], ntypes = hdr[1], G = hdr[2], nf = hdr[3], nwalk = hdr[4];
  const int64_t nodes = G * G * G;

  celltype_t *ct = calloc(ntypes, sizeof(celltype_t));
  for (int k = 0; k < ntypes; k++) {
    double t[5];
    rd(f, t, 8, 5);
    ct[k].size = t[0];
    ct[k].h = t[1];
    ct[k].h2 = t[2];
    ct[k].h3 = t[3];
    ct[k].h4 = t[4];
  }
  double g4[4];
  rd(f, g4, 8, 4);

  cellsinfo_t ci;
  memset(&ci, 0, sizeof(ci));
  ci.localcount.n = (uint64_t)n;
  ci.dimension = 3;
  ci.cells = calloc(n, sizeof(celldata_t));
  ci.forces = calloc(n, sizeof(double3dv_t));
  double *pos = malloc(sizeof(double) * 4 * n);
  float *young = malloc(sizeof(float) * n);
  int32_t *ctype = malloc(sizeof(int32_t) * n), *phase = malloc(sizeof(int32_t) * n);
  rd(f, pos, 8, 4 * n);
  rd(f, young, 4, n);
  rd(f, ctype, 4, n);
  rd(f, phase, 4, n);
  for (int64_t i = 0; i < n; i++) {
    ci.cells[i].x = pos[4 * i];
    ci.cells[i].y = pos[4 * i + 1];
    ci.cells[i].z = pos[4 * i + 2];
    ci.cells[i].size = pos[4 * i + 3];
    ci.cells[i].young = young[i];
    ci.cells[i].ctype = ctype[i];
    ci.cells[i].phase = phase[i];
    ci.cells[i].density = -1.0; /* overwritten by computepotential */
  }

  grid_t grid;
  memset(&grid, 0, sizeof(grid));
  grid.lowercorner.x = g4[0];
  grid.lowercorner.y = g4[1];
  grid.lowercorner.z = g4[2];
  grid.resolution = (float)g4[3];
  grid.globalsize.x = grid.globalsize.y = grid.globalsize.z = G;
  grid.localsize = grid.globalsize;
  grid.loweridx = calloc(1, sizeof(int643dv_t));
  grid.upperidx = calloc(1, sizeof(int643dv_t));
  grid.upperidx[0].x = grid.upperidx[0].y = grid.upperidx[0].z = G - 1;

  environment_t *env = calloc(nf, sizeof(environment_t));
  for (int64_t k = 0; k < nf; k++) {
    env[k].data = malloc(sizeof(double) * nodes);
    env[k].gradient = malloc(sizeof(double) * 3 * nodes);
    env[k].production = malloc(sizeof(double) * nodes);
    rd(f, env[k].data, 8, nodes);
    rd(f, env[k].gradient, 8, 3 * nodes);
    for (int64_t i = 0; i < nodes; i++) env[k].production[i] = 12345.0;
  }
  fclose(f);

  systeminfo_t si;
  memset(&si, 0, sizeof(si));
  si.rank = 0;
  si.size = 1;
  settings_t st;
  memset(&st, 0, sizeof(st));
  st.numberoffields = (int)nf;
  st.numberofcelltypes = (int)ntypes;
  st.dimension = 3;
  st.step = 1;
  cellcommdata_t cc;
  memset(&cc, 0, sizeof(cc));
  cellenvdata_t **ced = NULL;

  /* STEP_HARNESS_PREBUILD=free|held: first build the octree of a moved
   * copy of the cells (released with octfree before the step, or held
   * until the end), so the step's octbuild reuses the layer's kept octree
   * buffer or, while the copy holds it, gets a plain allocation */
  const char *pre = getenv("STEP_HARNESS_PREBUILD");
  cellsinfo_t ci2;
  memset(&ci2, 0, sizeof(ci2));
  if (pre && *pre) {
    ci2 = ci;
    ci2.cells = calloc(n, sizeof(celldata_t));
    for (int64_t i = 0; i < n; i++) {
      ci2.cells[i] = ci.cells[i];
      ci2.cells[i].x = 0.9 * ci.cells[i].x + 1.0;
      ci2.cells[i].y = 0.8 * ci.cells[i].y - 2.0;
    }
    octbuild(si, &ci2, ct);
    if (!strcmp(pre, "free")) octfree(&ci2);
  }

  /* src/main.c:70 then src/step.c:47-97 */
  octbuild(si, &ci, ct);
  patches_t patches;
  patches_alloc(si, st, &patches, &ci, &grid);
  patches_cells2envinit(si, st, &patches, &ci, &grid);
  computepotential(si, &ci, ct, cc);
  if (st.step > 0) patches_cells2envwait(si, st, &patches, &grid, &env);
  computeremotepotential(si, &ci, ct, cc);
  computegradient(si, &ci, ct, cc);
  patches_env2cellsinit(si, st, &patches, &grid, &env);
  patches_env2cellswait(si, st, &patches, &ci, &grid, &ced);
  patches_free(&patches);
  fieldgradient(si, st, &env, &grid);
  computeremotegradient(si, &ci, ct, cc);

  /* the reference's traversal over cellsinfo->octree (src/potential.c:108-133) */
  int64_t *woff = calloc(nwalk + 1, sizeof(int64_t));
  int64_t wcap = 1024, wn = 0;
  int32_t *wc = malloc(sizeof(int32_t) * wcap);
  for (int64_t p = 0; p < nwalk; p++) {
    uint3dv_t mn, mx;
    octheap_t h;
    woff[p] = wn;
    octheapinit(si, &h);
    octcomputebox(p, &mn, &mx, ci, ct);
    int start = octlocateregion(mn, mx, ci);
    for (int s = 0; s < 8; s++) {
      int64_t idx = ci.octree[start].child[s];
      if (idx != -1 && intersects(&ci.octree[idx], mn, mx)) octheappush(si, &h, (int)idx);
    }
    while (h.count > 0) {
      int idx = octheappop(&h);
      int cell = ci.octree[idx].data;
      if (cell >= 0) {
        if (cell != p) {
          if (wn == wcap) wc = realloc(wc, sizeof(int32_t) * (wcap *= 2));
          wc[wn++] = cell;
        }
      } else {
        for (int s = 0; s < 8; s++) {
          int64_t ni = ci.octree[idx].child[s];
          if (ni != -1 && intersects(&ci.octree[ni], mn, mx)) octheappush(si, &h, (int)ni);
        }
      }
    }
    octheapfree(&h);
  }
  woff[nwalk] = wn;

  FILE *o = fopen(argv[2], "wb");
  double a[5] = {affShift.x, affShift.y, affShift.z, affScale, (double)ci.octsize};
  fwrite(a, 8, 5, o);
  for (int64_t i = 0; i < n; i++) {
    double r[6] = {ci.cells[i].density, ci.cells[i].v, ci.cells[i].mindist,
                   ci.forces[i].x, ci.forces[i].y, ci.forces[i].z};
    fwrite(r, 8, 6, o);
  }
  for (int64_t k = 0; k < nf; k++) fwrite(env[k].production, 8, nodes, o);
  for (int64_t k = 0; k < nf; k++) fwrite(ced[k], sizeof(cellenvdata_t), n, o);
  for (int64_t k = 0; k < nf; k++) fwrite(env[k].gradient, 8, 3 * nodes, o);
  fwrite(woff, 8, nwalk + 1, o);
  fwrite(wc, 4, wn, o);
  fclose(o);

  if (argc == 6) {
    srand((unsigned)atoi(argv[3]));
    for (int k = 0; k < ntypes; k++) ct[k].randommove = (float)atof(argv[4]);
    unsigned char *removecell = calloc(n, 1);
    int64_t removecount = 0;
    updatepositions(st, ct, &ci, removecell, &removecount);
    const int64_t next = rand();
    FILE *m = fopen(argv[5], "wb");
    for (int64_t i = 0; i < n; i++) fwrite(&ci.cells[i].x, 8, 1, m);
    for (int64_t i = 0; i < n; i++) fwrite(&ci.cells[i].y, 8, 1, m);
    for (int64_t i = 0; i < n; i++) fwrite(&ci.cells[i].z, 8, 1, m);
    fwrite(removecell, 1, n, m);
    fwrite(&removecount, 8, 1, m);
    fwrite(&next, 8, 1, m);
    fclose(m);
    free(removecell);
    /* src/main.c's output after the update: vtk/step%08d.vtk in the working directory */
    mkdir("vtk", 0755);
    writevtk(si, st, ci);
  }

  octfree(&ci);
  if (pre && !strcmp(pre, "held")) octfree(&ci2);
  free(ci2.cells);
  for (int64_t k = 0; k < nf; k++) free(ced[k]);
  free(ced);
  tm_compat_shutdown();
  return 0;
}

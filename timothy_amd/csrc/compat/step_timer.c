/*
 * step_timer.c -- wall-clock time of one cell-mechanics step driven through
 * the reference's C API (timothy_compat.h over libtimothy_compat.so), in the
 * order of the reference's step_compute (src/step.c:47-97, one rank,
 * step > 0) after octbuild (src/main.c:70).  Host arrays in, host arrays
 * out, as a reference build linking these functions would run: the times
 * include the PCIe copies the compat layer makes.  One field, filled with a
 * smooth function (the values do not change the work).
 *
 * usage: step_timer <input.bin> <reps>
 * input:  int64 n, ntypes, G;  ntypes x {size,h,h2,h3,h4} double;
 *         grid {lower[3], res} double;  n x {x,y,z,size} double;
 *         n x young float;  n x ctype int32;  n x phase int32
 * output (stdout, one JSON line): n, G, reps, the median step time and the
 * median time of each call, in ms; precision from TM_PRECISION as for the
 * compat layer.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "timothy_compat.h"

static void rd(FILE *f, void *p, size_t sz, size_t n) {
  if (fread(p, sz, n, f) != n) {
    fprintf(stderr, "short read\n");
    exit(2);
  }
}

static double now_ms(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return 1e3 * (double)t.tv_sec + 1e-6 * (double)t.tv_nsec;
}

static int cmp_d(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

static double median(double *v, int n) {
  qsort(v, (size_t)n, sizeof(double), cmp_d);
  return n % 2 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

enum { C_BUILD, C_ALLOC, C_C2EINIT, C_POT, C_C2EWAIT, C_GRAD, C_E2CINIT, C_E2CWAIT, C_PFREE, C_FGRAD, C_OCTFREE, NC };
static const char *names[NC] = {"octbuild", "patches_alloc", "patches_cells2envinit", "computepotential",
                                "patches_cells2envwait", "computegradient", "patches_env2cellsinit",
                                "patches_env2cellswait", "patches_free", "fieldgradient", "octfree"};

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: step_timer <input.bin> <reps>\n");
    return 1;
  }
  const int reps = atoi(argv[2]);
  if (reps < 1 || reps > 1000) return 1;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 1;
  int64_t hdr[3];
  rd(f, hdr, 8, 3);
  const int64_t n = hdr[0], ntypes = hdr[1], G = hdr[2];
  if (n < 1 || ntypes < 1 || ntypes > 16 || G < 2 || G > 2048) return 1;
  const int64_t nodes = G * G * G;
  celltype_t *ct = calloc((size_t)ntypes, sizeof(celltype_t));
  for (int64_t k = 0; k < ntypes; k++) {
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
  ci.cells = calloc((size_t)n, sizeof(celldata_t));
  ci.forces = calloc((size_t)n, sizeof(double3dv_t));
  double *pos = malloc(sizeof(double) * 4 * (size_t)n);
  float *young = malloc(sizeof(float) * (size_t)n);
  int32_t *ctype = malloc(sizeof(int32_t) * (size_t)n), *phase = malloc(sizeof(int32_t) * (size_t)n);
  rd(f, pos, 8, 4 * (size_t)n);
  rd(f, young, 4, (size_t)n);
  rd(f, ctype, 4, (size_t)n);
  rd(f, phase, 4, (size_t)n);
  fclose(f);
  for (int64_t i = 0; i < n; i++) {
    ci.cells[i].x = pos[4 * i];
    ci.cells[i].y = pos[4 * i + 1];
    ci.cells[i].z = pos[4 * i + 2];
    ci.cells[i].size = pos[4 * i + 3];
    ci.cells[i].young = young[i];
    ci.cells[i].ctype = ctype[i];
    ci.cells[i].phase = phase[i];
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
  environment_t env;
  env.data = malloc(sizeof(double) * (size_t)nodes);
  env.gradient = malloc(sizeof(double) * 3 * (size_t)nodes);
  env.production = malloc(sizeof(double) * (size_t)nodes);
  for (int64_t i = 0; i < nodes; i++) {
    const int64_t a = i / (G * G), b = (i / G) % G, c = i % G;
    env.data[i] = sin(0.05 * (double)a) * cos(0.07 * (double)b) + 0.01 * (double)c;
  }
  environment_t *envp = &env;
  systeminfo_t si;
  memset(&si, 0, sizeof(si));
  si.rank = 0;
  si.size = 1;
  settings_t st;
  memset(&st, 0, sizeof(st));
  st.numberoffields = 1;
  st.numberofcelltypes = (int)ntypes;
  st.dimension = 3;
  st.step = 1;
  cellcommdata_t cc;
  memset(&cc, 0, sizeof(cc));
  cellenvdata_t **ced = NULL;

  double *t = calloc((size_t)(reps * (NC + 1)), sizeof(double));
  for (int r = -1; r < reps; r++) { /* r = -1: warm-up */
    double c[NC], t0 = now_ms(), tp = t0, tn;
    patches_t patches;
#define TIMED(k, call) \
  call;                \
  tn = now_ms();       \
  c[k] = tn - tp;      \
  tp = tn
    TIMED(C_BUILD, octbuild(si, &ci, ct));
    TIMED(C_ALLOC, patches_alloc(si, st, &patches, &ci, &grid));
    TIMED(C_C2EINIT, patches_cells2envinit(si, st, &patches, &ci, &grid));
    TIMED(C_POT, computepotential(si, &ci, ct, cc));
    TIMED(C_C2EWAIT, patches_cells2envwait(si, st, &patches, &grid, &envp));
    TIMED(C_GRAD, computegradient(si, &ci, ct, cc));
    TIMED(C_E2CINIT, patches_env2cellsinit(si, st, &patches, &grid, &envp));
    TIMED(C_E2CWAIT, patches_env2cellswait(si, st, &patches, &ci, &grid, &ced));
    TIMED(C_PFREE, patches_free(&patches));
    TIMED(C_FGRAD, fieldgradient(si, st, &envp, &grid));
    TIMED(C_OCTFREE, octfree(&ci));
#undef TIMED
    if (r < 0) continue;
    for (int k = 0; k < NC; k++) t[(size_t)k * reps + r] = c[k];
    t[(size_t)NC * reps + r] = tp - t0;
  }
  printf("{\"n\": %lld, \"G\": %lld, \"reps\": %d, \"step_ms\": %.3f, \"calls_ms\": {", (long long)n, (long long)G,
         reps, median(t + (size_t)NC * reps, reps));
  for (int k = 0; k < NC; k++) printf("%s\"%s\": %.3f", k ? ", " : "", names[k], median(t + (size_t)k * reps, reps));
  printf("}}\n");
  return 0;
}

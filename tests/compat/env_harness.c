/* Drives the reference's environment API (src/environment.h) through
 * libtimothy_compat.so in the order of src/main.c:58-61 and
 * src/step.c:60-62: environment_allocate, environment_init (rand()),
 * environment_initsystem, then environment_compute at step >= 1.
 *
 * usage: env_harness IN OUT
 * IN:  int64 lx, ly, lz, nfields, seed, step; float secondsperstep, gfdt,
 *      resolution, pad; per field double D, bc, mean, variance, lambda
 * OUT: int64 numberofiters, enviter, next rand(), CG iterations; double dt[0];
 *      every field's data after environment_init, then every field's solution */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "timothy_compat.h"

static void rd(FILE *f, void *p, size_t sz, size_t n) {
  if (fread(p, sz, n, f) != n) {
    fprintf(stderr, "short input\n");
    exit(1);
  }
}

int main(int argc, char **argv) {
  if (argc != 3) return 1;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 1;
  int64_t h[6];
  rd(f, h, 8, 6);
  float fl[4];
  rd(f, fl, 4, 4);
  const int64_t lx = h[0], ly = h[1], lz = h[2], nf = h[3];
  environment_t *env = calloc(nf, sizeof(environment_t));
  for (int64_t k = 0; k < nf; k++) {
    double p[5];
    rd(f, p, 8, 5);
    env[k].diffusioncoefficient = p[0];
    env[k].boundarycondition = p[1];
    env[k].initialconditionmean = p[2];
    env[k].initialconditionvariance = p[3];
    env[k].lambdadelay = p[4];
  }
  fclose(f);

  systeminfo_t si;
  memset(&si, 0, sizeof(si));
  si.size = 1;
  settings_t st;
  memset(&st, 0, sizeof(st));
  st.numberoffields = (int)nf;
  st.secondsperstep = fl[0];
  st.gfdt = fl[1];
  st.dimension = 3;
  st.step = (int)h[5];
  grid_t grid;
  memset(&grid, 0, sizeof(grid));
  grid.resolution = fl[2];
  grid.globalsize.x = lx;
  grid.globalsize.y = ly;
  grid.globalsize.z = lz;
  grid.localsize = grid.globalsize;
  grid.loweridx = calloc(1, sizeof(int643dv_t));
  grid.upperidx = calloc(1, sizeof(int643dv_t));
  grid.upperidx[0].x = lx - 1;
  grid.upperidx[0].y = ly - 1;
  grid.upperidx[0].z = lz - 1;
  solverdata_t sd;
  solversettings_t ss;
  memset(&sd, 0, sizeof(sd));
  memset(&ss, 0, sizeof(ss));

  srand((unsigned)h[4]);
  environment_allocate(si, st, grid, &env, &sd, &ss);
  environment_init(si, st, grid, &env);
  environment_initsystem(si, st, &grid, &env, &sd, &ss);
  environment_compute(si, st, &grid, &env, &sd, &ss);
  const int64_t next = rand();

  const int64_t nodes = lx * ly * lz;
  FILE *o = fopen(argv[2], "wb");
  const int64_t w[4] = {ss.numberofiters, ss.enviter, next, tm_compat_environment_iterations()};
  fwrite(w, 8, 4, o);
  fwrite(ss.dt, 8, 1, o);
  for (int64_t k = 0; k < nf; k++) fwrite(env[k].data, 8, nodes, o);
  double *sol = malloc(sizeof(double) * nodes);
  for (int64_t k = 0; k < nf; k++) {
    if (tm_compat_environment_solution((int)k, sol) != 0) return 2;
    fwrite(sol, 8, nodes, o);
  }
  fclose(o);
  tm_compat_shutdown();
  return 0;
}

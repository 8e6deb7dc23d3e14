/*
 * timothy_compat.h -- the reference's cell-mechanics C API, served by the
 * MI355X path (libtimothy_compat.so over libtimothy_hip.so).
 *
 * The structures below must have the reference's memory layout (they cross
 * the boundary by pointer and by value), so their members are those of
 * mcytu/timothy src/global.h; the functions have the signatures of
 * src/octree.h, src/potential.h, src/patches.h and src/fieldgradient.h.  A
 * program built against the reference headers links this library in place of
 * octree.c, potential.c, patches.c and fieldgradient.c.
 *
 * MPI types: when <mpi.h> is included before this header the structures use
 * MPI_Comm / MPI_Request; otherwise they use int, which is MPICH's layout.
 * Only one rank is served (systeminfo.size == 1): the single-GPU build has no
 * halo cells, so the remote-contribution calls are no-ops when
 * cellcommdata.numimp <= 0 and an error otherwise.
 *
 * Drop-in contract (same as the reference's octree): octbuild snapshots the
 * cells' positions, sizes, types, young's moduli and phases; the following
 * computepotential / computegradient / patches_cells2env* calls of the step
 * use that snapshot.  computepotential evaluates density, v, mindist and the
 * forces in one pass; computegradient returns those forces.
 */
#ifndef TIMOTHY_COMPAT_H
#define TIMOTHY_COMPAT_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(MPI_VERSION)
typedef MPI_Comm tm_mpi_comm_t;
typedef MPI_Request tm_mpi_request_t;
#else
typedef int tm_mpi_comm_t;
typedef int tm_mpi_request_t;
#endif

#ifndef ZOLTAN_ID_TYPE
#define ZOLTAN_ID_TYPE int
#endif

/* src/global.h:59-61: edges of the computational box */
#ifndef BOXSIZEX
#define BOXSIZEX 2048
#define BOXSIZEY 2048
#define BOXSIZEZ 2048
#endif

/* src/global.h:90-98 */
typedef struct cellcount_t {
  uint64_t n, g0phase, g1phase, sphase, g2phase, mphase, necroticphase;
} cellcount_t;

/* src/global.h:100-120: 112 bytes on x86-64 */
typedef struct celldata_t {
  int lifetime, phase, age, death;
  float phasetime, g1, s, g2, m, young;
  ZOLTAN_ID_TYPE gid;
  double x, y, z, size, v, density, mindist;
  int ctype;
} celldata_t;

/* src/global.h:122-151 */
typedef struct cellenvdata_t { double value, gx, gy, gz; } cellenvdata_t;
typedef struct double3dv_t { double x, y, z; } double3dv_t;
typedef struct float3dv_t { float x, y, z; } float3dv_t;
typedef struct int643dv_t { int64_t x, y, z; } int643dv_t;
typedef struct uint3dv_t { unsigned int x, y, z; } uint3dv_t;

/* src/global.h:153-164: 112 bytes */
typedef struct octnode_t {
  unsigned int xcode, ycode, zcode, xlimit, ylimit, zlimit, level;
  int64_t father;
  int64_t child[8];
  int data;
} octnode_t;

/* src/global.h:166-179 */
typedef struct cellsinfo_t {
  cellcount_t localcount, globalcount;
  cellcount_t *localtypecount, *globaltypecount;
  uint64_t *cellsperproc;
  celldata_t *cells;
  double3dv_t *forces;
  octnode_t *octree;
  int64_t octsize, octmaxsize;
  int dimension;
  double3dv_t *velocity;
} cellsinfo_t;

/* src/global.h:182-190 */
typedef struct grid_t {
  int643dv_t globalsize;
  double3dv_t lowercorner, uppercorner;
  int643dv_t localsize;
  float resolution;
  int643dv_t *loweridx, *upperidx;
  double3dv_t *lowleftnear, *uprightfar;
} grid_t;

/* src/global.h:212-215, 247-276 */
typedef struct explist_t { int64_t cell; int proc; } explist_t;
typedef struct cellindata_t { double x, y, z, size, young; int ctype; } cellindata_t;
typedef struct celloutdata_t { double v, density; } celloutdata_t;
typedef struct cellcommdata_t {
  explist_t *explist;
  int explistmaxsize;
  tm_mpi_request_t *reqsend, *reqrecv;
  int64_t *sendoffset, *recvoffset;
  int *recvcount, *sendcount;
  cellindata_t *sendcellindata, *recvcellindata;
  celloutdata_t *sendcelloutdata, *recvcelloutdata;
  int numexp, numimp;
} cellcommdata_t;

/* src/global.h:278-298 */
typedef struct settings_t {
  int64_t maxcells, maxlocalcells;
  int numberofsteps;
  float secondsperstep;
  int numberofcelltypes, numberoffields, dimension, restart;
  char rstfilename[128], outdir[128];
  int visoutstep, statoutstep, rstoutstep;
  float maxspeed, gfdt, gfh;
  int simulationstart;
  unsigned int rseed;
  int step;
} settings_t;

/* src/global.h:300-312 */
typedef struct systeminfo_t {
  int rank, size, nthreads;
  int dim[3];
  int noderank, nodesize, memperproc;
  tm_mpi_comm_t MPI_CART_COMM;
  int **coords;
  int endian;
} systeminfo_t;

/* src/global.h:315-335 */
typedef struct celltype_t {
  char name[128];
  float g1, s, g2, m, v, rd;
  double size, h, h2, h3, h4;
  char inputfile[128];
  float criticaldensity;
  float *production, *consumption, *criticallevel1, *criticallevel2;
  float randommove;
} celltype_t;

/* src/global.h:337-352 */
typedef struct patches_t {
  double **data;
  int *intersect, *receiver, *sender;
  double **recvdata;
  tm_mpi_request_t *reqsend, *reqrecv;
  int643dv_t *lowercorner, *uppercorner, *lowercornerR, *uppercornerR, *size;
  double **commbuff, **buff;
} patches_t;

/* src/global.h:354-364 */
typedef struct environment_t {
  char name[128];
  double diffusioncoefficient, boundarycondition, initialconditionmean, initialconditionvariance, lambdadelay;
  double *data, *gradient, *production;
} environment_t;

/* src/global.h:366-393 (built with HYPRE): its handles as opaque pointers,
 * HYPRE_Int and HYPRE_SStructVariable as int (HYPRE's default 32-bit ints) */
typedef struct solversettings_t {
  double *z;
  double *dt;
  int enviter;
  int numberofiters;
  int envobjecttype;
  int *vartypes;
  void **stencil;
  void *precond;
  void *grid;
  void *graph;
  int lower[3];
  int upper[3];
  int bclower[3];
  int bcupper[3];
} solversettings_t;

typedef struct solverdata_t {
  void *A, *b, *x, *parA, *parb, *parx, *solver;
} solverdata_t;

/* src/global.h:408-412 */
typedef struct octheap_t { int size, count; int *data; } octheap_t;

/* Layout guards (LP64: x86-64, AArch64).  A program built against the
 * reference's src/global.h passes these structures by pointer and by value,
 * so their sizes and the offsets of the members this library reads or writes
 * must be the reference's: any drift fails the build instead of corrupting
 * cells at run time. */
#if defined(__LP64__) && (defined(__x86_64__) || defined(__aarch64__))
#include <stddef.h>
#ifdef __cplusplus
#define TM_LAYOUT_ASSERT(c, m) static_assert(c, m)
#else
#define TM_LAYOUT_ASSERT(c, m) _Static_assert(c, m)
#endif
/* celldata_t, src/global.h:100-120 */
TM_LAYOUT_ASSERT(sizeof(celldata_t) == 112, "celldata_t is 112 bytes");
TM_LAYOUT_ASSERT(offsetof(celldata_t, phase) == 4, "celldata_t.phase at 4");
TM_LAYOUT_ASSERT(offsetof(celldata_t, young) == 36, "celldata_t.young at 36");
TM_LAYOUT_ASSERT(offsetof(celldata_t, x) == 48, "celldata_t.x at 48");
TM_LAYOUT_ASSERT(offsetof(celldata_t, size) == 72, "celldata_t.size at 72");
TM_LAYOUT_ASSERT(offsetof(celldata_t, v) == 80, "celldata_t.v at 80");
TM_LAYOUT_ASSERT(offsetof(celldata_t, density) == 88, "celldata_t.density at 88");
TM_LAYOUT_ASSERT(offsetof(celldata_t, mindist) == 96, "celldata_t.mindist at 96");
TM_LAYOUT_ASSERT(offsetof(celldata_t, ctype) == 104, "celldata_t.ctype at 104");
/* octnode_t, src/global.h:153-164 */
TM_LAYOUT_ASSERT(sizeof(octnode_t) == 112, "octnode_t is 112 bytes");
TM_LAYOUT_ASSERT(offsetof(octnode_t, level) == 24, "octnode_t.level at 24");
TM_LAYOUT_ASSERT(offsetof(octnode_t, father) == 32, "octnode_t.father at 32");
TM_LAYOUT_ASSERT(offsetof(octnode_t, child) == 40, "octnode_t.child at 40");
TM_LAYOUT_ASSERT(offsetof(octnode_t, data) == 104, "octnode_t.data at 104");
/* cellsinfo_t, src/global.h:166-179 */
TM_LAYOUT_ASSERT(sizeof(cellsinfo_t) == 192, "cellsinfo_t is 192 bytes");
TM_LAYOUT_ASSERT(offsetof(cellsinfo_t, cells) == 136, "cellsinfo_t.cells at 136");
TM_LAYOUT_ASSERT(offsetof(cellsinfo_t, forces) == 144, "cellsinfo_t.forces at 144");
TM_LAYOUT_ASSERT(offsetof(cellsinfo_t, octree) == 152, "cellsinfo_t.octree at 152");
TM_LAYOUT_ASSERT(offsetof(cellsinfo_t, octsize) == 160, "cellsinfo_t.octsize at 160");
TM_LAYOUT_ASSERT(offsetof(cellsinfo_t, dimension) == 176, "cellsinfo_t.dimension at 176");
/* grid_t, src/global.h:182-190 */
TM_LAYOUT_ASSERT(sizeof(grid_t) == 136, "grid_t is 136 bytes");
TM_LAYOUT_ASSERT(offsetof(grid_t, localsize) == 72, "grid_t.localsize at 72");
TM_LAYOUT_ASSERT(offsetof(grid_t, resolution) == 96, "grid_t.resolution at 96");
TM_LAYOUT_ASSERT(offsetof(grid_t, loweridx) == 104, "grid_t.loweridx at 104");
/* celltype_t, src/global.h:315-335 */
TM_LAYOUT_ASSERT(sizeof(celltype_t) == 368, "celltype_t is 368 bytes");
TM_LAYOUT_ASSERT(offsetof(celltype_t, size) == 152, "celltype_t.size at 152");
TM_LAYOUT_ASSERT(offsetof(celltype_t, h) == 160, "celltype_t.h at 160");
TM_LAYOUT_ASSERT(offsetof(celltype_t, h4) == 184, "celltype_t.h4 at 184");
TM_LAYOUT_ASSERT(offsetof(celltype_t, randommove) == 360, "celltype_t.randommove at 360");
/* environment_t, src/global.h:354-364; cellenvdata_t, src/global.h:122-127 */
TM_LAYOUT_ASSERT(sizeof(environment_t) == 192, "environment_t is 192 bytes");
TM_LAYOUT_ASSERT(offsetof(environment_t, data) == 168, "environment_t.data at 168");
TM_LAYOUT_ASSERT(offsetof(environment_t, production) == 184, "environment_t.production at 184");
TM_LAYOUT_ASSERT(sizeof(cellenvdata_t) == 32, "cellenvdata_t is 32 bytes");
#undef TM_LAYOUT_ASSERT
#endif

/* ---- src/octree.h ---- */
extern double3dv_t affShift;
extern double affScale;
void octbuild(systeminfo_t systeminfo, cellsinfo_t *cellsinfo, celltype_t *celltype);
void octfree(cellsinfo_t *cellsinfo);
void octheapinit(systeminfo_t systeminfo, octheap_t *ttheap);
void octcomputebox(int64_t c, uint3dv_t *minLocCode, uint3dv_t *maxLocCode, cellsinfo_t cellsinfo,
                   celltype_t *celltype);
int octlocateregion(uint3dv_t minLocCode, uint3dv_t maxLocCode, cellsinfo_t cellsinfo);
void octheappush(systeminfo_t systeminfo, octheap_t *ttheap, int idx);
int octheappop(octheap_t *ttheap);
void octheapfree(octheap_t *ttheap);
void octcomputeboxr(int64_t c, uint3dv_t *minLocCode, uint3dv_t *maxLocCode, cellcommdata_t cellcommdata,
                    celltype_t *celltype);

/* ---- src/potential.h ---- */
void computepotential(systeminfo_t systeminfo, cellsinfo_t *cellsinfo, celltype_t *celltype,
                      cellcommdata_t cellcommdata);
void computeremotepotential(systeminfo_t systeminfo, cellsinfo_t *cellsinfo, celltype_t *celltype,
                            cellcommdata_t cellcommdata);
void computegradient(systeminfo_t systeminfo, cellsinfo_t *cellsinfo, celltype_t *celltype,
                     cellcommdata_t cellcommdata);
void computeremotegradient(systeminfo_t systeminfo, cellsinfo_t *cellsinfo, celltype_t *celltype,
                           cellcommdata_t cellcommdata);

/* ---- src/patches.h ---- */
void patches_alloc(systeminfo_t systeminfo, settings_t settings, patches_t *patches, cellsinfo_t *cellsinfo,
                   grid_t *grid);
void patches_free(patches_t *patches);
void patches_env2cellsinit(systeminfo_t systeminfo, settings_t settings, patches_t *patches, grid_t *grid,
                           environment_t **environment);
void patches_env2cellswait(systeminfo_t systeminfo, settings_t settings, patches_t *patches,
                           cellsinfo_t *cellsinfo, grid_t *grid, cellenvdata_t ***cellenvdata);
void patches_cells2envwait(systeminfo_t systeminfo, settings_t settings, patches_t *patches, grid_t *grid,
                           environment_t **environment);
void patches_cells2envinit(systeminfo_t systeminfo, settings_t settings, patches_t *patches,
                           cellsinfo_t *cellsinfo, grid_t *grid);

/* ---- src/cells.c:342-365 (defined in cells.c, not declared in cells.h:
 * remove it from cells.c when linking this library) ---- */
void updatepositions(settings_t settings, celltype_t *celltype, cellsinfo_t *cellsinfo, unsigned char *removecell,
                     int64_t *removecount);

/* ---- src/io.h: writevtk (src/io.c:372-516), one rank; writes
 * vtk/step%08d.vtk (settings.step) like the reference.  The rest of io.c
 * stays; remove writevtk from it when linking this library. ---- */
void writevtk(systeminfo_t systeminfo, settings_t settings, cellsinfo_t cellsinfo);

/* ---- src/environment.h (src/environment.c): the environment solve on the
 * GPU instead of HYPRE.  environment_compute assembles each field's system as
 * the reference does (right-hand side from initialconditionmean, the six
 * Dirichlet faces and the point source of environment_pcfunction) and solves
 * it to the reference's tolerance (1e-12, 50 iterations, numberofiters
 * steps).  Like the reference it leaves environment[f].data unchanged; the
 * solution is kept and tm_compat_environment_solution returns it. ---- */
void environment_allocate(systeminfo_t systeminfo, settings_t settings, grid_t grid, environment_t **environment,
                          solverdata_t *solverdata, solversettings_t *solversettings);
void environment_init(systeminfo_t systeminfo, settings_t settings, grid_t grid, environment_t **environment);
void environment_initsystem(systeminfo_t systeminfo, settings_t settings, grid_t *grid, environment_t **environment,
                            solverdata_t *solverdata, solversettings_t *solversettings);
void environment_compute(systeminfo_t systeminfo, settings_t settings, grid_t *grid, environment_t **environment,
                         solverdata_t *solverdata, solversettings_t *solversettings);

/* ---- src/fieldgradient.h ---- */
void fieldgradient(systeminfo_t systeminfo, settings_t settings, environment_t **environment, grid_t *grid);

/* ---- configuration of the shim (no counterpart in the reference) ---- */
/* 0 = FP64 drop-in (default, or TM_PRECISION=fp64), 1 = mixed (TM_PRECISION=mixed) */
void tm_compat_set_precision(int precision);
/* HIP device of the implicit context (default 0, or TM_DEVICE) */
void tm_compat_set_device(int device);
/* 1 (default, or TM_COMPAT_HOST_OCTREE=1): octbuild fills cellsinfo->octree with
 * the PR-octree; 0: leaves cellsinfo->octree NULL for callers that never read it */
void tm_compat_set_host_octree(int enable);
/* the implicit context (a tm_ctx*, see timothy_hip.h) */
void *tm_compat_context(void);
/* the last environment_compute's solution of field f (localsize nodes,
 * [i][j][k] like environment_t.data) copied to out; 0 on success, -1 if there
 * is none.  The reference computes this solution and does not keep it. */
int tm_compat_environment_solution(int field, double *out);
/* CG iterations of the last environment_compute (all fields and steps) */
int tm_compat_environment_iterations(void);
/* release the implicit context */
void tm_compat_shutdown(void);

#ifdef __cplusplus
}
#endif
#endif

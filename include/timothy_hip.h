/*
 * timothy_hip.h -- native C ABI of the MI355X cell-mechanics path.
 *
 * One context per HIP device and stream.  Device-resident colonies (one, or
 * many independent colonies batched in one set of launches) go through:
 *
 *   tm_upload / tm_bind_device      cells, structure of arrays
 *   tm_build                        per-colony bounding box and affine map,
 *                                   29-bit location codes, Morton sort,
 *                                   PR-octree leaf octants, radix tree
 *   tm_neighbors                    tree walk -> neighbour lists, mindist,
 *                                   the last candidate visited and v (the
 *                                   potential with it, FP64), LDS tiles
 *   tm_sph                          density and forces: TM_FP64 (bit-exact
 *                                   drop-in) or TM_MIXED
 *   tm_cic_deposit / tm_field_gradient / tm_cic_interp
 *                                   cells <-> grid transfers
 *
 * Plain pointers and sizes only.  Every call returns a tm_status; TM_OK is 0.
 * Calls are stream-ordered on the context's stream; a context is not
 * thread-safe.  Result getters synchronise the stream.
 *
 * Reference interfaces these entry points replace (mcytu/timothy):
 *   tm_build          octbuild                        src/octree.c:145-207
 *   tm_neighbors      the tree walks of computepotential / computegradient
 *                     (octcomputebox, octlocateregion, octheap*, DFS)
 *                                                     src/potential.c:87-138,209-259
 *   tm_sph            computepotential + computegradient pair sums
 *                                                     src/potential.c:39-259, src/inline.h:40-83
 *   tm_cic_deposit    patches_cells2envinfo + patches_applycells2env
 *                                                     src/patches.c:156-234,321-367
 *   tm_cic_interp     patches_env2cellsinfo + patches_applyenv2cells
 *                                                     src/patches.c:380-439,512-593
 *   tm_field_gradient fieldgradient_compute            src/fieldgradient.c:149-224
 *   tm_get_octree     the octnode_t array octbuild leaves in cellsinfo->octree
 *                                                     src/global.h:153-164
 * The reference's own signatures are exported by libtimothy_compat
 * (timothy_compat.h).
 */
#ifndef TIMOTHY_HIP_H
#define TIMOTHY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tm_status {
  TM_OK = 0,
  TM_ERR_INVALID = 1,     /* bad argument or size */
  TM_ERR_HIP = 2,         /* HIP runtime error */
  TM_ERR_NOMEM = 3,       /* device or host allocation failed */
  TM_ERR_DEGENERATE = 4,  /* two cells share location codes down to bit 1, or more than 64 cells
                            lie within 2^-16 of their colony's extent on every axis */
  TM_ERR_STATE = 5,       /* call out of order (e.g. tm_sph before tm_neighbors) */
  TM_ERR_UNSUPPORTED = 6  /* e.g. more than one MPI rank in the compat layer */
} tm_status;

typedef enum tm_precision {
  TM_FP64 = 0,  /* reference arithmetic, bit-exact with the CPU path */
  TM_MIXED = 1  /* FP32 pair evaluation, FP64 predicate and FP64 per-cell sums */
} tm_precision;

typedef enum tm_stage {
  TM_STAGE_BUILD = 0,
  TM_STAGE_NEIGHBORS = 1,
  TM_STAGE_SPH = 2,
  TM_STAGE_CIC = 3,
  TM_STAGE_STEP = 4,
  TM_STAGE_TILES = 5 /* the LDS tiles alone (part of tm_neighbors) */
} tm_stage;

/* mechanics fields of celltype_t (src/global.h:315-335), as given:
 * h2..h4 are NOT recomputed from h (the reference reads them as stored). */
typedef struct tm_celltype {
  double size, h, h2, h3, h4;
} tm_celltype;

/* grid_t restricted to one rank (src/global.h:182-190, src/grid.c:37-136):
 * node (i,j,k) of the local block sits at lower + (loweridx + (i,j,k))*resolution,
 * arrays are [i][j][k], k fastest; gradients are 3-interleaved. */
typedef struct tm_grid {
  double lower[3];
  float resolution;
  int64_t loweridx[3];
  int64_t upperidx[3];
  int64_t localsize[3];
} tm_grid;

typedef struct tm_ctx tm_ctx;

/* -------- context -------- */
tm_status tm_create(tm_ctx **ctx, int device);
void tm_destroy(tm_ctx *ctx);
const char *tm_status_name(tm_status s);
/* last error message of this context ("" if none) */
const char *tm_last_error(const tm_ctx *ctx);
/* the context's HIP stream (hipStream_t) */
void *tm_stream(tm_ctx *ctx);
tm_status tm_sync(tm_ctx *ctx);

/* -------- configuration -------- */
/* dimension 2 or 3 (src/global.h:177); ntypes <= 16 */
tm_status tm_set_celltypes(tm_ctx *ctx, const tm_celltype *types, int32_t ntypes, int32_t dimension);
/* tunables: "nbr_cap" neighbour slots per cell (grows on overflow),
 * "sph_block" / "nbr_block" workgroup sizes (64, 128, 256), "sph_unroll_fp64" /
 * "sph_unroll_mixed" neighbours fetched per trip of the pair kernels (1, 2, 4,
 * 8), "sph_lanes" lanes per cell of the mixed kernel (1 = one lane per cell),
 * "sph_balance" (1 = the FP64 kernel ranks the cells of a workgroup by list
 * length), "sph_tiles" (1 = tm_neighbors also builds LDS tiles and the 3-D
 * pair kernels read neighbours from them; default 1), "tile_cap4" (1 = stage
 * at most 1279 records per tile when at most one tile in twenty needs more,
 * so four pair-kernel workgroups fit per CU; default 1), "walk_window" slots
 * of the pre-ordered tree a 256-cell walk workgroup stages in LDS around its
 * cells' leaves (0, 512, 1024, 1280, 2048; default 1280), "walk_f32" (1 = the walk
 * screens candidates with FP32 positions relative to the colony's corner and
 * a proved error bound, reading FP64 positions only where that bound cannot
 * decide; results are identical; default 1), "tile_window" (1 = the tile
 * build orders the distinct neighbours within 8192 cells of a tile's middle
 * by bitmap and hashes only the rest; 0 = hashes all of them; results are
 * identical; default 1), "tile_width" cells per LDS tile (128, 256, 512;
 * default 256; takes effect at the next tm_neighbors; results are
 * identical), "walk_coop" (1 = each wave of 64 cells walks the tree once for
 * the union of its query boxes, 64 slots per load, and each lane picks its
 * own candidates from the union's; waves spanning two colonies or whose
 * union box exceeds "walk_coop_ratio" (default 64) times their mean box
 * volume walk per lane; results are identical; default 0); a change of "nbr_cap" or "tile_width" retires a captured graph;
 * read-only "nbr_alloc" (slots per cell of the
 * current lists), "tile_lds" (records staged per tile, 0 without tiles) and
 * "tiles_over" (tiles with more than 1536 distinct neighbours, which take
 * the 32-bit lists, at the last tm_neighbors) */
tm_status tm_set_param(tm_ctx *ctx, const char *key, int64_t value);
tm_status tm_get_param(tm_ctx *ctx, const char *key, int64_t *value);

/* -------- cells -------- */
/* Host arrays (reference order).  ncolonies >= 1 independent colonies are
 * concatenated; colony_offsets has ncolonies+1 entries (NULL when
 * ncolonies == 1).  phase may be NULL (treated as 1), young may be NULL
 * (treated as 0, like the reference's uninitialised file cells). */
tm_status tm_upload(tm_ctx *ctx, int64_t ncells, int32_t ncolonies, const int64_t *colony_offsets,
                    const double *x, const double *y, const double *z, const double *size,
                    const float *young, const int32_t *ctype, const int32_t *phase);
/* The same with device pointers already in HBM (no copy; they must stay valid
 * and unchanged until the next tm_upload/tm_bind_device). colony_offsets is a
 * HOST array. */
tm_status tm_bind_device(tm_ctx *ctx, int64_t ncells, int32_t ncolonies, const int64_t *colony_offsets,
                         const double *x, const double *y, const double *z, const double *size,
                         const float *young, const int32_t *ctype, const int32_t *phase);

/* -------- path stages -------- */
tm_status tm_build(tm_ctx *ctx);
/* neighbour lists, mindist, v (src/potential.c:87-138 walk 1) */
tm_status tm_neighbors(tm_ctx *ctx);
/* density and forces over the lists (the pair sums of both walks) */
tm_status tm_sph(tm_ctx *ctx, tm_precision precision);
/* build + neighbors + sph */
tm_status tm_step(tm_ctx *ctx, tm_precision precision);

/* -------- results (host, reference cell order) -------- */
/* any pointer may be NULL */
tm_status tm_get_results(tm_ctx *ctx, double *density, double *v, double *mindist, double *fx, double *fy,
                         double *fz);
/* results in Morton-sorted order plus the permutation (sorted -> reference
 * index), device pointers valid until the next stage call; the layout the
 * stages keep in HBM */
tm_status tm_device_results(tm_ctx *ctx, const double **density, const double **v, const double **mindist,
                            const double **fx, const double **fy, const double **fz, const int32_t **perm);
tm_status tm_get_affine(tm_ctx *ctx, int32_t colony, double shift[3], double *scale);
/* 3 location codes per cell (x,y,z), reference order */
tm_status tm_get_codes(tm_ctx *ctx, uint32_t *loc);
/* perm[i] = reference index of the i-th cell in Morton order; span_log2[c] =
 * log2 of the side (in code units) of cell c's PR-octree leaf octant */
tm_status tm_get_order(tm_ctx *ctx, int32_t *perm, int32_t *span_log2);
/* in-cutoff neighbour lists in the reference's accumulation order (descending
 * Morton key), reference indices; offsets has ncells+1 entries; nbr may be
 * NULL to get the total only */
tm_status tm_get_neighbors(tm_ctx *ctx, int64_t *offsets, int32_t *nbr, int64_t *total);
/* first-visited-last candidate (minimum Morton key) per cell, -1 if none */
tm_status tm_get_qstar(tm_ctx *ctx, int32_t *qstar);
/* P = ordered in-cutoff pairs; max_nbr = largest neighbour count */
tm_status tm_pair_count(tm_ctx *ctx, int64_t *pairs, int64_t *max_nbr);
/* canonical PR-octree of one colony as octnode_t records (112 bytes each,
 * src/global.h:153-164): internal nodes first (root = 0), then one leaf per
 * cell; data holds the reference cell index, -2 for internal nodes.  Pass
 * nodes == NULL to get the count. */
tm_status tm_get_octree(tm_ctx *ctx, int32_t colony, void *nodes, int64_t *count);

/* -------- grid transfers (single colony; device arrays) -------- */
/* weight_mode 0: reference weights by phase (5 -> 0, 0 -> 0.75, else 1.0);
 * weight_mode 1: cell volume 4/3*pi*size^3.  production is overwritten
 * (every node written, like patches_applycells2env's zeroing and sum).
 * Each node holds the reference's terms summed in another order.  Grid node
 * indices (loweridx .. upperidx) must lie in (-2^29, 2^29): TM_ERR_INVALID
 * otherwise. */
tm_status tm_cic_deposit(tm_ctx *ctx, const tm_grid *grid, int32_t weight_mode, double *production);
/* gradient of a field with the reference's stencil (interior (f[i+1]-f[i-1])/res,
 * one-sided faces); FP64 or FP32 field */
tm_status tm_field_gradient(tm_ctx *ctx, const tm_grid *grid, const double *field, double *gradient);
tm_status tm_field_gradient_f32(tm_ctx *ctx, const tm_grid *grid, const float *field, float *gradient);
/* out[4*c + {0,1,2,3}] = value, gx, gy, gz at cell c (reference order) */
tm_status tm_cic_interp(tm_ctx *ctx, const tm_grid *grid, const double *field, const double *gradient,
                        double *out);
tm_status tm_cic_interp_f32(tm_ctx *ctx, const tm_grid *grid, const float *field, const float *gradient,
                            double *out);

/* -------- device memory for callers without their own allocator -------- */
tm_status tm_malloc(tm_ctx *ctx, void **ptr, size_t bytes);
tm_status tm_free(tm_ctx *ctx, void *ptr);
tm_status tm_memcpy_htod(tm_ctx *ctx, void *dst, const void *src, size_t bytes);
tm_status tm_memcpy_dtoh(tm_ctx *ctx, void *dst, const void *src, size_t bytes);
tm_status tm_memset(tm_ctx *ctx, void *dst, int value, size_t bytes);
/* the same copies without waiting: they run on the context's stream, in
 * order with its kernels; tm_sync waits for them */
tm_status tm_memcpy_htod_async(tm_ctx *ctx, void *dst, const void *src, size_t bytes);
tm_status tm_memcpy_dtoh_async(tm_ctx *ctx, void *dst, const void *src, size_t bytes);
/* page-locked host memory: copies between it and the device run at the
 * full link rate (the compat layer stages cells and the emitted octree in
 * it).  tm_host_free takes only pointers tm_host_alloc returned. */
tm_status tm_host_alloc(tm_ctx *ctx, void **ptr, size_t bytes);
tm_status tm_host_free(tm_ctx *ctx, void *ptr);
/* device-to-device copy on the context's stream (also the known-size
 * reference transfer used to calibrate memory-traffic counters) */
tm_status tm_memcpy_dtod(tm_ctx *ctx, void *dst, const void *src, size_t bytes);

/* -------- moving the cells -------- */
/* updatepositions (src/cells.c:342-365) on the uploaded cells, in HBM: each
 * cell moves by the force of the last tm_sph, then per axis (z only in 3-D)
 * by randommove[ctype] * (u * 2.0 - 1.0), in the reference's order and
 * FP64 arithmetic; a cell whose sphere reaches a face of the box
 * [-box/2, box/2] (outsidethebox, src/cells.c:41-58) gets removecell = 1.
 * u: draws, when not NULL, is a DEVICE array of dimension*ncells floats in
 * the reference's order (cell-major; x, y, then z), e.g. (float)rand()/RAND_MAX;
 * otherwise u is word a of Philox4x32-10 with counter (cell, step) and key
 * seed, as (float)(w >> 1) / 2147483648.0f.
 * removecell: DEVICE array of ncells bytes (reference order) or NULL; cells
 * already marked are not counted again.  *removecount (if not NULL) is
 * increased by the number of cells newly marked (with NULL removecell: by the
 * number outside).  Needs cells from tm_upload (bound cells are read-only)
 * and tm_sph results; afterwards the tree is stale until the next tm_build. */
typedef struct tm_move {
  float randommove[16]; /* per cell type, celltype_t.randommove (src/global.h:334) */
  double box[3];        /* box edges; the reference's is 2048^3 (BOXSIZEX/Y/Z, src/global.h:59-61) */
  uint64_t seed;        /* Philox key */
  uint64_t step;        /* Philox counter, high half (one value per step) */
} tm_move;
tm_status tm_update_positions(tm_ctx *ctx, const tm_move *move, const float *draws, uint8_t *removecell,
                              int64_t *removecount);
/* The same update on plain DEVICE arrays in cell order (no colony needed):
 * x, y, z moved in place by the forces fx, fy, fz; size and ctype (0..15)
 * per cell; randommove indexed by ctype.  Synchronises; *removecount as
 * above. */
tm_status tm_move_cells(tm_ctx *ctx, int64_t ncells, int32_t dimension, double *x, double *y, double *z,
                        const double *size, const int32_t *ctype, const double *fx, const double *fy,
                        const double *fz, const tm_move *move, const float *draws, uint8_t *removecell,
                        int64_t *removecount);
/* current positions of the cells, reference order, into host arrays (any may
 * be NULL) */
tm_status tm_get_positions(tm_ctx *ctx, double *x, double *y, double *z);

/* -------- the environment solve -------- */
/* Backward-Euler steps of du/dt = D lap(u) - lambda u + s on the field grid,
 * assembled as the reference assembles its HYPRE system
 * (src/environment.c:131-511): z = D dt / res^2 (res^2 in FP32),
 * (1 + 6z + lambda dt) u_i - z sum of the neighbours inside the grid = b_i,
 * b_i = u_old_i + z bc for each grid face the node lies on + s_i (Dirichlet
 * faces: the field outside the grid is bc).  Solved by conjugate gradients
 * on the GPU to ||b - A u|| <= tol ||b|| or maxiter iterations per step
 * (the reference: tol 1e-12, 50 iterations). */
typedef struct tm_diffusion {
  double diffusion; /* environment_t.diffusioncoefficient */
  double lambda;    /* environment_t.lambdadelay */
  double boundary;  /* environment_t.boundarycondition */
  double dt;        /* settings.gfdt */
  double tol;       /* relative residual */
  int32_t maxiter;  /* per step */
  int32_t substeps; /* backward-Euler steps in this call (the reference's numberofiters) */
} tm_diffusion;
/* field: DEVICE array of the grid's nodes ([i][j][k], k fastest, like
 * environment_t.data), u at t in, u at t + substeps dt out.  source: DEVICE
 * array of the same shape added to every step's right-hand side (e.g. the
 * CIC deposit's production), or NULL.  *iterations: CG iterations over all
 * steps; *residual: the last step's final relative residual. Synchronises. */
tm_status tm_diffusion_step(tm_ctx *ctx, const tm_grid *grid, const tm_diffusion *params, const double *source,
                            double *field, int32_t *iterations, double *residual);
/* No counterpart in the reference.  The same solve for callers that cannot
 * allocate device memory (tight stepping loops, harnesses that forbid
 * hipMalloc): identical kernels and results, but the work arrays (right-hand
 * side, residual, two directions, A p, partial sums, solver state) live in
 * one static device buffer of the library, sized for grids of at most
 * TM_NOALLOC_MAX_NODES nodes (about 80 MiB, held on every device the library
 * runs on until the process ends).  The call allocates nothing; larger grids
 * are refused with TM_ERR_INVALID.  Every context on one device shares that
 * buffer: calls that overlap on one device (two host threads, or two
 * contexts' streams) race and must be serialised by the caller.  Nothing in
 * the buffer needs to start at zero; each call writes every word it reads. */
#define TM_NOALLOC_MAX_NODES 2097152
tm_status tm_diffusion_step_noalloc(tm_ctx *ctx, const tm_grid *grid, const tm_diffusion *params,
                                    const double *source, double *field, int32_t *iterations, double *residual);

/* -------- the VTK writer -------- */
/* writevtk (src/io.c:372-516), one rank: the reference's legacy binary VTK
 * file -- POINTS (positions as big-endian float), CELL_TYPES (1), SCALARS
 * rank, SCALARS type (ctype), SCALARS density (the reference's constant 1.0)
 * and VECTORS force.  tm_write_vtk writes the uploaded cells' current
 * positions with the forces of the last tm_sph (also after tm_update_positions,
 * before the next tm_build), in the cells' order; the body is packed on the
 * GPU and written by the host. */
tm_status tm_write_vtk(tm_ctx *ctx, const char *path, int32_t rank);
/* the same from DEVICE arrays in cell order */
tm_status tm_write_vtk_arrays(tm_ctx *ctx, const char *path, int64_t ncells, const double *x, const double *y,
                              const double *z, const int32_t *ctype, const double *fx, const double *fy,
                              const double *fz, int32_t rank);

/* -------- measurement -------- */
/* runs `warmup` untimed then `iters` timed launches of a stage on the
 * context's stream, bracketed by HIP events on that stream; ms_per_iter is
 * the event time / iters.  For TM_STAGE_CIC the grid/production of the last
 * tm_cic_deposit call are reused. */
tm_status tm_time_stage(tm_ctx *ctx, tm_stage stage, tm_precision precision, int32_t warmup, int32_t iters,
                        double *ms_per_iter);
/* per-kernel average duration (ms) of the SPH kernel over the last
 * tm_time_stage(TM_STAGE_SPH) call */
tm_status tm_last_kernel_ms(tm_ctx *ctx, double *ms);

/* -------- captured step -------- */
/* capture build + neighbors + sph into a hipGraph (no allocation inside);
 * tm_graph_launch replays it and checks the overflow / degeneracy flags */
tm_status tm_graph_capture(tm_ctx *ctx, tm_precision precision);
/* the same step followed by tm_update_positions with Philox draws: each
 * replay of the graph is one mechanics step of the resident colony (sort,
 * tree, lists, sums, move), the first with move->step, then step + 1, ...
 * A replay whose neighbour lists overflow moves nothing on the device;
 * tm_graph_launch then grows the lists, redoes that step eagerly and
 * recaptures. */
tm_status tm_graph_capture_loop(tm_ctx *ctx, tm_precision precision, const tm_move *move, uint8_t *removecell);
/* cells newly marked by captured-loop replays since the last call */
tm_status tm_graph_removed(tm_ctx *ctx, int64_t *count);
tm_status tm_graph_launch(tm_ctx *ctx);
tm_status tm_graph_destroy(tm_ctx *ctx);

/* -------- autotuner -------- */
/* Times candidate shapes for the current colonies and keeps the fastest (the
 * current shape stays unless a candidate is more than 2% faster): the
 * walk (nbr_block x walk_window, then walk_coop, then walk_f32), then with
 * LDS tiles (3-D) tile_width (128, 256, 512) x sph_unroll timed as tile
 * build + pair kernel, else sph_block x sph_unroll (x sph_lanes, mixed).
 * No candidate changes a result.  Leaves the lists and tiles built with the chosen shapes.  Report
 * into buf (may be NULL): one line "trial <stage> key=value ... ms=<t>" per
 * candidate, then "chosen key=value ...".  Replaces the reference's only
 * knob, the OpenMP schedule of the walks (src/potential.c:90-93). */
tm_status tm_autotune(tm_ctx *ctx, tm_precision precision, char *buf, size_t buflen);

#ifdef __cplusplus
}
#endif
#endif

/*
 * oracle.h -- CPU restatement of Timothy's per-step cell-mechanics path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (timothy_amd/, include/)
 * includes, links or calls this code; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, as the checker and as the timed
 * CPU baseline.  It restates the reference's algorithm (mcytu/timothy, C99 +
 * OpenMP) from its source semantics; every function cites the reference file
 * and lines it follows.  Built with -ffp-contract=off and without
 * -ffast-math, so each floating-point expression is evaluated in the
 * order the reference source writes it, with IEEE division and sqrt.
 *
 * Pinning (what checks each function against something other than itself):
 *   or_octbuild          leaf levels vs brute-force min-msb (tests/test_oracle.py),
 *                        canonical node set vs GPU octree
 *   or_walk_* / or_computepotential / or_computegradient
 *                        neighbour sets vs brute force O(N^2); candidate sets vs
 *                        brute-force leaf-octant intersection; forces vs the
 *                        reference-produced VTK (new_run/vtk/step00000000.vtk);
 *                        force == -grad of the unrounded kernel sum (finite
 *                        differences); kernel normalisation (integral == 1);
 *                        lens term == textbook sphere-sphere lens volume
 *   or_potential Hertz term: pinned only by its young->inf limit and by xc->0
 *                        continuity -- the 5/2 integer division and the D
 *                        parenthesisation are source readings (parity partial)
 *   or_cic_deposit / or_cic_interp
 *                        partition of unity, mass conservation, exact
 *                        interpolation of trilinear fields, hand-computed indices
 *   or_fieldgradient     linear / quadratic fields (closed form)
 *   or_philox4x32_10     the known-answer vectors published with the
 *                        generator (Random123's kat_vectors)
 *   or_diffusion_*       the operator against an independent stencil and
 *                        right-hand side written with array shifts; a
 *                        constant field with the boundary at the same value is
 *                        reproduced; a Dirichlet sine eigenmode is divided by
 *                        its closed-form eigenvalue
 *   or_writevtk          byte for byte against the reference's own step-0
 *                        output (new_run/vtk/step00000000.vtk)
 *   or_updatepositions   hand-computed moves for chosen draws (u = 0, 1/2,
 *                        1), glibc's documented rand() sequence for srand(1),
 *                        cells placed exactly on the box faces, draw order
 *                        in 2-D and 3-D
 */
#ifndef TM_ORACLE_H
#define TM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_celltype {
  double size, h, h2, h3, h4;
} or_celltype;

/* one colony, structure-of-arrays, reference cell order */
typedef struct or_cells {
  int64_t n;
  const double *x, *y, *z, *size;
  const float *young;
  const int32_t *ctype;
  const int32_t *phase;
} or_cells;

/* node of the pointer PR-octree: same fields and meaning as octnode_t
 * (src/global.h:153-164), children as int64 indices, data >=0 cell, -1
 * empty, -2 internal. */
typedef struct or_node {
  uint32_t xcode, ycode, zcode, xlimit, ylimit, zlimit, level;
  int64_t father;
  int64_t child[8];
  int32_t data;
} or_node;

typedef struct or_tree {
  double shift[3];
  double scale;
  uint32_t *loc;       /* 3*n location codes, x,y,z interleaved */
  or_node *nodes;
  int64_t nnodes, cap;
} or_tree;

/* src/octree.c:145-207 (octbuild) with octinsertcell :84-140 and
 * octemptynode :40-79; compact growable node pool instead of 64*N.
 * Returns 0, or -1 if two cells share codes down to bit 1 (the reference
 * recurses with negative shift counts there: undefined behaviour). */
int or_octbuild(const or_cells *c, const or_celltype *types, or_tree *t);
void or_treefree(or_tree *t);

/* src/octree.c:246-266 */
void or_computebox(const or_cells *c, const or_celltype *types, const or_tree *t,
                   int64_t p, uint32_t mn[3], uint32_t mx[3]);
/* src/octree.c:271-316; returns a node index */
int64_t or_locateregion(const or_tree *t, const uint32_t mn[3], const uint32_t mx[3]);

/* src/potential.c:39-78 (potential) for one ordered pair; updates *density and
 * *mindist like the reference's side effects and returns the potential. */
double or_potential(int dimension, double px, double py, double pz, double psize, float pyoung,
                    double qx, double qy, double qz, double qsize, float qyoung,
                    const or_celltype *ptype, double *density, double *mindist);
/* src/inline.h:40-51 */
float or_sphkernel(int dimension, double dist, double h, double h2, double h3);
/* src/inline.h:59-83 (p = c1, q = c2) */
void or_sphgradient(int dimension, double px, double py, double pz, double qx, double qy, double qz,
                    double h, double h3, double h4, double f[3]);

/* src/potential.c:87-138 (computepotential): density, v, mindist per cell.
 * Leaves outputs untouched when n <= 1 (src/potential.c:89). */
void or_computepotential(const or_cells *c, const or_celltype *types, const or_tree *t, int dimension,
                         double *density, double *v, double *mindist);
/* src/potential.c:209-259 (computegradient) */
void or_computegradient(const or_cells *c, const or_celltype *types, const or_tree *t, int dimension,
                        double *fx, double *fy, double *fz);

/* The candidate walk of src/potential.c:108-133 with its visit order: for each
 * cell p, every leaf cell q != p the DFS pops, in pop order.  cand_off has
 * n+1 entries; cand is sized by the caller (pass NULL to count only).
 * in_h[k] = 1 when dist(p,q) <= h_p (the neighbour set N(p)).
 * Returns the total candidate count. */
int64_t or_walk_candidates(const or_cells *c, const or_celltype *types, const or_tree *t,
                           int64_t *cand_off, int32_t *cand, uint8_t *in_h);

/* O(N^2) neighbour sets without any tree: q != p with
 * sqrt((dx*dx+dy*dy)+dz*dz) <= h_p, ascending q.  Same calling convention. */
int64_t or_brute_neighbors(const or_cells *c, const or_celltype *types, int64_t *off, int32_t *nbr);

/* leaf level of every cell read off the pointer octree (level of the node whose
 * data == cell), -1 if absent */
void or_leaf_levels(const or_tree *t, int64_t n, int32_t *level);

/* CIC deposit, src/patches.c:156-234 (single rank: patch == local block).
 * weight_mode 0: reference (phase 5 -> none, phase 0 -> 0.75, else 1.0);
 * weight_mode 1: cell volume 4/3*pi*size^3 (configuration extension).
 * production is (lx*ly*lz), k fastest, zeroed here (src/patches.c:327-328). */
void or_cic_deposit(const or_cells *c, int weight_mode, const double lower[3], float resolution,
                    const int64_t loweridx[3], const int64_t upperidx[3], const int64_t localsize[3],
                    double *production);
/* CIC interpolation, src/patches.c:512-593: value,gx,gy,gz per cell from the
 * field and its 3-interleaved gradient (single rank). out is 4*n. */
void or_cic_interp(const or_cells *c, const double lower[3], float resolution,
                   const int64_t loweridx[3], const int64_t upperidx[3], const int64_t localsize[3],
                   const double *field, const double *gradient, double *out);
/* src/fieldgradient.c:149-224, no neighbour ranks (all faces one-sided) */
void or_fieldgradient(const int64_t localsize[3], float resolution, const double *field, double *gradient);

/* Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC 2011): ten rounds over
 * a 4-word counter with a 2-word key bumped by the Weyl constants between
 * rounds. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* the draws of the device path's counter-based moves: cell c, axis a takes
 * word a of philox(ctr = (c lo, c hi, step lo, step hi), key = (seed lo,
 * seed hi)) as (float)(w >> 1) / 2147483648.0f, the reference's
 * (float)rand()/RAND_MAX applied to a 31-bit value.  draws is dim*n,
 * cell-major. */
void or_philox_draws(int64_t n, int dimension, uint64_t seed, uint64_t step, float *draws);
/* src/cells.c:342-365 (updatepositions) with outsidethebox, src/cells.c:41-58:
 * x += F, then x += randommove[type] * (u * 2.0 - 1.0) per axis (z only in
 * 3-D), u in the reference's order (cell-major; x, y, then z); cells whose
 * sphere reaches a face of the box [-box/2, box/2] get removecell = 1.
 * draws == NULL takes u from the C library's rand() like the reference,
 * otherwise dim*n values.  box: edge lengths (the reference's BOXSIZEX/Y/Z,
 * src/global.h:59-61).  Returns the number of cells newly marked. */
int64_t or_updatepositions(int64_t n, int dimension, double *x, double *y, double *z, const double *size,
                           const int32_t *ctype, const float *randommove, const double *fx, const double *fy,
                           const double *fz, const float *draws, const double box[3], uint8_t *removecell);

/* The environment solve's linear system, src/environment.c:131-511 (the
 * HYPRE assembly in environment_initsystem / environment_initboundary /
 * environment_solve), on a local grid ls[0] x ls[1] x ls[2] stored [i][j][k]
 * with k fastest: z = D dt / (res * res) with res * res in FP32,
 * diag = 1 + 6 z + lambda dt; A u = diag u - z (sum of the neighbours inside
 * the grid); b = u_old + z bc for each face the node lies on (x-, x+, y-, y+,
 * z-, z+ in that order) + src (src may be NULL). */
void or_diffusion_rhs(const int64_t ls[3], float res, double D, double lambda, double bc, double dt,
                      const double *src, const double *uold, double *b);
void or_diffusion_apply(const int64_t ls[3], float res, double D, double lambda, double dt, const double *u,
                        double *Au);
/* substeps backward-Euler steps: each assembles b from the current field and
 * solves A u = b by conjugate gradients from u_old to ||b - A u|| <= tol ||b||
 * (at most maxiter iterations).  Returns the iterations over all steps. */
int or_diffusion_solve(const int64_t ls[3], float res, double D, double lambda, double bc, double dt,
                       const double *src, double *field, double tol, int maxiter, int substeps);

/* src/io.c:372-516 (writevtk), one rank: the legacy binary VTK file of the
 * cells' positions, types and forces (big-endian; density written as the
 * constant 1.0 like the reference).  Returns 0, or -1 if the file cannot be
 * written. */
int or_writevtk(const char *path, int64_t n, const double *x, const double *y, const double *z,
                const int32_t *ctype, const double *fx, const double *fy, const double *fz, int rank);

/* omp_get_wtime wrapper and thread control for the timed baseline */
double or_wtime(void);
int or_max_threads(void);
void or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif

// The environment solve: one or more backward-Euler steps of
//   du/dt = D lap(u) - lambda u + s
// on the field grid, as the reference assembles it for HYPRE
// (src/environment.c:131-511):
//   z = D dt / res^2   (res^2 formed in FP32, grid_t.resolution being float),
//   (1 + 6 z + lambda dt) u_i - z sum_{neighbours j inside the grid} u_j = b_i,
//   b_i = u_old_i + z bc per grid face the node lies on (the six faces in the
//         order x-, x+, y-, y+, z-, z+) + s_i,
// i.e. Dirichlet faces with the field outside the grid held at bc.  The
// reference solves with HYPRE's PCG (BoomerAMG-preconditioned) to a relative
// residual of 1e-12 in at most 50 iterations; the matrix is symmetric,
// positive definite and strongly diagonally dominant with a constant
// diagonal, so Jacobi-preconditioned CG (plain CG scaled by that constant,
// with the same stopping norm) reaches the same solution to that tolerance.
//
// Matrix-free, one lane per node, nodes in the field's layout ([i][j][k], k
// fastest).  An iteration is two streaming kernels and two one-block
// reductions:
//   k_cg_dir     p_new = r + beta p_old at the node and its 6 neighbours,
//                Ap = A p_new, partial sums of p_new . Ap      (reads r, p;
//                writes p_new, Ap: 32 B/node when neighbours come from cache)
//   k_cg_alpha   alpha = rr / pAp (fixed-order sum of the partials)
//   k_cg_update  x += alpha p, r -= alpha Ap, partial sums of r . r
//                                                             (48 B/node)
//   k_cg_beta    beta = rr_new / rr, convergence test, iteration count
// The scalars live in device memory; once converged, every kernel of the
// remaining iterations returns at once, so the host launches iterations in
// batches and reads the state once per batch.  The partial sums are added
// in a fixed order: results do not depend on timing.
#include <mutex>

#include "tm_internal.h"

namespace {

constexpr int BD = 256;
constexpr int MAXBLK = 4096;  // partial sums per reduction

struct CGState {
  double rr, bb, pAp, alpha, beta, tol2;
  unsigned long long iters;
  int done;
  int pad;
};

struct Op {
  int64_t lx, ly, lz;
  double diag, z;
};

__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < BD / 64; ++i) s += sh[i];
  return s;
}

// nodes [chunk * b, chunk * (b + 1)) belong to block b (contiguous, so a
// block's neighbour planes are mostly its own or its neighbours')
__device__ __forceinline__ void block_range(int64_t total, int64_t chunk, int64_t &lo, int64_t &hi) {
  lo = (int64_t)blockIdx.x * chunk;
  hi = min(lo + chunk, total);
}

// right-hand side, residual of the initial guess, first direction
__global__ __launch_bounds__(BD) void k_cg_init(Op A, double zbc, const double *__restrict__ src,
                                                const double *__restrict__ x, double *__restrict__ b,
                                                double *__restrict__ r, double *__restrict__ p, int64_t chunk,
                                                double *__restrict__ part_rr, double *__restrict__ part_bb) {
  __shared__ double sh[BD / 64];
  const int64_t total = A.lx * A.ly * A.lz, sx = A.ly * A.lz, sy = A.lz;
  int64_t lo, hi;
  block_range(total, chunk, lo, hi);
  double rr = 0.0, bb = 0.0;
  for (int64_t node = lo + threadIdx.x; node < hi; node += BD) {
    const int64_t k = node % A.lz, j = (node / A.lz) % A.ly, i = node / sx;
    const double xc = x[node];
    // b in the reference's order: u_old, the faces x-, x+, y-, y+, z-, z+, then s
    double bv = xc;
    if (i == 0) bv += zbc;
    if (i == A.lx - 1) bv += zbc;
    if (j == 0) bv += zbc;
    if (j == A.ly - 1) bv += zbc;
    if (k == 0) bv += zbc;
    if (k == A.lz - 1) bv += zbc;
    if (src) bv += src[node];
    double nb = 0.0;
    if (i > 0) nb += x[node - sx];
    if (i < A.lx - 1) nb += x[node + sx];
    if (j > 0) nb += x[node - sy];
    if (j < A.ly - 1) nb += x[node + sy];
    if (k > 0) nb += x[node - 1];
    if (k < A.lz - 1) nb += x[node + 1];
    const double rv = bv - (A.diag * xc - A.z * nb);
    b[node] = bv;
    r[node] = rv;
    p[node] = rv;
    rr += rv * rv;
    bb += bv * bv;
  }
  const double s1 = block_sum(rr, sh);
  __syncthreads();
  const double s2 = block_sum(bb, sh);
  if (threadIdx.x == 0) {
    part_rr[blockIdx.x] = s1;
    part_bb[blockIdx.x] = s2;
  }
}

__global__ void k_cg_start(const double *__restrict__ part_rr, const double *__restrict__ part_bb, int nblk,
                           double tol, int maxiter, CGState *st) {
  __shared__ double sh[BD / 64];
  double a = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nblk; i += BD) {
    a += part_rr[i];
    c += part_bb[i];
  }
  const double rr = block_sum(a, sh);
  __syncthreads();
  const double bb = block_sum(c, sh);
  if (threadIdx.x == 0) {
    st->rr = rr;
    st->bb = bb;
    st->tol2 = tol * tol;
    st->beta = 0.0;
    st->iters = 0;
    st->done = (rr <= st->tol2 * bb || maxiter <= 0) ? 1 : 0;
  }
}

// p_new = r + beta p_old (at the node and its neighbours), Ap = A p_new
__global__ __launch_bounds__(BD) void k_cg_dir(Op A, const double *__restrict__ r, const double *__restrict__ pold,
                                               double *__restrict__ pnew, double *__restrict__ Ap, int64_t chunk,
                                               const CGState *__restrict__ st, double *__restrict__ part) {
  __shared__ double sh[BD / 64];
  if (st->done) return;
  const double beta = st->beta;
  const int64_t total = A.lx * A.ly * A.lz, sx = A.ly * A.lz, sy = A.lz;
  int64_t lo, hi;
  block_range(total, chunk, lo, hi);
  double acc = 0.0;
  for (int64_t node = lo + threadIdx.x; node < hi; node += BD) {
    const int64_t k = node % A.lz, j = (node / A.lz) % A.ly, i = node / sx;
    const double pc = r[node] + beta * pold[node];
    double nb = 0.0;
    if (i > 0) nb += r[node - sx] + beta * pold[node - sx];
    if (i < A.lx - 1) nb += r[node + sx] + beta * pold[node + sx];
    if (j > 0) nb += r[node - sy] + beta * pold[node - sy];
    if (j < A.ly - 1) nb += r[node + sy] + beta * pold[node + sy];
    if (k > 0) nb += r[node - 1] + beta * pold[node - 1];
    if (k < A.lz - 1) nb += r[node + 1] + beta * pold[node + 1];
    const double ap = A.diag * pc - A.z * nb;
    pnew[node] = pc;
    Ap[node] = ap;
    acc += pc * ap;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// The same, marching along i: a workgroup owns an MJ x MK block of (j, k)
// columns over a run of planes [i0, i1).  Per plane it forms p_new once per
// node (and once per node of the block's edge ring, which the neighbouring
// blocks also form) into LDS; the i-1 and i+1 values of its own column stay
// in registers.  r and p_old are read about once instead of once per
// neighbour through the caches.  Each node's p_new and Ap are the values
// k_cg_dir forms (the same expressions, the neighbours added in the same
// order); only the partial sums of p_new . Ap group differently.
constexpr int MJ = 4, MK = 64, MT = MJ * MK;
constexpr int MHALO = 2 * MJ + 2 * MK;  // the edge ring without its corners

__global__ __launch_bounds__(MT) void k_cg_dir_march(Op A, const double *__restrict__ r,
                                                     const double *__restrict__ pold, double *__restrict__ pnew,
                                                     double *__restrict__ Ap, int nseg,
                                                     const CGState *__restrict__ st, double *__restrict__ part) {
  __shared__ double pl[2][MJ + 2][MK + 2];
  __shared__ double sh[MT / 64];
  if (st->done) return;
  const double beta = st->beta;
  const int64_t sx = A.ly * A.lz, sy = A.lz;
  const int ntk = (int)((A.lz + MK - 1) / MK), ntj = (int)((A.ly + MJ - 1) / MJ);
  int b = blockIdx.x;
  const int kt = b % ntk;
  b /= ntk;
  const int jt = b % ntj, seg = b / ntj;
  const int64_t i0 = A.lx * seg / nseg, i1 = A.lx * (seg + 1) / nseg;
  const int tk = (int)threadIdx.x % MK, tj = (int)threadIdx.x / MK;
  const int64_t j = (int64_t)jt * MJ + tj, k = (int64_t)kt * MK + tk;
  const bool own = j < A.ly && k < A.lz;
  // this thread's place on the edge ring, if any: (a, c) in the (MJ+2) x (MK+2) frame
  int ha = -1, hc = -1;
  {
    const int t = (int)threadIdx.x;
    if (t < MK) {
      ha = 0;
      hc = 1 + t;
    } else if (t < 2 * MK) {
      ha = MJ + 1;
      hc = 1 + t - MK;
    } else if (t < 2 * MK + MJ) {
      ha = 1 + t - 2 * MK;
      hc = 0;
    } else if (t < MHALO) {
      ha = 1 + t - 2 * MK - MJ;
      hc = MK + 1;
    }
  }
  const int64_t hj = (int64_t)jt * MJ + ha - 1, hk = (int64_t)kt * MK + hc - 1;
  const bool halo = ha >= 0 && hj >= 0 && hj < A.ly && hk >= 0 && hk < A.lz;
  // plane i into LDS buffer q; returns this column's value
  auto plane = [&](int64_t i, int q) -> double {
    double v = 0.0;
    if (own) {
      const int64_t node = i * sx + j * sy + k;
      v = r[node] + beta * pold[node];
      pl[q][tj + 1][tk + 1] = v;
    }
    if (halo) {
      const int64_t node = i * sx + hj * sy + hk;
      pl[q][ha][hc] = r[node] + beta * pold[node];
    }
    return v;
  };
  double acc = 0.0;
  double pprev = 0.0, pcur = 0.0, pnext = 0.0;
  if (i0 < i1) {
    if (i0 > 0 && own) {
      const int64_t node = (i0 - 1) * sx + j * sy + k;
      pprev = r[node] + beta * pold[node];
    }
    pcur = plane(i0, (int)(i0 & 1));
  }
  for (int64_t i = i0; i < i1; ++i) {
    __syncthreads();  // plane i is in its buffer; the reads of plane i - 1 are done
    const int q = (int)(i & 1);
    if (i + 1 < A.lx) pnext = plane(i + 1, q ^ 1);
    if (own) {
      double nb = 0.0;
      if (i > 0) nb += pprev;
      if (i < A.lx - 1) nb += pnext;
      if (j > 0) nb += pl[q][tj][tk + 1];
      if (j < A.ly - 1) nb += pl[q][tj + 2][tk + 1];
      if (k > 0) nb += pl[q][tj + 1][tk];
      if (k < A.lz - 1) nb += pl[q][tj + 1][tk + 2];
      const double ap = A.diag * pcur - A.z * nb;
      const int64_t node = i * sx + j * sy + k;
      pnew[node] = pcur;
      Ap[node] = ap;
      acc += pcur * ap;
    }
    pprev = pcur;
    pcur = pnext;
  }
  __syncthreads();
  double v = acc;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < MT / 64; ++w) s += sh[w];
    part[blockIdx.x] = s;
  }
}

__global__ void k_cg_alpha(const double *__restrict__ part, int nblk, CGState *st) {
  __shared__ double sh[BD / 64];
  if (st->done) return;
  double a = 0.0;
  for (int i = threadIdx.x; i < nblk; i += BD) a += part[i];
  const double pAp = block_sum(a, sh);
  if (threadIdx.x == 0) {
    st->pAp = pAp;
    if (!(pAp > 0.0)) {
      st->done = 1;  // breakdown: the residual is zero (or the operator is not positive definite)
      st->alpha = 0.0;
    } else {
      st->alpha = st->rr / pAp;
    }
  }
}

__global__ __launch_bounds__(BD) void k_cg_update(int64_t total, double *__restrict__ x, double *__restrict__ r,
                                                  const double *__restrict__ p, const double *__restrict__ Ap,
                                                  int64_t chunk, const CGState *__restrict__ st,
                                                  double *__restrict__ part) {
  __shared__ double sh[BD / 64];
  if (st->done) return;
  const double alpha = st->alpha;
  int64_t lo, hi;
  block_range(total, chunk, lo, hi);
  double acc = 0.0;
  for (int64_t node = lo + threadIdx.x; node < hi; node += BD) {
    x[node] += alpha * p[node];
    const double rv = r[node] - alpha * Ap[node];
    r[node] = rv;
    acc += rv * rv;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_cg_beta(const double *__restrict__ part, int nblk, int maxiter, CGState *st) {
  __shared__ double sh[BD / 64];
  if (st->done) return;
  double a = 0.0;
  for (int i = threadIdx.x; i < nblk; i += BD) a += part[i];
  const double rr = block_sum(a, sh);
  if (threadIdx.x == 0) {
    st->beta = rr / st->rr;
    st->rr = rr;
    st->iters += 1;
    if (rr <= st->tol2 * st->bb || st->iters >= (unsigned long long)maxiter) st->done = 1;
  }
}

}  // namespace

// doubles of work memory for a grid of `total` nodes: b, r, two directions,
// Ap; two arrays of partial sums; the state
static constexpr int64_t work_doubles(int64_t total) {
  return 5 * total + 2 * MAXBLK + (int64_t)((sizeof(CGState) + 7) / 8);
}

// The allocation-free entry point's work memory: one buffer per device, in
// the code object (the loader places it on every device the library runs
// on).  Host code never passes `scratch` itself to a kernel: on the host the
// name is a stand-in, not the device address.  The address is looked up
// once per device (tm_noalloc_init) and kept in scratch_addr.
constexpr int64_t SCRATCH_DOUBLES = work_doubles(TM_NOALLOC_MAX_NODES);
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
static std::mutex scratch_mu;
static double *scratch_addr[MAX_DEVICES];

// the solve with its work arrays at `work` (work_doubles(total) doubles)
static tm_status run_diffusion(tm_ctx *c, double *work, const tm_grid *g, const tm_diffusion *d, const double *src,
                               double *field, int32_t *iterations, double *residual) {
  const int64_t lx = g->localsize[0], ly = g->localsize[1], lz = g->localsize[2];
  const int64_t total = lx * ly * lz;
  double *b = work, *r = b + total, *p0 = r + total, *p1 = p0 + total, *Ap = p1 + total;
  double *part = Ap + total, *part2 = part + MAXBLK;
  CGState *st = (CGState *)(part2 + MAXBLK);
  const int64_t nblk0 = std::min<int64_t>(MAXBLK, (total + BD - 1) / BD);
  const int64_t chunk = (((total + nblk0 - 1) / nblk0) + BD - 1) / BD * BD;
  const int nblk = (int)((total + chunk - 1) / chunk);
  // the marching direction kernel: (j, k) blocks times runs of planes, about
  // 2,048 workgroups, at most MAXBLK (else the one-lane-per-node kernel)
  const int64_t ncol = ((ly + MJ - 1) / MJ) * ((lz + MK - 1) / MK);
  const int nseg = (int)std::max<int64_t>(1, std::min<int64_t>(lx, std::max<int64_t>(1, 2048 / ncol)));
  const bool march = ncol * nseg <= MAXBLK;
  const int nblk_dir = march ? (int)(ncol * nseg) : nblk;
  Op A;
  A.lx = lx;
  A.ly = ly;
  A.lz = lz;
  const float res2 = g->resolution * g->resolution;  // FP32 product, as the reference forms it
  A.z = d->diffusion * d->dt / (double)res2;
  A.diag = 1 + 6.0 * A.z + d->lambda * d->dt;
  const double zbc = A.z * d->boundary;
  int32_t it_total = 0;
  double last_res = 0.0;
  for (int s = 0; s < d->substeps; ++s) {
    k_cg_init<<<nblk, BD, 0, c->stream>>>(A, zbc, src, field, b, r, p0, chunk, part, part2);
    k_cg_start<<<1, BD, 0, c->stream>>>(part, part2, nblk, d->tol, d->maxiter, st);
    TM_HIP(c, hipGetLastError());
    // p0 holds r for the first direction (beta = 0); the directions then alternate
    double *pold = p0, *pnew = p1;
    int32_t launched = 0;
    // the state is read back into pinned host memory set aside at tm_create
    CGState &h = *reinterpret_cast<CGState *>(c->h_flags + TM_HFLAGS_CG);
    for (;;) {
      const int32_t batch = std::min<int32_t>(8, d->maxiter - launched);
      for (int32_t q = 0; q < batch; ++q) {
        if (march) k_cg_dir_march<<<nblk_dir, MT, 0, c->stream>>>(A, r, pold, pnew, Ap, nseg, st, part);
        else k_cg_dir<<<nblk, BD, 0, c->stream>>>(A, r, pold, pnew, Ap, chunk, st, part);
        k_cg_alpha<<<1, BD, 0, c->stream>>>(part, nblk_dir, st);
        k_cg_update<<<nblk, BD, 0, c->stream>>>(total, field, r, pnew, Ap, chunk, st, part);
        k_cg_beta<<<1, BD, 0, c->stream>>>(part, nblk, d->maxiter, st);
        std::swap(pold, pnew);
      }
      launched += batch;
      TM_HIP(c, hipGetLastError());
      TM_HIP(c, hipMemcpyAsync(&h, st, sizeof(CGState), hipMemcpyDeviceToHost, c->stream));
      TM_HIP(c, hipStreamSynchronize(c->stream));
      if (h.done || launched >= d->maxiter || batch == 0) break;
    }
    it_total += (int32_t)h.iters;
    last_res = h.bb > 0.0 ? sqrt(h.rr / h.bb) : sqrt(h.rr);
  }
  if (iterations) *iterations = it_total;
  if (residual) *residual = last_res;
  return TM_OK;
}

static_assert(sizeof(CGState) <= 4 * (TM_HFLAGS_WORDS - TM_HFLAGS_CG), "solver state fits the pinned words");

tm_status tm_run_diffusion(tm_ctx *c, const tm_grid *g, const tm_diffusion *d, const double *src, double *field,
                           int32_t *iterations, double *residual) {
  const int64_t total = g->localsize[0] * g->localsize[1] * g->localsize[2];
  TM_TRY(tm_ensure(c, &c->d_cg, c->cap_cg, work_doubles(total), "diffusion work arrays"));
  return run_diffusion(c, c->d_cg, g, d, src, field, iterations, residual);
}

tm_status tm_noalloc_init(tm_ctx *c) {
  // called by tm_create with c->device current
  if (c->device < 0 || c->device >= MAX_DEVICES) return tm_fail(c, TM_ERR_UNSUPPORTED, "device index too large");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_addr[c->device]) return TM_OK;
  void *p = nullptr;
  TM_HIP(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return tm_fail(c, TM_ERR_HIP, "no device address for the static work buffer");
  scratch_addr[c->device] = (double *)p;
  return TM_OK;
}

tm_status tm_run_diffusion_noalloc(tm_ctx *c, const tm_grid *g, const tm_diffusion *d, const double *src,
                                   double *field, int32_t *iterations, double *residual) {
  const int64_t total = g->localsize[0] * g->localsize[1] * g->localsize[2];
  if (total > TM_NOALLOC_MAX_NODES)
    return tm_fail(c, TM_ERR_INVALID, "tm_diffusion_step_noalloc: the grid has " + std::to_string(total) +
                                          " nodes, more than TM_NOALLOC_MAX_NODES (" +
                                          std::to_string((long long)TM_NOALLOC_MAX_NODES) +
                                          "); use tm_diffusion_step");
  double *work;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    work = scratch_addr[c->device];
  }
  if (!work) return tm_fail(c, TM_ERR_STATE, "static work buffer not set up on this device");
  return run_diffusion(c, work, g, d, src, field, iterations, residual);
}

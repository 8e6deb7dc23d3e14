// Grid transfers between cells and the field grid (one rank's block).
//
// Cloud-in-cell deposit  src/patches.c:156-234 (+ patches_applycells2env :321-367)
// Cloud-in-cell interp   src/patches.c:512-593
// Field gradient         src/fieldgradient.c:149-224
//
// Index and weight arithmetic follows the reference exactly:
//   cellIdx = (int64)((x - lower) / (double)resolution)        truncation
//   cicCoord = lower + (double)((float)cellIdx * resolution)   the product is
//              int64 * float, evaluated in float by C's usual conversions
//   d = (x - cicCoord) / resolution, t = 1 - d, corner weight a*d + (1-a)*t
// so corner indices match bit for bit and the interpolation (one lane per
// cell, corners in the reference's ax, ay, az order) matches bit for bit.
// The deposit sums the reference's terms into each node through LDS partial
// sums and FP64 atomics: the same terms in a different order (<= ~1e-15
// relative for positive weights).
#include "tm_internal.h"

namespace {

struct GridDev {
  double lower[3];
  float res;
  int64_t lo[3], hi[3], ls[3];
};

GridDev to_dev(const tm_grid *g) {
  GridDev d;
  for (int a = 0; a < 3; ++a) {
    d.lower[a] = g->lower[a];
    d.lo[a] = g->loweridx[a];
    d.hi[a] = g->upperidx[a];
    d.ls[a] = g->localsize[a];
  }
  d.res = g->resolution;
  return d;
}

__device__ __forceinline__ void cic_setup(const GridDev &g, double px, double py, double pz, int64_t ci[3],
                                          double d[3], double t[3]) {
  const double pc[3] = {px, py, pz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    ci[a] = (int64_t)((pc[a] - g.lower[a]) / (double)g.res);
    const float off = (float)ci[a] * g.res;
    const double cic = g.lower[a] + (double)off;
    d[a] = (pc[a] - cic) / (double)g.res;
    t[a] = 1.0 - d[a];
  }
}

__device__ __forceinline__ bool cic_inrange(const GridDev &g, const int64_t ci[3], int ax, int ay, int az) {
  return ci[0] + ax >= g.lo[0] && ci[1] + ay >= g.lo[1] && ci[2] + az >= g.lo[2] && ci[0] + ax <= g.hi[0] &&
         ci[1] + ay <= g.hi[1] && ci[2] + az <= g.hi[2];
}

__global__ void k_fill_f64(double *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0;
}

// Deposit, one workgroup per 256 cells of the Morton order.  Those cells lie
// in a small region, so their corners usually fit a box of at most DEP_LDS
// nodes: the workgroup adds its terms into that box in LDS, then adds each
// touched node to the grid with one global atomic (a few per cell instead of
// eight).  A workgroup whose box is larger adds its terms to the grid
// directly.
constexpr int DEP_LDS = 4096;  // 32 KB of FP64 partial sums

__global__ __launch_bounds__(256) void k_deposit(const double4 *__restrict__ pos, const int32_t *__restrict__ phase_s,
                                                 int64_t n, GridDev g, int weight_mode,
                                                 double *__restrict__ prod) {
  __shared__ double acc[DEP_LDS];
  __shared__ int box[6];  // min x, y, z of the cells' first corners; max x, y, z
  const int64_t i = tm_xcd_block() * 256 + threadIdx.x;
  if (threadIdx.x < 3) box[threadIdx.x] = 0x7FFFFFFF;
  else if (threadIdx.x < 6) box[threadIdx.x] = -0x7FFFFFFF;
  double wt = 0.0;
  bool dep = false;
  double4 P = make_double4(0, 0, 0, 0);
  if (i < n) {
    P = pos[i];
    if (weight_mode == 1) {
      wt = ((4.0 / 3.0) * M_PI) * (P.w * P.w * P.w);
      dep = true;
    } else {
      const int ph = phase_s[i];
      dep = ph != 5;                 // necrotic: no deposit
      wt = ph == 0 ? 0.75 : 1.0;     // G0: lower consumption
    }
  }
  int64_t ci[3] = {0, 0, 0};
  double d[3] = {0, 0, 0}, t[3] = {0, 0, 0};
  if (dep) cic_setup(g, P.x, P.y, P.z, ci, d, t);
  // the box only matters for corners in range: clamp to one node beyond it
  int cl[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) cl[a] = (int)max(min(ci[a], g.hi[a] + 1), g.lo[a] - 1);
  // the box over the depositing lanes: reduced across the wave first, then
  // one LDS atomic per wave and bound (not one per lane on the same word)
  int bmin[3], bmax[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    bmin[a] = dep ? cl[a] : 0x7FFFFFFF;
    bmax[a] = dep ? cl[a] : -0x7FFFFFFF;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      bmin[a] = min(bmin[a], __shfl_xor(bmin[a], o, 64));
      bmax[a] = max(bmax[a], __shfl_xor(bmax[a], o, 64));
    }
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      atomicMin(&box[a], bmin[a]);
      atomicMax(&box[3 + a], bmax[a]);
    }
  }
  __syncthreads();
  const bool any = box[0] <= box[3];  // some lane deposits
  const int b0 = box[0], b1 = box[1], b2 = box[2];
  // extent of the corners ci and ci + 1
  const int e0 = any ? box[3] - b0 + 2 : 0, e1 = any ? box[4] - b1 + 2 : 0, e2 = any ? box[5] - b2 + 2 : 0;
  const bool local = any && (int64_t)e0 * e1 * e2 <= DEP_LDS;
  if (local) {
    const int vol = e0 * e1 * e2;
    for (int k = threadIdx.x; k < vol; k += 256) acc[k] = 0.0;
    __syncthreads();
  }
  if (dep) {
#pragma unroll
    for (int ax = 0; ax < 2; ++ax)
#pragma unroll
      for (int ay = 0; ay < 2; ++ay)
#pragma unroll
        for (int az = 0; az < 2; ++az) {
          if (!cic_inrange(g, ci, ax, ay, az)) continue;
          const double val = wt * (ax * d[0] + (1 - ax) * t[0]) * (ay * d[1] + (1 - ay) * t[1]) *
                             (az * d[2] + (1 - az) * t[2]);
          if (local) {
            const int lx = (int)ci[0] + ax - b0, ly = (int)ci[1] + ay - b1, lz = (int)ci[2] + az - b2;
            atomicAdd(&acc[(lx * e1 + ly) * e2 + lz], val);
          } else {
            const int64_t gx = ci[0] + ax - g.lo[0], gy = ci[1] + ay - g.lo[1], gz = ci[2] + az - g.lo[2];
            unsafeAtomicAdd(prod + (gx * g.ls[1] + gy) * g.ls[2] + gz, val);
          }
        }
  }
  if (!local) return;
  __syncthreads();
  const int vol = e0 * e1 * e2;
  for (int k = threadIdx.x; k < vol; k += 256) {
    const double v = acc[k];
    if (v == 0.0) continue;  // untouched (only in-range corners were added)
    const int lz = k % e2, ly = (k / e2) % e1, lx = k / (e1 * e2);
    const int64_t gx = b0 + lx - g.lo[0], gy = b1 + ly - g.lo[1], gz = b2 + lz - g.lo[2];
    unsafeAtomicAdd(prod + (gx * g.ls[1] + gy) * g.ls[2] + gz, v);
  }
}

template <typename F>
__global__ __launch_bounds__(256) void k_interp(const double4 *__restrict__ pos, const int32_t *__restrict__ perm,
                                                int64_t n, GridDev g, const F *__restrict__ field,
                                                const F *__restrict__ grad, double *__restrict__ out) {
  const int64_t i = tm_xcd_block() * 256 + threadIdx.x;
  if (i >= n) return;
  const double4 P = pos[i];
  int64_t ci[3];
  double d[3], t[3];
  cic_setup(g, P.x, P.y, P.z, ci, d, t);
  double val = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
  for (int ax = 0; ax < 2; ++ax)
#pragma unroll
    for (int ay = 0; ay < 2; ++ay)
#pragma unroll
      for (int az = 0; az < 2; ++az) {
        if (!cic_inrange(g, ci, ax, ay, az)) continue;
        const int64_t node = ((ci[0] + ax - g.lo[0]) * g.ls[1] + (ci[1] + ay - g.lo[1])) * g.ls[2] +
                             (ci[2] + az - g.lo[2]);
        const double wx = (ax * d[0] + (1 - ax) * t[0]);
        const double wy = (ay * d[1] + (1 - ay) * t[1]);
        const double wz = (az * d[2] + (1 - az) * t[2]);
        val += (double)field[node] * wx * wy * wz;
        gx += (double)grad[3 * node + 0] * wx * wy * wz;
        gy += (double)grad[3 * node + 1] * wx * wy * wz;
        gz += (double)grad[3 * node + 2] * wx * wy * wz;
      }
  const int64_t o = perm[i];
  double4 r = make_double4(val, gx, gy, gz);
  reinterpret_cast<double4 *>(out)[o] = r;
}

// one lane per node, k fastest; faces one-sided (no neighbouring ranks)
template <typename F>
__global__ __launch_bounds__(256) void k_gradient(const F *__restrict__ f, F *__restrict__ g, int64_t lx, int64_t ly,
                                                  int64_t lz, float res) {
  const int64_t node = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t total = lx * ly * lz;
  if (node >= total) return;
  const int64_t k = node % lz;
  const int64_t j = (node / lz) % ly;
  const int64_t i = node / (lz * ly);
  const int64_t sx = ly * lz, sy = lz;
  // interior (f[i+1]-f[i-1])/res; faces (f[1]-f[0])/res, (f[L-1]-f[L-2])/res
  const int64_t ip = (i == lx - 1) ? node : node + sx, im = (i == 0) ? node : node - sx;
  const int64_t jp = (j == ly - 1) ? node : node + sy, jm = (j == 0) ? node : node - sy;
  const int64_t kp = (k == lz - 1) ? node : node + 1, km = (k == 0) ? node : node - 1;
  F r = (F)res;
  if (sizeof(F) == 8) {
    g[3 * node + 0] = (F)(((double)f[ip] - (double)f[im]) / (double)res);
    g[3 * node + 1] = (F)(((double)f[jp] - (double)f[jm]) / (double)res);
    g[3 * node + 2] = (F)(((double)f[kp] - (double)f[km]) / (double)res);
  } else {
    g[3 * node + 0] = (f[ip] - f[im]) / r;
    g[3 * node + 1] = (f[jp] - f[jm]) / r;
    g[3 * node + 2] = (f[kp] - f[km]) / r;
  }
}

}  // namespace

tm_status tm_run_deposit(tm_ctx *c, const tm_grid *grid, int weight_mode, double *prod) {
  GridDev g = to_dev(grid);
  const int64_t total = g.ls[0] * g.ls[1] * g.ls[2];
  k_fill_f64<<<(unsigned)std::min<int64_t>(tm_blocks(total, 256), 65536), 256, 0, c->stream>>>(prod, total);
  if (c->n > 0)
    k_deposit<<<tm_blocks(c->n, 256), 256, 0, c->stream>>>(c->d_pos, c->d_phase_s, c->n, g, weight_mode, prod);
  TM_HIP(c, hipGetLastError());
  return TM_OK;
}

static tm_status check_grid(tm_ctx *c, const tm_grid *g) {
  if (!g) return tm_fail(c, TM_ERR_INVALID, "grid is NULL");
  for (int a = 0; a < 3; ++a) {
    if (g->localsize[a] < 2) return tm_fail(c, TM_ERR_INVALID, "grid localsize must be >= 2 on every axis");
    if (g->upperidx[a] - g->loweridx[a] + 1 != g->localsize[a])
      return tm_fail(c, TM_ERR_INVALID, "grid upperidx - loweridx + 1 must equal localsize");
  }
  if (!(g->resolution > 0.0f)) return tm_fail(c, TM_ERR_INVALID, "grid resolution must be > 0");
  return TM_OK;
}

static tm_status check_cells_grid(tm_ctx *c, const tm_grid *g) {
  TM_TRY(check_grid(c, g));
  if (c->stage < 2) return tm_fail(c, TM_ERR_STATE, "grid transfer before tm_build");
  if (c->ncol != 1) return tm_fail(c, TM_ERR_UNSUPPORTED, "grid transfers take a single colony");
  return TM_OK;
}

extern "C" tm_status tm_cic_deposit(tm_ctx *c, const tm_grid *grid, int32_t weight_mode, double *production) {
  if (!c) return TM_ERR_INVALID;
  TM_TRY(check_cells_grid(c, grid));
  if (!production) return tm_fail(c, TM_ERR_INVALID, "production is NULL");
  if (weight_mode != 0 && weight_mode != 1) return tm_fail(c, TM_ERR_INVALID, "weight_mode must be 0 or 1");
  c->last_grid = *grid;
  c->last_weight = weight_mode;
  c->last_prod = production;
  return tm_run_deposit(c, grid, weight_mode, production);
}

template <typename F>
static tm_status gradient_impl(tm_ctx *c, const tm_grid *grid, const F *field, F *gradient) {
  if (!c) return TM_ERR_INVALID;
  TM_TRY(check_grid(c, grid));
  if (!field || !gradient) return tm_fail(c, TM_ERR_INVALID, "field or gradient is NULL");
  const int64_t total = grid->localsize[0] * grid->localsize[1] * grid->localsize[2];
  k_gradient<F><<<tm_blocks(total, 256), 256, 0, c->stream>>>(field, gradient, grid->localsize[0], grid->localsize[1],
                                                               grid->localsize[2], grid->resolution);
  TM_HIP(c, hipGetLastError());
  return TM_OK;
}

extern "C" tm_status tm_field_gradient(tm_ctx *c, const tm_grid *grid, const double *field, double *gradient) {
  return gradient_impl<double>(c, grid, field, gradient);
}
extern "C" tm_status tm_field_gradient_f32(tm_ctx *c, const tm_grid *grid, const float *field, float *gradient) {
  return gradient_impl<float>(c, grid, field, gradient);
}

template <typename F>
static tm_status interp_impl(tm_ctx *c, const tm_grid *grid, const F *field, const F *gradient, double *out) {
  if (!c) return TM_ERR_INVALID;
  TM_TRY(check_cells_grid(c, grid));
  if (!field || !gradient || !out) return tm_fail(c, TM_ERR_INVALID, "NULL array");
  if (c->n == 0) return TM_OK;
  k_interp<F><<<tm_blocks(c->n, 256), 256, 0, c->stream>>>(c->d_pos, c->d_perm, c->n, to_dev(grid), field, gradient,
                                                            out);
  TM_HIP(c, hipGetLastError());
  return TM_OK;
}

extern "C" tm_status tm_cic_interp(tm_ctx *c, const tm_grid *grid, const double *field, const double *gradient,
                                   double *out) {
  return interp_impl<double>(c, grid, field, gradient, out);
}
extern "C" tm_status tm_cic_interp_f32(tm_ctx *c, const tm_grid *grid, const float *field, const float *gradient,
                                       double *out) {
  return interp_impl<float>(c, grid, field, gradient, out);
}

// Moving the cells: updatepositions (src/cells.c:342-365) on the resident
// colony.
//
// Each cell moves by its force, then by a random step of randommove[type] *
// (u * 2 - 1) per axis (z only in 3-D), and is marked for removal when its
// sphere reaches a face of the box (outsidethebox, src/cells.c:41-58).  The
// reference draws u as (float)rand()/RAND_MAX from the program's one rand()
// stream, cell after cell, x then y then z.  Two sources of u:
//   draws     the caller's dim*n floats in that order (the reference-API
//             layer fills them from rand() so the program's stream advances
//             exactly as the reference's would), or
//   Philox    a counter-based Philox4x32-10 word per axis, counter (cell,
//             step), key seed, scaled like the reference's 31-bit rand()
//             values: independent of launch shape and of the sort order.
// The kernel runs over the Morton order: the sorted records, types and
// forces are read coalesced, and only the new x, y, z (and the removal mark)
// are scattered back into the caller's cell order.
#include "tm_internal.h"

namespace {

constexpr int BM = 256;

struct MoveArgs {
  float rm[TM_MAX_TYPES];  // randommove per cell type
  int ntypes;
  double hx, hy, hz;       // half box edges
};

MoveArgs move_args(const tm_ctx *c, const tm_move *m, int ntypes) {
  MoveArgs A;
  for (int t = 0; t < TM_MAX_TYPES; ++t) A.rm[t] = t < ntypes ? m->randommove[t] : 0.0f;
  A.ntypes = ntypes;
  A.hx = m->box[0] / 2.0;
  A.hy = m->box[1] / 2.0;
  A.hz = m->box[2] / 2.0;
  return A;
}

__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1, uint32_t out[4]) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    const uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
    const uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// SORTED: thread i moves cell perm[i] of the context's colony (records,
// types and forces in Morton order); otherwise thread i moves cell i of plain
// arrays in cell order.
template <bool SORTED>
__global__ __launch_bounds__(BM) void k_move(int64_t n, int dim, const int32_t *__restrict__ perm,
                                             const double4 *__restrict__ pos, const double *__restrict__ size,
                                             const int32_t *__restrict__ ctype, const double *__restrict__ fx,
                                             const double *__restrict__ fy, const double *__restrict__ fz,
                                             MoveArgs A, const float *__restrict__ draws, uint64_t seed,
                                             uint64_t step, const unsigned long long *__restrict__ step_dev,
                                             const unsigned int *__restrict__ maxcount, uint32_t cap,
                                             double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
                                             uint8_t *__restrict__ removecell, unsigned long long *__restrict__ removed,
                                             unsigned int *__restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * BM + threadIdx.x;
  if (maxcount && *maxcount > cap) return;  // this step's forces are incomplete: the host redoes it
  bool newly = false;
  if (i < n) {
    int64_t c = i;
    double4 P;
    if constexpr (SORTED) {
      c = perm[i];
      if (c < 0 || c >= n) c = 0;  // unreachable for a valid permutation; bounds the scatter
      P = pos[i];
    } else {
      P = make_double4(x[i], y[i], dim == 3 ? z[i] : 0.0, size[i]);
    }
    int32_t t = ctype[i];
    if (t < 0 || t >= A.ntypes) {
      atomicOr(flags, TM_FLAG_CTYPE);
      t = 0;
    }
    const double rm = (double)A.rm[t];
    float u[3];
    if (draws) {
#pragma unroll
      for (int a = 0; a < 3; ++a) u[a] = a < dim ? draws[dim * c + a] : 0.5f;
    } else {
      const uint64_t st = step_dev ? *step_dev : step;
      uint32_t w[4];
      philox4x32_10((uint32_t)c, (uint32_t)((uint64_t)c >> 32), (uint32_t)st, (uint32_t)(st >> 32), (uint32_t)seed,
                    (uint32_t)(seed >> 32), w);
#pragma unroll
      for (int a = 0; a < 3; ++a) u[a] = (float)(w[a] >> 1) / 2147483648.0f;
    }
    // the reference's order: x += F, then x += randommove * (u * 2.0 - 1.0)
    double nx = P.x + fx[i];
    double ny = P.y + fy[i];
    double nz = dim == 3 ? P.z + fz[i] : P.z;
    nx = nx + rm * ((double)u[0] * 2.0 - 1.0);
    ny = ny + rm * ((double)u[1] * 2.0 - 1.0);
    if (dim == 3) nz = nz + rm * ((double)u[2] * 2.0 - 1.0);
    x[c] = nx;
    y[c] = ny;
    if (dim == 3) z[c] = nz;
    const double r = P.w;
    const bool out = nx - r <= -A.hx || nx + r >= A.hx || ny - r <= -A.hy || ny + r >= A.hy ||
                     (dim == 3 && (nz - r <= -A.hz || nz + r >= A.hz));
    if (out) {
      if (!removecell) {
        newly = true;
      } else if (removecell[c] == 0) {
        removecell[c] = 1;
        newly = true;
      }
    }
  }
  // one atomic per wave for the count of newly marked cells
  const unsigned long long b = __ballot(newly);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(removed, (unsigned long long)__popcll(b));
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }
__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }
__global__ void k_inc_u64(unsigned long long *p, const unsigned int *maxcount, uint32_t cap) {
  if (*maxcount <= cap) *p += 1ull;
}

}  // namespace

tm_status tm_run_move(tm_ctx *c, const tm_move *m, const float *draws, uint8_t *removecell, bool graph) {
  const int64_t n = c->n;
  const MoveArgs A = move_args(c, m, c->ntypes);
  unsigned long long *cnt = graph ? tm_graph_removed_word(c) : tm_removed_word(c);
  unsigned long long *st = graph ? tm_graph_step_word(c) : nullptr;
  if (!graph) k_zero_u64<<<1, 1, 0, c->stream>>>(cnt);
  if (n > 0)
    k_move<true><<<tm_blocks(n, BM), BM, 0, c->stream>>>(
        n, c->dim, c->d_perm, c->d_pos, nullptr, c->d_ctype_s, c->d_fx, c->d_fy, c->d_fz, A, draws, m->seed, m->step,
        st, graph ? c->d_flags + 1 : nullptr, (uint32_t)c->nbr_alloc, c->own_x, c->own_y, c->own_z, removecell, cnt,
        (unsigned int *)tm_move_error_word(c));
  if (graph) k_inc_u64<<<1, 1, 0, c->stream>>>(st, c->d_flags + 1, (uint32_t)c->nbr_alloc);
  TM_HIP(c, hipGetLastError());
  return TM_OK;
}

tm_status tm_run_move_arrays(tm_ctx *c, int64_t n, int dim, double *x, double *y, double *z, const double *size,
                             const int32_t *ctype, const double *fx, const double *fy, const double *fz,
                             const tm_move *m, const float *draws, uint8_t *removecell) {
  const MoveArgs A = move_args(c, m, TM_MAX_TYPES);
  k_zero_u64<<<1, 1, 0, c->stream>>>(tm_removed_word(c));
  k_zero_u64<<<1, 1, 0, c->stream>>>(tm_move_error_word(c));
  if (n > 0)
    k_move<false><<<tm_blocks(n, BM), BM, 0, c->stream>>>(n, dim, nullptr, nullptr, size, ctype, fx, fy, fz, A, draws,
                                                          m->seed, m->step, nullptr, nullptr, 0u, x, y, z, removecell,
                                                          tm_removed_word(c), (unsigned int *)tm_move_error_word(c));
  TM_HIP(c, hipGetLastError());
  return TM_OK;
}

tm_status tm_set_graph_step(tm_ctx *c, uint64_t step) {
  k_set_u64<<<1, 1, 0, c->stream>>>(tm_graph_step_word(c), step);
  k_zero_u64<<<1, 1, 0, c->stream>>>(tm_graph_removed_word(c));
  TM_HIP(c, hipGetLastError());
  return TM_OK;
}

tm_status tm_take_graph_removed(tm_ctx *c, int64_t *count) {
  unsigned long long v = 0;
  TM_HIP(c, hipMemcpyAsync(&v, tm_graph_removed_word(c), sizeof(v), hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  k_zero_u64<<<1, 1, 0, c->stream>>>(tm_graph_removed_word(c));
  TM_HIP(c, hipGetLastError());
  if (count) *count = (int64_t)v;
  return TM_OK;
}

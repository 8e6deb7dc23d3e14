// SPH stage: density, v and kernel-gradient force per cell over its
// neighbour list, one lane per cell, pairs in the reference's accumulation
// order (descending Morton key).
//
// TM_FP64 restates computepotential/computegradient's pair arithmetic
// (src/potential.c:39-78,193-202, src/inline.h:40-83) operation by operation:
// no contraction (built with -ffp-contract=off), IEEE division and sqrt, the
// kernel value rounded to float before it is added (sphkernel returns float),
// so results equal the CPU path bit for bit.  Density and force share one
// distance and one u = d/h, which the reference computes twice identically.
//
// TM_MIXED keeps the exact neighbour set (built with the FP64 predicate),
// forms each pair's separation in FP64 and rounds it once to FP32, evaluates
// the kernel polynomials, potential and force vector in FP32 with the
// hardware reciprocal square root, and accumulates each cell's sums in FP64.
// Its v uses the closed form of the overlap-lens volume (algebraically equal
// to the reference's cap-difference expression, free of its cancellation).
#include "tm_internal.h"

namespace {

// src/potential.c:39-78 without the density/mindist side effects
__device__ __forceinline__ double pot_fp64(double4 P, float yp, double4 Q, float yq, double h) {
  const double dx = P.x - Q.x, dy = P.y - Q.y, dz = P.z - Q.z;
  const double dist = sqrt((dx * dx + dy * dy) + dz * dz);
  if (!(dist <= h)) return 0.0;
  const double s1 = P.w, s2 = Q.w;
  const double xc = s1 + s2 - dist;
  if (xc <= 0.0) return 0.0;
  const double poisson = 0.33;
  const double D = 0.75 * ((1.0 - poisson * poisson) / (double)yp + (1.0 - poisson * poisson / (double)yq));
  const double r01 = (s1 * s1 - s2 * s2 + dist * dist) / (2.0 * dist);
  const double r02 = dist - r01;
  const double area = M_PI * ((s1 * s1 * (s1 - r01) - (s1 * s1 * s1 - r01 * r01 * r01) / 3.0) +
                              (s2 * s2 * (s2 - r02) - (s2 * s2 * s2 - r02 * r02 * r02) / 3.0));
  // pow(xc, 5/2) with integer 5/2 == 2 is xc*xc
  return (2.0 * (xc * xc) / (5.0 * D)) * sqrt((s1 * s2) / (s1 + s2)) + area * 0.1;
}

// The overlap xc = (s1+s2) - d is formed as ((s1+s2)^2 - d^2) / ((s1+s2) + d)
// with the difference of squares in FP64, so it keeps its relative accuracy
// for barely touching cells.
__device__ __forceinline__ float pot_fp32(double4 P, float yp, double4 Q, float yq, double cut2) {
  const double ex = Q.x - P.x, ey = Q.y - P.y, ez = Q.z - P.z;
  const double d2d = (ex * ex + ey * ey) + ez * ez;
  if (!(d2d <= cut2)) return 0.0f;
  const float d2 = (float)d2d;
  const float d = d2 * __builtin_amdgcn_rsqf(d2);
  const float s1 = (float)P.w, s2 = (float)Q.w;
  const double sum = P.w + Q.w;
  const double gap2 = sum * sum - d2d;
  if (gap2 <= 0.0) return 0.0f;
  const float xc = (float)gap2 * __builtin_amdgcn_rcpf((s1 + s2) + d);
  const float pp = 0.1089f;
  const float D = 0.75f * ((1.0f - pp) * __builtin_amdgcn_rcpf(yp) + (1.0f - pp * __builtin_amdgcn_rcpf(yq)));
  const float ds = s1 - s2;
  // lens volume pi xc^2 (d^2 + 2 d (s1+s2) - 3 (s1-s2)^2) / (12 d)
  const float lens = 3.14159265f * xc * xc * (d2 + 2.0f * d * (s1 + s2) - 3.0f * ds * ds) *
                     __builtin_amdgcn_rcpf(12.0f * d);
  const float red = s1 * s2 * __builtin_amdgcn_rcpf(s1 + s2);
  return (0.4f * xc * xc) * __builtin_amdgcn_rcpf(D) * (red * __builtin_amdgcn_rsqf(red)) + 0.1f * lens;
}

// one pair of the FP64 drop-in: sphkernel (src/inline.h:40-51) into the
// density and sphgradient (src/inline.h:59-83) into the force, expressions
// and evaluation order as in the reference; 0 <= u <= 1 for listed neighbours
__device__ __forceinline__ void pair_fp64(const double4 &P, const double4 &Q, double h, double sc, double csize,
                                          double cW, double cF, bool dim3, double &de, double &f0, double &f1,
                                          double &f2) {
  const double dx = P.x - Q.x, dy = P.y - Q.y, dz = P.z - Q.z;
  const double dist = sqrt((dx * dx + dy * dy) + dz * dz);
  const double u = dist / h;
  double W;
  if (u <= 0.5) W = cW * (1 - 6 * u * u + 6 * u * u * u);
  else W = cW * (2 * (1 - u) * (1 - u) * (1 - u));
  de += sc * (Q.w / csize) * (double)(float)W;
  double w;
  if (u <= 0.5) w = -2.0 * u + 3.0 * u * u;
  else w = -(1.0 - u) * (1.0 - u);
  f0 += w * cF * (Q.x - P.x) / dist;
  f1 += w * cF * (Q.y - P.y) / dist;
  if (dim3) f2 += w * cF * (Q.z - P.z) / dist;
}

// one pair of the mixed form (FP32 evaluation, FP64 separation and sums)
__device__ __forceinline__ void pair_mixed(const double4 &P, const double4 &Q, float hinv, float hf, float hhalf,
                                           double hh, float cW, float cF, float scs, bool dim3, double &de,
                                           double &f0, double &f1, double &f2) {
  // separation and squared distance in FP64, rounded once to FP32
  const double ex = Q.x - P.x, ey = Q.y - P.y, ez = Q.z - P.z;
  const double d2d = __builtin_fma(ez, ez, __builtin_fma(ey, ey, ex * ex));
  const float dx = (float)ex, dy = (float)ey, dz = (float)ez;
  const float d2 = (float)d2d;
  const float rinv = __builtin_amdgcn_rsqf(d2);
  const float d = d2 * rinv;
  float W, w;
  if (d <= hhalf) {
    const float u = d * hinv;
    W = __builtin_fmaf(u * u, __builtin_fmaf(6.0f, u, -6.0f), 1.0f);
    w = u * __builtin_fmaf(3.0f, u, -2.0f);
  } else {
    // 1 - u = (h^2 - d^2) / ((h + d) h): the difference of squares in FP64
    // keeps the outer branch accurate up to the cutoff
    const float om = fmaxf((float)(hh - d2d), 0.0f) * __builtin_amdgcn_rcpf((hf + d) * hf);
    W = 2.0f * om * om * om;
    w = -om * om;
  }
  de += (double)((scs * (float)Q.w) * (cW * W));
  const float coef = w * cF * rinv;
  f0 += (double)(coef * dx);
  f1 += (double)(coef * dy);
  if (dim3) f2 += (double)(coef * dz);
}

// Pair kernels.  L consecutive lanes share one cell and take its list slots
// sub, sub+L, sub+2L, ... (L = 1 for the FP64 drop-in, which must add the
// pairs in list order).  Each trip issues U position gathers for slots whose
// indices were loaded one trip earlier, then loads the next trip's indices,
// then evaluates the U pairs in order: the index round trip to memory overlaps
// the gathers and the arithmetic instead of preceding them.  With L > 1 the
// lanes' FP64 partial sums are added across the group at the end.
template <bool MIXED, int U, int L>
__global__ __launch_bounds__(256) void k_sph(const double4 *__restrict__ pos, const float *__restrict__ young_s,
                                             const int32_t *__restrict__ ctype_s, const TypeDev *__restrict__ types,
                                             int64_t n, int64_t cap, const uint32_t *__restrict__ nbr,
                                             const uint32_t *__restrict__ count, const int32_t *__restrict__ qstar,
                                             int dim, int balance, double *__restrict__ dens,
                                             double *__restrict__ vout, double *__restrict__ fx,
                                             double *__restrict__ fy, double *__restrict__ fz) {
  const int64_t blk0 = tm_xcd_block() * blockDim.x;
  int64_t t = blk0 + threadIdx.x;
  if (!MIXED && balance) {
    // Rank the workgroup's cells by list length (counting sort in LDS) and
    // give wave w the w-th quarter, so the lanes of a wave run loops of
    // similar length.  Each cell is still summed by one lane in list order.
    __shared__ uint32_t bins[128];
    __shared__ uint32_t order[256];
    const uint32_t b = t < n ? min(min((uint32_t)cap, count[t]), 127u) : 0u;
    if (threadIdx.x < 128) bins[threadIdx.x] = 0;
    __syncthreads();
    atomicAdd(&bins[b], 1u);
    __syncthreads();
    if (threadIdx.x < 64) {
      const uint32_t lane = threadIdx.x;
      const uint32_t a0 = bins[2 * lane], a1 = bins[2 * lane + 1];
      uint32_t inc = a0 + a1;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += y;
      }
      const uint32_t exc = inc - (a0 + a1);
      bins[2 * lane] = exc;
      bins[2 * lane + 1] = exc + a0;
    }
    __syncthreads();
    order[atomicAdd(&bins[b], 1u)] = threadIdx.x;
    __syncthreads();
    t = blk0 + order[threadIdx.x];
  }
  const int64_t cell = t / L;
  const uint32_t sub = (uint32_t)(t % L);
  const bool live = cell < n;
  if (L == 1 && !live) return;
  const int64_t p = live ? cell : n - 1;  // idle lanes of the last group stay for the reduction
  const bool dim3 = dim == 3;
  const double4 P = pos[p];
  const TypeDev &T = types[ctype_s[p]];
  const double h = T.h, sc = T.sc, csize = T.size, cW = T.cW, cF = T.cF;
  const float hinv = T.hinv_f, cWf = T.cW_f, cFf = T.cF_f, scs = T.sc_over_size_f;
  const float hf = (float)T.h, hhalf = 0.5f * (float)T.h;
  const double hh = T.h * T.h;
  const uint32_t cnt = live ? min((uint32_t)cap, count[p]) : 0u;
  double de = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
  const uint32_t *col = nbr + tm_slot(p, 0, cap);  // slot k of this cell at col[64 k]
  constexpr uint32_t STEP = (uint32_t)(U * L);
  uint32_t nxt[U];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    const uint32_t k = sub + (uint32_t)(L * j);
    nxt[j] = k < cnt ? col[(int64_t)k << 6] : (uint32_t)p;
  }
  for (uint32_t k0 = sub; k0 < cnt; k0 += STEP) {
    double4 Q[U];
#pragma unroll
    for (int j = 0; j < U; ++j) Q[j] = pos[nxt[j]];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t k = k0 + STEP + (uint32_t)(L * j);
      nxt[j] = k < cnt ? col[(int64_t)k << 6] : (uint32_t)p;
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      if (k0 + (uint32_t)(L * j) < cnt) {
        if (MIXED) pair_mixed(P, Q[j], hinv, hf, hhalf, hh, cWf, cFf, scs, dim3, de, f0, f1, f2);
        else pair_fp64(P, Q[j], h, sc, csize, cW, cF, dim3, de, f0, f1, f2);
      }
    }
  }
  if (L > 1) {
#pragma unroll
    for (int o = 1; o < L; o <<= 1) {
      de += __shfl_xor(de, o, 64);
      f0 += __shfl_xor(f0, o, 64);
      f1 += __shfl_xor(f1, o, 64);
      f2 += __shfl_xor(f2, o, 64);
    }
    if (sub != 0 || !live) return;
  }
  const int32_t qs = qstar[p];
  dens[p] = de;
  if (MIXED) vout[p] = qs >= 0 ? (double)pot_fp32(P, young_s[p], pos[qs], young_s[qs], T.cut2) : 0.0;
  else vout[p] = qs >= 0 ? pot_fp64(P, young_s[p], pos[qs], young_s[qs], h) : 0.0;
  fx[p] = f0;
  fy[p] = f1;
  fz[p] = f2;
}

__global__ void k_unpermute(const int32_t *__restrict__ perm, int64_t n, const double *__restrict__ a0,
                            const double *__restrict__ a1, const double *__restrict__ a2,
                            const double *__restrict__ a3, const double *__restrict__ a4,
                            const double *__restrict__ a5, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t o = perm[i];
  out[o] = a0[i];
  out[n + o] = a1[i];
  out[2 * n + o] = a2[i];
  out[3 * n + o] = a3[i];
  out[4 * n + o] = a4[i];
  out[5 * n + o] = a5[i];
}

template <bool MIXED, int U, int L>
void launch_ul(tm_ctx *c) {
  const int64_t n = c->n;
  const int bs = c->sph_block;
  k_sph<MIXED, U, L><<<tm_blocks(n * L, bs), bs, 0, c->stream>>>(
      c->d_pos, c->d_young_s, c->d_ctype_s, c->d_types, n, c->nbr_alloc, c->d_nbr, c->d_ncount, c->d_qstar, c->dim,
      c->sph_balance, c->d_dens, c->d_v, c->d_fx, c->d_fy, c->d_fz);
}

template <bool MIXED, int L>
void launch_u(tm_ctx *c) {
  switch (MIXED ? c->sph_unroll_mixed : c->sph_unroll_fp64) {
    case 1: launch_ul<MIXED, 1, L>(c); break;
    case 2: launch_ul<MIXED, 2, L>(c); break;
    case 8: launch_ul<MIXED, 8, L>(c); break;
    default: launch_ul<MIXED, 4, L>(c); break;
  }
}

template <bool MIXED>
void launch_sph(tm_ctx *c) {
  if (!MIXED) {
    launch_u<false, 1>(c);  // the FP64 drop-in adds pairs in list order: one lane per cell
    return;
  }
  switch (c->sph_lanes) {
    case 2: launch_u<true, 2>(c); break;
    case 4: launch_u<true, 4>(c); break;
    case 8: launch_u<true, 8>(c); break;
    default: launch_u<true, 1>(c); break;
  }
}

}  // namespace

tm_status tm_run_sph(tm_ctx *c, tm_precision prec) {
  if (c->stage < 3) return tm_fail(c, TM_ERR_STATE, "tm_sph before tm_neighbors");
  if (c->n == 0) {
    c->stage = 4;
    return TM_OK;
  }
  if (prec == TM_MIXED) launch_sph<true>(c);
  else launch_sph<false>(c);
  TM_HIP(c, hipGetLastError());
  c->stage = 4;
  c->last_prec = prec;
  return TM_OK;
}

tm_status tm_run_unpermute(tm_ctx *c) {
  const int64_t n = c->n;
  if (n == 0) return TM_OK;
  k_unpermute<<<tm_blocks(n, 256), 256, 0, c->stream>>>(c->d_perm, n, c->d_dens, c->d_v, c->d_mind, c->d_fx,
                                                         c->d_fy, c->d_fz, c->d_out);
  TM_HIP(c, hipGetLastError());
  return TM_OK;
}

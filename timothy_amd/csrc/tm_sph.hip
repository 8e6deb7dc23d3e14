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

template <int BS, int DIM>
__global__ __launch_bounds__(BS) void k_sph_fp64(const double4 *__restrict__ pos, const float *__restrict__ young_s,
                                                 const int32_t *__restrict__ ctype_s,
                                                 const TypeDev *__restrict__ types, int64_t n, int64_t cap,
                                                 const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ count,
                                                 const int32_t *__restrict__ qstar, double *__restrict__ dens,
                                                 double *__restrict__ vout, double *__restrict__ fx,
                                                 double *__restrict__ fy, double *__restrict__ fz) {
  const int64_t p = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (p >= n) return;
  const double4 P = pos[p];
  const TypeDev &T = types[ctype_s[p]];
  const double h = T.h, sc = T.sc, csize = T.size, cW = T.cW, cF = T.cF;
  const uint32_t cnt = min((uint32_t)cap, count[p]);
  double de = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
  const uint32_t *col = nbr + p;
  for (uint32_t k = 0; k < cnt; ++k) {
    const uint32_t q = col[(int64_t)k * n];
    const double4 Q = pos[q];
    const double dx = P.x - Q.x, dy = P.y - Q.y, dz = P.z - Q.z;
    const double dist = sqrt((dx * dx + dy * dy) + dz * dz);
    const double u = dist / h;
    // sphkernel (src/inline.h:40-51): 0 <= u <= 1 for every listed neighbour
    double W;
    if (u <= 0.5) W = cW * (1 - 6 * u * u + 6 * u * u * u);
    else W = cW * (2 * (1 - u) * (1 - u) * (1 - u));
    de += sc * (Q.w / csize) * (double)(float)W;
    // sphgradient (src/inline.h:59-83)
    double w;
    if (u <= 0.5) w = -2.0 * u + 3.0 * u * u;
    else w = -(1.0 - u) * (1.0 - u);
    f0 += w * cF * (Q.x - P.x) / dist;
    f1 += w * cF * (Q.y - P.y) / dist;
    if (DIM == 3) f2 += w * cF * (Q.z - P.z) / dist;
  }
  const int32_t qs = qstar[p];
  dens[p] = de;
  vout[p] = qs >= 0 ? pot_fp64(P, young_s[p], pos[qs], young_s[qs], h) : 0.0;
  fx[p] = f0;
  fy[p] = f1;
  fz[p] = f2;
}

template <int BS, int DIM>
__global__ __launch_bounds__(BS) void k_sph_mixed(const double4 *__restrict__ pos, const float *__restrict__ young_s,
                                                  const int32_t *__restrict__ ctype_s,
                                                  const TypeDev *__restrict__ types, int64_t n, int64_t cap,
                                                  const uint32_t *__restrict__ nbr,
                                                  const uint32_t *__restrict__ count,
                                                  const int32_t *__restrict__ qstar, double *__restrict__ dens,
                                                  double *__restrict__ vout, double *__restrict__ fx,
                                                  double *__restrict__ fy, double *__restrict__ fz) {
  const int64_t p = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (p >= n) return;
  const double4 P = pos[p];
  const TypeDev &T = types[ctype_s[p]];
  const float hinv = T.hinv_f, cW = T.cW_f, cF = T.cF_f, scs = T.sc_over_size_f;
  const float hf = (float)T.h, hhalf = 0.5f * (float)T.h;
  const double hh = T.h * T.h;
  const uint32_t cnt = min((uint32_t)cap, count[p]);
  double de = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
  const uint32_t *col = nbr + p;
  for (uint32_t k = 0; k < cnt; ++k) {
    const uint32_t q = col[(int64_t)k * n];
    const double4 Q = pos[q];
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
    if (DIM == 3) f2 += (double)(coef * dz);
  }
  const int32_t qs = qstar[p];
  dens[p] = de;
  vout[p] = qs >= 0 ? (double)pot_fp32(P, young_s[p], pos[qs], young_s[qs], T.cut2) : 0.0;
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

template <int BS>
void launch_sph(tm_ctx *c, tm_precision prec) {
  const int64_t n = c->n;
  const unsigned g = tm_blocks(n, BS);
#define TM_SPH_ARGS                                                                                         \
  c->d_pos, c->d_young_s, c->d_ctype_s, c->d_types, n, c->nbr_alloc, c->d_nbr, c->d_ncount, c->d_qstar, \
      c->d_dens, c->d_v, c->d_fx, c->d_fy, c->d_fz
  if (prec == TM_FP64) {
    if (c->dim == 3) k_sph_fp64<BS, 3><<<g, BS, 0, c->stream>>>(TM_SPH_ARGS);
    else k_sph_fp64<BS, 2><<<g, BS, 0, c->stream>>>(TM_SPH_ARGS);
  } else {
    if (c->dim == 3) k_sph_mixed<BS, 3><<<g, BS, 0, c->stream>>>(TM_SPH_ARGS);
    else k_sph_mixed<BS, 2><<<g, BS, 0, c->stream>>>(TM_SPH_ARGS);
  }
#undef TM_SPH_ARGS
}

}  // namespace

tm_status tm_run_sph(tm_ctx *c, tm_precision prec) {
  if (c->stage < 3) return tm_fail(c, TM_ERR_STATE, "tm_sph before tm_neighbors");
  if (c->n == 0) {
    c->stage = 4;
    return TM_OK;
  }
  switch (c->sph_block) {
    case 64: launch_sph<64>(c, prec); break;
    case 128: launch_sph<128>(c, prec); break;
    case 512: launch_sph<512>(c, prec); break;
    default: launch_sph<256>(c, prec); break;
  }
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

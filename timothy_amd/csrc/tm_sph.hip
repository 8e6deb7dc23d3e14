// SPH stage: density and kernel-gradient force per cell over its neighbour
// list, pairs in the reference's accumulation order (descending Morton key).
// v and mindist come from the neighbour walk, which visits the candidates.
//
// TM_FP64 restates computepotential/computegradient's pair arithmetic
// (src/potential.c:39-78,193-202, src/inline.h:40-83) operation by operation:
// no contraction (built with -ffp-contract=off), correctly rounded division
// and sqrt, the kernel value rounded to float before it is added (sphkernel
// returns float), so results equal the CPU path bit for bit.  Density and
// force share one distance and one u = d/h, which the reference computes
// twice identically.
//
// TM_MIXED keeps the exact neighbour set (built with the FP64 predicate),
// forms each pair's separation in FP64 and rounds it once to FP32, evaluates
// the kernel polynomials and force vector in FP32 with the hardware
// reciprocal square root, and accumulates each cell's sums in FP64.
#include "tm_internal.h"

namespace {

// a / b correctly rounded from y = RN(1/b): q0 = RN(a y) is within one ulp of
// a / b, the remainder a - b q0 is exact in one FMA, and RN(q0 + r y) is the
// correctly rounded quotient (Markstein, IBM J. Res. Dev. 34(1), 1990,
// Theorem 1), i.e. the same double as the IEEE division a / b.
__device__ __forceinline__ double div_rn(double a, double b, double y) {
  const double q0 = a * y;
  const double r = __builtin_fma(-q0, b, a);
  return __builtin_fma(r, y, q0);
}

// one pair of the FP64 drop-in: sphkernel (src/inline.h:40-51) into the
// density and sphgradient (src/inline.h:59-83) into the force, expressions
// and evaluation order as in the reference; 0 <= u <= 1 for listed
// neighbours.  The five divisions of the reference (d/h, r_q/size and the
// three force components by d) give the same doubles through div_rn: 1/h and
// 1/size are per-type constants, 1/d is one IEEE division shared by the
// three components.
struct TypeFP64 {
  double h, hinv, sc, size, sinv, cW, cF;
};

__device__ __forceinline__ void pair_fp64(const double4 &P, const double4 &Q, const TypeFP64 &T, bool dim3,
                                          double &de, double &f0, double &f1, double &f2) {
  const double dx = P.x - Q.x, dy = P.y - Q.y, dz = P.z - Q.z;
  const double dist = sqrt((dx * dx + dy * dy) + dz * dz);
  const double u = div_rn(dist, T.h, T.hinv);
  const double cW = T.cW, cF = T.cF;
  double W;
  if (u <= 0.5) W = cW * (1 - 6 * u * u + 6 * u * u * u);
  else W = cW * (2 * (1 - u) * (1 - u) * (1 - u));
  de += T.sc * div_rn(Q.w, T.size, T.sinv) * (double)(float)W;
  double w;
  if (u <= 0.5) w = -2.0 * u + 3.0 * u * u;
  else w = -(1.0 - u) * (1.0 - u);
  const double rd = 1.0 / dist;
  f0 += div_rn(w * cF * (Q.x - P.x), dist, rd);
  f1 += div_rn(w * cF * (Q.y - P.y), dist, rd);
  if (dim3) f2 += div_rn(w * cF * (Q.z - P.z), dist, rd);
}

// One pair of the mixed form without its per-type constants, which are
// applied once per cell: the density weight size_q * W(u) and the force
// coefficient w(u) / d in FP32, with the FP32 separation.  The separation and
// d^2 are formed in FP64 and rounded once.  With om = 1 - u and
// m = max(1 - 2u, 0), the two pieces of the cubic spline combine into
//   W = 2 om^3 - m^3   and   w = -(dW/du)/6 = m^2 - om^2,
// exact for both u <= 1/2 (where 2 om^3 - (1-2u)^3 = 1 - 6u^2 + 6u^3) and
// u > 1/2 (m = 0), so no lane of a wave branches.  The force coefficient is
// evaluated as
//   w / d = -min(1/h, om/d) (om + m),
// the same function (om - m = u below 1/2, m = 0 above, and u/d = 1/h)
// without the cancellation of m^2 - om^2 as u -> 0, so a very close pair
// keeps its relative accuracy.  om comes from the FP64 difference of squares,
// om = (h^2 - d^2) / ((h + d) h), so a near-cutoff term keeps its relative
// accuracy too.  Beyond the cutoff both vanish exactly (om = m = 0), which is
// what makes the far stand-in record a zero term.
// Per-cell constants of the mixed pairs: h (FP32), h^2 (FP32 and FP64), -2/h,
// 1/h.
struct MixedC {
  float hf, hhf, m2hinv, hinv;
  double hh;
};

__device__ __forceinline__ MixedC mixed_consts(const TypeDev &T) {
  const float hf = (float)T.h;
  return MixedC{hf, hf * hf, -2.0f * T.hinv_f, T.hinv_f, T.h * T.h};
}

// One pair of the mixed form, added into the cell's sums: the separation
// and its square in FP64, the spline in FP32 with d from one reciprocal
// square root, om = 1 - d/h formed as (h^2 - d^2) / (h (h + d)) with h^2 - d^2
// in FP64 (so om keeps its relative accuracy where it vanishes at the
// cutoff, the terms there scaling as om^2 and om^3) and m = 1 - 2 d/h, both
// clamped at 0 (W = 2 om^3 - m^3 covers both branches of the M4 kernel and
// the far stand-in gives exact zeros); sw += size_q * W in FP32.  The force
// coefficient w / d (w = m^2 - om^2, the kernel's slope) is formed without
// the cancellation of m^2 - om^2 at small d: w = -(om + m) u for u <= 1/2
// and -(om + m) om beyond (m = 0), so w / d = -min(1/h, om / d) (om + m),
// the minimum picking the branch (om / d >= 1/h exactly when u <= 1/2); it
// multiplies the FP64 separation, added in FP64.
template <bool DIM3>
__device__ __forceinline__ void pair_mixed(const double4 &P, double qx, double qy, double qz, float sq,
                                           const MixedC &K, float &sw, double &f0, double &f1, double &f2) {
  const double ex = qx - P.x, ey = qy - P.y, ez = qz - P.z;
  const double d2d = __builtin_fma(ez, ez, __builtin_fma(ey, ey, ex * ex));
  const float d2 = (float)d2d;
  const float rinv = __builtin_amdgcn_rsqf(d2);
  const float d = d2 * rinv;
  const float om = fmaxf((float)(K.hh - d2d), 0.0f) * __builtin_amdgcn_rcpf(__builtin_fmaf(d, K.hf, K.hhf));
  const float m = fmaxf(__builtin_fmaf(d, K.m2hinv, 1.0f), 0.0f);
  const float om2 = om * om, m2 = m * m;
  sw = __builtin_fmaf(sq, __builtin_fmaf(om2 + om2, om, -(m2 * m)), sw);
  const double coef = (double)(-fminf(K.hinv, om * rinv) * (om + m));
  f0 = __builtin_fma(coef, ex, f0);
  f1 = __builtin_fma(coef, ey, f1);
  if (DIM3) f2 = __builtin_fma(coef, ez, f2);
}

typedef unsigned int tm_u32x4 __attribute__((ext_vector_type(4)));
typedef double tm_f64x2 __attribute__((ext_vector_type(2)));

// Position record q.  With BUF the array is addressed through a buffer
// resource and a 32-bit byte offset (one shift per gather instead of 64-bit
// address arithmetic; reads past the resource's end return zeros); this needs
// 32 n < 2^32.
template <bool BUF>
__device__ __forceinline__ double4 gather(const double4 *__restrict__ pos, __amdgpu_buffer_rsrc_t rs, uint32_t q) {
  if constexpr (BUF) {
    const uint32_t o = q << 5;
    const tm_f64x2 a = __builtin_bit_cast(tm_f64x2, __builtin_amdgcn_raw_buffer_load_b128(rs, o, 0, 0));
    const tm_f64x2 b = __builtin_bit_cast(tm_f64x2, __builtin_amdgcn_raw_buffer_load_b128(rs, o + 16u, 0, 0));
    return make_double4(a.x, a.y, b.x, b.y);
  } else {
    return pos[q];
  }
}

// Per-cell sums over one cell's list, shared by both pair kernels.  L
// consecutive lanes share one cell and take its list slots sub, sub+L, ...
// (L = 1 for the FP64 drop-in, which must add the pairs in list order).  The
// trips of a wave run to the longest list among its cells (wmax, wave-
// uniform).  slot(kb) returns the record number in slot kb + sub (kb wave-
// uniform), fetch(q) the record.  Each trip fetches U records whose numbers
// were read one trip earlier, then reads the next trip's numbers, then
// evaluates the U pairs in list order.  The mixed form adds a trip's U terms
// of the kernel value in FP32 and the trip's sum into the FP64 per-cell
// density, the force terms straight into FP64; its padding slots hold the
// far stand-in record, whose terms are exactly zero.
// The FP64 form evaluates only the slots below the cell's own count.
template <bool MIXED, int U, int L, bool DIM3, class Slot, class Fetch>
__device__ __forceinline__ void cell_sums(const double4 &P, const TypeFP64 &C64, const MixedC &K, uint32_t cnt, uint32_t wmax, uint32_t far, Slot slot,
                                          Fetch fetch, double &de, double &f0, double &f1, double &f2) {
  constexpr uint32_t STEP = (uint32_t)(U * L);
  uint32_t nxt[U];
#pragma unroll
  for (int j = 0; j < U; ++j) nxt[j] = (uint32_t)(L * j) < wmax ? slot((uint32_t)(L * j)) : far;
  for (uint32_t k0 = 0; k0 < wmax; k0 += STEP) {
    double4 Q[U];
#pragma unroll
    for (int j = 0; j < U; ++j) Q[j] = fetch(nxt[j]);
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t kb = k0 + STEP + (uint32_t)(L * j);
      nxt[j] = kb < wmax ? slot(kb) : far;
    }
    if constexpr (MIXED) {
      float sw = 0.0f;
#pragma unroll
      for (int j = 0; j < U; ++j)
        pair_mixed<DIM3>(P, Q[j].x, Q[j].y, Q[j].z, (float)Q[j].w, K, sw, f0, f1, f2);
      de += (double)sw;
    } else {
#pragma unroll
      for (int j = 0; j < U; ++j)
        if (k0 + (uint32_t)j < cnt) pair_fp64(P, Q[j], C64, DIM3, de, f0, f1, f2);
    }
  }
}

// wave-wide maximum (every lane of the wave takes part)
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return __builtin_amdgcn_readfirstlane(v);
}

// the per-type constants of the mixed sums, applied once per cell
template <bool MIXED>
__device__ __forceinline__ void finish_cell(const TypeDev &T, double &de, double &f0, double &f1, double &f2) {
  if (MIXED) {
    de *= T.sc / T.size * T.cW;
    f0 *= T.cF;
    f1 *= T.cF;
    f2 *= T.cF;
  }
}

// Sums of the cells whose 32-bit lists live in the slot groups from g0 on,
// records gathered from the position array: the pair kernel for tiles that
// are not staged, and for launches without tiles.  lane_off is this lane's
// byte offset in the list descriptor rl.
template <bool MIXED, int U, int L, bool DIM3, bool BUF>
__device__ __forceinline__ void direct_sums(const double4 *__restrict__ pos, int64_t n, int64_t cap,
                                            __amdgpu_buffer_rsrc_t rl, uint32_t lane_off, uint32_t sub,
                                            const double4 &P, const TypeFP64 &C64, const MixedC &K, uint32_t cnt,
                                            uint32_t wmax, double &de,
                                            double &f0, double &f1, double &f2) {
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc((void *)pos, 0, BUF ? (int)(uint32_t)(32 * (n + 1)) : 0, 0x00020000);
  const uint32_t far = (uint32_t)n;
  // Slots from a cell's count up to the longest list of its 64-cell group
  // hold the stand-in n.  With one lane per cell in Morton order a wave is
  // such a group, so every slot below wmax is a neighbour or the stand-in:
  // no per-lane test.  Otherwise a lane's slots past its own count are
  // replaced by the stand-in (with several lanes per cell the slot is first
  // clamped into the cell's rows).
  auto slot = [&](uint32_t kb) -> uint32_t {
    if constexpr (MIXED && L == 1) {
      return __builtin_amdgcn_raw_buffer_load_b32(rl, lane_off, (int)(256 * kb), 0);
    } else {
      const uint32_t k = kb + sub;
      const uint32_t v =
          __builtin_amdgcn_raw_buffer_load_b32(rl, lane_off + 256u * min(k, (uint32_t)cap - 1u), 0, 0);
      return k < cnt ? v : far;
    }
  };
  auto fetch = [&](uint32_t q) -> double4 { return gather<BUF>(pos, rs, q); };
  cell_sums<MIXED, U, L, DIM3>(P, C64, K, cnt, wmax, far, slot, fetch, de, f0, f1, f2);
}

// descriptor over a list array from element e0 on (reads past its end return 0)
template <class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t list_rsrc(const T *a, int64_t e0, int64_t elems) {
  const int64_t rem = (int64_t)sizeof(T) * (elems - e0);
  const uint32_t nrec = __builtin_amdgcn_readfirstlane((uint32_t)(rem > 0xFFFFFFFFll ? 0xFFFFFFFFll : rem));
  return __builtin_amdgcn_make_buffer_rsrc((void *)(a + e0), 0, (int)nrec, 0x00020000);
}

// Pair kernel without tiles: one workgroup per blockDim/L consecutive cells;
// the FP64 form may rank a workgroup's cells by list length (balance).
template <bool MIXED, int U, int L, bool DIM3, bool BUF>
__global__ __launch_bounds__(256) void k_sph(const double4 *__restrict__ pos, const int32_t *__restrict__ ctype_s,
                                             const TypeDev *__restrict__ types, int64_t n, int64_t cap,
                                             const uint32_t *__restrict__ nbr, int64_t nbr_elems,
                                             const uint32_t *__restrict__ count, int balance,
                                             double *__restrict__ dens, double *__restrict__ fx,
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
  const int64_t p = live ? cell : n - 1;  // idle lanes take part in the wave-wide steps only
  const double4 P = pos[p];
  const TypeDev &T = types[ctype_s[p]];
  const TypeFP64 C64{T.h, T.hinv, T.sc, T.size, T.sinv, T.cW, T.cF};
  const MixedC K = mixed_consts(T);
  const uint32_t cnt = live ? min((uint32_t)cap, count[p]) : 0u;
  const uint32_t wmax = wave_max(cnt);
  // The lists of this workgroup's cells lie in the slot groups from the one
  // holding its first cell on: one descriptor based there serves the whole
  // workgroup, a lane's slot k being at byte lane_off + 256 k.
  const int64_t g0 = (blk0 / L) >> 6;
  const __amdgpu_buffer_rsrc_t rl = list_rsrc(nbr, g0 * cap * 64, nbr_elems);
  const uint32_t lane_off = (uint32_t)((((p >> 6) - g0) * cap * 64 + (p & 63)) * 4);
  double de = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
  direct_sums<MIXED, U, L, DIM3, BUF>(pos, n, cap, rl, lane_off, sub, P, C64, K, cnt, wmax, de,
                                      f0, f1, f2);
  if (L > 1) {
#pragma unroll
    for (int o = 1; o < L; o <<= 1) {
      de += __shfl_xor(de, o, 64);
      f0 += __shfl_xor(f0, o, 64);
      f1 += __shfl_xor(f1, o, 64);
      f2 += __shfl_xor(f2, o, 64);
    }
  }
  if (!live || sub != 0) return;
  finish_cell<MIXED>(T, de, f0, f1, f2);
  dens[p] = de;
  fx[p] = f0;
  fy[p] = f1;
  fz[p] = f2;
}

// The size half of a staged record: FP64 for the drop-in; for the mixed form
// the FP32 size in the low word (the pair reads it without a conversion).
template <bool MIXED>
__device__ __forceinline__ double stage_size(double size) {
  if constexpr (MIXED) return __longlong_as_double((long long)__float_as_uint((float)size));
  else return size;
}

// Pair kernel over LDS tiles: one workgroup of TS lanes per tile of TS cells.  The
// tile's distinct neighbours (and the far stand-in after them) are copied to
// LDS once, as two arrays of 16-byte halves (x, y and z, size: consecutive
// records start in different bank groups, which halves bank conflicts
// against 32-byte records).  Lane r takes the cell of rank r and reads its
// 16-bit list 8 slots per load, every pair's record coming from LDS.  A tile
// with more distinct neighbours than the LDS holds takes the 32-bit lists and
// global records instead.
template <bool MIXED, int U, int TS>
__global__ __launch_bounds__(TS) void k_sph_tile(
    const double4 *__restrict__ pos, const int32_t *__restrict__ ctype_s, const TypeDev *__restrict__ types,
    int64_t n, int64_t cap, const uint32_t *__restrict__ nbr, int64_t nbr_elems, const uint32_t *__restrict__ count,
    const uint32_t *__restrict__ tile_idx, const uint32_t *__restrict__ tile_cnt,
    const uint16_t *__restrict__ tile_cell, const uint16_t *__restrict__ nbr16, int64_t nbr16_elems,
    uint32_t lds_cap, double *__restrict__ dens, double *__restrict__ fx, double *__restrict__ fy,
    double *__restrict__ fz) {
  extern __shared__ double2 lds[];
  double2 *const lxy = lds;
  double2 *const lzs = lds + lds_cap;
  const int64_t g = tm_xcd_block();
  const uint32_t t = threadIdx.x;
  // The tile's size, the first staging indices and the rank map are read
  // together, before the size is known (all lie inside their arrays; values
  // of an over-full tile are never used), so staging waits on one load
  // round trip for the indices instead of two.
  const uint32_t *ti = tile_idx + g * TM_TILE_MAX;
  constexpr int SPEC = 3;  // 3 TS <= TM_TILE_MAX indices ahead
  static_assert(SPEC * TS <= TM_TILE_MAX, "staging indices read ahead");
  uint32_t si[SPEC];
#pragma unroll
  for (int j = 0; j < SPEC; ++j) si[j] = ti[t + TS * j];
  const uint32_t rank_cell = tile_cell[g * TS + t];
  const uint32_t m = __builtin_amdgcn_readfirstlane(tile_cnt[g]);
  const bool staged = m < lds_cap;  // m records and the stand-in fit (TM_TILE_OVER never does)
  if (staged) {
#pragma unroll
    for (int j = 0; j < SPEC; ++j) {
      const uint32_t i = t + TS * j;
      if (i < m) {
        const double4 q = pos[si[j]];
        lxy[i] = make_double2(q.x, q.y);
        lzs[i] = make_double2(q.z, stage_size<MIXED>(q.w));
      }
    }
    for (uint32_t i = t + TS * SPEC; i < m; i += TS) {
      const double4 q = pos[ti[i]];
      lxy[i] = make_double2(q.x, q.y);
      lzs[i] = make_double2(q.z, stage_size<MIXED>(q.w));
    }
    if (t == 0) {
      lxy[m] = make_double2(TM_FAR, TM_FAR);
      lzs[m] = make_double2(TM_FAR, 0.0);
    }
  }
  const uint32_t off = staged ? (uint32_t)rank_cell : t;
  const int64_t cell = g * TS + off;
  const bool live = cell < n;
  const int64_t p = live ? cell : n - 1;
  const double4 P = pos[p];
  const TypeDev &T = types[ctype_s[p]];
  const TypeFP64 C64{T.h, T.hinv, T.sc, T.size, T.sinv, T.cW, T.cF};
  const MixedC K = mixed_consts(T);
  const uint32_t cnt = live ? min((uint32_t)cap, count[p]) : 0u;
  const uint32_t wmax = wave_max(cnt);
  double de = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
  __syncthreads();
  if (staged) {
    // every chunk below the wave's longest list is written (the padding
    // holds m, the stand-in's place)
    const __amdgpu_buffer_rsrc_t r16 = list_rsrc(nbr16, tm_nbr16_at(g * TS, 0, cap), nbr16_elems);
    const uint32_t lane_off = (uint32_t)(tm_nbr16_at(t, 0, cap) * 2);
    const uint32_t chunks = (wmax + 7) >> 3;
    tm_u32x4 nx = {0u, 0u, 0u, 0u};
    if (chunks) nx = __builtin_amdgcn_raw_buffer_load_b128(r16, lane_off, 0, 0);
    for (uint32_t ch = 0; ch < chunks; ++ch) {
      const tm_u32x4 cur = nx;
      if (ch + 1 < chunks) nx = __builtin_amdgcn_raw_buffer_load_b128(r16, lane_off, (int)(1024 * (ch + 1)), 0);
      const uint32_t idx[8] = {cur.x & 0xFFFFu, cur.x >> 16, cur.y & 0xFFFFu, cur.y >> 16,
                               cur.z & 0xFFFFu, cur.z >> 16, cur.w & 0xFFFFu, cur.w >> 16};
      float sw = 0.0f;
#pragma unroll
      for (int j0 = 0; j0 < 8; j0 += U) {
        if (j0 > 0 && 8 * ch + (uint32_t)j0 >= wmax) break;  // past every list of the wave
        double2 A[U], B[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
          A[j] = lxy[idx[j0 + j]];
          B[j] = lzs[idx[j0 + j]];
        }
#pragma unroll
        for (int j = 0; j < U; ++j) {
          if constexpr (MIXED) {
            const float sq = __uint_as_float((uint32_t)__double_as_longlong(B[j].y));
            pair_mixed<true>(P, A[j].x, A[j].y, B[j].x, sq, K, sw, f0, f1, f2);
          } else {
            if (8 * ch + (uint32_t)(j0 + j) < cnt)
              pair_fp64(P, make_double4(A[j].x, A[j].y, B[j].x, B[j].y), C64, true, de, f0, f1, f2);
          }
        }
      }
      if constexpr (MIXED) de += (double)sw;
    }
  } else {
    const int64_t g0 = g * (TS / 64);
    const __amdgpu_buffer_rsrc_t rl = list_rsrc(nbr, g0 * cap * 64, nbr_elems);
    const uint32_t lane_off = (uint32_t)((((p >> 6) - g0) * cap * 64 + (p & 63)) * 4);
    direct_sums<MIXED, U, 1, true, true>(pos, n, cap, rl, lane_off, 0, P, C64, K, cnt, wmax, de, f0, f1, f2);
  }
  // Results go out in cell order: each lane parks its cell's sums in LDS
  // (the staged records are no longer needed) and lane t writes cell t, so
  // every store of a wave is one contiguous run.
  finish_cell<MIXED>(T, de, f0, f1, f2);
  __syncthreads();
  lxy[2 * off] = make_double2(de, f0);
  lxy[2 * off + 1] = make_double2(f1, f2);
  __syncthreads();
  const int64_t c = g * TS + t;
  if (c >= n) return;
  const double2 a = lxy[2 * t], b = lxy[2 * t + 1];
  dens[c] = a.x;
  fx[c] = a.y;
  fy[c] = b.x;
  fz[c] = b.y;
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

template <bool MIXED, int U, int L, bool DIM3, bool BUF>
void launch_k(tm_ctx *c) {
  const int64_t n = c->n;
  const int bs = c->sph_block;
  k_sph<MIXED, U, L, DIM3, BUF><<<tm_blocks(n * L, bs), bs, 0, c->stream>>>(
      c->d_pos, c->d_ctype_s, c->d_types, n, c->nbr_alloc, c->d_nbr, c->nbr_elems, c->d_ncount, c->sph_balance,
      c->d_dens, c->d_fx, c->d_fy, c->d_fz);
}

template <bool MIXED, int U, int TS>
void launch_tile_ts(tm_ctx *c) {
  const int64_t n = c->n;
  const int64_t nt = (n + TS - 1) / TS;
  k_sph_tile<MIXED, U, TS><<<(unsigned)nt, TS, 2 * sizeof(double2) * c->tile_lds, c->stream>>>(
      c->d_pos, c->d_ctype_s, c->d_types, n, c->nbr_alloc, c->d_nbr, c->nbr_elems, c->d_ncount, c->d_tile_idx,
      c->d_tile_cnt, c->d_tile_cell, c->d_nbr16, c->cap_nbr16, c->tile_lds, c->d_dens, c->d_fx, c->d_fy, c->d_fz);
}

// the tiles as built (their width, not the parameter's current value)
template <bool MIXED, int U>
void launch_tile(tm_ctx *c) {
  // the pair kernel parks its TS cells' sums in the staging LDS
  if (c->tile_lds < (uint32_t)c->tiles_ts) c->tile_lds = TM_TILE_MAX + 1;
  switch (c->tiles_ts) {
    case 128: launch_tile_ts<MIXED, U, 128>(c); break;
    case 512: launch_tile_ts<MIXED, U, 512>(c); break;
    default: launch_tile_ts<MIXED, U, 256>(c); break;
  }
}

template <bool MIXED, int L>
void launch_u(tm_ctx *c) {
  switch (MIXED ? c->sph_unroll_mixed : c->sph_unroll_fp64) {
    case 1: launch_k<MIXED, 1, L, true, true>(c); break;
    case 2: launch_k<MIXED, 2, L, true, true>(c); break;
    case 8: launch_k<MIXED, 8, L, true, true>(c); break;
    default: launch_k<MIXED, 4, L, true, true>(c); break;
  }
}

template <bool MIXED>
void launch_sph(tm_ctx *c) {
  // 2-D colonies and colonies too large for 32-bit gather offsets take one
  // general configuration; 3-D ones the tiles or the tuned direct one
  const bool buf = 32 * (c->n + 1) <= (int64_t)0xFFFFFFFF;
  if (c->dim != 3 || !buf) {
    if (c->dim == 3) launch_k<MIXED, 4, 1, true, false>(c);
    else if (buf) launch_k<MIXED, 4, 1, false, true>(c);
    else launch_k<MIXED, 4, 1, false, false>(c);
    return;
  }
  if (c->sph_tiles && c->tile_lds > 0 && c->tiles_ts > 0) {
    switch (MIXED ? c->sph_unroll_mixed : c->sph_unroll_fp64) {
      case 1:
      case 2: launch_tile<MIXED, 2>(c); break;
      case 8: launch_tile<MIXED, 8>(c); break;
      default: launch_tile<MIXED, 4>(c); break;
    }
    return;
  }
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
    c->sums_valid = true;
    return TM_OK;
  }
  if (prec == TM_MIXED) launch_sph<true>(c);
  else launch_sph<false>(c);
  TM_HIP(c, hipGetLastError());
  c->stage = 4;
  c->sums_valid = true;
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

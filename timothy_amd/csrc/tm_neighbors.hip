// Neighbour stage: one lane per cell walks its colony's radix tree with the
// reference's query box and strict overlap test, visiting leaf octants in
// descending Morton key -- the order the reference's LIFO DFS pops them
// (children pushed 0..7, src/potential.c:111-131).  The tree is read in the
// walk's own pre-order (built by k_wnodes): one 16-byte record per visited
// node, the next slot after a hit and the escape slot after a miss.
//
// Per cell p this yields:
//   N(p)    candidates q != p with sqrt((dx*dx+dy*dy)+dz*dz) <= h_p, stored in
//           visit order at nbr[tm_slot(p, k, cap)] (slot-major inside each
//           group of 64 cells, coalesced for the one-lane-per-cell pair
//           kernels), count[p]
//   mindist min distance over all candidates (src/potential.c:49), computed as
//           sqrt of the minimum squared distance (sqrt is monotone and
//           correctly rounded, so this is the minimum of the rounded roots)
//   qstar   the last candidate visited = minimum key, the one whose potential
//           the reference leaves in v (src/potential.c:124 assigns, not adds)
//   v       that potential, in the reference's FP64 expressions
// The cutoff test d <= h is evaluated exactly as d2 <= cut2, cut2 being the
// largest double whose correctly rounded root is <= h.
#include <type_traits>

#include "tm_internal.h"

namespace {

__device__ __forceinline__ uint32_t box_code(double v, double shift, double scale) {
  return (uint32_t)((v - shift) / scale * 536870912.0);
}

// potential() (src/potential.c:39-78) without its density/mindist side effects,
// in the reference's expressions and order
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

// WIN > 0: the workgroup first copies WIN consecutive slots of the
// pre-ordered tree, centred on the leaves of its own cells, into LDS; its
// walks read the nodes inside that window from LDS and the rest from memory.
// The cells of a workgroup are consecutive in Morton order, so their walks
// share many nodes near their own leaves (about half of all visits fall in
// 1024 slots around 256 cells), and a read from LDS spares the texture path
// the per-lane gather that bounds the walk.
// reference squared distance, src/potential.c:49 order (contraction is off)
__device__ __forceinline__ double dist2_fp64(const double4 &P, const double4 &Q) {
  const double dx = P.x - Q.x, dy = P.y - Q.y, dz = P.z - Q.z;
  return (dx * dx + dy * dy) + dz * dz;
}

// F32: candidates are screened with their FP32 positions relative to the
// colony's origin (16 bytes per candidate instead of 32); only a candidate
// whose FP32 distance cannot decide the cutoff, or that may hold the minimum
// distance, has its FP64 position read.  The results are those of the FP64
// walk (the bounds below are proved, not tuned):
//   stored values  |p~ - (P - A)| <= W (2^-24 + 2^-53) + 2^-150 per axis,
//                  W = the colony's largest extent >= |P - A|, so the FP32
//                  difference vector is within K = sqrt(3) * 1.2e-7 W
//                  (+1e-44) of the exact one after its own 2^-24 rounding,
//                  which is taken relative (1e-7) to the FP32 distance
//   FP32 sum       s~ = fma(dz, dz, fma(dy, dy, dx dx)): relative 3e-7,
//                  absolute 1e-44 from underflow
//   FP64 sum       the reference's d2 is within 1e-15 (relative) of the
//                  exact squared distance
// so with r~ = sqrt(s~), the reference's d2 lies in
//   [((r~ (1 - 1e-7) - K)+)^2 (1 - 1e-15), (r~ (1 + 1e-7) + K)^2 (1 + 1e-15)]
// (r~ widened by 3e-7 and 1e-44 as above).  tin / tout are the FP32 values
// of s~ below which d2 <= cut2 holds and above which it fails, each rounded
// inward by 1e-6; between them the candidate is read in FP64.  For the
// minimum, a candidate c cannot hold it when s~_c > 1.0002 s~_b + beta for
// a candidate b already seen: the bound above on both sides, with the cross
// term 4.02 K r~ <= 1e-4 r~^2 + 40401 K^2 (AM-GM), gives that threshold.
// The smallest FP32 distance is kept with its cell; the candidate that loses
// that place (the newcomer, or the one it displaces) is read in FP64 at once
// when it is still within reach of the smallest (rare), and the smallest is
// read in FP64 at the end.
// COOP: each wave (BS = 64) first walks the tree once for the union of its
// cells' query boxes, reading the pre-ordered slots 64 at a time (one
// coalesced load for all lanes, where the per-lane walk gathers one record
// per lane per node).  A slot is visited when no missed internal node before
// it in the window skips it (a prefix maximum of the missed nodes' escape
// slots); the walk continues after the window or at the largest escape.
// The leaves it finds that overlap the union box are, in visit order, the
// wave's union candidates: every leaf whose octant overlaps a cell's own box
// overlaps the union box, and its ancestors contain its octant, so a cell's
// candidates are exactly the union candidates that overlap its own box, in
// the same order.  They are staged in LDS (CB at a time with their FP32 or
// FP64 records) and each lane takes its own from there -- its box test over
// 64 staged candidates at a time as a bit mask, then the set bits in order.
// A wave whose cells span two colonies, or whose union box is more than
// coop_ratio times its cells' mean box volume (a Morton jump: the union would
// hold far more candidates than any cell), walks per lane as without COOP.
constexpr int CB = 128;  // union candidates staged per wave

template <int BS, int WIN, bool F32, bool COOP>
__global__ __launch_bounds__(BS) void k_neighbors(const double4 *__restrict__ pos, const float4 *__restrict__ posf,
                                                  const uint4 *__restrict__ code,
                                                  const int32_t *__restrict__ ctype_s,
                                                  const float *__restrict__ young_s,
                                                  const uint4 *__restrict__ wn,
                                                  const int64_t *__restrict__ col_off,
                                                  const double4 *__restrict__ aff,
                                                  const TypeDev *__restrict__ types, int64_t n, int64_t cap,
                                                  uint32_t *__restrict__ nbr,
                                                  uint32_t *__restrict__ count, double *__restrict__ mindist,
                                                  int32_t *__restrict__ qstar, double *__restrict__ vout,
                                                  unsigned int *__restrict__ flags, float coop_ratio) {
  static_assert(!COOP || (BS == 64 && WIN == 0), "a cooperative walk takes one wave per workgroup, no window");
  const int64_t cell = tm_xcd_block() * BS + threadIdx.x;
  const bool live = cell < n;
  const int64_t p = live ? cell : n - 1;  // idle lanes only join the padding step below
  const double4 P = pos[p];
  const uint32_t cid = code[p].w;
  const double4 A = aff[cid];
  const TypeDev &T = types[ctype_s[p]];
  const double h = T.h;
  const double cut2 = T.cut2;
  const uint4 *winp = nullptr;
  // query box, src/octree.c:246-266: ((x - h) - shift) / scale * 2^29
  const uint32_t mn0 = box_code(P.x - h, A.x, A.w), mx0 = box_code(P.x + h, A.x, A.w);
  const uint32_t mn1 = box_code(P.y - h, A.y, A.w), mx1 = box_code(P.y + h, A.y, A.w);
  const uint32_t mn2 = box_code(P.z - h, A.z, A.w), mx2 = box_code(P.z + h, A.z, A.w);
  // The strict overlap of [mn, mx] with a node's [lo, lo + s) on one axis,
  // mx > lo and mn < lo + s, is (mx - 1 - lo) < (mx - mn - 1) + s in 32-bit
  // unsigned arithmetic: the left side wraps past every right side when
  // mx <= lo, and otherwise the two compares agree (mx >= mn: both codes
  // are the same monotone map of x - h and x + h).
  const uint32_t ex0 = mx0 - 1u, ex1 = mx1 - 1u, ex2 = mx2 - 1u;
  const uint32_t dm0 = mx0 - mn0 - 1u, dm1 = mx1 - mn1 - 1u, dm2 = mx2 - mn2 - 1u;
  // the window of slots staged in LDS: around the leaf slot of the
  // workgroup's middle cell, 2 s + 2 (e - k) for cell k of colony [s, e]
  // (up to its ancestor count), within the tree's 2 n - 1 slots
  uint32_t wlo = 0;
  if constexpr (WIN > 0) {
    __shared__ uint4 win[WIN];
    const int64_t km = min((int64_t)(tm_xcd_block() * BS + BS / 2), n - 1);
    const uint32_t cm = code[km].w;
    const int64_t cs = col_off[cm], ce = col_off[cm + 1] - 1;
    const int64_t nslot = 2 * n - 1;
    int64_t lo = 2 * cs + 2 * (ce - km) - WIN / 2;
    lo = max((int64_t)0, min(lo, nslot - WIN));
    wlo = (uint32_t)max((int64_t)0, lo);
    for (int i = threadIdx.x; i < WIN; i += BS) {
      const int64_t sl = (int64_t)wlo + i;
      win[i] = sl < nslot ? wn[sl] : make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();
    winp = win;
  }
  // the colony's slots of the pre-ordered tree
  uint32_t node = live ? (uint32_t)(2 * col_off[cid]) : 0u;
  const uint32_t end = live ? (uint32_t)(2 * col_off[cid + 1] - 1) : 0u;
  uint32_t cnt = 0;
  int32_t qs = -1;
  double md2 = 1.0 / 0.0;
  // candidate q (its leaf octant overlaps the box, q != p) with its position
  // Q: distance, mindist, the last-visited candidate, and the list entry when
  // within the cutoff (returned as its slot, stored by the caller; -1 if none)
  using QRec = typename std::conditional<F32, float4, double4>::type;
  // FP32 screening state: the cell's FP32 relative position, the cutoff
  // thresholds, and the smallest FP32 distance so far with its cell
  float pfx = 0.f, pfy = 0.f, pfz = 0.f, tin = -1.f, tout = 0.f, beta = 0.f;
  float s1 = 1.0f / 0.0f;
  uint32_t c1 = 0xFFFFFFFFu;
  if constexpr (F32) {
    pfx = (float)(P.x - A.x);
    pfy = (float)(P.y - A.y);
    pfz = (float)(P.z - A.z);
    const double K = 1.7321 * (A.w * 1.2e-7 + 1e-44);
    const double hc = sqrt(cut2);
    double ti = (hc * (1.0 - 1e-15) - K) / (1.0 + 1e-7);
    ti = ti > 0.0 ? ti * ti / (1.0 + 3e-7) - 1e-44 : -1.0;
    // a threshold outside FP32's normal range decides nothing
    tin = (ti >= 1e-30 && ti <= 1e38) ? (float)(ti * (1.0 - 1e-6)) : -1.0f;
    double to = (hc * (1.0 + 1e-15) + K) / (1.0 - 1e-7);
    to = to * to * (1.0 + 3e-7) + 1e-44;
    // (overflow gives inf and NaN stays NaN: both decide nothing; below
    // FP32's normal range the threshold is raised, which only widens the
    // band read in FP64)
    tout = to < 1e-30 ? 1e-30f : (float)(to * (1.0 + 1e-6));
    beta = (float)((40500.0 * K * K + 1e-37) * (1.0 + 1e-6));
  }
  // the reference's d2 of candidate q, read in FP64, into the minimum
  auto exact = [&](uint32_t q) -> double {
    const double d2 = dist2_fp64(P, pos[q]);
    md2 = d2 < md2 ? d2 : md2;
    return d2;
  };
  auto visit = [&](uint32_t q, const QRec &Q) -> int64_t {
    bool in;
    if constexpr (F32) {
      const float dx = pfx - Q.x, dy = pfy - Q.y, dz = pfz - Q.z;
      const float sq = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
      if (sq <= tin) in = true;
      else if (sq > tout) in = false;
      else in = exact(q) <= cut2;
      // the smallest (selects); the one that loses it
      const bool lt = sq < s1;
      const float so = lt ? s1 : sq;
      const uint32_t co = lt ? c1 : q;
      s1 = lt ? sq : s1;
      c1 = lt ? q : c1;
      if (so <= fmaf(1.0002f, s1, beta)) {
        if (co != 0xFFFFFFFFu) (void)exact(co);
      }
    } else {
      const double d2 = dist2_fp64(P, Q);
      md2 = d2 < md2 ? d2 : md2;
      in = d2 <= cut2;
    }
    qs = (int32_t)q;
    int64_t slot = -1;
    if (in) {
      if (cnt < cap) slot = tm_slot(p, cnt, cap);
      ++cnt;
    }
    return slot;
  };
  // One record per node: after a hit the next slot (the first child, or for
  // a leaf the next node), after a miss the escape slot; slots only grow, so
  // the walk ends at the colony's last slot.  A candidate's record load is
  // issued when its leaf is found and consumed one node later, under the
  // next node's load; its list entry is stored after that node's loads are
  // issued, so waiting for them does not also wait for the store (vector
  // memory counters retire in order).  Candidates are still taken in visit
  // order.
  bool coop = false;
  if constexpr (COOP) {
    // the wave's union box, over its live cells of one colony
    const uint32_t lane = threadIdx.x;
    uint32_t u[6] = {live ? mn0 : ~0u, live ? mn1 : ~0u, live ? mn2 : ~0u, live ? mx0 : 0u, live ? mx1 : 0u,
                     live ? mx2 : 0u};
    uint32_t cmin = live ? cid : ~0u, cmax = live ? cid : 0u;
    float vl = live ? (float)(mx0 - mn0) * (float)(mx1 - mn1) * (float)(mx2 - mn2) : 0.f;
    float nl = live ? 1.f : 0.f;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
      for (int a = 0; a < 3; ++a) u[a] = min(u[a], (uint32_t)__shfl_xor((int)u[a], o, 64));
#pragma unroll
      for (int a = 3; a < 6; ++a) u[a] = max(u[a], (uint32_t)__shfl_xor((int)u[a], o, 64));
      cmin = min(cmin, (uint32_t)__shfl_xor((int)cmin, o, 64));
      cmax = max(cmax, (uint32_t)__shfl_xor((int)cmax, o, 64));
      vl += __shfl_xor(vl, o, 64);
      nl += __shfl_xor(nl, o, 64);
    }
    const float vu = (float)(u[3] - u[0]) * (float)(u[4] - u[1]) * (float)(u[5] - u[2]);
    coop = nl > 0.f && cmin == cmax && vu <= coop_ratio * (vl / nl);
    if (coop) {
      __shared__ uint4 crec[CB];  // lo codes x, y, z and the cell
      __shared__ uint4 csz[CB];   // octant sizes x, y, z
      __shared__ QRec cq[CB];     // the cell's record
      const uint32_t uex0 = u[3] - 1u, uex1 = u[4] - 1u, uex2 = u[5] - 1u;
      const uint32_t udm0 = u[3] - u[0] - 1u, udm1 = u[4] - u[1] - 1u, udm2 = u[5] - u[2] - 1u;
      const uint32_t wbeg = (uint32_t)(2 * col_off[cmin]), wend = (uint32_t)(2 * col_off[cmin + 1] - 1);
      uint32_t nb = 0;
      // each lane's candidates among the staged ones, in order
      auto drain = [&]() {
        for (uint32_t e = lane; e < nb; e += 64) {
          if constexpr (F32) cq[e] = posf[crec[e].w];
          else cq[e] = pos[crec[e].w];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t b0 = 0; b0 < nb; b0 += 64) {
          const uint32_t m = min(64u, nb - b0);
          uint64_t mask = 0;
          for (uint32_t j = 0; j < m; ++j) {
            const uint4 r = crec[b0 + j], z = csz[b0 + j];
            const bool hit = (ex0 - r.x < dm0 + z.x) & (ex1 - r.y < dm1 + z.y) & (ex2 - r.z < dm2 + z.z);
            mask |= (uint64_t)(hit && live && r.w != (uint32_t)p) << j;
          }
          while (mask) {
            const uint32_t j = (uint32_t)__builtin_ctzll(mask);
            mask &= mask - 1;
            const uint32_t q = crec[b0 + j].w;
            const int64_t slot = visit(q, cq[b0 + j]);
            if (slot >= 0) nbr[slot] = q;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        nb = 0;
      };
      // the window after this one is loaded with it: taken when the walk
      // goes on there (no escape past it)
      uint4 nxt = wbeg + lane < wend ? wn[wbeg + lane] : make_uint4(0u, 0u, 0u, 0u);
      for (uint32_t cur = wbeg; cur < wend;) {
        const uint32_t sl = cur + lane;
        const bool valid = sl < wend;
        const uint4 a = nxt;
        nxt = sl + 64u < wend ? wn[sl + 64u] : make_uint4(0u, 0u, 0u, 0u);
        const uint32_t L = (a.x >> 29) | (((a.y >> 29) & 3u) << 3);
        const uint32_t lx = a.x & 0x1FFFFFFFu, ly = a.y & 0x1FFFFFFFu, lz = a.z & 0x1FFFFFFFu;
        const uint32_t sx = 1u << L, sy = sx >> ((a.z >> 30) & 1u), sz = sx >> ((a.z >> 29) & 1u);
        const bool hitu = (uex0 - lx < udm0 + sx) & (uex1 - ly < udm1 + sy) & (uex2 - lz < udm2 + sz);
        const bool leaf = (a.y & 0x80000000u) != 0u;
        const bool miss = valid && !hitu && !leaf;
        // slots skipped by a missed node earlier in the window
        uint32_t pm = miss ? a.w : 0u;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = (uint32_t)__shfl_up((int)pm, o, 64);
          if (lane >= (uint32_t)o) pm = max(pm, y);
        }
        uint32_t before = (uint32_t)__shfl_up((int)pm, 1, 64);
        if (lane == 0) before = 0u;
        const uint32_t mall = (uint32_t)__shfl((int)pm, 63, 64);
        const bool visited = valid && before <= sl;
        // a malformed tree: a visited miss that does not move forward, or a
        // leaf naming no cell
        if (__ballot(visited && ((miss && a.w <= sl) || (leaf && a.w >= (uint32_t)n)))) {
          if (lane == 0) atomicOr(flags, (unsigned)TM_FLAG_WALK);
          break;
        }
        const bool cand = visited && leaf && hitu;
        const uint64_t cm = __ballot(cand);
        const uint32_t k = (uint32_t)__builtin_popcountll(cm);
        if (k) {
          if (nb + k > (uint32_t)CB) drain();
          if (cand) {
            const uint32_t e = nb + (uint32_t)__builtin_popcountll(cm & tm_lanemask_lt());
            crec[e] = make_uint4(lx, ly, lz, a.w);
            csz[e] = make_uint4(sx, sy, sz, 0u);
          }
          nb += k;
        }
        if (mall > cur + 64u) {
          cur = mall;
          nxt = cur + lane < wend ? wn[cur + lane] : make_uint4(0u, 0u, 0u, 0u);
        } else {
          cur += 64u;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (nb) drain();
      node = end;  // the per-lane walk below has nothing left to do
    }
  }
  uint32_t pq = 0xFFFFFFFFu;  // the candidate whose record is in flight
  QRec Qp{};
  while (node < end) {
    uint4 a;
    if constexpr (WIN > 0) {
      const uint32_t wi = node - wlo;
      a = wi < (uint32_t)WIN ? winp[wi] : wn[node];
    } else {
      a = wn[node];
    }
    int64_t slot = -1;
    const uint32_t q = pq;
    if (pq != 0xFFFFFFFFu) {
      slot = visit(pq, Qp);
      pq = 0xFFFFFFFFu;
    }
    const uint32_t L = (a.x >> 29) | (((a.y >> 29) & 3u) << 3);
    const uint32_t lx = a.x & 0x1FFFFFFFu, ly = a.y & 0x1FFFFFFFu, lz = a.z & 0x1FFFFFFFu;
    const uint32_t sx = 1u << L, sy = sx >> ((a.z >> 30) & 1u), sz = sx >> ((a.z >> 29) & 1u);
    // mx > lo and mn < lo + s, as one unsigned compare per axis
    const bool hit = (ex0 - lx < dm0 + sx) & (ex1 - ly < dm1 + sy) & (ex2 - lz < dm2 + sz);
    const bool leaf = (a.y & 0x80000000u) != 0u;
    const uint32_t next = (hit || leaf) ? node + 1 : a.w;
    if (next <= node || (leaf && a.w >= (uint32_t)n)) {  // a malformed tree: a well-formed one only moves forward
      if (slot >= 0) nbr[slot] = q;
      atomicOr(flags, (unsigned)TM_FLAG_WALK);
      break;
    }
    if (leaf && hit && a.w != (uint32_t)p) {
      if constexpr (F32) Qp = posf[a.w];
      else Qp = pos[a.w];
      pq = a.w;
    }
    if (slot >= 0) nbr[slot] = q;
    node = next;
  }
  if (pq != 0xFFFFFFFFu) {
    const int64_t slot = visit(pq, Qp);
    if (slot >= 0) nbr[slot] = pq;
  }
  // Pad each list to the longest one of its 64-cell slot group (this wave)
  // with cell index n, the far-away stand-in record the build places after
  // the cells: the pair kernels then read whole slot rows without per-lane
  // tests, and a stand-in pair contributes exactly zero to the mixed sums.
  // (rows are written wave-wide, one 256-byte row per step)
  const uint32_t used = min(cnt, (uint32_t)cap);
  uint32_t g = used, lo = live ? used : 0xFFFFFFFFu;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    g = max(g, (uint32_t)__shfl_xor((int)g, o, 64));
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, o, 64));
  }
  for (uint32_t k = lo; k < g; ++k)
    if (live && k >= used) nbr[tm_slot(p, k, cap)] = (uint32_t)n;
  if (!live) return;
  if constexpr (F32) {
    // the smallest
    if (c1 != 0xFFFFFFFFu) (void)exact(c1);
  }
  count[p] = cnt;
  mindist[p] = qs >= 0 ? sqrt(md2) : 1.7976931348623157e308;
  qstar[p] = qs;
  // v: the potential with the last candidate visited (src/potential.c:124
  // assigns it per candidate; 0 beyond the cutoff and with no candidate)
  // the record of the last candidate, read again here rather than kept
  // through the walk
  const double4 Qs = qs >= 0 ? pos[qs] : P;
  vout[p] = qs >= 0 ? pot_fp64(P, young_s[p], Qs, young_s[qs], h) : 0.0;
  if (cnt > cap) atomicMax(flags + 1, cnt);
}

template <int BS, int WIN, bool COOP = false>
void launch_walk(tm_ctx *c) {
  auto kern = c->walk_f32 ? k_neighbors<BS, WIN, true, COOP> : k_neighbors<BS, WIN, false, COOP>;
  kern<<<tm_blocks(c->n, BS), BS, 0, c->stream>>>(
      c->d_pos, c->d_posf, c->d_code, c->d_ctype_s, c->d_young_s, c->d_wn, c->d_col_off, c->d_aff, c->d_types, c->n,
      c->nbr_alloc, c->d_nbr, c->d_ncount, c->d_mind, c->d_qstar, c->d_v, c->d_flags, (float)c->walk_coop_ratio);
}

}  // namespace

// grows the neighbour lists and repeats the walk when a cell overflowed them
tm_status tm_run_neighbors(tm_ctx *c, bool check) {
  if (c->stage < 2) return tm_fail(c, TM_ERR_STATE, "tm_neighbors before tm_build");
  const int64_t n = c->n;
  if (n == 0) {
    c->stage = 3;
    return TM_OK;
  }
  for (int attempt = 0; attempt < 2; ++attempt) {
    const int64_t need = tm_slot_elems(n, c->nbr_cap);
    if (c->nbr_elems < need || !c->d_nbr) {
      if (c->d_nbr) (void)hipFree(c->d_nbr);
      c->d_nbr = nullptr;
      c->nbr_elems = 0;
      if (hipMalloc((void **)&c->d_nbr, sizeof(uint32_t) * (size_t)need) != hipSuccess) {
        c->d_nbr = nullptr;
        return tm_fail(c, TM_ERR_NOMEM, "hipMalloc failed: neighbour lists");
      }
      c->nbr_elems = need;
    }
    c->nbr_alloc = c->nbr_cap;
    const int bs = c->nbr_block;
    if (c->walk_coop) launch_walk<64, 0, true>(c);
    else if (bs == 64) launch_walk<64, 0>(c);
    else if (bs == 128) launch_walk<128, 0>(c);
    else if (c->walk_window == 2048 && 2 * n - 1 >= 2048) launch_walk<256, 2048>(c);
    else if (c->walk_window == 1280 && 2 * n - 1 >= 1280) launch_walk<256, 1280>(c);
    else if (c->walk_window == 1024 && 2 * n - 1 >= 1024) launch_walk<256, 1024>(c);
    else if (c->walk_window == 512 && 2 * n - 1 >= 512) launch_walk<256, 512>(c);
    else launch_walk<256, 0>(c);
    TM_HIP(c, hipGetLastError());
    if (!check) break;
    TM_HIP(c, hipMemcpyAsync(c->h_flags, c->d_flags, 4 * sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
    TM_HIP(c, hipStreamSynchronize(c->stream));
    if (c->h_flags[0] & TM_FLAG_DEGENERATE)
      return tm_fail(c, TM_ERR_DEGENERATE,
                     "two cells share location codes down to bit 1, or more than 64 lie within 2^-16 of their colony's "
                     "extent on every axis (duplicate or near-duplicate positions)");
    if (c->h_flags[0] & TM_FLAG_ROOT) return tm_fail(c, TM_ERR_STATE, "colony root not found in the radix tree");
    if (c->h_flags[0] & TM_FLAG_WALK) return tm_fail(c, TM_ERR_STATE, "malformed tree: walk left its node range");
    if (c->h_flags[0] & TM_FLAG_CTYPE) return tm_fail(c, TM_ERR_INVALID, "ctype out of range");
    const int64_t mx = c->h_flags[1];
    if (mx <= c->nbr_alloc) break;
    c->nbr_cap = (mx + 7) & ~7ll;  // grow to the largest count seen, rounded up
  }
  TM_TRY(tm_run_tiles(c, check));
  c->stage = 3;
  return TM_OK;
}

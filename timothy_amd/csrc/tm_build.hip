// Build stage: per-colony affine map, 29-bit location codes, Morton order,
// PR-octree leaf octants and the radix tree the neighbour walk descends.
//
// Reference: octbuild (src/octree.c:145-207).  The reference inserts cells one
// by one into a pointer PR-octree (one cell per leaf).  The octree is unique
// for a set of codes, so it is rebuilt here from the Morton order instead:
// a cell's leaf octant has side 2^s with s the highest bit in which its codes
// differ from its closest Morton neighbour, and the binary radix tree over the
// sorted keys (Karras 2012) has every subtree's leaf octants inside the
// subtree's key-prefix region, which is what the walk tests.
#include "tm_internal.h"

namespace {

constexpr int BT = 256;

__device__ __forceinline__ unsigned long long ord_key(double v) {
  unsigned long long u = (unsigned long long)__double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double unord_key(unsigned long long u) {
  return __longlong_as_double((long long)((u >> 63) ? (u & 0x7fffffffffffffffull) : ~u));
}

__device__ __forceinline__ int colony_of(const int64_t *__restrict__ off, int ncol, int64_t i) {
  int lo = 0, hi = ncol;  // off[lo] <= i < off[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// (uint32)(((x - shift)/scale) * MAXVALUE), src/octree.c:188-190
__device__ __forceinline__ uint32_t loc_code(double x, double shift, double scale) {
  return (uint32_t)(((x - shift) / scale) * 536870912.0);
}

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// top 64 bits of the 87-bit Morton key (level 28 first, z,y,x within a level)
__device__ __forceinline__ uint64_t morton_prefix64(uint32_t cx, uint32_t cy, uint32_t cz) {
  uint64_t k63 = spread21(cx >> 8) | (spread21(cy >> 8) << 1) | (spread21(cz >> 8) << 2);
  return (k63 << 1) | ((cz >> 7) & 1u);
}

// full-key order of two cells (same colony): true if a < b
__device__ __forceinline__ bool morton_less(uint4 a, uint4 b) {
  if (a.w != b.w) return a.w < b.w;
  int bit = tm_diffbit(a, b);
  if (bit < 0) return false;
  uint32_t m = 1u << bit;
  if ((a.z ^ b.z) & m) return (b.z & m) != 0;
  if ((a.y ^ b.y) & m) return (b.y & m) != 0;
  return (b.x & m) != 0;
}

// common prefix length of (colony id : 16 bits, Morton key : 87 bits)
__device__ __forceinline__ int key_delta(const uint4 *__restrict__ code, int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  uint4 a = code[i], b = code[j];
  if (a.w != b.w) return __clz((int)(a.w ^ b.w)) - 16;
  int bit = tm_diffbit(a, b);
  if (bit < 0) return TM_CID_BITS + 87 + __clz((int)((uint32_t)i ^ (uint32_t)j));  // duplicate codes (flagged)
  uint32_t m = 1u << bit;
  int r = ((a.z ^ b.z) & m) ? 0 : (((a.y ^ b.y) & m) ? 1 : 2);
  return TM_CID_BITS + 3 * (28 - bit) + r;
}

// The same prefix length from the sorted 64-bit sort keys where they differ
// (cb top bits the colony id, then the Morton key's leading 64 - cb bits,
// z before y before x in each triple as in the code comparison above), the
// codes only where the keys tie: half the bytes per probe of the searches.
// ki and kj are the keys of cells i and j, both in [0, n).
__device__ __forceinline__ int key_delta64(uint64_t ki, uint64_t kj, const uint4 *__restrict__ code, int64_t n,
                                           int cb, int64_t i, int64_t j) {
  const uint64_t x = ki ^ kj;
  if (x) {
    const int p = __clzll((long long)x);
    return p < cb ? TM_CID_BITS - cb + p : TM_CID_BITS + p - cb;
  }
  return key_delta(code, n, i, j);
}

// bounding box of x -+ (h_type + 0.01) per colony (src/octree.c:161-174): one
// workgroup per chunk of one colony, reduced in registers and LDS into the
// chunk's six order-preserving integer images of the doubles (k_affine
// reduces a colony's chunks; no atomics on six shared words)
__global__ __launch_bounds__(BT) void k_bbox(const double *__restrict__ x, const double *__restrict__ y,
                                             const double *__restrict__ z, const int32_t *__restrict__ ctype,
                                             const int64_t *__restrict__ chunk_start,
                                             const int32_t *__restrict__ chunk_cid,
                                             const TypeDev *__restrict__ types, int ntypes,
                                             unsigned long long *__restrict__ cbox, unsigned int *__restrict__ flags) {
  __shared__ unsigned long long red[BT / 64][6];
  const int64_t s = chunk_start[blockIdx.x], e = chunk_start[blockIdx.x + 1];
  unsigned long long k[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
  for (int64_t i = s + threadIdx.x; i < e; i += BT) {
    int t = ctype[i];
    if (t < 0 || t >= ntypes) {
      atomicOr(flags, (unsigned)TM_FLAG_CTYPE);
      t = 0;
    }
    const double ext = types[t].h + 0.01;  // src/octree.c:167
    const double px = x[i], py = y[i], pz = z[i];
    const unsigned long long v[6] = {ord_key(px - ext), ord_key(py - ext), ord_key(pz - ext),
                                     ord_key(px + ext), ord_key(py + ext), ord_key(pz + ext)};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      k[a] = v[a] < k[a] ? v[a] : k[a];
      k[a + 3] = v[a + 3] > k[a + 3] ? v[a + 3] : k[a + 3];
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      unsigned long long t = __shfl_xor(k[a], o, 64);
      k[a] = t < k[a] ? t : k[a];
      unsigned long long u = __shfl_xor(k[a + 3], o, 64);
      k[a + 3] = u > k[a + 3] ? u : k[a + 3];
    }
  }
  const int w = threadIdx.x >> 6;
  if (__lane_id() == 0)
    for (int a = 0; a < 6; ++a) red[w][a] = k[a];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < BT / 64; ++q)
      for (int a = 0; a < 3; ++a) {
        k[a] = red[q][a] < k[a] ? red[q][a] : k[a];
        k[a + 3] = red[q][a + 3] > k[a + 3] ? red[q][a + 3] : k[a + 3];
      }
    unsigned long long *b = cbox + 6 * (int64_t)blockIdx.x;
    for (int a = 0; a < 6; ++a) b[a] = k[a];
  }
}

// src/octree.c:175-180: one wave per colony reduces its chunks' boxes
__global__ __launch_bounds__(64) void k_affine(const unsigned long long *__restrict__ cbox,
                                               const int64_t *__restrict__ col_chunk, int ncol,
                                               double4 *__restrict__ aff) {
  const int c = blockIdx.x;
  if (c >= ncol) return;
  unsigned long long k[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
  for (int64_t q = col_chunk[c] + threadIdx.x; q < col_chunk[c + 1]; q += 64) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const unsigned long long lo = cbox[6 * q + a], hi = cbox[6 * q + 3 + a];
      k[a] = lo < k[a] ? lo : k[a];
      k[a + 3] = hi > k[a + 3] ? hi : k[a + 3];
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const unsigned long long t = __shfl_xor(k[a], o, 64), u = __shfl_xor(k[a + 3], o, 64);
      k[a] = t < k[a] ? t : k[a];
      k[a + 3] = u > k[a + 3] ? u : k[a + 3];
    }
  }
  if (threadIdx.x != 0) return;
  double mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = unord_key(k[a]);
    mx[a] = unord_key(k[3 + a]);
  }
  double s = mx[0] - mn[0];
  s = (s > mx[1] - mn[1] ? s : mx[1] - mn[1]);
  s = (s > mx[2] - mn[2] ? s : mx[2] - mn[2]);
  aff[c] = make_double4(mn[0], mn[1], mn[2], s);
}

// Morton keys, and the cell inputs packed in reference order (one 32-byte and
// one 16-byte record per cell) so the Morton gather below reads two records
// per cell instead of seven scattered arrays.
__global__ __launch_bounds__(BT) void k_keys(const double *__restrict__ x, const double *__restrict__ y,
                                             const double *__restrict__ z, const double *__restrict__ size,
                                             const float *__restrict__ young, const int32_t *__restrict__ ctype,
                                             const int32_t *__restrict__ phase, int64_t n,
                                             const int64_t *__restrict__ off, int ncol, int cb, int ntypes,
                                             const double4 *__restrict__ aff, uint64_t *__restrict__ key,
                                             int32_t *__restrict__ val, double4 *__restrict__ pos_o,
                                             uint4 *__restrict__ aux_o) {
  int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
  if (i >= n) return;
  const int cid = ncol > 1 ? colony_of(off, ncol, i) : 0;
  const double4 a = aff[cid];
  const double px = x[i], py = y[i], pz = z[i];
  const uint32_t cx = loc_code(px, a.x, a.w), cy = loc_code(py, a.y, a.w), cz = loc_code(pz, a.z, a.w);
  uint64_t k = morton_prefix64(cx, cy, cz);
  if (cb > 0) k = ((uint64_t)cid << (64 - cb)) | (k >> cb);
  key[i] = k;
  val[i] = (int32_t)i;
  pos_o[i] = make_double4(px, py, pz, size[i]);
  const int32_t t = ctype[i];
  aux_o[i] = make_uint4(__float_as_uint(young ? young[i] : 0.0f),
                        (uint32_t)((t < 0 || t >= ntypes) ? 0 : t),  // out-of-range types were flagged by k_bbox
                        (uint32_t)(phase ? phase[i] : 1), (uint32_t)cid);
}

__global__ __launch_bounds__(BT) void k_gather(const int32_t *__restrict__ perm, int64_t n,
                                               const double4 *__restrict__ pos_o, const uint4 *__restrict__ aux_o,
                                               const double4 *__restrict__ aff, uint4 *__restrict__ code,
                                               double4 *__restrict__ pos, float4 *__restrict__ posf,
                                               float *__restrict__ young_s, int32_t *__restrict__ ctype_s,
                                               int32_t *__restrict__ phase_s) {
  int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
  if (i == 0) pos[n] = make_double4(TM_FAR, TM_FAR, TM_FAR, 0.0);  // stand-in neighbour padding the lists
  if (i >= n) return;
  int64_t o = perm[i];
  if (o < 0 || o >= n) o = 0;  // unreachable for a valid permutation; bounds the gathers
  const double4 P = pos_o[o];
  const uint4 X = aux_o[o];
  const double4 a = aff[X.w];
  code[i] = make_uint4(loc_code(P.x, a.x, a.w), loc_code(P.y, a.y, a.w), loc_code(P.z, a.z, a.w), X.w);
  pos[i] = P;
  // the walk's screening copy: x, y, z relative to the colony's lower
  // corner, formed in FP64 and rounded to FP32
  posf[i] = make_float4((float)(P.x - a.x), (float)(P.y - a.y), (float)(P.z - a.z), 0.0f);
  young_s[i] = __uint_as_float(X.x);
  ctype_s[i] = (int32_t)X.y;
  phase_s[i] = (int32_t)X.z;
}

// The sort orders the keys by their bits from SORT_LO_BIT up.  Cells whose
// keys tie there are put in full 87-bit key order here (and their keys with
// them); the first thread of each run sorts it (runs are a cell or two at
// most: a tie needs two cells within 2^-16 of the colony's extent on every
// axis, a colony id in the key's top bits leaving fewer Morton bits).
constexpr int SORT_LO_BIT = 16;

__global__ __launch_bounds__(BT) void k_fixties(uint64_t *key, int64_t n, int32_t *perm, uint4 *code,
                                                double4 *pos, float4 *posf, float *young_s, int32_t *ctype_s,
                                                int32_t *phase_s, unsigned int *flags) {
  int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
  if (i >= n - 1) return;
  const uint64_t ki = key[i] >> SORT_LO_BIT;
  if (key[i + 1] >> SORT_LO_BIT != ki) return;
  if (i > 0 && key[i - 1] >> SORT_LO_BIT == ki) return;
  int64_t e = i + 1;
  while (e < n && key[e] >> SORT_LO_BIT == ki && e - i <= 64) ++e;
  if (e - i > 64) {  // > 64 cells that close together: treated as degenerate
    atomicOr(flags, (unsigned)TM_FLAG_DEGENERATE);
    return;
  }
  for (int64_t a = i + 1; a < e; ++a) {
    const uint64_t ka = key[a];
    uint4 c = code[a];
    int32_t pm = perm[a];
    double4 ps = pos[a];
    float4 pf = posf[a];
    float yg = young_s[a];
    int32_t ct = ctype_s[a], ph = phase_s[a];
    int64_t b = a - 1;
    while (b >= i && morton_less(c, code[b])) {
      key[b + 1] = key[b];
      code[b + 1] = code[b];
      perm[b + 1] = perm[b];
      pos[b + 1] = pos[b];
      posf[b + 1] = posf[b];
      young_s[b + 1] = young_s[b];
      ctype_s[b + 1] = ctype_s[b];
      phase_s[b + 1] = phase_s[b];
      --b;
    }
    key[b + 1] = ka;
    code[b + 1] = c;
    perm[b + 1] = pm;
    pos[b + 1] = ps;
    posf[b + 1] = pf;
    young_s[b + 1] = yg;
    ctype_s[b + 1] = ct;
    phase_s[b + 1] = ph;
  }
}

// leaf octant exponent: min over the same-colony Morton neighbours of the
// highest differing code bit (the level at which the PR-octree separates the
// cell); 29 (the root) for a colony of one.  Flags duplicate codes and splits
// below node level 3, where the reference shifts by negative counts
// (src/octree.c:105-107).
__global__ __launch_bounds__(BT) void k_leaf(const uint4 *__restrict__ code, int64_t n, int32_t *__restrict__ span,
                                             uint32_t *__restrict__ chain, unsigned int *__restrict__ flags) {
  int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
  if (i >= n) return;
  chain[i] = 0;  // counted by k_karras
  uint4 me = code[i];
  int best = 99;
  bool bad = false;
  if (i > 0) {
    uint4 o = code[i - 1];
    if (o.w == me.w) {
      int b = tm_diffbit(me, o);
      if (b < 0) bad = true;
      best = min(best, b);
    }
  }
  if (i + 1 < n) {
    uint4 o = code[i + 1];
    if (o.w == me.w) {
      int b = tm_diffbit(me, o);
      if (b < 0) bad = true;
      best = min(best, b);
    }
  }
  if (best == 99) best = 29;
  if (best <= 1) bad = true;
  if (bad) atomicOr(flags, (unsigned)TM_FLAG_DEGENERATE);
  span[i] = best < 0 ? 0 : best;
}

// Karras (2012) radix tree over (colony, Morton) keys: internal node i's key
// range, common-prefix length and split.  chain[v] counts the nodes inside
// one colony (common prefix at least the colony id) whose range ends at v.
__global__ __launch_bounds__(BT) void k_karras(const uint4 *__restrict__ code, const uint64_t *__restrict__ key,
                                               int64_t n, int cb, int2 *__restrict__ range,
                                               int32_t *__restrict__ delta, int32_t *__restrict__ split,
                                               uint32_t *__restrict__ chain) {
  // the keys of this workgroup's cells and of BT cells on either side, in
  // LDS: the searches of nodes spanning up to about BT cells (nearly all of
  // them) probe only there
  __shared__ uint64_t kw[3 * BT];
  const int64_t base = (int64_t)blockIdx.x * BT;
  for (int u = threadIdx.x; u < 3 * BT; u += BT) {
    const int64_t jj = base - BT + u;
    kw[u] = jj >= 0 && jj < n ? key[jj] : 0ull;
  }
  __syncthreads();
  const int64_t i = base + threadIdx.x;
  if (i >= n - 1) return;
  const uint64_t ki = kw[BT + threadIdx.x];
  auto dl = [&](int64_t jj) -> int {
    if (jj < 0 || jj >= n) return -1;
    const uint64_t u = (uint64_t)(jj - (base - BT));
    return key_delta64(ki, u < (uint64_t)(3 * BT) ? kw[u] : key[jj], code, n, cb, i, jj);
  };
  int d = (dl(i + 1) - dl(i - 1)) >= 0 ? 1 : -1;
  int dmin = dl(i - d);
  int64_t lmax = 2;
  while (dl(i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (dl(i + (l + t) * d) > dmin) l += t;
  int64_t j = i + l * d;
  int dnode = dl(j);
  int64_t s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (dl(i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
  range[i] = make_int2((int)min(i, j), (int)max(i, j));
  delta[i] = dnode;
  split[i] = (int32_t)gamma;
  if (dnode >= TM_CID_BITS) atomicAdd(&chain[max(i, j)], 1u);
}

__global__ void k_colroot(const int64_t *__restrict__ off, int ncol, const int2 *__restrict__ range, int64_t n,
                          int32_t *__restrict__ root, unsigned int *__restrict__ flags) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  int64_t s = off[c], e = off[c + 1] - 1;
  int32_t r = -1;
  if (e == s) {
    r = (int32_t)((n - 1) + s);
  } else if (e > s) {
    if (s <= n - 2 && range[s].x == s && range[s].y == e) r = (int32_t)s;
    else if (e <= n - 2 && range[e].x == s && range[e].y == e) r = (int32_t)e;
    else atomicOr(flags, (unsigned)TM_FLAG_ROOT);
  }
  root[c] = r;
}

// The walk's tree.  Each colony's radix tree is laid out in the order the
// descending walk visits it (the first-visited child -- the one holding the
// higher keys -- right after its parent), so a node that overlaps the query
// box is followed by the next record and a node that does not by its escape
// link: one 16-byte record per visit, and runs of hits read consecutive
// records.  Colony c (cells [s, e]) takes the slots [2s, 2e + 1).  In that
// order a node covering cells [f, l] sits at
//   2s + 2 (e - l) + (its ancestors up to the colony root of which it lies in
//                     the first-visited child)
// -- the nodes before it are those ancestors and, for each ancestor it lies
// in the second child of, that ancestor's first child with its 2 k - 1 nodes
// over the k cells above l -- and its subtree ends 2 (l - f + 1) - 1 slots on.
// A record holds the node's region as codes (lower corner; the size level in
// the spare top bits of x and y; in z's, whether the y and z sides are half
// the x side; bit 31 of y set for a leaf) and, in .w,
// the escape slot of an internal node or the cell of a leaf (whose escape is
// always the next slot).

// the record of node k (a leaf when k >= n - 1) covering cells [f, l], placed
// at slot pos
__device__ __forceinline__ void put_wnode(const uint4 *__restrict__ code, const int32_t *__restrict__ span,
                                          const int32_t *__restrict__ delta, int64_t n, int64_t k, int64_t f,
                                          int64_t l, int64_t pos, uint4 *__restrict__ wn) {
  const bool leaf = k >= n - 1;
  const uint4 cf = code[f];
  uint32_t L, rr = 0, sx, sy, sz;
  if (leaf) {
    L = (uint32_t)min(span[f], 29);
    sx = sy = sz = 1u << L;
  } else if (delta[k] >= TM_CID_BITS) {
    const int dm = delta[k] - TM_CID_BITS;  // duplicate codes (already flagged) give dm >= 87: level 0
    rr = (uint32_t)(dm % 3);
    L = dm / 3 >= 29 ? 0u : 29u - (uint32_t)(dm / 3);
    sx = 1u << L;
    sy = sx >> (rr >= 2 ? 1 : 0);
    sz = sx >> (rr >= 1 ? 1 : 0);
  } else {
    L = 29u;
    sx = sy = sz = 0u;  // the whole code space (lower corner 0)
  }
  const uint32_t lx = sx ? cf.x & ~(sx - 1u) : 0u, ly = sy ? cf.y & ~(sy - 1u) : 0u, lz = sz ? cf.z & ~(sz - 1u) : 0u;
  const uint32_t w = leaf ? (uint32_t)f : (uint32_t)(pos + 2 * (l - f + 1) - 1);
  // z's spare bits: bit 29 set when the z side is halved, bit 30 when the y side is
  const uint32_t halved = (rr >= 1 ? 1u : 0u) | (rr >= 2 ? 2u : 0u);
  wn[pos] = make_uint4(lx | ((L & 7u) << 29), ly | ((L >> 3) << 29) | (leaf ? 0x80000000u : 0u), lz | (halved << 29), w);
}

// One thread per node k (a leaf when k >= n - 1) covering cells [f, l] of
// colony [s, e]; internal nodes above the colony roots are not placed.  The
// ancestors of the node in whose first-visited child it lies are the
// colony's internal nodes B with split(B) < f and l <= last(B): the
// colony's boundaries (p - 1, p) for s < p <= f are each the split of one
// internal node of the colony, and the ranges nest, so of those nodes the
// ones that are not such ancestors are the ones ending before f -- the
// P[f] - P[s] nodes counted at cells s .. f - 1 (P the exclusive sums of the
// chain counts).  So the node sits at
//   2s + 2 (e - l) + (f - s) - (P[f] - P[s]).
__global__ __launch_bounds__(BT) void k_wnodes(const uint4 *__restrict__ code, const int32_t *__restrict__ span,
                                               const int2 *__restrict__ range, const int32_t *__restrict__ delta,
                                               const int64_t *__restrict__ off, const uint32_t *__restrict__ P,
                                               int64_t n, uint4 *__restrict__ wn, unsigned int *__restrict__ flags) {
  const int64_t k = (int64_t)blockIdx.x * BT + threadIdx.x;
  if (k >= 2 * n - 1) return;
  int64_t f, l;
  if (k < n - 1) {
    if (delta[k] < TM_CID_BITS) return;
    const int2 r = range[k];
    f = r.x;
    l = r.y;
  } else {
    f = l = k - (n - 1);
  }
  const uint32_t cid = code[f].w;
  const int64_t s = off[cid], e = off[cid + 1] - 1;
  const int64_t pos = 2 * s + 2 * (e - l) + (f - s) - (int64_t)(P[f] - P[s]);
  if (l > e || pos < 2 * s || pos > 2 * e) {  // a malformed tree: a slot outside the colony's
    atomicOr(flags, (unsigned)TM_FLAG_WALK);
    return;
  }
  put_wnode(code, span, delta, n, k, f, l, pos, wn);
}

__global__ void k_zero_u32(unsigned int *p, int n) {
  int i = threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

tm_status tm_run_build(tm_ctx *c) {
  const int64_t n = c->n;
  if (c->stage < 1) return tm_fail(c, TM_ERR_STATE, "tm_build before tm_upload");
  c->moved = false;  // the sorted records are read afresh from the cells
  c->sums_valid = false;
  if (n == 0) {
    c->stage = 2;
    return TM_OK;
  }
  hipStream_t st = c->stream;
  k_zero_u32<<<1, 64, 0, st>>>(c->d_flags, 4);
  // bounding boxes and affine maps
  if (c->nchunks > 0)
    k_bbox<<<(unsigned)c->nchunks, BT, 0, st>>>(c->x, c->y, c->z, c->ctype, c->d_chunk_start, c->d_chunk_cid,
                                                 c->d_types, c->ntypes, c->d_cbox, c->d_flags);
  k_affine<<<(unsigned)c->ncol, 64, 0, st>>>(c->d_cbox, c->d_col_chunk, c->ncol, c->d_aff);
  TM_HIP(c, hipGetLastError());
  // Morton keys and the sort
  int cb = 0;
  while ((1ll << cb) < c->ncol) ++cb;
  k_keys<<<tm_blocks(n, BT), BT, 0, st>>>(c->x, c->y, c->z, c->size, c->young, c->ctype, c->phase, n, c->d_col_off,
                                           c->ncol, cb, c->ntypes, c->d_aff, c->d_key, c->d_perm, c->d_pos_o,
                                           c->d_aux_o);
  TM_HIP(c, hipGetLastError());
  TM_TRY(tm_sort_pairs(c, c->d_key, c->d_perm, n, SORT_LO_BIT));
  k_gather<<<tm_blocks(n, BT), BT, 0, st>>>(c->d_perm, n, c->d_pos_o, c->d_aux_o, c->d_aff, c->d_code, c->d_pos,
                                             c->d_posf, c->d_young_s, c->d_ctype_s, c->d_phase_s);
  k_fixties<<<tm_blocks(n, BT), BT, 0, st>>>(c->d_key, n, c->d_perm, c->d_code, c->d_pos, c->d_posf,
                                              c->d_young_s, c->d_ctype_s, c->d_phase_s, c->d_flags);
  k_leaf<<<tm_blocks(n, BT), BT, 0, st>>>(c->d_code, n, c->d_span, c->d_wchain, c->d_flags);
  TM_HIP(c, hipGetLastError());
  // radix tree
  if (n > 1)
    k_karras<<<tm_blocks(n - 1, BT), BT, 0, st>>>(c->d_code, c->d_key, n, cb, c->d_range, c->d_delta, c->d_split,
                                                   c->d_wchain);
  k_colroot<<<tm_blocks(c->ncol, 64), 64, 0, st>>>(c->d_col_off, c->ncol, c->d_range, n, c->d_colroot, c->d_flags);
  TM_TRY(tm_scan_u32(c, c->d_wchain, c->d_wchain, n));
  k_wnodes<<<tm_blocks(2 * n - 1, BT), BT, 0, st>>>(c->d_code, c->d_span, c->d_range, c->d_delta, c->d_col_off,
                                                     c->d_wchain, n, c->d_wn, c->d_flags);
  TM_HIP(c, hipGetLastError());
  c->stage = 2;
  return TM_OK;
}

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
// The deposit sums the reference's terms into each node owned by a grid
// tile in LDS (k_dep_tile): the same terms in a different order (<= ~1e-15
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

// Deposit by grid tile (owner computes; no global atomics, no separate
// clearing).
//
// The grid's nodes are cut into tiles of DT^3 nodes; tile t owns the nodes
// lo + DT t + [0, DT) per axis.  A cell with corner index ci reaches nodes
// ci and ci + 1 per axis; it is keyed to the tile holding max(ci, lo) (the
// clamp takes cells whose ci is lo - 1, which reach only node lo), so all
// of its in-grid corners lie in that tile's nodes or, where its local index
// is DT - 1, in the next tile's first plane.  Its layer bits mark those
// axes.  Counting-sorting the cells by (tile, layer bits) puts, for each
// tile, the cells that reach a neighbour's first plane in whole groups, so
// tile t sums (into LDS) its own cells and, from each of its up to 7 lower
// neighbours t - d, only the groups whose layer bits include d.  Every
// owned node is then complete and written once with a plain store.
// Within a node, terms are added in LDS atomic order (the reference adds
// them in ascending cell order: the same terms, a different order).
//
// Integer work runs in 32 bits: the host admits grids whose node indices
// lie in (-2^29, 2^29), and a cell whose scaled offset from the grid's lower
// corner is 2^30 or more in magnitude reaches no node of such a grid.
#ifndef TM_DEP_DT
#define TM_DEP_DT 16
#endif
#ifndef TM_DEP_THREADS
#define TM_DEP_THREADS 256
#endif
#ifndef TM_DEP_HEAVY
#define TM_DEP_HEAVY 1024
#endif
constexpr int DT = TM_DEP_DT;              // tile edge in nodes (8 or 16)
constexpr int DEP_T = TM_DEP_THREADS;      // threads per tile workgroup
constexpr uint32_t DEP_HEAVY = TM_DEP_HEAVY;  // tiles with this many own cells are launched first
static_assert(DT == 8 || DT == 16, "deposit tile edge");
constexpr int DEP_IDX_LIMIT = 1 << 29;

struct DepGrid {
  double lower[3];
  float res;
  int lo[3], hi[3];  // node index range of this rank's block
  int nt[3];         // tiles per axis
};

// LDS slot of owned node (lx, ly, lz): k fastest, with lz XOR-swizzled by
// (lx, ly) so that the nodes of a small block of cells (what one wave's
// Morton-consecutive cells reach) fall in different banks
__device__ __forceinline__ int dep_slot(int lx, int ly, int lz) {
  const int sw = DT == 16 ? (((ly >> 1) + 3 * lx) * 4) & 15 : ((lx + (ly >> 2)) * 4) & 7;
  return (lx * DT + ly) * DT + (lz ^ sw);
}

// One axis of the reference's cloud-in-cell set-up (src/patches.c:172-190):
// corner index ci = trunc((p - lower) / res), corner coordinate
// lower + (double)((float)ci * res), d = (p - corner) / res, and the weights
// t = 1 - d of node ci and d of node ci + 1 (the reference's a d + (1 - a) t
// for a = 0, 1).  False when the cell is too far from the grid to reach it.
__device__ __forceinline__ bool dep_axis(const DepGrid &g, int a, double p, int &ci, double &w0, double &w1) {
  const double q = (p - g.lower[a]) / (double)g.res;
  if (!(fabs(q) < 1073741824.0)) return false;
  ci = (int)q;
  const float off = (float)ci * g.res;
  const double corner = g.lower[a] + (double)off;
  w1 = (p - corner) / (double)g.res;
  w0 = 1.0 - w1;
  return true;
}

// the deposit weight (src/patches.c:213-227): volume mode, or 0.75 for G0
// cells and 1 otherwise; necrotic cells (phase 5) deposit nothing
__device__ __forceinline__ double dep_wt(const double4 &P, int ph, int weight_mode) {
  return weight_mode == 1 ? ((4.0 / 3.0) * M_PI) * (P.w * P.w * P.w) : (ph == 0 ? 0.75 : 1.0);
}

// key of a cell: (tile, layer bits), or nkeys when it deposits nothing
__device__ __forceinline__ uint32_t dep_key(const double4 &P, bool deposits, const DepGrid &g, uint32_t nkeys) {
  if (!deposits) return nkeys;
  const double pc[3] = {P.x, P.y, P.z};
  uint32_t tc[3], layer = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int ci;
    double w0, w1;
    if (!dep_axis(g, a, pc[a], ci, w0, w1) || ci < g.lo[a] - 1 || ci > g.hi[a]) return nkeys;
    const uint32_t c = (uint32_t)(max(ci, g.lo[a]) - g.lo[a]);
    tc[a] = c / DT;
    layer |= (c % DT == DT - 1 ? 1u : 0u) << a;
  }
  return ((tc[0] * (uint32_t)g.nt[1] + tc[1]) * (uint32_t)g.nt[2] + tc[2]) * 8u + layer;
}

// Lanes of the wave holding the same key are served by one atomic on
// ctr[key] (consecutive cells of the Morton order share keys); a lane's rank
// among the key's cells is ctr[key]'s old value plus its rank among those
// lanes.  wave_rank_issue has each key's first lane issue its atomic and
// keep the old value; wave_rank_read, after every issue, has each lane read
// its key's value from that lane, so the wave waits once for all of them.
struct WaveRank {
  uint32_t held, from, below;
};

__device__ __forceinline__ WaveRank wave_rank_issue(uint32_t key, bool active, uint32_t *__restrict__ ctr) {
  WaveRank w{0u, 0u, 0u};
  uint64_t left = __ballot(active);
  while (left) {
    const int lead = __builtin_ctzll(left);
    const uint32_t lk = (uint32_t)__shfl((int)key, lead, 64);
    const uint64_t same = __ballot(active && key == lk);
    if (__lane_id() == (unsigned)lead) w.held = atomicAdd(ctr + lk, (uint32_t)__builtin_popcountll(same));
    if (active && key == lk) {
      w.from = (uint32_t)lead;
      w.below = (uint32_t)__builtin_popcountll(same & tm_lanemask_lt());
    }
    left &= ~same;
  }
  return w;
}

__device__ __forceinline__ uint32_t wave_rank_read(const WaveRank &w) {
  return (uint32_t)__shfl((int)w.held, (int)w.from, 64) + w.below;
}

__global__ void k_fill_u32(uint32_t *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// Counting sort of the cells by key: count (each cell's rank within its key
// from the count's atomics), exclusive scan, then each cell placed at its
// key's offset plus its rank.  Within a key the order is that of the
// atomics.  Each thread takes DEP_CU cells, 256 apart, all loads issued
// first.
constexpr int DEP_CU = 2;

// kr[i] = key | rank << 32 | (G0 cell under weight mode 0) << 63 (key
// nkeys: deposits nothing).  Block 0 also clears the two launch-order
// counters ctr.
__global__ __launch_bounds__(256) void k_dep_count(const double4 *__restrict__ pos,
                                                   const int32_t *__restrict__ phase_s, int64_t n, DepGrid g,
                                                   int weight_mode, uint32_t nkeys, uint64_t *__restrict__ kr,
                                                   uint32_t *__restrict__ cnt, uint32_t *__restrict__ ctr) {
  const int64_t i0 = (int64_t)blockIdx.x * (256 * DEP_CU) + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x < 2) ctr[threadIdx.x] = 0u;
  double4 P[DEP_CU];
  int ph[DEP_CU];
#pragma unroll
  for (int u = 0; u < DEP_CU; ++u) {
    const int64_t i = i0 + 256 * u;
    P[u] = i < n ? pos[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    ph[u] = i < n && weight_mode != 1 ? phase_s[i] : 1;
  }
  uint32_t k[DEP_CU];
  WaveRank w[DEP_CU];
#pragma unroll
  for (int u = 0; u < DEP_CU; ++u) {
    const int64_t i = i0 + 256 * u;
    k[u] = i < n ? dep_key(P[u], ph[u] != 5, g, nkeys) : nkeys;
    w[u] = wave_rank_issue(k[u], k[u] < nkeys, cnt);
  }
#pragma unroll
  for (int u = 0; u < DEP_CU; ++u) {
    const int64_t i = i0 + 256 * u;
    const uint32_t r = wave_rank_read(w[u]);
    if (i < n) kr[i] = (uint64_t)k[u] | (uint64_t)r << 32 | (uint64_t)(weight_mode != 1 && ph[u] == 0) << 63;
  }
}

// Writes each keyed cell's index at its place in key order, with bit 31
// set for a G0 cell under weight mode 0 (its weight is 0.75).  The rank-0
// cell of each key clears the key's count (the scan has read it), so the
// counts are zero after the deposit.
__global__ __launch_bounds__(256) void k_dep_place(const uint64_t *__restrict__ kr, int64_t n, uint32_t nkeys,
                                                   const uint32_t *__restrict__ offs, uint32_t *__restrict__ cnt,
                                                   uint32_t *__restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint64_t e = kr[i];
  const uint32_t k = (uint32_t)e, r = (uint32_t)(e >> 32) & 0x7FFFFFFFu;
  if (k >= nkeys) return;
  if (r == 0) cnt[k] = 0u;
  idx[offs[k] + r] = (uint32_t)i | (uint32_t)(e >> 63) << 31;
}

// Exclusive scan of n words by one workgroup, out[n] the total, for the
// key counts of grids of up to DEP_SCAN1_MAX keys (one launch instead of
// three): chunks of SCAN1_C words staged through LDS (padded one word per
// 32), the next chunk's loads issued before the current one is scanned.
constexpr int SCAN1_T = 1024, SCAN1_E = 16, SCAN1_C = SCAN1_T * SCAN1_E;
constexpr int64_t DEP_SCAN1_MAX = 1 << 16;
__global__ __launch_bounds__(SCAN1_T) void k_dep_scan1(const uint32_t *__restrict__ in, uint32_t n,
                                                       uint32_t *__restrict__ out) {
  __shared__ uint32_t buf[SCAN1_C + SCAN1_C / 32];
  __shared__ uint32_t wtot[SCAN1_T / 64];
  const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
  uint32_t nxt[SCAN1_E];
#pragma unroll
  for (int e = 0; e < SCAN1_E; ++e) {
    const uint32_t i = e * SCAN1_T + t;
    nxt[e] = i < n ? in[i] : 0u;
  }
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += SCAN1_C) {
#pragma unroll
    for (int e = 0; e < SCAN1_E; ++e) {
      const uint32_t i = e * SCAN1_T + t;
      buf[i + (i >> 5)] = nxt[e];
    }
#pragma unroll
    for (int e = 0; e < SCAN1_E; ++e) {
      const uint32_t i = base + SCAN1_C + e * SCAN1_T + t;
      nxt[e] = i < n ? in[i] : 0u;
    }
    __syncthreads();
    uint32_t v[SCAN1_E], sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN1_E; ++k) {
      const uint32_t j = SCAN1_E * t + k;
      v[k] = buf[j + (j >> 5)];
      sum += v[k];
    }
    uint32_t inc = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o, 64);
      if (lane >= (uint32_t)o) inc += y;
    }
    if (lane == 63) wtot[wv] = inc;
    __syncthreads();
    uint32_t wb = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < SCAN1_T / 64; ++q) {
      const uint32_t x = wtot[q];
      if ((uint32_t)q < wv) wb += x;
      tot += x;
    }
    uint32_t run = carry + wb + inc - sum;
#pragma unroll
    for (int k = 0; k < SCAN1_E; ++k) {
      const uint32_t j = SCAN1_E * t + k;
      buf[j + (j >> 5)] = run;
      run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < SCAN1_E; ++e) {
      const uint32_t i = e * SCAN1_T + t;
      if (base + i < n) out[base + i] = buf[i + (i >> 5)];
    }
    carry += tot;
    __syncthreads();
  }
  if (t == 0) out[n] = carry;
}

// The runs of the sorted order a tile reads: its own run (all 8 layer
// groups), then from each lower neighbour t - dd the groups whose layer bits
// include dd, adjacent groups merged.  Entry: dd | first group << 3 | end
// group << 6.
constexpr int DEP_R = 14;
__constant__ uint16_t dep_runs[DEP_R] = {
    0 | 0 << 3 | 8 << 6, 1 | 1 << 3 | 2 << 6, 1 | 3 << 3 | 4 << 6, 1 | 5 << 3 | 6 << 6, 1 | 7 << 3 | 8 << 6,
    2 | 2 << 3 | 4 << 6, 2 | 6 << 3 | 8 << 6, 3 | 3 << 3 | 4 << 6, 3 | 7 << 3 | 8 << 6, 4 | 4 << 3 | 8 << 6,
    5 | 5 << 3 | 6 << 6, 5 | 7 << 3 | 8 << 6, 6 | 6 << 3 | 8 << 6, 7 | 7 << 3 | 8 << 6};

// A launch slot of the tile pass: a record of DEP_SLOT words, [0] the tile,
// [1 + r] the start of run r and [1 + DEP_R + r] the inclusive prefix of run
// lengths through run r.
constexpr int DEP_SLOT = 32;
static_assert(1 + 2 * DEP_R <= DEP_SLOT, "slot record");

// One thread per tile: its runs, and its slot in launch order.  Tiles with
// at least DEP_HEAVY own cells take slots from the front, so that the
// longest workgroups start first; the rest from the back.  ctr[0..1]
// (front, back) start at zero.
__global__ __launch_bounds__(256) void k_dep_order(const uint32_t *__restrict__ offs, DepGrid g, uint32_t ntiles,
                                                   uint32_t *__restrict__ ctr, uint32_t *__restrict__ slots) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= ntiles) return;
  const int tz = (int)(t % (uint32_t)g.nt[2]), ty = (int)((t / (uint32_t)g.nt[2]) % (uint32_t)g.nt[1]),
            tx = (int)(t / ((uint32_t)g.nt[2] * (uint32_t)g.nt[1]));
  uint32_t rb[DEP_R], rp[DEP_R], tot = 0;
#pragma unroll
  for (int r = 0; r < DEP_R; ++r) {
    const uint32_t e = dep_runs[r], dd = e & 7u;
    const int nx = tx - (int)(dd & 1u), ny = ty - (int)((dd >> 1) & 1u), nz = tz - (int)((dd >> 2) & 1u);
    uint32_t b = 0, len = 0;
    if (nx >= 0 && ny >= 0 && nz >= 0) {
      const uint32_t nk = (((uint32_t)nx * (uint32_t)g.nt[1] + (uint32_t)ny) * (uint32_t)g.nt[2] + (uint32_t)nz) * 8u;
      b = offs[nk + ((e >> 3) & 7u)];
      len = offs[nk + (e >> 6)] - b;
    }
    rb[r] = b;
    tot += len;
    rp[r] = tot;
  }
  const uint32_t slot = rp[0] >= DEP_HEAVY ? atomicAdd(ctr, 1u) : ntiles - 1 - atomicAdd(ctr + 1, 1u);
  uint32_t *w = slots + (size_t)slot * DEP_SLOT;
  w[0] = t;
#pragma unroll
  for (int r = 0; r < DEP_R; ++r) {
    w[1 + r] = rb[r];
    w[1 + DEP_R + r] = rp[r];
  }
}

// adds the corners of a cell that fall in the tile's owned in-grid nodes:
// local corner index l (relative to the tile's first node) per axis, node
// l + o owned when 0 <= l + o <= vhi; weights t = 1 - d and d as the
// reference forms them, multiplied ((wt wx) wy) wz (src/patches.c:172-227)
__device__ __forceinline__ void dep_cell(const double4 &R, const int l[3], const int vhi[3], double *acc) {
  const double w[3][2] = {{1.0 - R.x, R.x}, {1.0 - R.y, R.y}, {1.0 - R.z, R.z}};
#pragma unroll
  for (int ax = 0; ax < 2; ++ax) {
    if (l[0] + ax < 0 || l[0] + ax > vhi[0]) continue;
    const double wx = R.w * w[0][ax];
#pragma unroll
    for (int ay = 0; ay < 2; ++ay) {
      if (l[1] + ay < 0 || l[1] + ay > vhi[1]) continue;
      const double wxy = wx * w[1][ay];
#pragma unroll
      for (int az = 0; az < 2; ++az) {
        if (l[2] + az < 0 || l[2] + az > vhi[2]) continue;
        atomicAdd(&acc[dep_slot(l[0] + ax, l[1] + ay, l[2] + az)], wxy * w[2][az]);
      }
    }
  }
}

constexpr int DEP_U = 4;  // cells per thread and round in k_dep_tile

// One workgroup per tile, in launch-slot order.  The tile's runs are laid
// end to end as one index space that every thread strides through, DEP_U
// cells per round with all loads issued first.  A cell's corner index and
// weights are formed here from its position, as the key was (its corner
// index relative to this tile's first node is -1 for a cell of a lower
// neighbour's last plane).
__global__ __launch_bounds__(DEP_T) void k_dep_tile(const double4 *__restrict__ pos, int weight_mode,
                                                  const uint32_t *__restrict__ idx,
                                                  const uint32_t *__restrict__ slots, DepGrid g, int64_t ls1,
                                                  int64_t ls2, double *__restrict__ prod) {
  __shared__ double acc[DT * DT * DT];
  __shared__ uint32_t w[DEP_SLOT];  // the slot record; w[1 + r] run starts, w[1 + DEP_R + r] prefix
  if (threadIdx.x < DEP_SLOT) w[threadIdx.x] = slots[(size_t)blockIdx.x * DEP_SLOT + threadIdx.x];
  for (int k = threadIdx.x; k < DT * DT * DT; k += DEP_T) acc[k] = 0.0;
  __syncthreads();
  const uint32_t tile = w[0];
  const int tz = (int)(tile % (uint32_t)g.nt[2]), ty = (int)((tile / (uint32_t)g.nt[2]) % (uint32_t)g.nt[1]),
            tx = (int)(tile / ((uint32_t)g.nt[2] * (uint32_t)g.nt[1]));
  const int T[3] = {g.lo[0] + DT * tx, g.lo[1] + DT * ty, g.lo[2] + DT * tz};
  const int vhi[3] = {min(g.hi[0] - T[0], DT - 1), min(g.hi[1] - T[1], DT - 1), min(g.hi[2] - T[2], DT - 1)};
  const uint32_t total = w[2 * DEP_R];
  // the cells of round v0 (the next round's indices are read while this
  // round's positions are)
  auto cells_of = [&](uint32_t v0, uint32_t c[DEP_U]) {
#pragma unroll
    for (int u = 0; u < DEP_U; ++u) {
      const uint32_t v = v0 + DEP_T * u + threadIdx.x;
      c[u] = UINT32_MAX;
      if (v < total) {
        int r = 0;
        while (w[1 + DEP_R + r] <= v) ++r;  // ends at r < DEP_R since v < total
        c[u] = idx[w[1 + r] + (v - (r == 0 ? 0u : w[DEP_R + r]))];
      }
    }
  };
  uint32_t cn[DEP_U];
  cells_of(0, cn);
  for (uint32_t v0 = 0; v0 < total; v0 += DEP_U * DEP_T) {
    uint32_t c[DEP_U];
#pragma unroll
    for (int u = 0; u < DEP_U; ++u) c[u] = cn[u];
    double4 P[DEP_U];
#pragma unroll
    for (int u = 0; u < DEP_U; ++u)
      P[u] = c[u] != UINT32_MAX ? pos[c[u] & 0x7FFFFFFFu] : make_double4(0.0, 0.0, 0.0, 0.0);
    cells_of(v0 + DEP_U * DEP_T, cn);
#pragma unroll
    for (int u = 0; u < DEP_U; ++u) {
      if (c[u] == UINT32_MAX) continue;
      const double pc[3] = {P[u].x, P[u].y, P[u].z};
      double d[3];
      int l[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        int ci;
        double w0;
        (void)dep_axis(g, a, pc[a], ci, w0, d[a]);  // a keyed cell reaches the grid
        l[a] = ci - T[a];
      }
      const int ph = (c[u] >> 31) ? 0 : 1;  // G0 or not, as the key pass found it
      dep_cell(make_double4(d[0], d[1], d[2], dep_wt(P[u], ph, weight_mode)), l, vhi, acc);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < DT * DT * DT; k += DEP_T) {
    const int lx = k / (DT * DT), ly = (k / DT) % DT, lz = k % DT;
    if (lx > vhi[0] || ly > vhi[1] || lz > vhi[2]) continue;
    prod[((int64_t)(T[0] + lx - g.lo[0]) * ls1 + (T[1] + ly - g.lo[1])) * ls2 + (T[2] + lz - g.lo[2])] =
        acc[dep_slot(lx, ly, lz)];
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
  DepGrid g;
  for (int a = 0; a < 3; ++a) {
    if (grid->loweridx[a] <= -DEP_IDX_LIMIT || grid->upperidx[a] >= DEP_IDX_LIMIT)
      return tm_fail(c, TM_ERR_INVALID, "grid node indices must lie in (-2^29, 2^29) for the deposit");
    g.lower[a] = grid->lower[a];
    g.lo[a] = (int)grid->loweridx[a];
    g.hi[a] = (int)grid->upperidx[a];
    g.nt[a] = (int)((grid->localsize[a] + DT - 1) / DT);
  }
  g.res = grid->resolution;
  const int64_t ntiles = (int64_t)g.nt[0] * g.nt[1] * g.nt[2];
  const int64_t nkeys = ntiles * 8;  // the key nkeys marks a cell that deposits nothing
  if (nkeys >= (int64_t)UINT32_MAX) return tm_fail(c, TM_ERR_INVALID, "grid too large for the deposit's tiles");
  const int64_t n = c->n;
  if ((uint64_t)ntiles * DEP_SLOT >= UINT32_MAX) return tm_fail(c, TM_ERR_INVALID, "grid too large for the deposit's tiles");
  // per key a count (nkeys + 1, the last always 0), two order counters, per
  // key an exclusive offset (nkeys + 1: offs[nkeys] is the number of keyed
  // cells), the tiles' slot records
  const int64_t cap0 = c->cap_drange;
  TM_TRY(tm_ensure(c, &c->d_drange, c->cap_drange, 2 * nkeys + 4 + DEP_SLOT * ntiles, "deposit offsets"));
  if (c->cap_drange != cap0) c->dep_zero = 0;
  uint32_t *cnt = c->d_drange, *ctr = cnt + nkeys + 1, *offs = ctr + 2, *slots = offs + nkeys + 1;
  // counts left at zero by the last deposit need no clearing; the counters
  // are cleared by k_dep_count
  if (c->dep_zero < nkeys + 1 || n == 0) {
    k_fill_u32<<<(unsigned)std::min<int64_t>(tm_blocks(nkeys + 3, 256), 4096), 256, 0, c->stream>>>(cnt, nkeys + 3);
  }
  c->dep_zero = 0;
  if (n > 0) {
    TM_TRY(tm_ensure(c, &c->d_dkr, c->cap_dkr, n, "deposit keys"));
    TM_TRY(tm_ensure(c, &c->d_didx, c->cap_didx, n, "deposit cell order"));
    k_dep_count<<<tm_blocks(n, 256 * DEP_CU), 256, 0, c->stream>>>(c->d_pos, c->d_phase_s, n, g, weight_mode,
                                                          (uint32_t)nkeys, c->d_dkr, cnt, ctr);
  }
  if (nkeys <= DEP_SCAN1_MAX) {
    k_dep_scan1<<<1, SCAN1_T, 0, c->stream>>>(cnt, (uint32_t)nkeys, offs);
  } else {
    TM_TRY(tm_scan_u32(c, cnt, offs, nkeys + 1));
  }
  if (n > 0)
    k_dep_place<<<tm_blocks(n, 256), 256, 0, c->stream>>>(c->d_dkr, n, (uint32_t)nkeys, offs, cnt, c->d_didx);
  k_dep_order<<<tm_blocks(ntiles, 256), 256, 0, c->stream>>>(offs, g, (uint32_t)ntiles, ctr, slots);
  // with no cells every offset is 0 and no cell is read
  k_dep_tile<<<(unsigned)ntiles, DEP_T, 0, c->stream>>>(c->d_pos, weight_mode, c->d_didx, slots, g,
                                                        grid->localsize[1], grid->localsize[2], prod);
  TM_HIP(c, hipGetLastError());
  c->dep_zero = nkeys + 1;
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

// LDS tiles for the pair kernels, built from the neighbour lists.
//
// Cells are taken in tiles of tile_width (128, 256 or 512; TM_TILE by
// default) consecutive cells of the Morton order.
// A tile's neighbours overlap heavily (in the 2M-cell Gaussian colony about
// 600 distinct neighbours serve 4,100 pairs per tile), so the pair kernel
// stages the distinct ones in LDS once and reads every pair from there
// instead of gathering 32-byte records through the vector memory path per
// pair.  Per tile this kernel produces:
//   tile_idx   the distinct neighbour indices, ascending (so the staging
//              loads run along the Morton order), at most TM_TILE_MAX
//   tile_cnt   their number, or TM_TILE_OVER when there are more
//   tile_cell  a ranking of the tile's cells by list length, longest first,
//              so each wave of the pair kernel takes cells of similar length
//   nbr16      the lists as 16-bit positions in tile_idx, by rank, in the
//              same order as the 32-bit lists (the FP64 sums keep the
//              reference's order), in chunks of 8 slots: slot k of rank r at
//              nbr16_at(r, k, cap), so a lane reads 8 slots in one 16-byte
//              load and a wave one 1 KB row; each list is padded to the
//              longest list of its wave, rounded up to 8, with the position
//              tile_cnt, where the pair kernel places the far stand-in
// The 32-bit lists stay as they are: the getters and over-full tiles use
// them.
#include "tm_internal.h"

namespace {

// Open-addressing set of the distinct neighbours outside a tile's bitmap
// window.  A new key reserves a place before it is inserted, and a
// reservation is granted only while fewer than RESERVE places are held.
// Each lane holds at most one reservation that is not yet a key of the set
// (two lanes may race to insert the same key; the loser gives its place
// back), so the held places never exceed the keys inserted plus TS - 1.
// Hence: the set never holds more than RESERVE keys, so at least
// HASH - RESERVE slots stay empty and every probe ends; and a refused
// reservation means at least RESERVE - (TS - 1) = TM_TILE_MAX + 1 keys are
// in the set, i.e. the tile is truly over-full.  TS is the tile width (the
// workgroup size): 128, 256 or 512 cells.
template <int TS>
struct TileCfg {
  static constexpr int HASH = TS <= 256 ? 2048 : 4096;
  static constexpr int HBITS = TS <= 256 ? 11 : 12;
  static constexpr uint32_t RESERVE = TM_TILE_MAX + TS;
  static_assert(HASH == 1 << HBITS && HASH > (int)RESERVE, "hash slots");
};
constexpr int SORTN = 2048;  // power of two >= TM_TILE_MAX
constexpr int WB = 16384;    // cells around a tile's middle whose order a bitmap gives
constexpr uint32_t EMPTY = 0xFFFFFFFFu;

template <int HBITS>
__device__ __forceinline__ uint32_t hslot(uint32_t q) {
  return (q * 2654435761u) >> (32 - HBITS);
}

// Ascending bitonic sort of list[0, m) (m <= TS E) by the TS threads of
// the workgroup, each holding E consecutive elements (padding EMPTY sorts
// last).  A compare-exchange partner at distance j lies in the same thread
// (j < E), in another lane of the same wave (j < 64 E: a shuffle), or in
// another wave (through LDS, the only stages that need the workgroup's
// barrier).  Element i keeps the smaller of itself and i ^ j when its place
// in the pair (i & j) matches the direction of its run (i & size).
__device__ __forceinline__ uint32_t bitonic_pick(uint32_t x, uint32_t y, uint32_t i, uint32_t j, uint32_t size) {
  const bool lower = (i & j) == 0, asc = (i & size) == 0;
  return lower == asc ? min(x, y) : max(x, y);
}

template <int TS, int E>
__device__ void tile_sort(uint32_t *list, uint32_t m, uint32_t t) {
  constexpr uint32_t N = (uint32_t)TS * E;
  static_assert(N <= SORTN, "sort buffer");
  uint32_t a[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const uint32_t i = t * E + e;
    a[e] = i < m ? list[i] : 0xFFFFFFFFu;
  }
  for (uint32_t size = 2; size <= N; size <<= 1) {
    for (uint32_t j = size >> 1; j > 0; j >>= 1) {
      if (j >= 64u * E) {
        __syncthreads();  // the previous stage's readers are done
#pragma unroll
        for (int e = 0; e < E; ++e) list[t * E + e] = a[e];
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const uint32_t i = t * E + e;
          a[e] = bitonic_pick(a[e], list[i ^ j], i, j, size);
        }
      } else if (j >= (uint32_t)E) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const uint32_t i = t * E + e;
          const uint32_t y = (uint32_t)__shfl_xor((int)a[e], (int)(j / E), 64);
          a[e] = bitonic_pick(a[e], y, i, j, size);
        }
      } else {
        uint32_t b[E];
#pragma unroll
        for (int e = 0; e < E; ++e) b[e] = a[e ^ j];
#pragma unroll
        for (int e = 0; e < E; ++e) a[e] = bitonic_pick(a[e], b[e], t * E + e, j, size);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const uint32_t i = t * E + e;
    if (i < m) list[i] = a[e];
  }
  __syncthreads();
}

// exclusive prefix sum over the threads of the workgroup (wsum: one word
// per wave)
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t *wsum) {
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(inc, o, 64);
    if (lane >= (uint32_t)o) inc += y;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t i = 0; i < w; ++i) base += wsum[i];
  __syncthreads();
  return base + inc - v;
}

// win: WB, or 0 to send every key through the hash set (same outputs)
template <int TS>
__global__ __launch_bounds__(TS) void k_tiles(int64_t n, int64_t cap, uint32_t win, const uint32_t *__restrict__ nbr,
                                              const uint32_t *__restrict__ count, uint32_t *__restrict__ tile_idx,
                                              uint32_t *__restrict__ tile_cnt, uint16_t *__restrict__ tile_cell,
                                              uint16_t *__restrict__ nbr16, unsigned int *__restrict__ flags) {
  constexpr int HASH = TileCfg<TS>::HASH, HBITS = TileCfg<TS>::HBITS;
  constexpr uint32_t RESERVE = TileCfg<TS>::RESERVE;
  constexpr int WPT = WB / 32 / TS;  // bitmap words per thread
  static_assert(WPT * TS * 32 == WB, "bitmap words");
  __shared__ uint32_t keys[HASH];
  __shared__ uint32_t list[SORTN];
  // list's words are shared by two arrays whose lives do not meet its own:
  // the ranking's bins (dead before the barrier that ends the neighbour
  // pass, after which list is first written) and the hash set's positions
  // (written after every lane has read the sorted set into registers).  The
  // smaller footprint lets seven workgroups share a CU instead of five.
  static_assert(TS <= SORTN && HASH * sizeof(uint16_t) <= sizeof(list), "arrays inside list");
  uint32_t *const bins = list;
  uint16_t *const vals = reinterpret_cast<uint16_t *>(list);
  __shared__ uint32_t rcnt[TS];
  __shared__ uint16_t order[TS];
  __shared__ uint32_t wsum[TS / 64];
  __shared__ uint32_t bm[WB / 32], pre[WB / 32];
  __shared__ uint32_t distinct, nwin, nbelow, hfull;
  const int64_t g = blockIdx.x;
  const uint32_t t = threadIdx.x;
  const int64_t cell = g * TS + t;
  const uint32_t cnt = cell < n ? min(count[cell], (uint32_t)cap) : 0u;
  for (int i = t; i < HASH; i += TS) keys[i] = EMPTY;
  for (int i = t; i < WB / 32; i += TS) bm[i] = 0;
  bins[t] = 0;
  if (t == 0) distinct = nbelow = hfull = 0;
  __syncthreads();

  // rank the cells by list length, longest first (counting sort; the order
  // among equal lengths is immaterial: each cell is still summed by one
  // lane over its own list)
  const uint32_t b = (uint32_t)(TS - 1) - min(cnt, (uint32_t)(TS - 1));  // lists of TS - 1 or more share a bin
  atomicAdd(&bins[b], 1u);
  __syncthreads();
  const uint32_t start = block_scan(bins[t], wsum);
  bins[t] = start;
  __syncthreads();
  const uint32_t rank = atomicAdd(&bins[b], 1u);
  order[rank] = (uint16_t)t;
  rcnt[rank] = cnt;
  __syncthreads();
  // from here on thread t reads rank t's list, so the lists of a wave's
  // lanes have similar lengths; its first 16 slots stay in registers for the
  // second pass over it
  const uint32_t off = order[t];
  const int64_t src = g * TS + off;
  const uint32_t rc = rcnt[t];
  uint32_t p0[8], p1[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) p0[j] = (uint32_t)j < rc ? nbr[tm_slot(src, j, cap)] : EMPTY;
#pragma unroll
  for (int j = 0; j < 8; ++j) p1[j] = (uint32_t)(8 + j) < rc ? nbr[tm_slot(src, 8 + j, cap)] : EMPTY;

  // Distinct neighbours of the tile.  Most lie within WB / 2 cells of its
  // middle (84% in the 2M-cell Gaussian colony): those are marked in a
  // bitmap, whose popcount prefix later gives their ascending order and
  // their positions; the others go into the hash set and are sorted.  The
  // set's order is: the sorted ones below the window, the window's, the
  // sorted ones above it.  (Stop counting past TM_TILE_MAX; a lane reads 8
  // list slots per round so their loads overlap.)
  const int64_t wlo = g * TS + TS / 2 - (int64_t)(win / 2);  // may be negative
  auto insert8 = [&](const uint32_t (&qs)[8]) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t q = qs[j];
      if (q == EMPTY) continue;
      const int64_t r = (int64_t)q - wlo;
      if (r >= 0 && r < (int64_t)win) {
        atomicOr(&bm[r >> 5], 1u << (r & 31));
        continue;
      }
      // look it up first (most occurrences are repeats); slots are never
      // emptied, so at most TM_TILE_MAX of the HASH slots read are taken
      uint32_t h = hslot<HBITS>(q);
      uint32_t k = keys[h];
      while (k != q && k != EMPTY) {
        h = (h + 1) & (HASH - 1);
        k = keys[h];
      }
      if (k == q) continue;
      // a new key: reserve a place (the count of reserved places bounds the
      // set), then insert from the empty slot found
      if (atomicAdd(&distinct, 1u) >= RESERVE) {
        atomicSub(&distinct, 1u);
        hfull = 1u;  // more than TM_TILE_MAX keys: the tile keeps the 32-bit lists
        continue;
      }
      for (;;) {
        const uint32_t old = atomicCAS(&keys[h], EMPTY, q);
        if (old == EMPTY) break;
        if (old == q) {  // another lane inserted it meanwhile
          atomicSub(&distinct, 1u);
          break;
        }
        h = (h + 1) & (HASH - 1);
      }
    }
  };
  if (rc > 0) insert8(p0);
  if (rc > 8) insert8(p1);
  for (uint32_t k0 = 16; k0 < rc; k0 += 8) {
    if (*(volatile uint32_t *)&hfull) break;
    uint32_t qs[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) qs[j] = k0 + j < rc ? nbr[tm_slot(src, k0 + j, cap)] : EMPTY;
    insert8(qs);
  }
  __syncthreads();
  // the window's count and prefix (thread t holds words WPT t .. WPT t + WPT - 1)
  uint32_t bw[WPT], cw = 0;
#pragma unroll
  for (int e = 0; e < WPT; ++e) {
    bw[e] = bm[WPT * t + e];
    cw += __popc(bw[e]);
  }
  const uint32_t pw = block_scan(cw, wsum);
  {
    uint32_t run = pw;
#pragma unroll
    for (int e = 0; e < WPT; ++e) {
      pre[WPT * t + e] = run;
      run += __popc(bw[e]);
    }
  }
  if (t == TS - 1) nwin = pw + cw;
  __syncthreads();
  const uint32_t mo = distinct, mw = nwin, m = mo + mw;
  if (m > TM_TILE_MAX || hfull) {
    if (t == 0) {
      tile_cnt[g] = TM_TILE_OVER;
      atomicMax(flags + 4, m);
      atomicAdd(flags + 14, 1u);
    }
    return;
  }
  // compact the hash set in slot order, then sort it ascending
  uint32_t mine = 0, low = 0;
  for (int i = 0; i < HASH / TS; ++i) {
    const uint32_t k = keys[t * (HASH / TS) + i];
    if (k == EMPTY) continue;
    ++mine;
    low += (int64_t)k < wlo;
  }
  if (low) atomicAdd(&nbelow, low);
  uint32_t w = block_scan(mine, wsum);
  for (int i = 0; i < HASH / TS; ++i) {
    const uint32_t k = keys[t * (HASH / TS) + i];
    if (k != EMPTY) list[w++] = k;
  }
  __syncthreads();
  const uint32_t mb = nbelow;
  // mo <= TM_TILE_MAX <= SORTN: the largest sort below has TS E = SORTN
  static_assert(SORTN >= TM_TILE_MAX && SORTN % TS == 0, "tile_sort sizes");
  if (mo == 0) {
    // all in the window: nothing to sort
  } else if (mo <= TS) tile_sort<TS, 1>(list, mo, t);
  else if (mo <= 2 * TS) tile_sort<TS, 2>(list, mo, t);
  else if (TS * 4 <= SORTN && mo <= 4 * TS) tile_sort<TS, (TS * 4 <= SORTN ? 4 : 1)>(list, mo, t);
  else if (TS * 8 <= SORTN && mo <= 8 * TS) tile_sort<TS, (TS * 8 <= SORTN ? 8 : 1)>(list, mo, t);
  else tile_sort<TS, SORTN / TS>(list, mo, t);
  // positions: the sorted ones (read out of list before vals overwrites it) ...
  constexpr int EV = SORTN / TS;
  uint32_t qv[EV];
#pragma unroll
  for (int e = 0; e < EV; ++e) {
    const uint32_t i = t + (uint32_t)TS * e;
    qv[e] = i < mo ? list[i] : EMPTY;
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < EV; ++e) {
    const uint32_t i = t + (uint32_t)TS * e;
    if (i >= mo) break;
    const uint32_t q = qv[e];
    const uint32_t pos = i < mb ? i : i + mw;
    uint32_t h = hslot<HBITS>(q);
    while (keys[h] != q) h = (h + 1) & (HASH - 1);
    vals[h] = (uint16_t)pos;
    tile_idx[g * TM_TILE_MAX + pos] = q;
  }
  // ... and the window's, in order
  {
    uint32_t pos = mb + pw;
    for (int e = 0; e < WPT; ++e) {
      uint32_t bits = bw[e];
      const int64_t q0 = wlo + 32 * (WPT * (int64_t)t + e);
      while (bits) {
        const uint32_t bt = (uint32_t)__ffs(bits) - 1u;
        bits &= bits - 1u;
        tile_idx[g * TM_TILE_MAX + pos++] = (uint32_t)(q0 + bt);
      }
    }
  }
  if (t == 0) {
    tile_cnt[g] = m;
    atomicMax(flags + 4, m);
    if (m + 1 > TM_TILE_LDS4) atomicAdd(flags + 5, 1u);  // staged only without the occupancy cap
  }
  __syncthreads();

  // lists by rank: thread t writes rank t's list (rows of a wave coalesced)
  uint32_t pad = rc;  // longest list of this wave of ranks
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) pad = max(pad, (uint32_t)__shfl_xor((int)pad, o, 64));
  tile_cell[g * TS + t] = (uint16_t)off;
  const int64_t r = g * TS + t;
  const uint32_t pad8 = (pad + 7) & ~7u;  // <= cap, a multiple of 8
  // one 16-byte chunk of 8 slots per round: its 8 list loads overlap
  auto emit8 = [&](uint32_t k0, const uint32_t (&qs)[8]) {
    uint32_t v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      v[j] = m;
      const uint32_t q = qs[j];
      if (q != EMPTY) {
        const int64_t rr = (int64_t)q - wlo;
        if (rr >= 0 && rr < (int64_t)win) {
          const uint32_t wd = (uint32_t)(rr >> 5), bt = (uint32_t)(rr & 31);
          v[j] = mb + pre[wd] + __popc(bm[wd] & ((1u << bt) - 1u));
        } else {
          uint32_t h = hslot<HBITS>(q);
          while (keys[h] != q) h = (h + 1) & (HASH - 1);
          v[j] = vals[h];
        }
      }
    }
    *reinterpret_cast<uint4 *>(nbr16 + tm_nbr16_at(r, k0, cap)) =
        make_uint4(v[0] | (v[1] << 16), v[2] | (v[3] << 16), v[4] | (v[5] << 16), v[6] | (v[7] << 16));
  };
  if (pad8 > 0) emit8(0, p0);
  if (pad8 > 8) emit8(8, p1);
  for (uint32_t k0 = 16; k0 < pad8; k0 += 8) {
    uint32_t qs[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) qs[j] = k0 + j < rc ? nbr[tm_slot(src, k0 + j, cap)] : EMPTY;
    emit8(k0, qs);
  }
}

// clears the tile counters: flags[4], flags[5] and flags[14]
__global__ void k_zero_tile_flags(unsigned int *flags) {
  flags[4] = 0;
  flags[5] = 0;
  flags[14] = 0;
}

template <int TS>
void launch_tiles(tm_ctx *c, int64_t nt) {
  k_tiles<TS><<<(unsigned)nt, TS, 0, c->stream>>>(c->n, c->nbr_alloc, c->tile_window ? (uint32_t)WB : 0u, c->d_nbr,
                                                   c->d_ncount, c->d_tile_idx, c->d_tile_cnt, c->d_tile_cell,
                                                   c->d_nbr16, c->d_flags);
}

}  // namespace

tm_status tm_run_tiles(tm_ctx *c, bool check) {
  const int64_t n = c->n;
  if (!c->sph_tiles || n == 0) {
    c->tile_lds = 0;
    c->tiles_ts = 0;
    return TM_OK;
  }
  const int TS = c->tile_width;
  const int64_t nt = (n + TS - 1) / TS;
  TM_TRY(tm_ensure(c, &c->d_tile_idx, c->cap_tile_idx, nt * TM_TILE_MAX, "tile indices"));
  TM_TRY(tm_ensure(c, &c->d_tile_cnt, c->cap_tile_cnt, nt, "tile sizes"));
  TM_TRY(tm_ensure(c, &c->d_tile_cell, c->cap_tile_cell, nt * TS, "tile ranks"));
  TM_TRY(tm_ensure(c, &c->d_nbr16, c->cap_nbr16, tm_slot_elems(nt * TS, c->nbr_alloc), "tile lists"));
  k_zero_tile_flags<<<1, 1, 0, c->stream>>>(c->d_flags);
  switch (TS) {
    case 128: launch_tiles<128>(c, nt); break;
    case 512: launch_tiles<512>(c, nt); break;
    default: launch_tiles<256>(c, nt); break;
  }
  TM_HIP(c, hipGetLastError());
  c->tiles_ts = TS;
  if (check) {
    TM_HIP(c, hipMemcpyAsync(c->h_flags + 4, c->d_flags + 4, 2 * sizeof(unsigned int), hipMemcpyDeviceToHost,
                             c->stream));
    TM_HIP(c, hipMemcpyAsync(c->h_flags + 14, c->d_flags + 14, sizeof(unsigned int), hipMemcpyDeviceToHost,
                             c->stream));
    TM_HIP(c, hipStreamSynchronize(c->stream));
    c->tiles_over = c->h_flags[14];
    // records staged per tile: the largest tile plus the far stand-in, rounded
    // (at least TS: the pair kernel also parks a tile's sums there) ...
    int64_t lds = std::max<int64_t>(TS, std::min<int64_t>(TM_TILE_MAX + 1, ((int64_t)c->h_flags[4] + 1 + 63) & ~63ll));
    // ... but no more than four workgroups' worth per CU when at most one tile
    // in twenty is larger: those take the pair kernel's global-record path and
    // every other tile runs at the higher occupancy
    if (lds > TM_TILE_LDS4 && c->tile_cap4 && 20 * (int64_t)c->h_flags[5] <= nt) lds = TM_TILE_LDS4;
    c->tile_lds = (uint32_t)lds;
    c->tile_lds_ts = TS;
  } else if (c->tile_lds == 0 || c->tile_lds_ts != TS) {
    // unsized, or sized for another width (the sums of TS cells must fit):
    // stage the most a tile can hold
    c->tile_lds = TM_TILE_MAX + 1;
    c->tile_lds_ts = TS;
  }
  return TM_OK;
}

// Device-wide exclusive scan and stable LSD radix sort of (u64 key, i32 value)
// pairs, 8-bit digits, for the Morton ordering of cells.
//
// Each pass: histogram per 4096-key tile (LDS atomics) -> exclusive scan of
// the digit-major histogram -> stable scatter.  Stable ranks inside a tile
// come from 64-lane ballots: the lanes holding the same digit are found with
// eight ballots, the leader of each group advances a per-wave digit counter
// in LDS.
#include "tm_internal.h"

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 8;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;

// lanes of this wave that are valid and hold the same 8-bit digit
__device__ __forceinline__ unsigned long long tm_match8(uint32_t d, bool valid) {
  unsigned long long m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    bool bit = (d >> b) & 1u;
    unsigned long long bb = __ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *__restrict__ keys, int64_t n, int shift,
                                                        uint32_t *__restrict__ hist, int nblocks) {
  __shared__ uint32_t h[RS_WAVES][256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int j = tid; j < RS_WAVES * 256; j += RS_THREADS) (&h[0][0])[j] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE + (int64_t)w * (64 * RS_ITEMS);
  uint32_t dg[RS_ITEMS];  // all of this lane's digits are loaded before any is counted
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const int64_t i = base + it * 64 + lane;
    dg[it] = i < n ? (uint32_t)((keys[i] >> shift) & 255u) : 0u;
  }
  // counts only (the ranks are the scatter's), so one LDS atomic per key
  // into this wave's histogram
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it)
    if (base + it * 64 + lane < n) atomicAdd(&h[w][dg[it]], 1u);
  __syncthreads();
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < RS_WAVES; ++k) s += h[k][tid];
  hist[(int64_t)tid * nblocks + blockIdx.x] = s;
}

// Exclusive scan of each digit's row of per-tile counts (one workgroup per
// digit), the row's total stored after the rows: the scatter adds the
// totals of the smaller digits itself.
__global__ __launch_bounds__(RS_THREADS) void k_rs_rowscan(uint32_t *__restrict__ hist, int nblocks) {
  __shared__ uint32_t wsum[RS_WAVES];
  const int d = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  uint32_t *row = hist + (int64_t)d * nblocks;
  const int per = (nblocks + RS_THREADS - 1) / RS_THREADS;
  const int b0 = tid * per, b1 = min(b0 + per, nblocks);
  uint32_t sum = 0;
  for (int b = b0; b < b1; ++b) sum += row[b];
  uint32_t inc = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t run = inc - sum, all = 0;
#pragma unroll
  for (int q = 0; q < RS_WAVES; ++q) {
    if (q < w) run += wsum[q];
    all += wsum[q];
  }
  for (int b = b0; b < b1; ++b) {
    const uint32_t v = row[b];
    row[b] = run;
    run += v;
  }
  if (tid == 0) hist[(int64_t)256 * nblocks + d] = all;
}

// The tile's keys are ranked stably (by digit, then input order), placed in
// LDS in that order, and written out from there: consecutive lanes write
// consecutive slots of one digit's run instead of 64 scattered records.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *__restrict__ keys,
                                                           const int32_t *__restrict__ vals,
                                                           uint64_t *__restrict__ okeys, int32_t *__restrict__ ovals,
                                                           int64_t n, int shift, const uint32_t *__restrict__ offs,
                                                           int nblocks) {
  __shared__ uint32_t cnt[RS_WAVES][256];
  __shared__ uint32_t pre[RS_WAVES][256];
  __shared__ uint32_t gbase[256];
  __shared__ uint32_t tstart[256];  // first tile-local slot of each digit
  __shared__ uint32_t wsum[RS_WAVES];
  __shared__ uint64_t skey[RS_TILE];
  __shared__ int32_t sval[RS_TILE];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int j = tid; j < RS_WAVES * 256; j += RS_THREADS) (&cnt[0][0])[j] = 0;
  // this tile's first slot for digit tid: its row prefix (offs, per digit
  // over the tiles) plus the keys of all smaller digits (the row totals
  // after the rows, scanned here)
  {
    const uint32_t rt = offs[(int64_t)256 * nblocks + tid];
    uint32_t inc = rt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o, 64);
      if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t wb = 0;
#pragma unroll
    for (int q = 0; q < RS_WAVES; ++q)
      if (q < w) wb += wsum[q];
    gbase[tid] = offs[(int64_t)tid * nblocks + blockIdx.x] + wb + inc - rt;
    __syncthreads();
  }
  const int64_t tile0 = (int64_t)blockIdx.x * RS_TILE;
  const int64_t base = tile0 + (int64_t)w * (64 * RS_ITEMS);
  uint64_t k[RS_ITEMS];
  int32_t v[RS_ITEMS];
  uint32_t rank[RS_ITEMS];
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const int64_t i = base + it * 64 + lane;
    const bool valid = i < n;
    k[it] = valid ? keys[i] : 0ull;
    v[it] = valid ? vals[i] : 0;
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const bool valid = base + it * 64 + lane < n;
    uint32_t d = (uint32_t)((k[it] >> shift) & 255u);
    unsigned long long peers = tm_match8(d, valid);
    uint32_t before = (uint32_t)__popcll(peers & tm_lanemask_lt());
    uint32_t c0 = valid ? cnt[w][d] : 0u;
    __builtin_amdgcn_wave_barrier();
    if (valid && before == 0) cnt[w][d] = c0 + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    rank[it] = c0 + before;
  }
  __syncthreads();
  uint32_t total = 0;
  {
    uint32_t run = 0;
#pragma unroll
    for (int q = 0; q < RS_WAVES; ++q) {
      pre[q][tid] = run;
      run += cnt[q][tid];
    }
    total = run;
  }
  // exclusive scan of the digit totals over the 256 threads
  {
    uint32_t inc = total;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o, 64);
      if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t wb = 0;
#pragma unroll
    for (int q = 0; q < RS_WAVES; ++q)
      if (q < w) wb += wsum[q];
    tstart[tid] = wb + inc - total;
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    if (base + it * 64 + lane < n) {
      const uint32_t d = (uint32_t)((k[it] >> shift) & 255u);
      const uint32_t lp = tstart[d] + pre[w][d] + rank[it];  // < the tile's valid count
      skey[lp] = k[it];
      sval[lp] = v[it];
    }
  }
  __syncthreads();
  const int64_t valid_n = min((int64_t)RS_TILE, n - tile0);
  for (int j = tid; j < valid_n; j += RS_THREADS) {
    const uint64_t key = skey[j];
    const uint32_t d = (uint32_t)((key >> shift) & 255u);
    const int64_t pos = (int64_t)gbase[d] + (uint32_t)j - tstart[d];
    if (pos >= 0 && pos < n) {  // always true for a consistent histogram; keeps a bad scan from writing out of bounds
      okeys[pos] = key;
      ovals[pos] = sval[j];
    }
  }
}

// ---- scan ----------------------------------------------------------------

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const int lane = __lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x;
}

// exclusive block scan of one value per thread; returns the block total in *tot
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t *tot) {
  __shared__ uint32_t ws[SC_THREADS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = wave_incl_scan(x);
  if (lane == 63) ws[w] = inc;
  __syncthreads();
  uint32_t wbase = 0, all = 0;
#pragma unroll
  for (int k = 0; k < SC_THREADS / 64; ++k) {
    uint32_t s = ws[k];
    if (k < w) wbase += s;
    all += s;
  }
  __syncthreads();
  *tot = all;
  return wbase + inc - x;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_reduce(const uint32_t *__restrict__ in, int64_t n,
                                                            uint32_t *__restrict__ sums) {
  const int64_t base = (int64_t)blockIdx.x * SC_TILE + (int64_t)threadIdx.x * SC_ITEMS;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SC_ITEMS; ++k)
    if (base + k < n) s += in[base + k];
  uint32_t tot;
  (void)block_excl_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_apply(const uint32_t *__restrict__ in, int64_t n,
                                                           const uint32_t *__restrict__ sums_scanned,
                                                           uint32_t *__restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * SC_TILE + (int64_t)threadIdx.x * SC_ITEMS;
  uint32_t v[SC_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SC_ITEMS; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan(s, &tot) + (sums_scanned ? sums_scanned[blockIdx.x] : 0u);
#pragma unroll
  for (int k = 0; k < SC_ITEMS; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

}  // namespace

// exclusive scan; `out` may alias `in`.  Uses c->d_scan as scratch.
tm_status tm_scan_u32(tm_ctx *c, const uint32_t *in, uint32_t *out, int64_t n) {
  if (n <= 0) return TM_OK;
  int64_t nb = (n + SC_TILE - 1) / SC_TILE;
  if (nb == 1) {
    k_scan_apply<<<1, SC_THREADS, 0, c->stream>>>(in, n, nullptr, out);
    TM_HIP(c, hipGetLastError());
    return TM_OK;
  }
  // scratch layout: level sums for up to three levels
  int64_t nb2 = (nb + SC_TILE - 1) / SC_TILE;
  int64_t need = nb + nb2 + 1;
  TM_TRY(tm_ensure(c, &c->d_scan, c->cap_scan, need, "scan scratch"));
  uint32_t *sums = c->d_scan;
  uint32_t *sums2 = c->d_scan + nb;
  k_scan_reduce<<<(unsigned)nb, SC_THREADS, 0, c->stream>>>(in, n, sums);
  TM_HIP(c, hipGetLastError());
  if (nb2 == 1) {
    k_scan_apply<<<1, SC_THREADS, 0, c->stream>>>(sums, nb, nullptr, sums);
  } else {
    if (nb2 > SC_TILE) return tm_fail(c, TM_ERR_INVALID, "scan too large");
    k_scan_reduce<<<(unsigned)nb2, SC_THREADS, 0, c->stream>>>(sums, nb, sums2);
    k_scan_apply<<<1, SC_THREADS, 0, c->stream>>>(sums2, nb2, nullptr, sums2);
    k_scan_apply<<<(unsigned)nb2, SC_THREADS, 0, c->stream>>>(sums, nb, sums2, sums);
  }
  TM_HIP(c, hipGetLastError());
  k_scan_apply<<<(unsigned)nb, SC_THREADS, 0, c->stream>>>(in, n, sums, out);
  TM_HIP(c, hipGetLastError());
  return TM_OK;
}

// Stable sort of (keys, vals) by key bits [lo_bit, 64) (lo_bit a multiple
// of 16, so the passes are even in number); the result ends in the input
// arrays (the passes ping-pong through c->d_key_alt / d_val_alt).
tm_status tm_sort_pairs(tm_ctx *c, uint64_t *keys, int32_t *vals, int64_t n, int lo_bit) {
  if (n <= 1) return TM_OK;
  if (lo_bit < 0 || lo_bit >= 64 || lo_bit % 16) return tm_fail(c, TM_ERR_INVALID, "sort: lo_bit");
  int nblocks = (int)((n + RS_TILE - 1) / RS_TILE);
  TM_TRY(tm_ensure(c, &c->d_hist, c->cap_hist, (int64_t)256 * nblocks + 256, "radix histogram"));
  uint64_t *k0 = keys, *k1 = c->d_key_alt;
  int32_t *v0 = vals, *v1 = c->d_val_alt;
  for (int shift = lo_bit; shift < 64; shift += 8) {
    k_rs_hist<<<nblocks, RS_THREADS, 0, c->stream>>>(k0, n, shift, c->d_hist, nblocks);
    TM_HIP(c, hipGetLastError());
    k_rs_rowscan<<<256, RS_THREADS, 0, c->stream>>>(c->d_hist, nblocks);
    TM_HIP(c, hipGetLastError());
    k_rs_scatter<<<nblocks, RS_THREADS, 0, c->stream>>>(k0, v0, k1, v1, n, shift, c->d_hist, nblocks);
    TM_HIP(c, hipGetLastError());
    uint64_t *tk = k0;
    k0 = k1;
    k1 = tk;
    int32_t *tv = v0;
    v0 = v1;
    v1 = tv;
  }
  return TM_OK;
}

// Internal declarations of the MI355X cell-mechanics library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/timothy_hip.h"

#define TM_MAX_TYPES 16
#define TM_END 0xFFFFFFFFu
#define TM_CID_BITS 16
#define TM_FLAG_DEGENERATE 1
#define TM_FLAG_ROOT 2
#define TM_FLAG_CTYPE 4
#define TM_FLAG_WALK 8
#define TM_BBOX_CHUNK 2048
// LDS tiles: the pair kernels take cells in tiles of tile_width (128, 256 or
// 512; TM_TILE by default) consecutive (Morton-order) cells; the distinct
// neighbours of a tile, at most
// TM_TILE_MAX, are staged once in LDS and the lists hold 16-bit tile-local
// indices.  A tile with more distinct neighbours is marked TM_TILE_OVER and
// read through the 32-bit lists.
#ifndef TM_TILE
#define TM_TILE 256
#endif
#define TM_TILE_MAX 1536
#define TM_TILE_OVER 0xFFFFFFFFu
// staged records per tile at which four pair-kernel workgroups fill a CU's
// 160 KB of LDS (two 16-byte halves per record)
#define TM_TILE_LDS4 1280
// coordinate of the stand-in record pos[n] that pads neighbour lists: far
// enough that a pair with it lies beyond every cutoff, near enough that its
// squared distance stays finite in FP32
#define TM_FAR 1.0e12
// pinned host words of a context: [0, 16) copies of the device flags,
// [TM_HFLAGS_CG, TM_HFLAGS_WORDS) the environment solve's state read back
#define TM_HFLAGS_WORDS 32
#define TM_HFLAGS_CG 16

// Per cell-type constants, derived on the host with the reference's own
// expressions (src/inline.h:45-46,71-72) so the device reads identical values.
struct TypeDev {
  double size, h, h2, h3, h4;
  double sc;       // h3 in 3D, h2 in 2D: density scale (src/potential.c:58)
  double cW;       // kernel normalisation 8/(pi h3) (3D) or 40/(7 pi h2) (2D)
  double cF;       // gradient normalisation 48/(pi h4) (3D) or 320/(7 pi h3) (2D)
  double cut2;     // largest d2 with RN(sqrt(d2)) <= h: exact form of dist <= h
  double hinv, sinv;  // RN(1/h), RN(1/size)
  float hinv_f;
};

struct tm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // configuration
  int ntypes = 0;
  int dim = 3;
  tm_celltype types[TM_MAX_TYPES];
  TypeDev *d_types = nullptr;

  // tunables
  int64_t nbr_cap = 48;
  int sph_block = 128;
  int sph_unroll_fp64 = 4;   // neighbours fetched per trip in the FP64 pair kernel
  int sph_unroll_mixed = 4;  // the same for the mixed kernel
  int nbr_block = 256;
  int walk_window = 1280;  // slots of the tree a 256-cell walk workgroup stages in LDS (0, 512, 1024, 1280, 2048)
  int walk_f32 = 1;  // the walk screens candidates with FP32 positions and an error bound (results unchanged)
  int sph_lanes = 1;
  int sph_balance = 0;  // FP64 kernel without tiles: rank cells by list length inside each workgroup

  // cells (reference order)
  int64_t n = 0, cap_n = 0;
  int ncol = 0;
  std::vector<int64_t> col_off;
  bool external = false;
  const double *x = nullptr, *y = nullptr, *z = nullptr, *size = nullptr;
  const float *young = nullptr;
  const int32_t *ctype = nullptr, *phase = nullptr;
  double *own_x = nullptr, *own_y = nullptr, *own_z = nullptr, *own_size = nullptr;
  float *own_young = nullptr;
  int32_t *own_ctype = nullptr, *own_phase = nullptr;
  int64_t own_cap = 0;
  int64_t *d_col_off = nullptr;
  int64_t cap_col = 0;
  // bounding-box work split: chunks of at most TM_BBOX_CHUNK cells of one colony
  int64_t nchunks = 0, cap_chunks = 0;
  int64_t *d_chunk_start = nullptr;  // nchunks + 1
  int32_t *d_chunk_cid = nullptr;
  unsigned long long *d_cbox = nullptr;  // per chunk 6 ordered keys: its cells' bounding box
  // inputs packed in reference order by the key kernel
  double4 *d_pos_o = nullptr;  // x, y, z, size
  uint4 *d_aux_o = nullptr;    // young bits, ctype, phase, colony

  // per colony
  double4 *d_aff = nullptr;              // shift.xyz, scale
  int64_t *d_col_chunk = nullptr;  // per colony its first bounding-box chunk (ncol + 1)
  int32_t *d_colroot = nullptr;

  // sort scratch
  uint64_t *d_key = nullptr, *d_key_alt = nullptr;
  int32_t *d_val_alt = nullptr;
  uint32_t *d_hist = nullptr, *d_scan = nullptr;
  int64_t cap_hist = 0, cap_scan = 0;

  // Morton order
  int32_t *d_perm = nullptr;
  uint4 *d_code = nullptr;  // x, y, z codes, colony id
  double4 *d_pos = nullptr; // x, y, z, size
  float4 *d_posf = nullptr;  // x, y, z relative to the colony's origin, rounded to FP32 (the walk's candidate reads)
  float *d_young_s = nullptr;
  int32_t *d_ctype_s = nullptr, *d_phase_s = nullptr;
  int32_t *d_span = nullptr;

  // radix tree
  // the walk's tree: every colony's nodes in its descending pre-order (see
  // tm_build.hip), and the counts their slots are computed from
  uint4 *d_wn = nullptr;
  uint32_t *d_wchain = nullptr;  // per cell: the colony's internal nodes ending there, then their exclusive sums
  int2 *d_range = nullptr;
  int32_t *d_delta = nullptr;
  int32_t *d_split = nullptr;   // split position of each internal node
  int32_t *d_parent = nullptr;  // parent of every node (2n-1), -1 at roots
  // the emitted octree (tm_get_octree): chain counts and the node records,
  // kept between calls
  uint32_t *d_octcnt = nullptr;
  int64_t cap_octcnt = 0;
  char *d_octout = nullptr;
  int64_t cap_octout = 0;  // bytes
  int64_t cap_tree_aux = 0;

  // neighbours: slot k of cell p at nbr[tm_slot(p, k, nbr_alloc)]
  uint32_t *d_nbr = nullptr;
  int64_t nbr_alloc = 0;  // slots per cell of the current layout
  int64_t nbr_elems = 0;  // elements allocated (>= n * nbr_alloc)
  uint32_t *d_ncount = nullptr;
  int32_t *d_qstar = nullptr;

  // LDS tiles built from the lists (see TM_TILE)
  int sph_tiles = 1;
  int tile_cap4 = 1;  // cap the staged records at TM_TILE_LDS4 when few tiles are larger
  int tile_window = 1;  // k_tiles orders the neighbours near a tile's middle by bitmap (0: all by hash set)
  int64_t tiles_over = 0;  // over-full tiles at the last checked tile build
  uint32_t *d_tile_idx = nullptr;   // per tile: distinct neighbours, ascending (TM_TILE_MAX slots)
  uint32_t *d_tile_cnt = nullptr;   // per tile: their number, or TM_TILE_OVER
  uint16_t *d_tile_cell = nullptr;  // per tile: cell offset of each rank (longest list first)
  uint16_t *d_nbr16 = nullptr;      // slot k of rank r of tile g at tm_nbr16_at(g*TS + r, k, nbr_alloc)
  int tile_width = TM_TILE;         // cells per tile for the next tile build: 128, 256 or 512
  int walk_coop = 0;                // 1: one union walk per wave (see k_neighbors)
  int walk_coop_ratio = 64;         // union box volume / mean cell box volume above which a wave walks per lane
  int tiles_ts = 0;                 // cells per tile of the tiles as built (0: none)
  int tile_lds_ts = 0;              // the tile width tile_lds was sized for
  int64_t cap_tile_idx = 0, cap_tile_cnt = 0, cap_tile_cell = 0, cap_nbr16 = 0;

  // environment solve: right-hand side, residual, two directions, A p, partial sums, state
  double *d_cg = nullptr;
  int64_t cap_cg = 0;
  uint32_t tile_lds = 0;            // neighbour records the pair kernel stages per tile (0: no tiles built)

  // outputs, Morton order
  double *d_dens = nullptr, *d_v = nullptr, *d_mind = nullptr;
  double *d_fx = nullptr, *d_fy = nullptr, *d_fz = nullptr;
  // unpermuted outputs
  double *d_out = nullptr;

  // flags: [0] error bits, [1] max neighbour count, [2] pairs low, [3] pairs high,
  // [4] largest tile, [5] tiles above TM_TILE_LDS4 records, [6..7] cells newly
  // marked by the last eager move (u64),
  // [8..9] the same summed over captured-loop replays, [10..11] the captured
  // loop's next Philox step, [12..13] error bits of the moves, [14] over-full
  // tiles
  unsigned int *d_flags = nullptr;
  unsigned int *h_flags = nullptr;  // pinned, TM_HFLAGS_WORDS words: copies of d_flags, then the solver state

  int stage = 0;  // 0 empty, 1 cells, 2 built, 3 neighbours, 4 sph
  bool moved = false;  // a captured loop moved the cells after its sums (cleared by the next build)
  bool sums_valid = false;  // d_fx.. hold sums in the current d_perm order (set by the sums, cleared by a build)

  // VTK body (40 bytes per cell)
  uint32_t *d_vtk = nullptr;
  int64_t cap_vtk = 0;
  tm_precision last_prec = TM_FP64;

  // deposit by grid tile: each cell's (tile, layer) key and rank; the cells'
  // indices in key order; per key a count and an offset, the tiles' launch
  // slots.  The counts
  // are zero after a deposit: dep_zero words at the start of d_drange are
  // known to be zero.
  uint64_t *d_dkr = nullptr;  // key | rank within the key << 32
  uint32_t *d_didx = nullptr;  // the keyed cells in key order
  uint32_t *d_drange = nullptr;
  int64_t cap_dkr = 0, cap_didx = 0, cap_drange = 0;
  int64_t dep_zero = 0;

  // last deposit, for timing
  tm_grid last_grid{};
  int last_weight = 0;
  double *last_prod = nullptr;

  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_kernel_ms = 0.0;

  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  tm_precision graph_prec = TM_FP64;
  bool graph_loop = false;     // the captured step also moves the cells
  tm_move graph_move{};        // its parameters
  uint8_t *graph_removecell = nullptr;
};

// error helpers
tm_status tm_fail(tm_ctx *c, tm_status s, const std::string &msg);
#define TM_HIP(c, expr)                                                                  \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      return tm_fail((c), TM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)
#define TM_TRY(expr)                 \
  do {                               \
    tm_status _s = (expr);           \
    if (_s != TM_OK) return _s;      \
  } while (0)

template <typename T>
tm_status tm_ensure(tm_ctx *c, T **p, int64_t &cap, int64_t need, const char *what) {
  if (need <= cap && *p) return TM_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  cap = 0;
  if (need <= 0) return TM_OK;
  if (hipMalloc((void **)p, sizeof(T) * (size_t)need) != hipSuccess) {
    *p = nullptr;
    return tm_fail(c, TM_ERR_NOMEM, std::string("hipMalloc failed: ") + what);
  }
  cap = need;
  return TM_OK;
}

// stage entry points implemented per file
tm_status tm_sort_pairs(tm_ctx *c, uint64_t *keys, int32_t *vals, int64_t n, int lo_bit);
tm_status tm_scan_u32(tm_ctx *c, const uint32_t *in, uint32_t *out, int64_t n);
tm_status tm_run_build(tm_ctx *c);
tm_status tm_run_neighbors(tm_ctx *c, bool check);
tm_status tm_run_tiles(tm_ctx *c, bool check);
tm_status tm_run_sph(tm_ctx *c, tm_precision prec);
tm_status tm_run_unpermute(tm_ctx *c);
tm_status tm_run_deposit(tm_ctx *c, const tm_grid *g, int weight_mode, double *prod);
tm_status tm_run_octree(tm_ctx *c, int colony, void *nodes, int64_t *count);
// moves the cells by the forces of the last tm_run_sph; graph: Philox step from
// the device word, advanced afterwards, count summed into the loop's word, and
// nothing moves when the neighbour lists of this step overflowed
tm_status tm_run_move(tm_ctx *c, const tm_move *m, const float *draws, uint8_t *removecell, bool graph);
tm_status tm_set_graph_step(tm_ctx *c, uint64_t step);
tm_status tm_run_diffusion(tm_ctx *c, const tm_grid *g, const tm_diffusion *d, const double *src, double *field,
                           int32_t *iterations, double *residual);
tm_status tm_take_graph_removed(tm_ctx *c, int64_t *count);
// looks up the device address of the allocation-free solve's static work
// buffer on c->device (current); once per device, at tm_create
tm_status tm_noalloc_init(tm_ctx *c);
// the environment solve with its work arrays in that buffer (no allocation)
tm_status tm_run_diffusion_noalloc(tm_ctx *c, const tm_grid *g, const tm_diffusion *d, const double *src,
                                   double *field, int32_t *iterations, double *residual);
inline unsigned long long *tm_removed_word(tm_ctx *c) { return (unsigned long long *)(c->d_flags + 6); }
inline unsigned long long *tm_graph_removed_word(tm_ctx *c) { return (unsigned long long *)(c->d_flags + 8); }
inline unsigned long long *tm_graph_step_word(tm_ctx *c) { return (unsigned long long *)(c->d_flags + 10); }
inline unsigned long long *tm_move_error_word(tm_ctx *c) { return (unsigned long long *)(c->d_flags + 12); }
tm_status tm_run_move_arrays(tm_ctx *c, int64_t n, int dim, double *x, double *y, double *z, const double *size,
                             const int32_t *ctype, const double *fx, const double *fy, const double *fz,
                             const tm_move *m, const float *draws, uint8_t *removecell);

// small device helpers ------------------------------------------------------
__device__ __forceinline__ unsigned long long tm_lanemask_lt() {
  unsigned lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ int tm_msb32(uint32_t v) { return 31 - __clz((int)v); }

// highest differing bit of two code triples (-1 if equal)
__device__ __forceinline__ int tm_diffbit(uint4 a, uint4 b) {
  uint32_t o = (a.x ^ b.x) | (a.y ^ b.y) | (a.z ^ b.z);
  return o ? tm_msb32(o) : -1;
}

inline unsigned tm_blocks(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// Neighbour-list layout: slot k of cell p.  The slots of each group of 64
// consecutive cells are stored together, slot-major inside the group, so one
// wave reads slot k of its 64 cells as one contiguous 256-byte line pair and
// a group's whole list lies in one contiguous stretch of memory.
__host__ __device__ __forceinline__ int64_t tm_slot(int64_t p, int64_t k, int64_t cap) {
  return (((p >> 6) * cap + k) << 6) | (p & 63);
}
inline int64_t tm_slot_elems(int64_t n, int64_t cap) { return ((n + 63) >> 6) * cap * 64; }
// 16-bit tile lists: the slots of each rank in chunks of 8 (16 bytes), the
// chunks of a group of 64 ranks side by side (cap a multiple of 8)
__host__ __device__ __forceinline__ int64_t tm_nbr16_at(int64_t r, int64_t k, int64_t cap) {
  return (((r >> 6) * cap + (k & ~7ll)) << 6) + ((r & 63) << 3) + (k & 7);
}

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8), each
// with its own L2.  Cell-ordered kernels remap so that XCD x runs one
// contiguous slice of the blocks, i.e. one compact stretch of the Morton
// order, and its L2 holds that region's cells and tree nodes.  Bijective on
// [0, gridDim.x).
__device__ __forceinline__ int64_t tm_xcd_block() {
  const int64_t nb = gridDim.x, b = blockIdx.x;
  const int64_t x = b & 7, j = b >> 3, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + j;
}

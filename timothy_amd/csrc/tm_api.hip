// C ABI of the cell-mechanics library: context, memory, stage calls, results,
// measurement, graph capture and autotuning.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "tm_internal.h"

tm_status tm_fail(tm_ctx *c, tm_status s, const std::string &msg) {
  if (c) c->err = msg;
  return s;
}

namespace {

// largest double d2 with RN(sqrt(d2)) <= h: "dist <= h" as a test on d2
double cutoff_sq(double h) {
  double t = h * h;
  while (std::sqrt(t) > h) t = std::nextafter(t, 0.0);
  while (std::sqrt(std::nextafter(t, INFINITY)) <= h) t = std::nextafter(t, INFINITY);
  return t;
}

template <typename T>
void free_ptr(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

tm_status alloc_cells(tm_ctx *c, int64_t n) {
  if (n <= c->cap_n) return TM_OK;
  int64_t cap = n + (n >> 4) + 64;
  free_ptr(c->d_key);
  free_ptr(c->d_key_alt);
  free_ptr(c->d_val_alt);
  free_ptr(c->d_perm);
  free_ptr(c->d_code);
  free_ptr(c->d_pos);
  free_ptr(c->d_posf);
  free_ptr(c->d_young_s);
  free_ptr(c->d_ctype_s);
  free_ptr(c->d_phase_s);
  free_ptr(c->d_span);
  free_ptr(c->d_wn);
  free_ptr(c->d_wchain);
  free_ptr(c->d_range);
  free_ptr(c->d_delta);
  free_ptr(c->d_ncount);
  free_ptr(c->d_qstar);
  free_ptr(c->d_dens);
  free_ptr(c->d_v);
  free_ptr(c->d_mind);
  free_ptr(c->d_fx);
  free_ptr(c->d_fy);
  free_ptr(c->d_fz);
  free_ptr(c->d_out);
  free_ptr(c->d_split);
  free_ptr(c->d_parent);
  free_ptr(c->d_pos_o);
  free_ptr(c->d_aux_o);
  free_ptr(c->d_nbr);
  c->nbr_alloc = 0;
  c->nbr_elems = 0;
  c->cap_n = 0;
  struct A {
    void **p;
    size_t bytes;
  } list[] = {
      {(void **)&c->d_key, 8},      {(void **)&c->d_key_alt, 8},  {(void **)&c->d_val_alt, 4},
      {(void **)&c->d_perm, 4},     {(void **)&c->d_code, 16},    {(void **)&c->d_pos, 32},    {(void **)&c->d_posf, 16},
      {(void **)&c->d_young_s, 4},  {(void **)&c->d_ctype_s, 4},
      {(void **)&c->d_phase_s, 4},  {(void **)&c->d_span, 4},
      {(void **)&c->d_wn, 32},      {(void **)&c->d_wchain, 4},     {(void **)&c->d_range, 8},    {(void **)&c->d_delta, 4},    {(void **)&c->d_ncount, 4},
      {(void **)&c->d_qstar, 4},    {(void **)&c->d_dens, 8},     {(void **)&c->d_v, 8},
      {(void **)&c->d_mind, 8},     {(void **)&c->d_fx, 8},       {(void **)&c->d_fy, 8},
      {(void **)&c->d_fz, 8},       {(void **)&c->d_out, 48},     {(void **)&c->d_split, 4},
      {(void **)&c->d_parent, 8},   {(void **)&c->d_pos_o, 32},   {(void **)&c->d_aux_o, 16},
  };
  for (auto &a : list) {
    if (hipMalloc(a.p, a.bytes * (size_t)cap) != hipSuccess) {
      *a.p = nullptr;
      return tm_fail(c, TM_ERR_NOMEM, "hipMalloc failed for per-cell arrays");
    }
  }
  c->cap_n = cap;
  return TM_OK;
}

tm_status alloc_colonies(tm_ctx *c, int ncol) {
  if (ncol <= c->cap_col) return TM_OK;
  free_ptr(c->d_col_off);
  free_ptr(c->d_aff);
  free_ptr(c->d_col_chunk);
  free_ptr(c->d_colroot);
  c->cap_col = 0;
  if (hipMalloc((void **)&c->d_col_off, sizeof(int64_t) * (ncol + 1)) != hipSuccess ||
      hipMalloc((void **)&c->d_aff, sizeof(double4) * ncol) != hipSuccess ||
      hipMalloc((void **)&c->d_col_chunk, sizeof(int64_t) * (ncol + 1)) != hipSuccess ||
      hipMalloc((void **)&c->d_colroot, sizeof(int32_t) * ncol) != hipSuccess)
    return tm_fail(c, TM_ERR_NOMEM, "hipMalloc failed for colony arrays");
  c->cap_col = ncol;
  return TM_OK;
}

tm_status set_cells(tm_ctx *c, int64_t n, int32_t ncol, const int64_t *off, const double *x, const double *y,
                    const double *z, const double *size, const float *young, const int32_t *ctype,
                    const int32_t *phase, bool device) {
  if (!c) return TM_ERR_INVALID;
  if (n < 0 || ncol < 1) return tm_fail(c, TM_ERR_INVALID, "ncells < 0 or ncolonies < 1");
  if (n >= (int64_t)0x7fffffff) return tm_fail(c, TM_ERR_INVALID, "at most 2^31-1 cells per context");
  if (ncol > (1 << TM_CID_BITS)) return tm_fail(c, TM_ERR_INVALID, "at most 65536 colonies");
  if (c->ntypes == 0) return tm_fail(c, TM_ERR_STATE, "tm_set_celltypes first");
  if (n > 0 && (!x || !y || !z || !size || !ctype)) return tm_fail(c, TM_ERR_INVALID, "NULL cell array");
  std::vector<int64_t> o(ncol + 1);
  if (ncol == 1) {
    o[0] = 0;
    o[1] = n;
  } else {
    if (!off) return tm_fail(c, TM_ERR_INVALID, "colony_offsets required for ncolonies > 1");
    for (int i = 0; i <= ncol; ++i) o[i] = off[i];
    if (o[0] != 0 || o[ncol] != n) return tm_fail(c, TM_ERR_INVALID, "colony_offsets must run from 0 to ncells");
    for (int i = 0; i < ncol; ++i)
      if (o[i + 1] < o[i]) return tm_fail(c, TM_ERR_INVALID, "colony_offsets must be non-decreasing");
  }
  if (!device && n > 0) {
    for (int64_t i = 0; i < n; ++i)
      if (ctype[i] < 0 || ctype[i] >= c->ntypes) return tm_fail(c, TM_ERR_INVALID, "ctype out of range");
  }
  TM_TRY(alloc_cells(c, std::max<int64_t>(n, 1)));
  TM_TRY(alloc_colonies(c, ncol));
  c->n = n;
  c->ncol = ncol;
  c->col_off = o;
  TM_HIP(c, hipMemcpyAsync(c->d_col_off, o.data(), sizeof(int64_t) * (ncol + 1), hipMemcpyHostToDevice, c->stream));
  // bounding-box chunks: each covers at most TM_BBOX_CHUNK cells of one colony
  std::vector<int64_t> cs, cfirst;
  std::vector<int32_t> cc;
  for (int k = 0; k < ncol; ++k) {
    cfirst.push_back((int64_t)cc.size());
    for (int64_t s = o[k]; s < o[k + 1]; s += TM_BBOX_CHUNK) {
      cs.push_back(s);
      cc.push_back(k);
    }
  }
  cfirst.push_back((int64_t)cc.size());
  cs.push_back(n);
  c->nchunks = (int64_t)cc.size();
  if (c->nchunks + 1 > c->cap_chunks) {
    free_ptr(c->d_chunk_start);
    free_ptr(c->d_chunk_cid);
    free_ptr(c->d_cbox);
    c->cap_chunks = 0;
    if (hipMalloc((void **)&c->d_chunk_start, sizeof(int64_t) * (c->nchunks + 1)) != hipSuccess ||
        hipMalloc((void **)&c->d_chunk_cid, sizeof(int32_t) * (c->nchunks + 1)) != hipSuccess ||
        hipMalloc((void **)&c->d_cbox, sizeof(unsigned long long) * 6 * (c->nchunks + 1)) != hipSuccess)
      return tm_fail(c, TM_ERR_NOMEM, "hipMalloc failed for bounding-box chunks");
    c->cap_chunks = c->nchunks + 1;
  }
  TM_HIP(c, hipMemcpyAsync(c->d_col_chunk, cfirst.data(), sizeof(int64_t) * cfirst.size(), hipMemcpyHostToDevice,
                           c->stream));
  TM_HIP(c, hipMemcpyAsync(c->d_chunk_start, cs.data(), sizeof(int64_t) * cs.size(), hipMemcpyHostToDevice,
                           c->stream));
  if (c->nchunks > 0)
    TM_HIP(c, hipMemcpyAsync(c->d_chunk_cid, cc.data(), sizeof(int32_t) * cc.size(), hipMemcpyHostToDevice,
                             c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  if (device) {
    c->external = true;
    c->x = x;
    c->y = y;
    c->z = z;
    c->size = size;
    c->young = young;
    c->ctype = ctype;
    c->phase = phase;
  } else {
    c->external = false;
    // owned copies sized to the per-cell capacity
    auto realloc_own = [&](void **p, size_t elt) -> tm_status {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
      if (hipMalloc(p, elt * (size_t)c->cap_n) != hipSuccess) {
        *p = nullptr;
        return tm_fail(c, TM_ERR_NOMEM, "hipMalloc failed for cell inputs");
      }
      return TM_OK;
    };
    if (!c->own_x || c->own_cap < c->cap_n) {
      TM_TRY(realloc_own((void **)&c->own_x, 8));
      TM_TRY(realloc_own((void **)&c->own_y, 8));
      TM_TRY(realloc_own((void **)&c->own_z, 8));
      TM_TRY(realloc_own((void **)&c->own_size, 8));
      TM_TRY(realloc_own((void **)&c->own_young, 4));
      TM_TRY(realloc_own((void **)&c->own_ctype, 4));
      TM_TRY(realloc_own((void **)&c->own_phase, 4));
      c->own_cap = c->cap_n;
    }
    if (n > 0) {
      TM_HIP(c, hipMemcpyAsync(c->own_x, x, 8 * n, hipMemcpyHostToDevice, c->stream));
      TM_HIP(c, hipMemcpyAsync(c->own_y, y, 8 * n, hipMemcpyHostToDevice, c->stream));
      TM_HIP(c, hipMemcpyAsync(c->own_z, z, 8 * n, hipMemcpyHostToDevice, c->stream));
      TM_HIP(c, hipMemcpyAsync(c->own_size, size, 8 * n, hipMemcpyHostToDevice, c->stream));
      TM_HIP(c, hipMemcpyAsync(c->own_ctype, ctype, 4 * n, hipMemcpyHostToDevice, c->stream));
      if (young) TM_HIP(c, hipMemcpyAsync(c->own_young, young, 4 * n, hipMemcpyHostToDevice, c->stream));
      else TM_HIP(c, hipMemsetAsync(c->own_young, 0, 4 * n, c->stream));
      if (phase) TM_HIP(c, hipMemcpyAsync(c->own_phase, phase, 4 * n, hipMemcpyHostToDevice, c->stream));
    }
    c->x = c->own_x;
    c->y = c->own_y;
    c->z = c->own_z;
    c->size = c->own_size;
    c->young = c->own_young;
    c->ctype = c->own_ctype;
    c->phase = phase ? c->own_phase : nullptr;
    TM_HIP(c, hipStreamSynchronize(c->stream));  // host arrays may be freed after return
  }
  c->stage = 1;
  return TM_OK;
}

tm_status check_flags(tm_ctx *c) {
  TM_HIP(c, hipMemcpyAsync(c->h_flags, c->d_flags, 4 * sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  if (c->h_flags[0] & TM_FLAG_DEGENERATE)
    return tm_fail(c, TM_ERR_DEGENERATE,
                   "two cells share location codes down to bit 1, or more than 64 lie within 2^-16 of their colony's "
                     "extent on every axis (duplicate or near-duplicate positions)");
  if (c->h_flags[0] & TM_FLAG_ROOT) return tm_fail(c, TM_ERR_STATE, "colony root not found in the radix tree");
  if (c->h_flags[0] & TM_FLAG_CTYPE) return tm_fail(c, TM_ERR_INVALID, "ctype out of range");
  if (c->h_flags[0] & TM_FLAG_WALK) return tm_fail(c, TM_ERR_STATE, "malformed tree: walk left its node range");
  return TM_OK;
}

tm_status run_stage(tm_ctx *c, tm_stage s, tm_precision p) {
  switch (s) {
    case TM_STAGE_BUILD: return tm_run_build(c);
    case TM_STAGE_NEIGHBORS: return tm_run_neighbors(c, false);
    case TM_STAGE_SPH: return tm_run_sph(c, p);
    case TM_STAGE_CIC:
      if (!c->last_prod) return tm_fail(c, TM_ERR_STATE, "no tm_cic_deposit call to repeat");
      return tm_run_deposit(c, &c->last_grid, c->last_weight, c->last_prod);
    case TM_STAGE_STEP:
      TM_TRY(tm_run_build(c));
      TM_TRY(tm_run_neighbors(c, false));
      return tm_run_sph(c, p);
    case TM_STAGE_TILES:
      return tm_run_tiles(c, false);
  }
  return tm_fail(c, TM_ERR_INVALID, "unknown stage");
}

}  // namespace

extern "C" {

const char *tm_status_name(tm_status s) {
  switch (s) {
    case TM_OK: return "TM_OK";
    case TM_ERR_INVALID: return "TM_ERR_INVALID";
    case TM_ERR_HIP: return "TM_ERR_HIP";
    case TM_ERR_NOMEM: return "TM_ERR_NOMEM";
    case TM_ERR_DEGENERATE: return "TM_ERR_DEGENERATE";
    case TM_ERR_STATE: return "TM_ERR_STATE";
    case TM_ERR_UNSUPPORTED: return "TM_ERR_UNSUPPORTED";
  }
  return "TM_ERR_UNKNOWN";
}

tm_status tm_create(tm_ctx **out, int device) {
  if (!out) return TM_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TM_ERR_HIP;
  if (device < 0 || device >= ndev) return TM_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return TM_ERR_HIP;
  tm_ctx *c = new tm_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void **)&c->d_types, sizeof(TypeDev) * TM_MAX_TYPES) != hipSuccess ||
      hipMalloc((void **)&c->d_flags, 16 * sizeof(unsigned int)) != hipSuccess ||
      hipHostMalloc((void **)&c->h_flags, TM_HFLAGS_WORDS * sizeof(unsigned int), hipHostMallocDefault) !=
          hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    tm_destroy(c);
    return TM_ERR_HIP;
  }
  if (tm_noalloc_init(c) != TM_OK) {
    tm_destroy(c);
    return TM_ERR_HIP;
  }
  (void)hipMemset(c->d_flags, 0, 16 * sizeof(unsigned int));
  *out = c;
  return TM_OK;
}

void tm_destroy(tm_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  tm_graph_destroy(c);
  free_ptr(c->d_key);
  free_ptr(c->d_key_alt);
  free_ptr(c->d_val_alt);
  free_ptr(c->d_perm);
  free_ptr(c->d_code);
  free_ptr(c->d_pos);
  free_ptr(c->d_posf);
  free_ptr(c->d_young_s);
  free_ptr(c->d_ctype_s);
  free_ptr(c->d_phase_s);
  free_ptr(c->d_span);
  free_ptr(c->d_wn);
  free_ptr(c->d_wchain);
  free_ptr(c->d_range);
  free_ptr(c->d_delta);
  free_ptr(c->d_ncount);
  free_ptr(c->d_qstar);
  free_ptr(c->d_dens);
  free_ptr(c->d_v);
  free_ptr(c->d_mind);
  free_ptr(c->d_fx);
  free_ptr(c->d_fy);
  free_ptr(c->d_fz);
  free_ptr(c->d_out);
  free_ptr(c->d_nbr);
  free_ptr(c->d_hist);
  free_ptr(c->d_scan);
  free_ptr(c->d_col_off);
  free_ptr(c->d_aff);
  free_ptr(c->d_col_chunk);
  free_ptr(c->d_colroot);
  free_ptr(c->own_x);
  free_ptr(c->own_y);
  free_ptr(c->own_z);
  free_ptr(c->own_size);
  free_ptr(c->own_young);
  free_ptr(c->own_ctype);
  free_ptr(c->own_phase);
  free_ptr(c->d_types);
  free_ptr(c->d_flags);
  free_ptr(c->d_split);
  free_ptr(c->d_parent);
  free_ptr(c->d_pos_o);
  free_ptr(c->d_aux_o);
  free_ptr(c->d_chunk_start);
  free_ptr(c->d_chunk_cid);
  free_ptr(c->d_cbox);
  free_ptr(c->d_octcnt);
  free_ptr(c->d_octout);
  free_ptr(c->d_tile_idx);
  free_ptr(c->d_tile_cnt);
  free_ptr(c->d_tile_cell);
  free_ptr(c->d_nbr16);
  free_ptr(c->d_cg);
  free_ptr(c->d_dkr);
  free_ptr(c->d_didx);
  free_ptr(c->d_drange);
  free_ptr(c->d_vtk);
  if (c->h_flags) (void)hipHostFree(c->h_flags);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *tm_last_error(const tm_ctx *c) { return c ? c->err.c_str() : "no context"; }

void *tm_stream(tm_ctx *c) { return c ? (void *)c->stream : nullptr; }

tm_status tm_sync(tm_ctx *c) {
  if (!c) return TM_ERR_INVALID;
  TM_HIP(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

tm_status tm_set_celltypes(tm_ctx *c, const tm_celltype *types, int32_t ntypes, int32_t dimension) {
  if (!c) return TM_ERR_INVALID;
  if (!types || ntypes < 1 || ntypes > TM_MAX_TYPES)
    return tm_fail(c, TM_ERR_INVALID, "1 <= ntypes <= 16 cell types required");
  if (dimension != 2 && dimension != 3) return tm_fail(c, TM_ERR_INVALID, "dimension must be 2 or 3");
  TypeDev td[TM_MAX_TYPES];
  memset(td, 0, sizeof(td));
  for (int i = 0; i < ntypes; ++i) {
    const tm_celltype &t = types[i];
    if (!(t.h > 0.0) || !(t.size > 0.0)) return tm_fail(c, TM_ERR_INVALID, "cell type h and size must be > 0");
    c->types[i] = t;
    TypeDev &d = td[i];
    d.size = t.size;
    d.h = t.h;
    d.h2 = t.h2;
    d.h3 = t.h3;
    d.h4 = t.h4;
    // src/inline.h:45-46 (kernel), :71-72 (gradient), src/potential.c:58 (scale)
    if (dimension == 3) {
      d.sc = t.h3;
      d.cW = 8 / (M_PI * t.h3);
      d.cF = 48 / (M_PI * t.h4);
    } else {
      d.sc = t.h2;
      d.cW = 40 / (7 * M_PI * t.h2);
      d.cF = (40 * 8) / (7 * M_PI * t.h3);
    }
    d.cut2 = cutoff_sq(t.h);
    d.hinv = 1.0 / t.h;  // correctly rounded reciprocals: the divisor of u = d/h and of r_q/size
    d.sinv = 1.0 / t.size;
    d.hinv_f = (float)(1.0 / t.h);
  }
  c->ntypes = ntypes;
  c->dim = dimension;
  TM_HIP(c, hipMemcpyAsync(c->d_types, td, sizeof(td), hipMemcpyHostToDevice, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

tm_status tm_set_param(tm_ctx *c, const char *key, int64_t v) {
  if (!c || !key) return TM_ERR_INVALID;
  // a captured graph holds the buffers of the lists and tiles; a parameter
  // that changes their sizes retires it (tm_graph_launch then reports that
  // there is none until the next capture)
  if (!strcmp(key, "nbr_cap")) {
    if (v < 1 || v > 4096) return tm_fail(c, TM_ERR_INVALID, "nbr_cap in [1, 4096]");
    if (((v + 7) & ~7ll) != c->nbr_cap) tm_graph_destroy(c);
    c->nbr_cap = (v + 7) & ~7ll;  // lists are read in chunks of 8 slots
  } else if (!strcmp(key, "sph_block")) {
    if (v != 64 && v != 128 && v != 256) return tm_fail(c, TM_ERR_INVALID, "sph_block in {64,128,256}");
    c->sph_block = (int)v;
  } else if (!strcmp(key, "sph_unroll_fp64") || !strcmp(key, "sph_unroll_mixed")) {
    if (v != 1 && v != 2 && v != 4 && v != 8) return tm_fail(c, TM_ERR_INVALID, "sph_unroll in {1,2,4,8}");
    if (key[11] == 'f') c->sph_unroll_fp64 = (int)v;
    else c->sph_unroll_mixed = (int)v;
  } else if (!strcmp(key, "nbr_block")) {
    if (v != 64 && v != 128 && v != 256) return tm_fail(c, TM_ERR_INVALID, "nbr_block in {64,128,256}");
    c->nbr_block = (int)v;
  } else if (!strcmp(key, "sph_lanes")) {
    if (v != 1 && v != 2 && v != 4 && v != 8) return tm_fail(c, TM_ERR_INVALID, "sph_lanes in {1,2,4,8}");
    c->sph_lanes = (int)v;
  } else if (!strcmp(key, "sph_balance")) {
    c->sph_balance = v ? 1 : 0;
  } else if (!strcmp(key, "sph_tiles")) {
    c->sph_tiles = v ? 1 : 0;
  } else if (!strcmp(key, "tile_cap4")) {
    c->tile_cap4 = v ? 1 : 0;
  } else if (!strcmp(key, "walk_f32")) {
    c->walk_f32 = v ? 1 : 0;
  } else if (!strcmp(key, "tile_window")) {
    c->tile_window = v ? 1 : 0;
  } else if (!strcmp(key, "walk_coop")) {
    c->walk_coop = v ? 1 : 0;
  } else if (!strcmp(key, "walk_coop_ratio")) {
    if (v < 1 || v > 1000000) return tm_fail(c, TM_ERR_INVALID, "walk_coop_ratio in [1, 1e6]");
    c->walk_coop_ratio = (int)v;
  } else if (!strcmp(key, "tile_width")) {
    if (v != 128 && v != 256 && v != 512) return tm_fail(c, TM_ERR_INVALID, "tile_width in {128,256,512}");
    if (v != c->tile_width) tm_graph_destroy(c);
    c->tile_width = (int)v;
  } else if (!strcmp(key, "walk_window")) {
    if (v != 0 && v != 512 && v != 1024 && v != 1280 && v != 2048)
      return tm_fail(c, TM_ERR_INVALID, "walk_window in {0,512,1024,1280,2048}");
    c->walk_window = (int)v;
  } else {
    return tm_fail(c, TM_ERR_INVALID, std::string("unknown parameter ") + key);
  }
  return TM_OK;
}

tm_status tm_get_param(tm_ctx *c, const char *key, int64_t *v) {
  if (!c || !key || !v) return TM_ERR_INVALID;
  if (!strcmp(key, "nbr_cap")) *v = c->nbr_cap;
  else if (!strcmp(key, "sph_block")) *v = c->sph_block;
  else if (!strcmp(key, "sph_unroll_fp64")) *v = c->sph_unroll_fp64;
  else if (!strcmp(key, "sph_unroll_mixed")) *v = c->sph_unroll_mixed;
  else if (!strcmp(key, "nbr_block")) *v = c->nbr_block;
  else if (!strcmp(key, "sph_lanes")) *v = c->sph_lanes;
  else if (!strcmp(key, "sph_balance")) *v = c->sph_balance;
  else if (!strcmp(key, "sph_tiles")) *v = c->sph_tiles;
  else if (!strcmp(key, "tile_cap4")) *v = c->tile_cap4;
  else if (!strcmp(key, "walk_window")) *v = c->walk_window;
  else if (!strcmp(key, "walk_f32")) *v = c->walk_f32;
  else if (!strcmp(key, "tile_window")) *v = c->tile_window;
  else if (!strcmp(key, "tile_width")) *v = c->tile_width;
  else if (!strcmp(key, "walk_coop")) *v = c->walk_coop;
  else if (!strcmp(key, "walk_coop_ratio")) *v = c->walk_coop_ratio;
  else if (!strcmp(key, "tile_lds")) *v = c->tile_lds;
  else if (!strcmp(key, "tiles_over")) *v = c->tiles_over;
  else if (!strcmp(key, "nbr_alloc")) *v = c->nbr_alloc;
  else return tm_fail(c, TM_ERR_INVALID, std::string("unknown parameter ") + key);
  return TM_OK;
}

tm_status tm_upload(tm_ctx *c, int64_t n, int32_t ncol, const int64_t *off, const double *x, const double *y,
                    const double *z, const double *size, const float *young, const int32_t *ctype,
                    const int32_t *phase) {
  return set_cells(c, n, ncol, off, x, y, z, size, young, ctype, phase, false);
}

tm_status tm_bind_device(tm_ctx *c, int64_t n, int32_t ncol, const int64_t *off, const double *x, const double *y,
                         const double *z, const double *size, const float *young, const int32_t *ctype,
                         const int32_t *phase) {
  return set_cells(c, n, ncol, off, x, y, z, size, young, ctype, phase, true);
}

tm_status tm_build(tm_ctx *c) {
  if (!c) return TM_ERR_INVALID;
  TM_TRY(tm_run_build(c));
  return check_flags(c);
}

tm_status tm_neighbors(tm_ctx *c) {
  if (!c) return TM_ERR_INVALID;
  return tm_run_neighbors(c, true);
}

tm_status tm_sph(tm_ctx *c, tm_precision p) {
  if (!c) return TM_ERR_INVALID;
  if (p != TM_FP64 && p != TM_MIXED) return tm_fail(c, TM_ERR_INVALID, "precision must be TM_FP64 or TM_MIXED");
  return tm_run_sph(c, p);
}

tm_status tm_step(tm_ctx *c, tm_precision p) {
  if (!c) return TM_ERR_INVALID;
  TM_TRY(tm_run_build(c));
  TM_TRY(tm_run_neighbors(c, true));
  return tm_sph(c, p);
}

tm_status tm_get_results(tm_ctx *c, double *density, double *v, double *mindist, double *fx, double *fy, double *fz) {
  if (!c) return TM_ERR_INVALID;
  if (c->stage < 4) return tm_fail(c, TM_ERR_STATE, "no results: run tm_sph first");
  const int64_t n = c->n;
  if (n == 0) return TM_OK;
  TM_TRY(tm_run_unpermute(c));
  double *dst[6] = {density, v, mindist, fx, fy, fz};
  for (int j = 0; j < 6; ++j)
    if (dst[j]) TM_HIP(c, hipMemcpyAsync(dst[j], c->d_out + j * n, 8 * n, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

tm_status tm_device_results(tm_ctx *c, const double **density, const double **v, const double **mindist,
                            const double **fx, const double **fy, const double **fz, const int32_t **perm) {
  if (!c) return TM_ERR_INVALID;
  if (c->stage < 4) return tm_fail(c, TM_ERR_STATE, "no results: run tm_sph first");
  if (density) *density = c->d_dens;
  if (v) *v = c->d_v;
  if (mindist) *mindist = c->d_mind;
  if (fx) *fx = c->d_fx;
  if (fy) *fy = c->d_fy;
  if (fz) *fz = c->d_fz;
  if (perm) *perm = c->d_perm;
  return TM_OK;
}

tm_status tm_get_affine(tm_ctx *c, int32_t colony, double shift[3], double *scale) {
  if (!c) return TM_ERR_INVALID;
  if (c->stage < 2) return tm_fail(c, TM_ERR_STATE, "run tm_build first");
  if (colony < 0 || colony >= c->ncol) return tm_fail(c, TM_ERR_INVALID, "colony out of range");
  double4 a;
  TM_HIP(c, hipMemcpyAsync(&a, c->d_aff + colony, sizeof(a), hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  if (shift) {
    shift[0] = a.x;
    shift[1] = a.y;
    shift[2] = a.z;
  }
  if (scale) *scale = a.w;
  return TM_OK;
}

tm_status tm_get_order(tm_ctx *c, int32_t *perm, int32_t *span_log2) {
  if (!c) return TM_ERR_INVALID;
  if (c->stage < 2) return tm_fail(c, TM_ERR_STATE, "run tm_build first");
  const int64_t n = c->n;
  if (n == 0) return TM_OK;
  std::vector<int32_t> p(n), s(n);
  TM_HIP(c, hipMemcpyAsync(p.data(), c->d_perm, 4 * n, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipMemcpyAsync(s.data(), c->d_span, 4 * n, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  if (perm) memcpy(perm, p.data(), 4 * n);
  if (span_log2)
    for (int64_t i = 0; i < n; ++i) span_log2[p[i]] = s[i];
  return TM_OK;
}

tm_status tm_get_codes(tm_ctx *c, uint32_t *loc) {
  if (!c || !loc) return TM_ERR_INVALID;
  if (c->stage < 2) return tm_fail(c, TM_ERR_STATE, "run tm_build first");
  const int64_t n = c->n;
  std::vector<int32_t> p(n);
  std::vector<uint4> code(n);
  TM_HIP(c, hipMemcpyAsync(p.data(), c->d_perm, 4 * n, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipMemcpyAsync(code.data(), c->d_code, 16 * n, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < n; ++i) {
    loc[3 * (int64_t)p[i] + 0] = code[i].x;
    loc[3 * (int64_t)p[i] + 1] = code[i].y;
    loc[3 * (int64_t)p[i] + 2] = code[i].z;
  }
  return TM_OK;
}

tm_status tm_get_neighbors(tm_ctx *c, int64_t *offsets, int32_t *nbr, int64_t *total) {
  if (!c) return TM_ERR_INVALID;
  if (c->stage < 3) return tm_fail(c, TM_ERR_STATE, "run tm_neighbors first");
  const int64_t n = c->n;
  const int64_t cap = c->nbr_alloc;
  std::vector<int32_t> p(n);
  std::vector<uint32_t> cnt(n);
  TM_HIP(c, hipMemcpyAsync(p.data(), c->d_perm, 4 * n, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipMemcpyAsync(cnt.data(), c->d_ncount, 4 * n, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  // offsets in reference order
  std::vector<int64_t> cnt_ref(n + 1, 0);
  int64_t tot = 0;
  for (int64_t i = 0; i < n; ++i) {
    if ((int64_t)cnt[i] > cap) return tm_fail(c, TM_ERR_STATE, "neighbour lists overflowed; rerun tm_neighbors");
    cnt_ref[p[i]] = cnt[i];
  }
  if (offsets) {
    for (int64_t r = 0; r < n; ++r) {
      offsets[r] = tot;
      tot += cnt_ref[r];
    }
    offsets[n] = tot;
  } else {
    for (int64_t r = 0; r < n; ++r) tot += cnt_ref[r];
  }
  if (total) *total = tot;
  if (nbr && offsets && n > 0) {
    const int64_t elems = tm_slot_elems(n, cap);
    std::vector<uint32_t> slots((size_t)elems);
    TM_HIP(c, hipMemcpyAsync(slots.data(), c->d_nbr, 4 * (size_t)elems, hipMemcpyDeviceToHost, c->stream));
    TM_HIP(c, hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; ++i) {
      int64_t o = offsets[p[i]];
      for (uint32_t k = 0; k < cnt[i]; ++k) nbr[o + k] = p[slots[(size_t)tm_slot(i, k, cap)]];
    }
  }
  return TM_OK;
}

tm_status tm_get_qstar(tm_ctx *c, int32_t *qstar) {
  if (!c || !qstar) return TM_ERR_INVALID;
  if (c->stage < 3) return tm_fail(c, TM_ERR_STATE, "run tm_neighbors first");
  const int64_t n = c->n;
  std::vector<int32_t> p(n), q(n);
  TM_HIP(c, hipMemcpyAsync(p.data(), c->d_perm, 4 * n, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipMemcpyAsync(q.data(), c->d_qstar, 4 * n, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < n; ++i) qstar[p[i]] = q[i] >= 0 ? p[q[i]] : -1;
  return TM_OK;
}

tm_status tm_pair_count(tm_ctx *c, int64_t *pairs, int64_t *max_nbr) {
  if (!c) return TM_ERR_INVALID;
  if (c->stage < 3) return tm_fail(c, TM_ERR_STATE, "run tm_neighbors first");
  const int64_t n = c->n;
  std::vector<uint32_t> cnt(n);
  TM_HIP(c, hipMemcpyAsync(cnt.data(), c->d_ncount, 4 * n, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  int64_t s = 0, m = 0;
  for (int64_t i = 0; i < n; ++i) {
    s += cnt[i];
    m = std::max<int64_t>(m, cnt[i]);
  }
  if (pairs) *pairs = s;
  if (max_nbr) *max_nbr = m;
  return TM_OK;
}

tm_status tm_get_octree(tm_ctx *c, int32_t colony, void *nodes, int64_t *count) {
  if (!c || !count) return TM_ERR_INVALID;
  if (c->stage < 2) return tm_fail(c, TM_ERR_STATE, "run tm_build first");
  if (colony < 0 || colony >= c->ncol) return tm_fail(c, TM_ERR_INVALID, "colony out of range");
  return tm_run_octree(c, colony, nodes, count);
}

tm_status tm_malloc(tm_ctx *c, void **ptr, size_t bytes) {
  if (!c || !ptr) return TM_ERR_INVALID;
  TM_HIP(c, hipSetDevice(c->device));
  if (hipMalloc(ptr, bytes) != hipSuccess) {
    *ptr = nullptr;
    return tm_fail(c, TM_ERR_NOMEM, "hipMalloc failed");
  }
  return TM_OK;
}

tm_status tm_free(tm_ctx *c, void *ptr) {
  if (!c) return TM_ERR_INVALID;
  TM_HIP(c, hipStreamSynchronize(c->stream));
  if (ptr) TM_HIP(c, hipFree(ptr));
  return TM_OK;
}

tm_status tm_host_alloc(tm_ctx *c, void **ptr, size_t bytes) {
  if (!c || !ptr) return TM_ERR_INVALID;
  TM_HIP(c, hipSetDevice(c->device));
  if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess) {
    *ptr = nullptr;
    return tm_fail(c, TM_ERR_NOMEM, "hipHostMalloc failed");
  }
  return TM_OK;
}

tm_status tm_host_free(tm_ctx *c, void *ptr) {
  if (!c) return TM_ERR_INVALID;
  TM_HIP(c, hipStreamSynchronize(c->stream));
  if (ptr) TM_HIP(c, hipHostFree(ptr));
  return TM_OK;
}

tm_status tm_memcpy_htod_async(tm_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c) return TM_ERR_INVALID;
  TM_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return TM_OK;
}

tm_status tm_memcpy_dtoh_async(tm_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c) return TM_ERR_INVALID;
  TM_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  return TM_OK;
}

tm_status tm_memcpy_htod(tm_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c) return TM_ERR_INVALID;
  TM_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

tm_status tm_memcpy_dtoh(tm_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c) return TM_ERR_INVALID;
  TM_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

tm_status tm_memcpy_dtod(tm_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c) return TM_ERR_INVALID;
  TM_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

tm_status tm_memset(tm_ctx *c, void *dst, int value, size_t bytes) {
  if (!c) return TM_ERR_INVALID;
  TM_HIP(c, hipMemsetAsync(dst, value, bytes, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

tm_status tm_time_stage(tm_ctx *c, tm_stage s, tm_precision p, int32_t warmup, int32_t iters, double *ms) {
  if (!c || !ms || iters < 1 || warmup < 0) return TM_ERR_INVALID;
  // the neighbour lists must have been sized by one checked tm_neighbors call
  if (s == TM_STAGE_NEIGHBORS && (c->stage < 2 || c->nbr_alloc == 0))
    return tm_fail(c, TM_ERR_STATE, "run tm_neighbors once before timing");
  if ((s == TM_STAGE_SPH || s == TM_STAGE_STEP || s == TM_STAGE_TILES) && c->stage < 3)
    return tm_fail(c, TM_ERR_STATE, "run tm_neighbors once before timing");
  for (int i = 0; i < warmup; ++i) TM_TRY(run_stage(c, s, p));
  TM_HIP(c, hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < iters; ++i) TM_TRY(run_stage(c, s, p));
  TM_HIP(c, hipEventRecord(c->ev1, c->stream));
  TM_HIP(c, hipEventSynchronize(c->ev1));
  float e = 0.0f;
  TM_HIP(c, hipEventElapsedTime(&e, c->ev0, c->ev1));
  *ms = (double)e / iters;
  if (s == TM_STAGE_SPH) c->last_kernel_ms = *ms;
  if (s != TM_STAGE_CIC) TM_TRY(check_flags(c));
  if (c->h_flags[1] > (unsigned)c->nbr_alloc)
    return tm_fail(c, TM_ERR_STATE, "neighbour lists overflowed during timing; rerun tm_neighbors");
  return TM_OK;
}

tm_status tm_last_kernel_ms(tm_ctx *c, double *ms) {
  if (!c || !ms) return TM_ERR_INVALID;
  *ms = c->last_kernel_ms;
  return TM_OK;
}

tm_status tm_graph_destroy(tm_ctx *c) {
  if (!c) return TM_ERR_INVALID;
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  return TM_OK;
}

namespace {

tm_status check_move(tm_ctx *c, const tm_move *m) {
  if (!m) return tm_fail(c, TM_ERR_INVALID, "move parameters missing");
  if (c->external) return tm_fail(c, TM_ERR_STATE, "cells bound with tm_bind_device are read-only; use tm_upload");
  if (!(m->box[0] > 0.0) || !(m->box[1] > 0.0) || !(m->box[2] > 0.0))
    return tm_fail(c, TM_ERR_INVALID, "box edges must be positive");
  return TM_OK;
}

// capture build + neighbors + sph (+ the move) on the context's stream
tm_status capture(tm_ctx *c, tm_precision p) {
  tm_graph_destroy(c);
  TM_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  tm_status s = tm_run_build(c);
  if (s == TM_OK) s = tm_run_neighbors(c, false);
  if (s == TM_OK) s = tm_run_sph(c, p);
  if (s == TM_OK && c->graph_loop) s = tm_run_move(c, &c->graph_move, nullptr, c->graph_removecell, true);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (s != TM_OK) {
    if (g) (void)hipGraphDestroy(g);
    return s;
  }
  if (e != hipSuccess) return tm_fail(c, TM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  c->graph = g;
  TM_HIP(c, hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
  c->graph_prec = p;
  return TM_OK;
}

}  // namespace

tm_status tm_update_positions(tm_ctx *c, const tm_move *m, const float *draws, uint8_t *removecell,
                              int64_t *removecount) {
  if (!c) return TM_ERR_INVALID;
  TM_TRY(check_move(c, m));
  if (c->stage < 4) return tm_fail(c, TM_ERR_STATE, "no forces: run tm_sph first");
  if (c->moved) return tm_fail(c, TM_ERR_STATE, "the cells moved since the last sums: run tm_step first");
  TM_TRY(tm_run_move(c, m, draws, removecell, false));
  c->stage = 1;  // positions changed: the tree and the sums are stale
  if (removecount) {
    unsigned long long v = 0;
    TM_HIP(c, hipMemcpyAsync(&v, tm_removed_word(c), sizeof(v), hipMemcpyDeviceToHost, c->stream));
    TM_HIP(c, hipStreamSynchronize(c->stream));
    *removecount += (int64_t)v;
  }
  return TM_OK;
}

tm_status tm_move_cells(tm_ctx *c, int64_t n, int32_t dimension, double *x, double *y, double *z,
                        const double *size, const int32_t *ctype, const double *fx, const double *fy,
                        const double *fz, const tm_move *m, const float *draws, uint8_t *removecell,
                        int64_t *removecount) {
  if (!c) return TM_ERR_INVALID;
  if (!m) return tm_fail(c, TM_ERR_INVALID, "move parameters missing");
  if (n < 0 || (dimension != 2 && dimension != 3)) return tm_fail(c, TM_ERR_INVALID, "bad cell count or dimension");
  if (n > 0 && (!x || !y || (dimension == 3 && (!z || !fz)) || !size || !ctype || !fx || !fy))
    return tm_fail(c, TM_ERR_INVALID, "missing cell or force array");
  if (!(m->box[0] > 0.0) || !(m->box[1] > 0.0) || !(m->box[2] > 0.0))
    return tm_fail(c, TM_ERR_INVALID, "box edges must be positive");
  TM_TRY(tm_run_move_arrays(c, n, dimension, x, y, z, size, ctype, fx, fy, fz, m, draws, removecell));
  unsigned long long w[2] = {0, 0};  // newly marked, error bits
  TM_HIP(c, hipMemcpyAsync(&w[0], tm_removed_word(c), 8, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipMemcpyAsync(&w[1], tm_move_error_word(c), 8, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  if (w[1] & TM_FLAG_CTYPE) return tm_fail(c, TM_ERR_INVALID, "ctype out of range (0..15)");
  if (removecount) *removecount += (int64_t)w[0];
  return TM_OK;
}

static tm_status check_diffusion(tm_ctx *c, const tm_grid *g, const tm_diffusion *d, const double *field) {
  if (!g || !d || !field) return tm_fail(c, TM_ERR_INVALID, "grid, parameters and field are required");
  for (int a = 0; a < 3; ++a)
    if (g->localsize[a] < 1) return tm_fail(c, TM_ERR_INVALID, "grid localsize must be >= 1 on every axis");
  if (!(g->resolution > 0.0f)) return tm_fail(c, TM_ERR_INVALID, "grid resolution must be positive");
  if (d->substeps < 0 || d->maxiter < 0 || !(d->tol >= 0.0))
    return tm_fail(c, TM_ERR_INVALID, "substeps and maxiter must be >= 0, tol >= 0");
  if (!(d->diffusion >= 0.0) || !(d->dt >= 0.0) || !(d->lambda >= 0.0))
    return tm_fail(c, TM_ERR_INVALID, "diffusion, dt and lambda must be >= 0");
  return TM_OK;
}

tm_status tm_diffusion_step(tm_ctx *c, const tm_grid *g, const tm_diffusion *d, const double *source, double *field,
                            int32_t *iterations, double *residual) {
  if (!c) return TM_ERR_INVALID;
  TM_TRY(check_diffusion(c, g, d, field));
  return tm_run_diffusion(c, g, d, source, field, iterations, residual);
}

tm_status tm_diffusion_step_noalloc(tm_ctx *c, const tm_grid *g, const tm_diffusion *d, const double *source,
                                    double *field, int32_t *iterations, double *residual) {
  if (!c) return TM_ERR_INVALID;
  TM_TRY(check_diffusion(c, g, d, field));
  return tm_run_diffusion_noalloc(c, g, d, source, field, iterations, residual);
}

tm_status tm_get_positions(tm_ctx *c, double *x, double *y, double *z) {
  if (!c) return TM_ERR_INVALID;
  if (c->stage < 1) return tm_fail(c, TM_ERR_STATE, "no cells");
  const int64_t n = c->n;
  if (n == 0) return TM_OK;
  if (x) TM_HIP(c, hipMemcpyAsync(x, c->x, 8 * n, hipMemcpyDeviceToHost, c->stream));
  if (y) TM_HIP(c, hipMemcpyAsync(y, c->y, 8 * n, hipMemcpyDeviceToHost, c->stream));
  if (z) TM_HIP(c, hipMemcpyAsync(z, c->z, 8 * n, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

tm_status tm_graph_capture_loop(tm_ctx *c, tm_precision p, const tm_move *m, uint8_t *removecell) {
  if (!c) return TM_ERR_INVALID;
  if (c->stage < 1) return tm_fail(c, TM_ERR_STATE, "upload cells first");
  TM_TRY(check_move(c, m));
  // one eager step sizes every buffer, so the capture allocates nothing
  TM_TRY(tm_step(c, p));
  c->graph_loop = true;
  c->graph_move = *m;
  c->graph_removecell = removecell;
  TM_TRY(tm_set_graph_step(c, m->step));
  tm_status s = capture(c, p);
  if (s != TM_OK) c->graph_loop = false;
  return s;
}

tm_status tm_graph_removed(tm_ctx *c, int64_t *count) {
  if (!c) return TM_ERR_INVALID;
  return tm_take_graph_removed(c, count);
}

tm_status tm_graph_capture(tm_ctx *c, tm_precision p) {
  if (!c) return TM_ERR_INVALID;
  if (c->stage < 1) return tm_fail(c, TM_ERR_STATE, "upload cells first");
  // one eager step sizes every buffer, so the capture allocates nothing
  TM_TRY(tm_step(c, p));
  c->graph_loop = false;
  return capture(c, p);
}

tm_status tm_graph_launch(tm_ctx *c) {
  if (!c) return TM_ERR_INVALID;
  if (!c->gexec) return tm_fail(c, TM_ERR_STATE, "no captured graph");
  TM_HIP(c, hipGraphLaunch(c->gexec, c->stream));
  TM_TRY(check_flags(c));
  if (c->h_flags[1] > (unsigned)c->nbr_alloc) {
    // a cell outgrew the captured lists (and a captured move moved nothing):
    // grow, redo the step eagerly from this replay's tree, recapture
    TM_TRY(tm_run_neighbors(c, true));
    TM_TRY(tm_run_sph(c, c->graph_prec));
    if (c->graph_loop) TM_TRY(tm_run_move(c, &c->graph_move, nullptr, c->graph_removecell, true));
    TM_TRY(capture(c, c->graph_prec));
  }
  c->stage = 4;
  c->sums_valid = true;
  c->moved = c->graph_loop;
  c->last_prec = c->graph_prec;
  return TM_OK;
}

// Report: one line per trial, "trial <stage> key=value ... ms=<time>", then
// "chosen key=value ...".  Every candidate gives the same results (the FP64
// sums keep each list's order whatever the shapes).
tm_status tm_autotune(tm_ctx *c, tm_precision p, char *buf, size_t buflen) {
  if (!c) return TM_ERR_INVALID;
  if (c->stage < 3) TM_TRY(tm_step(c, p));
  std::string rep;
  char line[200];
  // A candidate replaces the current shape only when it is faster by more
  // than KEEP (run-to-run noise must not flip a near tie).
  constexpr double KEEP = 0.02;
  // neighbour walk workgroup and the tree window it stages in LDS
  const int nb0 = c->nbr_block, win0 = c->walk_window;
  int best_nb = nb0, best_win = win0;
  double best_t = 1e30, cur_t = 1e30;
  const int shapes[][2] = {{64, 0}, {128, 0}, {256, 0}, {256, 512}, {256, 1024}, {256, 1280}, {256, 2048}};
  for (const auto &sh : shapes) {
    c->nbr_block = sh[0];
    c->walk_window = sh[1];
    double ms = 0;
    TM_TRY(tm_time_stage(c, TM_STAGE_NEIGHBORS, p, 1, 3, &ms));
    snprintf(line, sizeof(line), "trial neighbors nbr_block=%d walk_window=%d ms=%.4f\n", sh[0], sh[1], ms);
    rep += line;
    if (sh[0] == nb0 && sh[1] == win0) cur_t = ms;
    if (ms < best_t) {
      best_t = ms;
      best_nb = sh[0];
      best_win = sh[1];
    }
  }
  if (cur_t <= best_t * (1.0 + KEEP)) {
    best_nb = nb0;
    best_win = win0;
  }
  c->nbr_block = best_nb;
  c->walk_window = best_win;
  // that shape against one union walk per wave (same results)
  {
    double tl = 0, tc = 0;
    TM_TRY(tm_time_stage(c, TM_STAGE_NEIGHBORS, p, 1, 3, &tl));
    const int c0 = c->walk_coop;
    c->walk_coop = 1;
    TM_TRY(tm_time_stage(c, TM_STAGE_NEIGHBORS, p, 1, 3, &tc));
    snprintf(line, sizeof(line), "trial neighbors walk_coop=0 ms=%.4f\ntrial neighbors walk_coop=1 walk_coop_ratio=%d ms=%.4f\n",
             tl, c->walk_coop_ratio, tc);
    rep += line;
    c->walk_coop = c0 ? (tl * (1.0 + KEEP) < tc ? 0 : 1) : (tc * (1.0 + KEEP) < tl ? 1 : 0);
  }
  // that shape with the FP32 candidate screening on and off (same results)
  {
    const int f0 = c->walk_f32;
    double t[2] = {0, 0};
    for (int k = 0; k < 2; ++k) {
      c->walk_f32 = k == 0 ? f0 : !f0;
      TM_TRY(tm_time_stage(c, TM_STAGE_NEIGHBORS, p, 1, 10, &t[k]));
      snprintf(line, sizeof(line), "trial neighbors walk_f32=%d ms=%.4f\n", c->walk_f32, t[k]);
      rep += line;
    }
    c->walk_f32 = t[1] * (1.0 + KEEP) < t[0] ? !f0 : f0;
  }
  int &unroll = (p == TM_MIXED) ? c->sph_unroll_mixed : c->sph_unroll_fp64;
  const bool tiles = c->sph_tiles && c->dim == 3;
  if (tiles) {
    // tile width x neighbours fetched per trip, timed as tile build + pairs
    const int w0 = c->tile_width, u0 = unroll;
    int best_w = w0, best_u = u0;
    best_t = 1e30;
    cur_t = 1e30;
    for (int w : {128, 256, 512}) {
      c->tile_width = w;
      TM_TRY(tm_run_neighbors(c, true));  // lists, and tiles of this width sized for LDS
      double tt = 0;
      TM_TRY(tm_time_stage(c, TM_STAGE_TILES, p, 1, 5, &tt));
      for (int u : {2, 4, 8}) {
        unroll = u;
        double ms = 0;
        TM_TRY(tm_time_stage(c, TM_STAGE_SPH, p, 2, 10, &ms));
        snprintf(line, sizeof(line), "trial tiles_sph tile_width=%d unroll=%d tiles_ms=%.4f sph_ms=%.4f ms=%.4f\n", w,
                 u, tt, ms, tt + ms);
        rep += line;
        if (w == w0 && u == u0) cur_t = tt + ms;
        if (tt + ms < best_t) {
          best_t = tt + ms;
          best_w = w;
          best_u = u;
        }
      }
    }
    if (cur_t <= best_t * (1.0 + KEEP)) {
      best_w = w0;
      best_u = u0;
    }
    if (best_w != w0) tm_graph_destroy(c);  // a captured graph holds tiles of the old width
    c->tile_width = best_w;
    unroll = best_u;
    TM_TRY(tm_run_neighbors(c, true));
  } else {
    // direct pair kernel: workgroup size x neighbours per trip (x lanes per cell)
    TM_TRY(tm_run_neighbors(c, true));
    int best_sb = c->sph_block, best_u = unroll, best_l = c->sph_lanes;
    best_t = 1e30;
    const std::vector<int> lanes = (p == TM_MIXED) ? std::vector<int>{1, 2, 4, 8} : std::vector<int>{1};
    for (int l : lanes) {
      for (int bs : {64, 128, 256}) {
        for (int u : {1, 2, 4, 8}) {
          c->sph_block = bs;
          unroll = u;
          if (p == TM_MIXED) c->sph_lanes = l;
          double ms = 0;
          TM_TRY(tm_time_stage(c, TM_STAGE_SPH, p, 2, 10, &ms));
          snprintf(line, sizeof(line), "trial sph lanes=%d sph_block=%d unroll=%d ms=%.4f\n", l, bs, u, ms);
          rep += line;
          if (ms < best_t) {
            best_t = ms;
            best_sb = bs;
            best_u = u;
            best_l = l;
          }
        }
      }
    }
    c->sph_block = best_sb;
    unroll = best_u;
    if (p == TM_MIXED) c->sph_lanes = best_l;
  }
  snprintf(line, sizeof(line),
           "chosen precision=%s nbr_block=%d walk_window=%d walk_coop=%d walk_f32=%d sph_tiles=%d tile_width=%d "
           "sph_block=%d unroll=%d lanes=%d\n",
           p == TM_MIXED ? "mixed" : "fp64", c->nbr_block, c->walk_window, c->walk_coop, c->walk_f32, tiles ? 1 : 0,
           c->tile_width, c->sph_block, unroll, c->sph_lanes);
  rep += line;
  if (buf && buflen) {
    strncpy(buf, rep.c_str(), buflen - 1);
    buf[buflen - 1] = 0;
  }
  return TM_OK;
}

}  // extern "C"

// The VTK writer: writevtk (src/io.c:372-516), one rank.
//
// The file is the reference's legacy binary VTK: text headers and six
// big-endian sections of n values each -- POINTS (x, y, z as float), CELL_TYPES
// (1), SCALARS rank, SCALARS type (ctype), SCALARS density (1.0: the reference
// writes the constant, its density line being commented out) and VECTORS
// force (float).  One kernel converts and byte-swaps the cells' values
// straight from HBM into the six sections (40 bytes per cell, in cell order);
// the host copies that body back once and writes it between the headers.
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <vector>

#include "tm_internal.h"

namespace {

constexpr int BV = 256;

__device__ __forceinline__ uint32_t be_f32(float v) { return __builtin_bswap32(__float_as_uint(v)); }
__device__ __forceinline__ uint32_t be_i32(int32_t v) { return __builtin_bswap32((uint32_t)v); }

// SORTED: slot i holds the forces of cell perm[i] (the context's Morton order);
// otherwise slot i is cell i.  Positions and types are in cell order.
template <bool SORTED>
__global__ __launch_bounds__(BV) void k_vtk_pack(int64_t n, const int32_t *__restrict__ perm,
                                                 const double *__restrict__ x, const double *__restrict__ y,
                                                 const double *__restrict__ z, const int32_t *__restrict__ ctype,
                                                 const double *__restrict__ fx, const double *__restrict__ fy,
                                                 const double *__restrict__ fz, int32_t rank,
                                                 uint32_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * BV + threadIdx.x;
  if (i >= n) return;
  int64_t c = i;
  if constexpr (SORTED) {
    c = perm[i];
    if (c < 0 || c >= n) c = 0;  // unreachable for a valid permutation; bounds the scatter
  }
  uint32_t *pts = out, *ctp = out + 3 * n, *rk = out + 4 * n, *ty = out + 5 * n, *de = out + 6 * n,
           *fo = out + 7 * n;
  // cell i's position and type (cell order, coalesced) ...
  pts[3 * i + 0] = be_f32((float)x[i]);
  pts[3 * i + 1] = be_f32((float)y[i]);
  pts[3 * i + 2] = be_f32((float)z[i]);
  ctp[i] = be_i32(1);
  rk[i] = be_i32(rank);
  ty[i] = be_i32(ctype[i]);
  de[i] = be_f32(1.0f);
  // ... and the force of slot i, which belongs to cell c
  fo[3 * c + 0] = be_f32((float)fx[i]);
  fo[3 * c + 1] = be_f32((float)fy[i]);
  fo[3 * c + 2] = be_f32((float)fz[i]);
}

tm_status write_file(tm_ctx *c, const char *path, int64_t n, const uint32_t *body) {
  FILE *f = fopen(path, "wb");
  if (!f) return tm_fail(c, TM_ERR_INVALID, std::string("cannot open ") + path);
  bool ok = true;
  auto put = [&](const void *p, size_t len) {
    if (len && fwrite(p, 1, len, f) != len) ok = false;
  };
  auto text = [&](const char *s) { put(s, strlen(s)); };
  char buf[256];
  const size_t sec = 4 * (size_t)n;
  text("# vtk DataFile Version 2.0\nTimothy output\nBINARY\nDATASET UNSTRUCTURED_GRID");
  snprintf(buf, sizeof(buf), "\nPOINTS %" PRId64 " float\n", n);
  text(buf);
  put(body, 3 * sec);
  snprintf(buf, sizeof(buf), "\nCELL_TYPES %" PRId64 "\n", n);
  text(buf);
  put(body + 3 * n, sec);
  snprintf(buf, sizeof(buf), "\nPOINT_DATA %" PRId64, n);
  text(buf);
  text("\nSCALARS rank integer 1\nLOOKUP_TABLE default\n");
  put(body + 4 * n, sec);
  text("\nSCALARS type integer 1\nLOOKUP_TABLE default\n");
  put(body + 5 * n, sec);
  text("\nSCALARS density float 1\nLOOKUP_TABLE default\n");
  put(body + 6 * n, sec);
  text("\nVECTORS force float\n");
  put(body + 7 * n, 3 * sec);
  if (fclose(f) != 0) ok = false;
  if (!ok) return tm_fail(c, TM_ERR_INVALID, std::string("short write to ") + path);
  return TM_OK;
}

tm_status pack_and_write(tm_ctx *c, const char *path, int64_t n, bool sorted, const int32_t *perm, const double *x,
                         const double *y, const double *z, const int32_t *ctype, const double *fx, const double *fy,
                         const double *fz, int32_t rank) {
  if (!path) return tm_fail(c, TM_ERR_INVALID, "path is NULL");
  std::vector<uint32_t> body((size_t)(10 * n));
  if (n > 0) {
    TM_TRY(tm_ensure(c, &c->d_vtk, c->cap_vtk, 10 * n, "VTK body"));
    if (sorted)
      k_vtk_pack<true><<<tm_blocks(n, BV), BV, 0, c->stream>>>(n, perm, x, y, z, ctype, fx, fy, fz, rank, c->d_vtk);
    else
      k_vtk_pack<false><<<tm_blocks(n, BV), BV, 0, c->stream>>>(n, nullptr, x, y, z, ctype, fx, fy, fz, rank,
                                                                c->d_vtk);
    TM_HIP(c, hipGetLastError());
    TM_HIP(c, hipMemcpyAsync(body.data(), c->d_vtk, 40 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    TM_HIP(c, hipStreamSynchronize(c->stream));
  }
  return write_file(c, path, n, body.data());
}

}  // namespace

extern "C" tm_status tm_write_vtk(tm_ctx *c, const char *path, int32_t rank) {
  if (!c) return TM_ERR_INVALID;
  if (!c->sums_valid) return tm_fail(c, TM_ERR_STATE, "no forces for the current order: run tm_sph first");
  return pack_and_write(c, path, c->n, true, c->d_perm, c->x, c->y, c->z, c->ctype, c->d_fx, c->d_fy, c->d_fz,
                        rank);
}

extern "C" tm_status tm_write_vtk_arrays(tm_ctx *c, const char *path, int64_t n, const double *x, const double *y,
                                         const double *z, const int32_t *ctype, const double *fx, const double *fy,
                                         const double *fz, int32_t rank) {
  if (!c) return TM_ERR_INVALID;
  if (n < 0) return tm_fail(c, TM_ERR_INVALID, "negative cell count");
  if (n > 0 && (!x || !y || !z || !ctype || !fx || !fy || !fz))
    return tm_fail(c, TM_ERR_INVALID, "missing cell or force array");
  return pack_and_write(c, path, n, false, nullptr, x, y, z, ctype, fx, fy, fz, rank);
}

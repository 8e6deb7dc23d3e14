// The PR-octree of one colony as octnode_t records (src/global.h:153-164),
// emitted from the radix tree so callers that read cellsinfo->octree see the
// node set the reference's octbuild (src/octree.c:40-207) would build.
//
// An octree internal node is an octant holding >= 2 cells.  Along a radix
// node k with key-prefix length d_k (Morton bits) and parent prefix d_P, the
// octants at depths floor(d_P/3)+1 .. floor(d_k/3) hold exactly k's cells:
// that is the chain of single-octant internal nodes k contributes (from depth
// 0 at the colony root).  A cell's leaf sits one depth below the deepest such
// octant containing it.  Node numbering: internal nodes chain by chain in
// radix-node order (root = 0), then one leaf per cell in Morton order -- the
// reference numbers by insertion order, which only the node set pins.
//
// Field conventions kept from the reference: level is 29 at the root and
// log2(side) - 1 below it (octinsertcell decrements before octemptynode);
// limit = code + side; data = cell index, -2 for internal nodes.
#include <cstring>

#include "tm_internal.h"

namespace {

struct OctNode {
  uint32_t xcode, ycode, zcode, xlimit, ylimit, zlimit, level;
  int64_t father;
  int64_t child[8];
  int32_t data;
};
static_assert(sizeof(OctNode) == 112, "octnode_t layout");

__device__ __forceinline__ int dm_of(const int32_t *delta, int32_t k) { return delta[k] - TM_CID_BITS; }

// first chain depth of radix node k (0 for the colony root)
__device__ __forceinline__ int chain_top(const int32_t *delta, const int32_t *parent, int32_t k, int32_t root) {
  return k == root ? 0 : dm_of(delta, parent[k]) / 3 + 1;
}

// octree node id of the octant at `depth` on the path through radix node P
__device__ int64_t owner_id(const int32_t *delta, const int32_t *parent, const uint32_t *base, int32_t P,
                            int32_t root, int64_t s, int depth) {
  for (int guard = 0; guard < 128; ++guard) {
    if (P < 0) return -1;
    int top = chain_top(delta, parent, P, root);
    int bot = dm_of(delta, P) / 3;
    if (depth >= top && depth <= bot) return (int64_t)base[P - s] + (depth - top);
    if (P == root) return -1;
    P = parent[P];
  }
  return -1;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// parent of every radix node (internal: k, leaves: n-1+i)
__global__ void k_parents(const int2 *__restrict__ range, const int32_t *__restrict__ split, int64_t n,
                          int32_t *__restrict__ parent) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n - 1) return;
  const int32_t g = split[k];
  const int2 r = range[k];
  // children [first, g] and [g + 1, last]: a leaf (n - 1 + cell) when one cell
  const int64_t left = (r.x == g) ? (n - 1) + g : g;
  const int64_t right = (r.y == g + 1) ? (n - 1) + g + 1 : g + 1;
  parent[left] = (int32_t)k;
  parent[right] = (int32_t)k;
}

__global__ void k_chain_counts(const int2 *__restrict__ range, const int32_t *__restrict__ delta,
                               const int32_t *__restrict__ parent, int64_t n, int64_t s, int64_t e, int32_t root,
                               uint32_t *__restrict__ counts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > e - s) return;
  int64_t k = s + i;
  uint32_t cnt = 0;
  if (k <= n - 2 && k < e + 1) {
    int2 r = range[k];
    if (r.x >= s && r.y <= e) cnt = (uint32_t)(dm_of(delta, (int32_t)k) / 3 - chain_top(delta, parent, (int32_t)k, root) + 1);
  }
  counts[i] = cnt;
  if (i == e - s) counts[i + 1] = 0;
}

__global__ void k_emit_internal(const int2 *__restrict__ range, const int32_t *__restrict__ delta,
                                const int32_t *__restrict__ parent, const uint4 *__restrict__ code,
                                const uint32_t *__restrict__ base, int64_t n, int64_t s, int64_t e, int32_t root,
                                OctNode *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > e - s) return;
  int64_t k = s + i;
  if (k > n - 2) return;
  int2 r = range[k];
  if (r.x < s || r.y > e) return;
  int top = chain_top(delta, parent, (int32_t)k, root);
  int bot = dm_of(delta, (int32_t)k) / 3;
  uint4 c = code[r.x];
  for (int d = top; d <= bot; ++d) {
    int64_t id = (int64_t)base[i] + (d - top);
    uint32_t side = 1u << (29 - d);
    OctNode o;
    o.xcode = c.x & ~(side - 1u);
    o.ycode = c.y & ~(side - 1u);
    o.zcode = c.z & ~(side - 1u);
    o.xlimit = o.xcode + side;
    o.ylimit = o.ycode + side;
    o.zlimit = o.zcode + side;
    o.level = d == 0 ? 29u : (uint32_t)(28 - d);
    if (d > top) o.father = id - 1;
    else if (k == root) o.father = -1;
    else o.father = owner_id(delta, parent, base, parent[k], root, s, d - 1);
    for (int j = 0; j < 8; ++j) o.child[j] = -1;
    o.data = -2;
    out[id] = o;
  }
}

__global__ void k_emit_leaves(const int32_t *__restrict__ delta, const int32_t *__restrict__ parent,
                              const uint4 *__restrict__ code, const int32_t *__restrict__ span,
                              const int32_t *__restrict__ perm, const uint32_t *__restrict__ base, int64_t n,
                              int64_t s, int64_t e, int32_t root, int64_t ninternal, OctNode *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > e - s) return;
  int64_t q = s + i;
  int sq = span[q];
  uint32_t side = 1u << sq;
  uint4 c = code[q];
  OctNode o;
  o.xcode = c.x & ~(side - 1u);
  o.ycode = c.y & ~(side - 1u);
  o.zcode = c.z & ~(side - 1u);
  o.xlimit = o.xcode + side;
  o.ylimit = o.ycode + side;
  o.zlimit = o.zcode + side;
  o.level = (uint32_t)(sq - 1);
  o.father = owner_id(delta, parent, base, parent[(n - 1) + q], root, s, 29 - sq - 1);
  for (int j = 0; j < 8; ++j) o.child[j] = -1;
  o.data = perm[q];
  out[ninternal + i] = o;
}

// child pointers: a node of side 2^b sits in slot bit_b(x) + 2 bit_b(y) + 4 bit_b(z) of its father
__global__ void k_link(OctNode *__restrict__ out, int64_t count) {
  int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= count) return;
  const OctNode &o = out[id];
  if (o.father < 0) return;
  uint32_t side = o.xlimit - o.xcode;
  int b = 31 - __clz((int)side);
  int slot = ((o.xcode >> b) & 1) + 2 * ((o.ycode >> b) & 1) + 4 * ((o.zcode >> b) & 1);
  out[o.father].child[slot] = id;
}

}  // namespace

tm_status tm_run_octree(tm_ctx *c, int colony, void *nodes, int64_t *count) {
  const int64_t s = c->col_off[colony], e = c->col_off[colony + 1] - 1;
  const int64_t nc = e - s + 1;
  if (nc <= 0) {
    *count = 0;
    return TM_OK;
  }
  if (nc == 1) {  // the root holds the only cell (src/octree.c:96-98)
    *count = 1;
    if (nodes) {
      int32_t p0 = 0;
      TM_HIP(c, hipMemcpyAsync(&p0, c->d_perm + s, 4, hipMemcpyDeviceToHost, c->stream));
      TM_HIP(c, hipStreamSynchronize(c->stream));
      OctNode o;
      memset(&o, 0, sizeof(o));
      o.xlimit = o.ylimit = o.zlimit = 1u << 29;
      o.level = 29;
      o.father = -1;
      for (int j = 0; j < 8; ++j) o.child[j] = -1;
      o.data = p0;
      memcpy(nodes, &o, sizeof(o));
    }
    return TM_OK;
  }
  const int64_t n = c->n;
  int32_t root = 0;
  TM_HIP(c, hipMemcpyAsync(&root, c->d_colroot + colony, 4, hipMemcpyDeviceToHost, c->stream));
  TM_HIP(c, hipStreamSynchronize(c->stream));
  if (root < 0 || root > n - 2) return tm_fail(c, TM_ERR_STATE, "colony has no internal root");
  k_fill_i32<<<tm_blocks(2 * n - 1, 256), 256, 0, c->stream>>>(c->d_parent, 2 * n - 1, -1);
  k_parents<<<tm_blocks(n - 1, 256), 256, 0, c->stream>>>(c->d_range, c->d_split, n, c->d_parent);
  TM_TRY(tm_ensure(c, &c->d_octcnt, c->cap_octcnt, nc + 1, "octree chain counts"));
  uint32_t *cnt = c->d_octcnt;
  k_chain_counts<<<tm_blocks(nc, 256), 256, 0, c->stream>>>(c->d_range, c->d_delta, c->d_parent, n, s, e, root, cnt);
  tm_status st = tm_scan_u32(c, cnt, cnt, nc + 1);
  uint32_t ninternal = 0;
  if (st == TM_OK) {
    hipError_t he = hipMemcpyAsync(&ninternal, cnt + nc, 4, hipMemcpyDeviceToHost, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) st = tm_fail(c, TM_ERR_HIP, hipGetErrorString(he));
  }
  const int64_t total = (int64_t)ninternal + nc;
  if (st == TM_OK) {
    *count = total;
    if (nodes) {
      TM_TRY(tm_ensure(c, &c->d_octout, c->cap_octout, (int64_t)sizeof(OctNode) * total, "octree nodes"));
      OctNode *d_out = reinterpret_cast<OctNode *>(c->d_octout);
      k_emit_internal<<<tm_blocks(nc, 256), 256, 0, c->stream>>>(c->d_range, c->d_delta, c->d_parent, c->d_code, cnt,
                                                                  n, s, e, root, d_out);
      k_emit_leaves<<<tm_blocks(nc, 256), 256, 0, c->stream>>>(c->d_delta, c->d_parent, c->d_code, c->d_span,
                                                                c->d_perm, cnt, n, s, e, root, ninternal, d_out);
      k_link<<<tm_blocks(total, 256), 256, 0, c->stream>>>(d_out, total);
      hipError_t he = hipGetLastError();
      if (he == hipSuccess)
        he = hipMemcpyAsync(nodes, d_out, sizeof(OctNode) * (size_t)total, hipMemcpyDeviceToHost, c->stream);
      if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
      if (he != hipSuccess) st = tm_fail(c, TM_ERR_HIP, hipGetErrorString(he));
    }
  }
  return st;
}

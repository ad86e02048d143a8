// 8-wide BVH for traversal, collapsed from dray's binary LBVH.
//
// Each wide node holds up to 8 children in the left-to-right order of the binary tree (an inner child
// is replaced in place by its two children, largest surface area first, until 8 children or only leaves
// remain).  Child boxes are the binary tree's exact float32 boxes and every box contains the boxes below
// it, so visiting the hit children of a wide node in slot order enumerates exactly the leaves the binary
// depth-first traversal of DeviceMesh::locate (device_mesh.hpp:247-354) would reach, in the same order.
//
// Node layout (64 floats, 256 B): lo_x[8] lo_y[8] lo_z[8] hi_x[8] hi_y[8] hi_z[8] child[8] pad[8];
// child >= 0: wide node index, child < 0: leaf -(child) - 1; empty slots have lo = +inf, hi = -inf.
// Wide nodes are numbered breadth-first, deterministically (prefix sums, no atomics).
#include "drx_common.hpp"

#include <algorithm>
#include <vector>

namespace drx
{

constexpr int kW = 8;

struct WideItem
{
  int32_t code;   // dray encoding: >= 0 binary inner float4 offset (4 * index), < 0 leaf
  float box[6];   // lo x y z, hi x y z
};

__device__ __forceinline__ float surf (const float *b)
{
  const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ void node_children (const float4 *__restrict__ nodes, int code, WideItem &l, WideItem &r)
{
  const float4 *nd = nodes + code;
  const float4 a = nd[0], b = nd[1], c = nd[2], ch = nd[3];
  l.box[0] = a.x; l.box[1] = a.y; l.box[2] = a.z; l.box[3] = a.w; l.box[4] = b.x; l.box[5] = b.y;
  r.box[0] = b.z; r.box[1] = b.w; r.box[2] = c.x; r.box[3] = c.y; r.box[4] = c.z; r.box[5] = c.w;
  l.code = __float_as_int (ch.x);
  r.code = __float_as_int (ch.y);
}

// Expand one frontier node into its ordered child list (<= 8), count the inner children.
__global__ void k_widen_expand (const float4 *__restrict__ nodes, const int32_t *__restrict__ f_bin, int n_f,
                                WideItem *__restrict__ items, int32_t *__restrict__ n_items,
                                int32_t *__restrict__ n_inner)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_f) return;
  WideItem it[kW];
  int cnt = 2;
  node_children (nodes, 4 * f_bin[i], it[0], it[1]);
  while (cnt < kW)
  {
    int best = -1;
    float bs = -1.f;
    for (int k = 0; k < cnt; ++k)
    {
      if (it[k].code < 0) continue;
      const float s = surf (it[k].box);
      if (best < 0 || s > bs)
      {
        best = k;
        bs = s;
      }
    }
    if (best < 0) break;
    WideItem l, r;
    node_children (nodes, it[best].code, l, r);
    for (int k = cnt; k > best + 1; --k) it[k] = it[k - 1];
    it[best] = l;
    it[best + 1] = r;
    ++cnt;
  }
  int inner = 0;
  for (int k = 0; k < kW; ++k)
  {
    if (k < cnt)
    {
      items[(int64_t)i * kW + k] = it[k];
      inner += it[k].code >= 0 ? 1 : 0;
    }
  }
  n_items[i] = cnt;
  n_inner[i] = inner;
}

__global__ void k_widen_write (const int32_t *__restrict__ f_wide, int n_f, const WideItem *__restrict__ items,
                               const int32_t *__restrict__ n_items, const int32_t *__restrict__ base,
                               int32_t next_first_id, float *__restrict__ out, int32_t *__restrict__ nf_bin,
                               int32_t *__restrict__ nf_wide)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_f) return;
  const int cnt = n_items[i];
  const int w = f_wide[i];
  int rank = 0;
  for (int k = 0; k < kW; ++k)
  {
    float box[6] = { INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY };
    int32_t child = 0;
    if (k < cnt)
    {
      const WideItem it = items[(int64_t)i * kW + k];
      for (int q = 0; q < 6; ++q) box[q] = it.box[q];
      if (it.code < 0)
        child = it.code;
      else
      {
        const int slot = base[i] + rank++;
        child = next_first_id + slot;
        if (nf_bin)
        {
          nf_bin[slot] = it.code / 4;
          nf_wide[slot] = child;
        }
      }
    }
    if (out)
    {
      float *o = out + (int64_t)w * 64;
      for (int q = 0; q < 6; ++q) o[q * 8 + k] = box[q];
      o[48 + k] = __int_as_float (child);
      o[56 + k] = 0.f;
    }
  }
}

static inline unsigned nblk (int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

} // namespace drx

using namespace drx;

extern "C" {

int drx_bvh_widen (const drx_bvh *bvh, float *out_wide, int64_t capacity_nodes, int64_t *out_count, void *stream)
{
  if (!bvh || !bvh->inner_nodes || !out_count) return fail (DRX_ERR_ARG, "drx_bvh_widen: null pointer");
  if (bvh->n_leaves < 2) return fail (DRX_ERR_ARG, "drx_bvh_widen: %d leaves", bvh->n_leaves);
  hipStream_t st = as_stream (stream);
  const float4 *nodes = (const float4 *)bvh->inner_nodes;
  // frontier buffers (binary index, wide id) x 2, items, counts
  int32_t n_f = 1;
  int64_t total = 1;       // wide ids handed out so far (root = 0)
  std::vector<int32_t> h0 = { 0, 0 };
  int32_t *fr = (int32_t *)arena (SLOT_TREE, 2 * sizeof (int32_t), st);
  if (!fr) return DRX_ERR_HIP;
  DRX_HIP (hipMemcpyAsync (fr, h0.data (), 2 * sizeof (int32_t), hipMemcpyHostToDevice, st));
  DRX_HIP (hipStreamSynchronize (st));
  int cur_slot = SLOT_TREE, nxt_slot = SLOT_TREE2;
  while (n_f > 0)
  {
    int32_t *f_bin = (int32_t *)arena (cur_slot, 2 * sizeof (int32_t) * (size_t)n_f, st);
    if (!f_bin) return DRX_ERR_HIP;
    int32_t *f_wide = f_bin + n_f;
    WideItem *items = (WideItem *)arena (SLOT_SORT_KEYS_A, sizeof (WideItem) * kW * (size_t)n_f, st);
    int32_t *cnts = (int32_t *)arena (SLOT_SORT_VALS_A, sizeof (int32_t) * 3 * (size_t)n_f + 64, st);
    if (!items || !cnts) return DRX_ERR_HIP;
    int32_t *n_items = cnts, *n_inner = cnts + n_f, *base = cnts + 2 * n_f;
    hipLaunchKernelGGL (k_widen_expand, dim3 (nblk (n_f, 256)), dim3 (256), 0, st, nodes, f_bin, n_f, items, n_items,
                        n_inner);
    DRX_LAUNCH_CHECK ();
    int32_t n_next = 0;
    int rc = exclusive_scan_i32 (n_inner, base, n_f, &n_next, st);
    if (rc) return rc;
    if (total + n_next > (1ll << 31) - 1)
      return fail (DRX_ERR_UNSUPPORTED, "drx_bvh_widen: more than 2^31 wide nodes");
    if (out_wide && total + n_next > capacity_nodes)
      return fail (DRX_ERR_ARG, "drx_bvh_widen: capacity %lld nodes is too small", (long long)capacity_nodes);
    int32_t *nf = nullptr;
    if (n_next > 0)
    {
      // the next frontier must not alias the current one: it lives in the other slot
      nf = (int32_t *)arena (nxt_slot, 2 * sizeof (int32_t) * (size_t)n_next, st);
      if (!nf) return DRX_ERR_HIP;
    }
    hipLaunchKernelGGL (k_widen_write, dim3 (nblk (n_f, 256)), dim3 (256), 0, st, f_wide, n_f, items, n_items, base,
                        (int32_t)total, out_wide, nf, nf ? nf + n_next : nullptr);
    DRX_LAUNCH_CHECK ();
    total += n_next;
    n_f = n_next;
    std::swap (cur_slot, nxt_slot);
  }
  DRX_HIP (hipStreamSynchronize (st));
  *out_count = total;
  return DRX_OK;
}

} // extern "C"

// Helpers shared by the query (drx_query.hip) and integration (drx_integrate.hip) translation units.
#pragma once

#include "drx_common.hpp"
#include "drx_element.hpp"

namespace drx
{

// ---- arithmetic that must round like dray's float32 CPU path (no FMA contraction) --------------
__device__ __forceinline__ float ray_point (float o, float t, float d)
{
#pragma clang fp contract(off)
  return o + t * d;
}

__device__ __forceinline__ float mul_sub_nc (float a, float b, float c)
{
#pragma clang fp contract(off)
  return a * b - c;
}

__device__ __forceinline__ void blend_nc (float f[4], const float4 b)   // rendering/colors.hpp:15-23
{
#pragma clang fp contract(off)
  const float a = b.w * (1.f - f[3]);
  f[0] = f[0] + b.x * a;
  f[1] = f[1] + b.y * a;
  f[2] = f[2] + b.z * a;
  f[3] = a + f[3];
}

// DeviceColorMap::color for Float = double (device_color_map.hpp:61-77): clamp and normalise in double,
// the normalised value rounded to float32 before it is scaled to an index.
__device__ __forceinline__ float4 color_lookup64 (const float4 *__restrict__ lut, int size, double mn, double mx,
                                                  double inv, double scalar)
{
#pragma clang fp contract(off)
  const double s = fmin (mx, fmax (mn, scalar));
  const double s2 = (scalar != scalar) ? mn : s;
  const float normalized = (float)((s2 - mn) * inv);
  int idx = (int)(normalized * (float)(size - 1));
  idx = min (max (idx, 0), size - 1);
  return lut[idx];
}

// DeviceColorMap::color (device_color_map.hpp:61-77)
__device__ __forceinline__ float4 color_lookup (const float4 *__restrict__ lut, int size, float mn, float mx, float inv,
                                                float scalar)
{
#pragma clang fp contract(off)
  const float s = fminf (mx, fmaxf (mn, scalar));
  const float s2 = (scalar != scalar) ? mn : s;   // std::max(min, NaN) keeps min
  const float normalized = (s2 - mn) * inv;
  int idx = (int)(normalized * (float)(size - 1));
  idx = min (max (idx, 0), size - 1);
  return lut[idx];
}

static inline unsigned nblk (int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

#ifndef DRX_WAVES_PER_EU
#define DRX_WAVES_PER_EU 0   // 0: per-kernel choice below; otherwise a fixed minimum waves per SIMD
#endif

// Minimum waves per SIMD for the locate / integrate kernels (sets the VGPR budget: 5 -> 96, 4 -> 128).
// Measured on MI355X: order <= 3 walks run fastest at 5 (1024^2 render of a 64^3 order-3 mesh, 512
// samples per ray: 19.97 ms vs 22.37 ms at 4, some spills included; with FP64 polish, whose correction is
// an out-of-line call, 14.36 ms vs 16.40 ms at 4); order 4 and the learned start need the 128-register
// budget (2048^2 render of a 128^3 order-4 mesh, 1000 samples per ray: 297.5 ms at 4 vs 323 ms at 5).
#ifndef DRX_LEARNED_WAVES
#define DRX_LEARNED_WAVES 4   // learned-start integrate kernels (mode 3); 3 (no spills) measured slower
#endif
__host__ __device__ constexpr int waves_per_simd (int P, bool polish, int mode)
{
  return DRX_WAVES_PER_EU > 0 ? DRX_WAVES_PER_EU
                              : (mode == 3 ? DRX_LEARNED_WAVES : ((P >= 1 && P <= 3) ? 5 : 4));
}

// Compute units of device `dev` (cached per device; 256 if the query fails).
inline int compute_units (int dev)
{
  static int cus[16];
  if (dev < 0 || dev >= 16) return 256;
  if (!cus[dev])
  {
    int v = 0;
    cus[dev] = (hipDeviceGetAttribute (&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) ? v : 256;
  }
  return cus[dev];
}

// Streamed batched-locate kernels (k_locate_stream, k_locate_mlp_stream): the integrate kernels' choice
// unless DRX_LOCATE_WAVES fixes it.
#ifndef DRX_LOCATE_WAVES
#define DRX_LOCATE_WAVES 0
#endif
__host__ __device__ constexpr int locate_waves_per_simd (int P, bool polish, int mode)
{
  return DRX_LOCATE_WAVES > 0 ? DRX_LOCATE_WAVES : (mode == 3 ? 4 : waves_per_simd (P, polish, mode));
}

static __global__ void k_fill_i32q (int32_t *p, int64_t n, int32_t v)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

static inline int check_mesh (const drx_mesh *m, int ncomp, const char *what)
{
  if (!m || !(m->data || (ncomp == 1 && m->data64))) return fail (DRX_ERR_ARG, "%s: null mesh", what);
  if (m->ncomp != ncomp) return fail (DRX_ERR_ARG, "%s: expected ncomp %d, got %d", what, ncomp, m->ncomp);
  if (ncomp == 3 && m->data64) return fail (DRX_ERR_ARG, "%s: FP64 blocks are for scalar fields only", what);
  if ((m->elem_order == nullptr) != (m->elem_offset == nullptr))
    return fail (DRX_ERR_ARG, "%s: mixed-order meshes need elem_order and elem_offset", what);
  if (!m->elem_order)
  {
    if (m->order < 1 || m->order > DRX_MAX_ORDER)
      return fail (DRX_ERR_ARG, "%s: order %d outside 1..%d", what, m->order, DRX_MAX_ORDER);
    if (m->stride != drx_mesh_stride (m->order, ncomp))
      return fail (DRX_ERR_ARG, "%s: stride %d does not match order %d", what, m->stride, m->order);
  }
  return DRX_OK;
}

static inline int check_bvh (const drx_bvh *b, const char *what)
{
  if (!b || !b->inner_nodes || !b->leaf_pack) return fail (DRX_ERR_ARG, "%s: bvh needs inner_nodes and leaf_pack", what);
  if (b->n_leaves < 2) return fail (DRX_ERR_ARG, "%s: bvh has %d leaves (dray pads to >= 2)", what, b->n_leaves);
  return DRX_OK;
}

static inline LocateCfg make_cfg (const drx_mesh *m, const drx_bvh *b, const drx_locate_opts *o)
{
  LocateCfg c;
  c.nodes = (const float4 *)b->inner_nodes;
  c.wide = (const float4 *)b->wide_nodes;
  c.leaves = (const float4 *)b->leaf_pack;
  c.qnodes = (const uint4 *)b->qnodes;
  c.leaf_boxes = (const float4 *)b->leaf_boxes;
  for (int d = 0; d < 3; ++d)
  {
    c.q_org[d] = b->qgrid[d];
    c.q_inv[d] = b->qgrid[3 + d];
  }
  if (c.qnodes && !c.leaf_boxes) c.qnodes = nullptr;
  c.centre = o && (o->init_mode == DRX_INIT_SUBREF_CENTRE || (o->init_mode == DRX_INIT_MLP && o->mlp_fallback == 1));
  c.starts = c.centre ? nullptr : (const float4 *)b->leaf_starts;
  c.mesh = m->data;
  c.offsets = m->elem_offset;
  c.orders = m->elem_order;
  c.stride = m->stride;
  c.polish = o ? (o->fp64_polish != 0) : false;
  c.cond_threshold = o ? o->cond_threshold : 1e4f;
  c.polished_out = o ? o->polished : nullptr;
  c.guess = (o && o->init_mode == DRX_INIT_GUESS) ? (const Location *)o->guess : nullptr;
  c.record = o ? (Location *)o->record : nullptr;
  c.samples_per_ray = o ? o->samples_per_ray : 0;
  return c;
}

} // namespace drx

// LBVH build for gfx950: per-element de Casteljau split into 33 sub-boxes, Morton codes, stable LSD
// radix sort, Karras tree, bottom-up refit, dray-layout emit.
//
// This file is compiled with -ffp-contract=off: every float operation here rounds exactly like dray's
// float32 CPU path, so the boxes, codes and the emitted tree are bit-identical to the reference's.
#include "drx_common.hpp"
#include "drx_wave.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace drx
{

// =============================================================================================
// Split: one 64-lane wavefront per element, the 33 control nets of the element in LDS.
// Semantics of construct_bvh (data_model/mesh_utils.cpp:557-670): 32 times pick the box of largest
// volume (surface area when the volume is 0; ties -> lowest index), split its SubRef at 0.5 along the
// longest reference axis (ties -> lowest axis, :482-506), de Casteljau-split the net (mother keeps the
// lower half, the copy at index `count` keeps the upper half; elem_ops.hpp:1346-1406), re-hull both.
// Then scale every box by 1.000001f about its centre (range.hpp:161-173).
// =============================================================================================
constexpr int kSplitBoxes = 33;
constexpr int kSplitGlobalBlocks = 512;   // workgroups of the order-7/8 split, each with its nets in global memory

__device__ __forceinline__ float box_measure (const float *b)   // b: xmin,xmax,ymin,ymax,zmin,zmax
{
  const float l0 = b[1] - b[0], l1 = b[3] - b[2], l2 = b[5] - b[4];
  float v = 1.f;
  v *= l0;
  v *= l1;
  v *= l2;
  if (v == 0.f)
  {
    float a = 2.f * l0 * l1;
    a += 2.f * l1 * l2;
    a += 2.f * l2 * l0;
    v = a;
  }
  return v;
}

// de Casteljau halving of every line of a net along `axis` (order P, nets of NPE_ stride per component):
// lane t handles (component t / lines, line t % lines), 64 lanes per pass.  Running min/max of both halves
// per component go to a/b mn/mx.
template <int P>
__device__ __forceinline__ void split_lines (float *mom, float *dau, int NPE_, int axis, int lane, float &lo_a,
                                             float &hi_a, float &lo_b, float &hi_b)
{
  // Row c of the wavefront (lanes 16c .. 16c+15) takes component c's lines, 16 at a time; row 3 idles.
  // Each lane's hulls of its lines then reduce within its row.
  constexpr int N1 = P + 1, LINES = N1 * N1;
  const int stride = (axis == 0) ? 1 : ((axis == 1) ? N1 : N1 * N1);
  const int c = lane >> 4;
  if (c > 2) return;
  for (int l = lane & 15; l < LINES; l += 16)
  {
    const int u = l % N1, v = l / N1;   // the two other indices, lowest axis first
    const int base = (axis == 0) ? (N1 * u + LINES * v) : ((axis == 1) ? (u + LINES * v) : (u + N1 * v));
    float *m = mom + c * NPE_ + base, *d = dau + c * NPE_ + base;
    float a[N1], b[N1];
#pragma unroll
    for (int ii = 0; ii < N1; ++ii) a[ii] = b[ii] = m[ii * stride];
#pragma unroll
    for (int r = 1; r <= P; ++r)
    {
      // mother: new[ii] = .5 old[ii-1] + .5 old[ii] for ii >= r (descending, so old[ii-1] is still old)
#pragma unroll
      for (int ii = P; ii >= r; --ii) a[ii] = a[ii - 1] * (1.0f - 0.5f) + a[ii] * 0.5f;
      // daughter: new[ii] = .5 old[ii] + .5 old[ii+1] for ii <= P - r (ascending)
#pragma unroll
      for (int ii = 0; ii <= P - r; ++ii) b[ii] = b[ii] * (1.0f - 0.5f) + b[ii + 1] * 0.5f;
    }
#pragma unroll
    for (int ii = 0; ii < N1; ++ii)
    {
      m[ii * stride] = a[ii];
      d[ii * stride] = b[ii];
      lo_a = fminf (lo_a, a[ii]);
      hi_a = fmaxf (hi_a, a[ii]);
      lo_b = fminf (lo_b, b[ii]);
      hi_b = fmaxf (hi_b, b[ii]);
    }
  }
}

template <int PMAX>
__global__ void __launch_bounds__ (64) k_split (const int32_t *__restrict__ ctrl_idx, const float *__restrict__ values,
                                                 const int32_t *__restrict__ orders, const int64_t *__restrict__ offsets,
                                                 int order, int n_el, float *__restrict__ out_aabbs,
                                                 int32_t *__restrict__ out_prims, float *__restrict__ out_subrefs)
{
  constexpr int NPE = (PMAX + 1) * (PMAX + 1) * (PMAX + 1);
  constexpr int DPL = (NPE + 63) / 64;   // dofs per lane
  __shared__ float nets[kSplitBoxes * 3 * NPE];

  const int e = blockIdx.x;
  if (e >= n_el) return;
  const int lane = threadIdx.x;
  const int p = orders ? orders[e] : order;
  const int npe = (p + 1) * (p + 1) * (p + 1);
  const int64_t off = offsets ? offsets[e] : (int64_t)e * npe;

  // net 0 <- element control points; box 0 = hull
  float mn[3] = { INFINITY, INFINITY, INFINITY }, mx[3] = { -INFINITY, -INFINITY, -INFINITY };
#pragma unroll
  for (int k = 0; k < DPL; ++k)
  {
    const int d = lane + 64 * k;
    if (d < npe)
    {
      const int64_t g = (int64_t)ctrl_idx[off + d] * 3;
#pragma unroll
      for (int c = 0; c < 3; ++c)
      {
        const float v = values[g + c];
        nets[c * NPE + d] = v;
        mn[c] = fminf (mn[c], v);
        mx[c] = fmaxf (mx[c], v);
      }
    }
  }
  // Per-box state lives in registers: lane b (< 33) holds box b's hull, SubRef and measure.
  float my_box[6], my_ref[6], my_meas = -INFINITY;
  {
    float b0[6];
#pragma unroll
    for (int c = 0; c < 3; ++c)
    {
      b0[2 * c] = wave_reduce_min (mn[c]);
      b0[2 * c + 1] = wave_reduce_max (mx[c]);
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) my_box[q] = b0[q];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      my_ref[d] = 0.f;
      my_ref[3 + d] = 1.f;
    }
    if (lane == 0) my_meas = box_measure (b0);
  }
  __syncthreads ();

  for (int count = 1; count <= kSplitBoxes - 1; ++count)
  {
    // pick the box: dray's scan keeps the first strictly larger measure (ties -> lowest index; a NaN
    // never compares larger, and a NaN first measure is never replaced)
    int max_id = 0;
    {
      const float m0 = __int_as_float (__builtin_amdgcn_readfirstlane (__float_as_int (my_meas)));
      if (m0 == m0)
      {
        // integer keys: NaN measures and lanes beyond count can never be the maximum (box 0's key is a
        // number here, so the maximum is at least that)
        const int32_t cand = (lane < count && my_meas == my_meas) ? float_key (my_meas) : INT32_MIN;
        const int32_t M = wave_reduce_max (cand);
        const uint64_t hit = __ballot (lane < count && cand == M);
        max_id = __ffsll ((unsigned long long)hit) - 1;
      }
    }
    float sr[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) sr[q] = __int_as_float (__builtin_amdgcn_readlane (__float_as_int (my_ref[q]), max_id));
    int axis = 0;
    float maxlen = sr[3] - sr[0];
    for (int d = 1; d < 3; ++d)
    {
      const float len = sr[3 + d] - sr[d];
      if (len > maxlen)
      {
        axis = d;
        maxlen = len;
      }
    }
    const float split_pt = sr[axis] * (1.0f - 0.5f) + sr[3 + axis] * 0.5f;
    if (lane == count)   // complement: lo[axis] = split
    {
#pragma unroll
      for (int q = 0; q < 6; ++q) my_ref[q] = (q == axis) ? split_pt : sr[q];
    }
    if (lane == max_id)   // in place: hi[axis] = split
    {
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (d == axis) my_ref[3 + d] = split_pt;
    }
    // split the mother's net along `axis`, one (line, component) per lane (component c in row c): a line is the p+1 dofs along
    // the axis with the other two indices fixed; each lane reads its line, runs the de Casteljau rounds in
    // registers (mother keeps [lo, mid], daughter [mid, hi]; the same operations in the same order as
    // split_inplace, elem_ops.hpp:1346-1406) and writes both halves.  Lines are disjoint, so no lane
    // reads what another lane writes in this count.
    float *mom = nets + max_id * 3 * NPE;
    float *dau = nets + count * 3 * NPE;
    float lo_a = INFINITY, hi_a = -INFINITY, lo_b = INFINITY, hi_b = -INFINITY;
    if constexpr (PMAX == 0)
      ;
    else if (p == PMAX)
      split_lines<PMAX> (mom, dau, NPE, axis, lane, lo_a, hi_a, lo_b, hi_b);
    else
      switch (p)   // mixed-order meshes
      {
      case 1: split_lines<1> (mom, dau, NPE, axis, lane, lo_a, hi_a, lo_b, hi_b); break;
      case 2: split_lines<(PMAX >= 2 ? 2 : 1)> (mom, dau, NPE, axis, lane, lo_a, hi_a, lo_b, hi_b); break;
      case 3: split_lines<(PMAX >= 3 ? 3 : 1)> (mom, dau, NPE, axis, lane, lo_a, hi_a, lo_b, hi_b); break;
      default: split_lines<(PMAX >= 4 ? 4 : 1)> (mom, dau, NPE, axis, lane, lo_a, hi_a, lo_b, hi_b); break;
      }
    // the lanes' hulls are never NaN (fminf / fmaxf from +-inf skip NaN values), so their integer keys
    // reduce exactly like the floats
    const int32_t klo_a = row_reduce_min (float_key (lo_a)), khi_a = row_reduce_max (float_key (hi_a));
    const int32_t klo_b = row_reduce_min (float_key (lo_b)), khi_b = row_reduce_max (float_key (hi_b));
    float ba[6], bb[6];
#pragma unroll
    for (int c = 0; c < 3; ++c)
    {
      ba[2 * c] = key_float (__builtin_amdgcn_readlane (klo_a, 16 * c));
      ba[2 * c + 1] = key_float (__builtin_amdgcn_readlane (khi_a, 16 * c));
      bb[2 * c] = key_float (__builtin_amdgcn_readlane (klo_b, 16 * c));
      bb[2 * c + 1] = key_float (__builtin_amdgcn_readlane (khi_b, 16 * c));
    }
    if (lane == max_id)
    {
#pragma unroll
      for (int q = 0; q < 6; ++q) my_box[q] = ba[q];
      my_meas = box_measure (ba);
    }
    if (lane == count)
    {
#pragma unroll
      for (int q = 0; q < 6; ++q) my_box[q] = bb[q];
      my_meas = box_measure (bb);
    }
    __syncthreads ();
  }

  if (lane < kSplitBoxes)
  {
    float b[6];
    for (int q = 0; q < 6; ++q) b[q] = my_box[q];
    const float s = (float)1.000001;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      float lo = b[2 * d], hi = b[2 * d + 1];
      if (!(lo > hi))
      {
        const float c = 0.5f * (lo + hi);
        const float delta = s * 0.5f * (hi - lo);
        const float v0 = c - delta, v1 = c + delta;
        lo = fminf (lo, v0);
        hi = fmaxf (hi, v0);
        lo = fminf (lo, v1);
        hi = fmaxf (hi, v1);
      }
      b[2 * d] = lo;
      b[2 * d + 1] = hi;
    }
    const int64_t k = (int64_t)e * kSplitBoxes + lane;
    for (int q = 0; q < 6; ++q)
    {
      out_aabbs[6 * k + q] = b[q];
      out_subrefs[6 * k + q] = my_ref[q];
    }
    out_prims[k] = e;
  }
}

// The same split with one wavefront per component (a 192-thread workgroup per element), for orders >= 4, where
// the 33 nets take 49.5 KB of LDS (order 4) and allow only 3 elements per CU: each wavefront splits only its own
// component's lines (one line per lane, one pass), reduces that component's two hulls over the wave, and the
// three wavefronts exchange the six values through LDS once per split.  A wavefront reads and writes only its
// own component's part of the nets, so the exchange is the only synchronisation.  Every wavefront keeps the
// same box state (hulls, SubRefs, measures) and takes the same decisions; wavefront 0 writes the result.
template <int P>
__device__ __forceinline__ void split_lines_comp (float *mom, float *dau, int axis, int lane, float &lo_a,
                                                  float &hi_a, float &lo_b, float &hi_b)
{
  constexpr int N1 = P + 1, LINES = N1 * N1;
  const int stride = (axis == 0) ? 1 : ((axis == 1) ? N1 : N1 * N1);
  for (int l = lane; l < LINES; l += 64)   // one pass up to order 7, two at order 8
  {
    const int u = l % N1, v = l / N1;
    const int base = (axis == 0) ? (N1 * u + LINES * v) : ((axis == 1) ? (u + LINES * v) : (u + N1 * v));
    float *m = mom + base, *d = dau + base;
    float a[N1], b[N1];
#pragma unroll
    for (int ii = 0; ii < N1; ++ii) a[ii] = b[ii] = m[ii * stride];
#pragma unroll
    for (int r = 1; r <= P; ++r)
    {
#pragma unroll
      for (int ii = P; ii >= r; --ii) a[ii] = a[ii - 1] * (1.0f - 0.5f) + a[ii] * 0.5f;
#pragma unroll
      for (int ii = 0; ii <= P - r; ++ii) b[ii] = b[ii] * (1.0f - 0.5f) + b[ii + 1] * 0.5f;
    }
#pragma unroll
    for (int ii = 0; ii < N1; ++ii)
    {
      m[ii * stride] = a[ii];
      d[ii * stride] = b[ii];
      lo_a = fminf (lo_a, a[ii]);
      hi_a = fmaxf (hi_a, a[ii]);
      lo_b = fminf (lo_b, b[ii]);
      hi_b = fmaxf (hi_b, b[ii]);
    }
  }
}

// GNETS (orders 7 and 8, whose nets exceed the LDS): the nets live in a global scratch area, one per
// workgroup, and each workgroup takes elements blockIdx.x, blockIdx.x + gridDim.x, ...
template <int PMAX, bool GNETS>
__global__ void __launch_bounds__ (192) k_split_cw (const int32_t *__restrict__ ctrl_idx, const float *__restrict__ values,
                                                    const int32_t *__restrict__ orders, const int64_t *__restrict__ offsets,
                                                    int order, int n_el, float *__restrict__ out_aabbs,
                                                    int32_t *__restrict__ out_prims, float *__restrict__ out_subrefs,
                                                    float *__restrict__ gnets)
{
  constexpr int NPE = (PMAX + 1) * (PMAX + 1) * (PMAX + 1);
  __shared__ float lnets[GNETS ? 1 : kSplitBoxes * 3 * NPE];   // box b, component c at (b * 3 + c) * NPE
  __shared__ float xch[2][3][4];                                // per split parity: each component's hull pair
  float *nets = GNETS ? gnets + (size_t)blockIdx.x * kSplitBoxes * 3 * NPE : lnets;
  const int c = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int e = blockIdx.x; e < n_el; e += gridDim.x)
  {
  const int p = orders ? orders[e] : order;
  const int npe = (p + 1) * (p + 1) * (p + 1);
  const int64_t off = offsets ? offsets[e] : (int64_t)e * npe;

  // net 0 <- component c of the element's control points
  float mn = INFINITY, mx = -INFINITY;
  for (int d = lane; d < npe; d += 64)
  {
    const float v = values[(int64_t)ctrl_idx[off + d] * 3 + c];
    nets[c * NPE + d] = v;
    mn = fminf (mn, v);
    mx = fmaxf (mx, v);
  }
  const int32_t kmn = wave_reduce_min (float_key (mn)), kmx = wave_reduce_max (float_key (mx));
  if (lane == 0)   // buffer 0: the first split (count 1) writes buffer 1, and count 2 writes buffer 0 only
  {                // after every wavefront has passed count 1's barrier, so after these reads
    xch[0][c][0] = key_float (kmn);
    xch[0][c][1] = key_float (kmx);
  }
  __syncthreads ();
  float my_box[6], my_ref[6], my_meas = -INFINITY;
  {
    float b0[6];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      b0[2 * d] = xch[0][d][0];
      b0[2 * d + 1] = xch[0][d][1];
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) my_box[q] = b0[q];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      my_ref[d] = 0.f;
      my_ref[3 + d] = 1.f;
    }
    if (lane == 0) my_meas = box_measure (b0);
  }

  for (int count = 1; count <= kSplitBoxes - 1; ++count)
  {
    int max_id = 0;
    {
      const float m0 = __int_as_float (__builtin_amdgcn_readfirstlane (__float_as_int (my_meas)));
      if (m0 == m0)
      {
        const int32_t cand = (lane < count && my_meas == my_meas) ? float_key (my_meas) : INT32_MIN;
        const int32_t M = wave_reduce_max (cand);
        const uint64_t hit = __ballot (lane < count && cand == M);
        max_id = __ffsll ((unsigned long long)hit) - 1;
      }
    }
    float sr[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) sr[q] = __int_as_float (__builtin_amdgcn_readlane (__float_as_int (my_ref[q]), max_id));
    int axis = 0;
    float maxlen = sr[3] - sr[0];
    for (int d = 1; d < 3; ++d)
    {
      const float len = sr[3 + d] - sr[d];
      if (len > maxlen)
      {
        axis = d;
        maxlen = len;
      }
    }
    const float split_pt = sr[axis] * (1.0f - 0.5f) + sr[3 + axis] * 0.5f;
    if (lane == count)
    {
#pragma unroll
      for (int q = 0; q < 6; ++q) my_ref[q] = (q == axis) ? split_pt : sr[q];
    }
    if (lane == max_id)
    {
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (d == axis) my_ref[3 + d] = split_pt;
    }
    float *mom = nets + (max_id * 3 + c) * NPE;
    float *dau = nets + (count * 3 + c) * NPE;
    float lo_a = INFINITY, hi_a = -INFINITY, lo_b = INFINITY, hi_b = -INFINITY;
    if (p == PMAX)
      split_lines_comp<PMAX> (mom, dau, axis, lane, lo_a, hi_a, lo_b, hi_b);
    else
      switch (p)   // mixed-order meshes
      {
      case 1: split_lines_comp<1> (mom, dau, axis, lane, lo_a, hi_a, lo_b, hi_b); break;
      case 2: split_lines_comp<(PMAX >= 2 ? 2 : 1)> (mom, dau, axis, lane, lo_a, hi_a, lo_b, hi_b); break;
      case 3: split_lines_comp<(PMAX >= 3 ? 3 : 1)> (mom, dau, axis, lane, lo_a, hi_a, lo_b, hi_b); break;
      default: split_lines_comp<(PMAX >= 4 ? 4 : 1)> (mom, dau, axis, lane, lo_a, hi_a, lo_b, hi_b); break;
      }
    // the lanes' hulls are never NaN (fminf / fmaxf from +-inf skip NaN values): integer keys reduce them
    const int32_t klo_a = wave_reduce_min (float_key (lo_a)), khi_a = wave_reduce_max (float_key (hi_a));
    const int32_t klo_b = wave_reduce_min (float_key (lo_b)), khi_b = wave_reduce_max (float_key (hi_b));
    const int par = count & 1;
    if (lane == 0)
    {
      xch[par][c][0] = key_float (klo_a);
      xch[par][c][1] = key_float (khi_a);
      xch[par][c][2] = key_float (klo_b);
      xch[par][c][3] = key_float (khi_b);
    }
    __syncthreads ();
    float ba[6], bb[6];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      ba[2 * d] = xch[par][d][0];
      ba[2 * d + 1] = xch[par][d][1];
      bb[2 * d] = xch[par][d][2];
      bb[2 * d + 1] = xch[par][d][3];
    }
    if (lane == max_id)
    {
#pragma unroll
      for (int q = 0; q < 6; ++q) my_box[q] = ba[q];
      my_meas = box_measure (ba);
    }
    if (lane == count)
    {
#pragma unroll
      for (int q = 0; q < 6; ++q) my_box[q] = bb[q];
      my_meas = box_measure (bb);
    }
  }

  if (c == 0 && lane < kSplitBoxes)
  {
    float b[6];
    for (int q = 0; q < 6; ++q) b[q] = my_box[q];
    const float s = (float)1.000001;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      float lo = b[2 * d], hi = b[2 * d + 1];
      if (!(lo > hi))
      {
        const float cc = 0.5f * (lo + hi);
        const float delta = s * 0.5f * (hi - lo);
        const float v0 = cc - delta, v1 = cc + delta;
        lo = fminf (lo, v0);
        hi = fmaxf (hi, v0);
        lo = fminf (lo, v1);
        hi = fmaxf (hi, v1);
      }
      b[2 * d] = lo;
      b[2 * d + 1] = hi;
    }
    const int64_t k = (int64_t)e * kSplitBoxes + lane;
    for (int q = 0; q < 6; ++q)
    {
      out_aabbs[6 * k + q] = b[q];
      out_subrefs[6 * k + q] = my_ref[q];
    }
    out_prims[k] = e;
  }
  __syncthreads ();   // the next element rewrites the nets and the exchange buffers
  }
}

// =============================================================================================
// Bounds reduction (linear_bvh_builder.cpp:19-58)
// =============================================================================================
constexpr int kRedBlock = 256;

__global__ void __launch_bounds__ (kRedBlock) k_bounds_partial (const float *__restrict__ aabbs, int64_t n,
                                                                float *__restrict__ partial)
{
  float mn[3] = { INFINITY, INFINITY, INFINITY }, mx[3] = { -INFINITY, -INFINITY, -INFINITY };
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
  {
    const float *a = aabbs + 6 * i;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      mn[d] = fminf (mn[d], a[2 * d]);
      mx[d] = fmaxf (mx[d], a[2 * d + 1]);
    }
  }
  __shared__ float s[kRedBlock / 64][6];
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    mn[d] = wave_reduce_min (mn[d]);
    mx[d] = wave_reduce_max (mx[d]);
  }
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 3; ++d)
    {
      s[threadIdx.x >> 6][2 * d] = mn[d];
      s[threadIdx.x >> 6][2 * d + 1] = mx[d];
    }
  __syncthreads ();
  if (threadIdx.x < 6)
  {
    float v = s[0][threadIdx.x];
    for (int w = 1; w < kRedBlock / 64; ++w)
      v = (threadIdx.x & 1) ? fmaxf (v, s[w][threadIdx.x]) : fminf (v, s[w][threadIdx.x]);
    partial[6 * blockIdx.x + threadIdx.x] = v;
  }
}

// The partial bounds of n_partial blocks -> the bounds (res.include(mins); res.include(maxs), as
// linear_bvh_builder.cpp:49-57) in out[0..5] and the Morton frame (mins, 1 / extent or 0) in out[6..11].
// One block; n_partial 0 leaves the empty box (+inf, -inf).
__global__ void __launch_bounds__ (kRedBlock) k_bounds_final (const float *__restrict__ partial, int n_partial,
                                                              float *__restrict__ out)
{
  __shared__ float s[kRedBlock / 64][6];
  float mn[3] = { INFINITY, INFINITY, INFINITY }, mx[3] = { -INFINITY, -INFINITY, -INFINITY };
  for (int b = threadIdx.x; b < n_partial; b += blockDim.x)
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      mn[d] = fminf (mn[d], partial[6 * b + 2 * d]);
      mx[d] = fmaxf (mx[d], partial[6 * b + 2 * d + 1]);
    }
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    mn[d] = wave_reduce_min (mn[d]);
    mx[d] = wave_reduce_max (mx[d]);
  }
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 3; ++d)
    {
      s[threadIdx.x >> 6][2 * d] = mn[d];
      s[threadIdx.x >> 6][2 * d + 1] = mx[d];
    }
  __syncthreads ();
  if (threadIdx.x < 3)
  {
    const int d = threadIdx.x;
    float a = s[0][2 * d], b = s[0][2 * d + 1];
    for (int w = 1; w < kRedBlock / 64; ++w)
    {
      a = fminf (a, s[w][2 * d]);
      b = fmaxf (b, s[w][2 * d + 1]);
    }
    float lo = INFINITY, hi = -INFINITY;   // no boxes: the empty box as it stands
    if (n_partial > 0)
    {
      lo = fminf (lo, a);
      hi = fmaxf (hi, a);
      lo = fminf (lo, b);
      hi = fmaxf (hi, b);
    }
    out[2 * d] = lo;
    out[2 * d + 1] = hi;
    const float ext = hi - lo;
    out[6 + d] = lo;
    out[9 + d] = (ext == .0f) ? 0.f : 1.f / ext;
  }
}

// =============================================================================================
// Morton codes (linear_bvh_builder.cpp:60-90, morton_codes.hpp:16-51)
// =============================================================================================
__device__ __forceinline__ uint32_t expand_bits32 (uint32_t x)
{
  x = (x | (x << 16)) & 0x030000FF;
  x = (x | (x << 8)) & 0x0300F00F;
  x = (x | (x << 4)) & 0x030C30C3;
  x = (x | (x << 2)) & 0x09249249;
  return x;
}

__global__ void k_morton (const float *__restrict__ aabbs, int n, const float *__restrict__ frame,
                          uint32_t *__restrict__ keys, int32_t *__restrict__ vals)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float *a = aabbs + 6 * (int64_t)i;
  const float mins[3] = { frame[0], frame[1], frame[2] }, inv[3] = { frame[3], frame[4], frame[5] };
  uint32_t q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    const float lo = a[2 * d], hi = a[2 * d + 1];
    const float c = (lo > hi) ? __builtin_nanf ("") : 0.5f * (lo + hi);
    float x = (c - mins[d]) * inv[d];
    x = fminf (fmaxf (x * 1024.0f, 0.0f), 1023.0f);
    q[d] = (uint32_t)x;
  }
  keys[i] = (expand_bits32 (q[2]) << 2) | (expand_bits32 (q[1]) << 1) | expand_bits32 (q[0]);
  vals[i] = i;
}

// =============================================================================================
// Stable LSD radix sort of (uint32 key, int32 value), 8-bit digits.
// Pass = per-tile histograms -> exclusive scan over (digit, tile) -> stable per-tile scatter.
// =============================================================================================
constexpr int kSortBlock = 256;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortBlock * kSortRounds;   // keys per tile

__global__ void __launch_bounds__ (kSortBlock) k_radix_hist (const uint32_t *__restrict__ keys, int n, int shift,
                                                              int n_tiles, int32_t *__restrict__ hist)
{
  __shared__ int32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads ();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortRounds; ++r)
  {
    const int64_t i = base + r * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd (&h[(keys[i] >> shift) & 0xFF], 1);
  }
  __syncthreads ();
  hist[threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__ (kSortBlock) k_radix_scatter (const uint32_t *__restrict__ keys_in,
                                                                 const int32_t *__restrict__ vals_in, int n, int shift,
                                                                 int n_tiles, const int32_t *__restrict__ offs,
                                                                 uint32_t *__restrict__ keys_out,
                                                                 int32_t *__restrict__ vals_out)
{
  constexpr int W = kSortBlock / 64;
  __shared__ int32_t running[256];
  __shared__ int32_t wcnt[W][256];
  running[threadIdx.x] = offs[threadIdx.x * n_tiles + blockIdx.x];
  const int w = threadIdx.x >> 6;
  const uint64_t lt = lanemask_lt ();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortRounds; ++r)
  {
    const int64_t i = base + r * kSortBlock + threadIdx.x;
    const bool valid = i < n;
    uint32_t key = 0;
    int32_t val = 0;
    if (valid)
    {
      key = keys_in[i];
      val = vals_in[i];
    }
    const uint32_t digit = (key >> shift) & 0xFF;
    // lanes of this wave with the same digit (invalid lanes form their own group)
    uint64_t peers = __ballot (valid);
    if (!valid) peers = ~peers;
#pragma unroll
    for (int b = 0; b < 8; ++b)
    {
      const uint64_t m = __ballot ((digit >> b) & 1u);
      peers &= ((digit >> b) & 1u) ? m : ~m;
    }
    const int rank = __popcll (peers & lt);
    const int cnt = __popcll (peers);
#pragma unroll
    for (int q = 0; q < W; ++q) wcnt[q][threadIdx.x] = 0;
    __syncthreads ();
    if (valid && rank == 0) wcnt[w][digit] = cnt;
    __syncthreads ();
    if (valid)
    {
      int32_t pos = running[digit] + rank;
      for (int q = 0; q < w; ++q) pos += wcnt[q][digit];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads ();
    int32_t add = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) add += wcnt[q][threadIdx.x];
    running[threadIdx.x] += add;
    __syncthreads ();
  }
}

// =============================================================================================
// Gathers, Karras tree, refit, emit
// =============================================================================================
__global__ void k_gather_boxes (const int32_t *__restrict__ ids, const float *__restrict__ aabbs,
                                const int32_t *__restrict__ prims, int n, float *__restrict__ s_aabbs,
                                int32_t *__restrict__ s_prims)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t j = ids[i];
#pragma unroll
  for (int q = 0; q < 6; ++q) s_aabbs[6 * (int64_t)i + q] = aabbs[6 * j + q];
  s_prims[i] = prims[j];
}

__device__ __forceinline__ int clz_u32 (uint32_t x) { return x == 0 ? 32 : __clz ((int)x); }

__device__ __forceinline__ int karras_delta (int a, int b, int inner_size, const uint32_t *__restrict__ mc)
{
  const bool oor = (b < 0 || b > inner_size);
  const int bb = oor ? 0 : b;
  uint32_t ex = mc[a] ^ mc[bb];
  const bool tie = (ex == 0);
  ex = tie ? ((uint32_t)a ^ (uint32_t)bb) : ex;
  int count = clz_u32 (ex);
  if (tie) count += 32;
  return oor ? -1 : count;
}

// Karras 2012 (linear_bvh_builder.cpp:168-251), including the float ceil of the split search.
__global__ void k_build_tree (const uint32_t *__restrict__ mc, int inner, int32_t *__restrict__ left,
                              int32_t *__restrict__ right, int32_t *__restrict__ parent, int32_t *__restrict__ range_lo,
                              int32_t *__restrict__ range_hi)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= inner) return;
  const int d = 0 > (karras_delta (i, i + 1, inner, mc) - karras_delta (i, i - 1, inner, mc)) ? -1 : 1;
  const int min_delta = karras_delta (i, i - d, inner, mc);
  int lmax = 2;
  while (karras_delta (i, i + lmax * d, inner, mc) > min_delta) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (karras_delta (i, i + (l + t) * d, inner, mc) > min_delta) l += t;
  const int j = i + l * d;
  const int delta_node = karras_delta (i, j, inner, mc);
  int s = 0;
  float div = 2.f;
  for (int t = (int)ceilf ((float)l / div);; div *= 2, t = (int)ceilf ((float)l / div))
  {
    if (karras_delta (i, i + (s + t) * d, inner, mc) > delta_node) s += t;
    if (t == 1) break;
  }
  const int split = i + s * d + min (d, 0);
  range_lo[i] = min (i, j);   // node i covers sorted leaves [min(i, j), max(i, j)]
  range_hi[i] = max (i, j);
  if (min (i, j) == split)
  {
    parent[split + inner] = i;
    left[i] = split + inner;
  }
  else
  {
    parent[split] = i;
    left[i] = split;
  }
  if (max (i, j) == split + 1)
  {
    parent[split + inner + 1] = i;
    right[i] = split + inner + 1;
  }
  else
  {
    parent[split + 1] = i;
    right[i] = split + 1;
  }
  if (i == 0) parent[0] = -1;
}

__global__ void k_fill_i32 (int32_t *p, int64_t n, int32_t v)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ __forceinline__ void box_union (float *o, const float *a, const float *b)
{
  // AABB::include of two non-empty ranges per axis (aabb.hpp:46-51): min/max are order independent.
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    float lo = INFINITY, hi = -INFINITY;
    if (!(a[2 * d] > a[2 * d + 1]))
    {
      lo = fminf (lo, a[2 * d]);
      hi = fmaxf (hi, a[2 * d]);
      lo = fminf (lo, a[2 * d + 1]);
      hi = fmaxf (hi, a[2 * d + 1]);
    }
    if (!(b[2 * d] > b[2 * d + 1]))
    {
      lo = fminf (lo, b[2 * d]);
      hi = fmaxf (hi, b[2 * d]);
      lo = fminf (lo, b[2 * d + 1]);
      hi = fmaxf (hi, b[2 * d + 1]);
    }
    o[2 * d] = lo;
    o[2 * d + 1] = hi;
  }
}

// propagate_aabbs (linear_bvh_builder.cpp:253-317) without atomics.  A node's box is the exact min/max
// union of the boxes of its contiguous range of sorted leaves (the value propagate_aabbs computes, since
// min/max do not depend on order).  A
// pyramid of pairwise unions over the sorted leaf boxes (level l entry i covers leaves [i 2^l, (i+1) 2^l))
// answers each node's range with at most 2 log2(n) reads, so every node is computed independently.
__device__ __forceinline__ void box_include (float *acc, const float *x)
{
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (!(x[2 * d] > x[2 * d + 1]))
    {
      acc[2 * d] = fminf (acc[2 * d], x[2 * d]);
      acc[2 * d + 1] = fmaxf (acc[2 * d + 1], x[2 * d]);
      acc[2 * d] = fminf (acc[2 * d], x[2 * d + 1]);
      acc[2 * d + 1] = fmaxf (acc[2 * d + 1], x[2 * d + 1]);
    }
}

__global__ void k_pyramid_level (const float *__restrict__ src, int n_src, float *__restrict__ dst)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (2 * i >= n_src) return;
  float acc[6] = { INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY };
  box_include (acc, src + 12 * (int64_t)i);
  if (2 * i + 1 < n_src) box_include (acc, src + 12 * (int64_t)i + 6);
#pragma unroll
  for (int q = 0; q < 6; ++q) dst[6 * (int64_t)i + q] = acc[q];
}

constexpr int kMaxLevels = 32;
struct Pyramid
{
  const float *level[kMaxLevels];
};

__global__ void k_refit_ranges (Pyramid pyr, const int32_t *__restrict__ range_lo, const int32_t *__restrict__ range_hi,
                                int inner, float *__restrict__ inner_aabbs)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= inner) return;
  float acc[6] = { INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY };
  int a = range_lo[i], b = range_hi[i] + 1;
  for (int l = 0; a < b; ++l, a >>= 1, b >>= 1)
  {
    if (a & 1) box_include (acc, pyr.level[l] + 6 * (int64_t)(a++));
    if (b & 1) box_include (acc, pyr.level[l] + 6 * (int64_t)(--b));
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) inner_aabbs[6 * (int64_t)i + q] = acc[q];
}

// emit (linear_bvh_builder.cpp:319-397)
__global__ void k_emit (const int32_t *__restrict__ left, const int32_t *__restrict__ right,
                        const float *__restrict__ leaf_aabbs, const float *__restrict__ inner_aabbs, int inner,
                        float4 *__restrict__ out)
{
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= inner) return;
  int lc = left[node], rc = right[node];
  const float *la, *ra;
  if (lc >= inner)
  {
    la = leaf_aabbs + 6 * (int64_t)(lc - inner);
    lc = -(lc - inner + 1);
  }
  else
  {
    la = inner_aabbs + 6 * (int64_t)lc;
    lc *= 4;
  }
  if (rc >= inner)
  {
    ra = leaf_aabbs + 6 * (int64_t)(rc - inner);
    rc = -(rc - inner + 1);
  }
  else
  {
    ra = inner_aabbs + 6 * (int64_t)rc;
    rc *= 4;
  }
  float4 *o = out + 4 * (int64_t)node;
  o[0] = make_float4 (la[0], la[2], la[4], la[1]);
  o[1] = make_float4 (la[3], la[5], ra[0], ra[2]);
  o[2] = make_float4 (ra[4], ra[1], ra[3], ra[5]);
  o[3] = make_float4 (__int_as_float (lc), __int_as_float (rc), 0.f, 0.f);
}

__global__ void k_pack_leaves (const int32_t *__restrict__ leaf_nodes, const int32_t *__restrict__ aabb_ids,
                               const float *__restrict__ subrefs, int n, float4 *__restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float *s = subrefs + 6 * (int64_t)aabb_ids[i];
  out[2 * (int64_t)i] = make_float4 (__int_as_float (leaf_nodes[i]), s[0], s[1], s[2]);
  out[2 * (int64_t)i + 1] = make_float4 (s[3], s[4], s[5], 0.f);
}

static inline unsigned nblk (int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

int radix_tiles (int n) { return (n + kSortTile - 1) / kSortTile; }

int radix_sort_pairs (uint32_t *kA, int32_t *vA, uint32_t *kB, int32_t *vB, int32_t *hist, int n, int bits,
                      hipStream_t st, uint32_t **keys_out, int32_t **vals_out, int32_t *scan_sums)
{
  const int n_tiles = radix_tiles (n);
  uint32_t *ks = kA, *kd = kB;
  int32_t *vs = vA, *vd = vB;
  int32_t *offs = hist + 256 * (size_t)n_tiles;
  for (int shift = 0; shift < bits; shift += 8)
  {
    hipLaunchKernelGGL (k_radix_hist, dim3 (n_tiles), dim3 (kSortBlock), 0, st, ks, n, shift, n_tiles, hist);
    DRX_LAUNCH_CHECK ();
    int rc = exclusive_scan_i32 (hist, offs, 256 * n_tiles, nullptr, st, scan_sums);
    if (rc != DRX_OK) return rc;
    hipLaunchKernelGGL (k_radix_scatter, dim3 (n_tiles), dim3 (kSortBlock), 0, st, ks, vs, n, shift, n_tiles, offs, kd,
                        vd);
    DRX_LAUNCH_CHECK ();
    std::swap (ks, kd);
    std::swap (vs, vd);
  }
  *keys_out = ks;
  *vals_out = vs;
  return DRX_OK;
}

// Morton codes of query points in the normalised frame of `bounds` (the LBVH's 30-bit quantisation),
// kept to their top kPointMortonBits bits: the order only groups nearby points for coherent walks (any
// order gives the same results), and 24 bits (256 cells per axis) sort in three radix passes, not four.
#ifndef DRX_POINT_MORTON_BITS
#define DRX_POINT_MORTON_BITS 24
#endif
constexpr int kPointMortonBits = DRX_POINT_MORTON_BITS;

__global__ void k_point_morton (const float *__restrict__ pts, int n, float minx, float miny, float minz, float ix,
                                float iy, float iz, uint32_t *__restrict__ keys, int32_t *__restrict__ vals)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float mins[3] = { minx, miny, minz }, inv[3] = { ix, iy, iz };
  uint32_t q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    float x = (pts[3 * (int64_t)i + d] - mins[d]) * inv[d];
    x = fminf (fmaxf (x * 1024.0f, 0.0f), 1023.0f);
    q[d] = (uint32_t)x;
  }
  keys[i] = ((expand_bits32 (q[2]) << 2) | (expand_bits32 (q[1]) << 1) | expand_bits32 (q[0])) >> (30 - kPointMortonBits);
  vals[i] = i;
}

int morton_order_points (const float *pts, int n, const float *bounds6, hipStream_t st, const int32_t **order_out,
                         const SortScratch *ws)
{
  uint32_t *kA, *kB;
  int32_t *vA, *vB, *hist, *sums = nullptr;
  if (ws)
  {
    kA = ws->kA, kB = ws->kB, vA = ws->vA, vB = ws->vB, hist = ws->hist, sums = ws->scan_sums;
  }
  else
  {
    kA = (uint32_t *)arena (SLOT_SORT_KEYS_A, sizeof (uint32_t) * (size_t)n, st);
    kB = (uint32_t *)arena (SLOT_SORT_KEYS_B, sizeof (uint32_t) * (size_t)n, st);
    vA = (int32_t *)arena (SLOT_SORT_VALS_A, sizeof (int32_t) * (size_t)n, st);
    vB = (int32_t *)arena (SLOT_SORT_VALS_B, sizeof (int32_t) * (size_t)n, st);
    hist = (int32_t *)arena (SLOT_SORT_HIST, sizeof (int32_t) * (size_t)radix_hist_len (n), st);
  }
  if (!kA || !kB || !vA || !vB || !hist) return DRX_ERR_HIP;
  float minc[3], inv[3];
  for (int d = 0; d < 3; ++d)
  {
    minc[d] = bounds6[2 * d];
    const float ext = bounds6[2 * d + 1] - bounds6[2 * d];
    inv[d] = (ext > 0.f) ? 1.f / ext : 0.f;
  }
  hipLaunchKernelGGL (k_point_morton, dim3 (nblk (n, 256)), dim3 (256), 0, st, pts, n, minc[0], minc[1], minc[2], inv[0],
                      inv[1], inv[2], kA, vA);
  DRX_LAUNCH_CHECK ();
  uint32_t *ks;
  int32_t *vs;
  int rc = radix_sort_pairs (kA, vA, kB, vB, hist, n, kPointMortonBits, st, &ks, &vs, sums);
  if (rc) return rc;
  *order_out = vs;
  return DRX_OK;
}

} // namespace drx

using namespace drx;

extern "C" {

int drx_bvh_split_elements (const int32_t *ctrl_idx, const float *values, const int32_t *orders, const int64_t *offsets,
                            int32_t order, int32_t n_el, float *out_aabbs, int32_t *out_prims, float *out_subrefs,
                            void *stream)
{
  if (n_el < 0) return fail (DRX_ERR_ARG, "drx_bvh_split_elements: n_el < 0");
  if (n_el == 0) return DRX_OK;
  if (!ctrl_idx || !values || !out_aabbs || !out_prims || !out_subrefs)
    return fail (DRX_ERR_ARG, "drx_bvh_split_elements: null pointer");
  if ((orders == nullptr) != (offsets == nullptr))
    return fail (DRX_ERR_ARG, "drx_bvh_split_elements: orders and offsets go together");
  const int pmax = orders ? DRX_MAX_MIXED_ORDER : order;
  if (pmax < 1 || pmax > DRX_MAX_ORDER)
    return fail (DRX_ERR_ARG, "drx_bvh_split_elements: order %d outside 1..%d", order, DRX_MAX_ORDER);
  hipStream_t st = as_stream (stream);
  float *gnets = nullptr;
  const int gn_blocks = n_el < kSplitGlobalBlocks ? n_el : kSplitGlobalBlocks;
  if (pmax >= 7)
  {
    const size_t npe = (size_t)(pmax + 1) * (pmax + 1) * (pmax + 1);
    gnets = (float *)arena (SLOT_TREE, sizeof (float) * kSplitBoxes * 3 * npe * gn_blocks, st);
    if (!gnets) return DRX_ERR_HIP;
  }
  switch (pmax)
  {
  case 1: hipLaunchKernelGGL (k_split<1>, dim3 (n_el), dim3 (64), 0, st, ctrl_idx, values, orders, offsets, order, n_el, out_aabbs, out_prims, out_subrefs); break;
  case 2: hipLaunchKernelGGL (k_split<2>, dim3 (n_el), dim3 (64), 0, st, ctrl_idx, values, orders, offsets, order, n_el, out_aabbs, out_prims, out_subrefs); break;
  case 3: hipLaunchKernelGGL (k_split<3>, dim3 (n_el), dim3 (64), 0, st, ctrl_idx, values, orders, offsets, order, n_el, out_aabbs, out_prims, out_subrefs); break;
  case 4: hipLaunchKernelGGL ((k_split_cw<4, false>), dim3 (n_el), dim3 (192), 0, st, ctrl_idx, values, orders, offsets, order, n_el, out_aabbs, out_prims, out_subrefs, nullptr); break;
  case 5: hipLaunchKernelGGL ((k_split_cw<5, false>), dim3 (n_el), dim3 (192), 0, st, ctrl_idx, values, orders, offsets, order, n_el, out_aabbs, out_prims, out_subrefs, nullptr); break;
  case 6: hipLaunchKernelGGL ((k_split_cw<6, false>), dim3 (n_el), dim3 (192), 0, st, ctrl_idx, values, orders, offsets, order, n_el, out_aabbs, out_prims, out_subrefs, nullptr); break;
  case 7: hipLaunchKernelGGL ((k_split_cw<7, true>), dim3 (gn_blocks), dim3 (192), 0, st, ctrl_idx, values, orders, offsets, order, n_el, out_aabbs, out_prims, out_subrefs, gnets); break;
  default: hipLaunchKernelGGL ((k_split_cw<8, true>), dim3 (gn_blocks), dim3 (192), 0, st, ctrl_idx, values, orders, offsets, order, n_el, out_aabbs, out_prims, out_subrefs, gnets); break;
  }
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

int drx_bvh_build (const float *aabbs_in, const int32_t *prims_in, int32_t n_in, float *out_inner,
                   int32_t *out_leaf_nodes, int32_t *out_aabb_ids, float *out_bounds, uint32_t *out_mcodes,
                   void *stream)
{
  if (n_in < 0) return fail (DRX_ERR_ARG, "drx_bvh_build: n < 0");
  if (!out_inner || !out_leaf_nodes || !out_aabb_ids || !out_bounds)
    return fail (DRX_ERR_ARG, "drx_bvh_build: null output");
  if (n_in > 0 && (!aabbs_in || !prims_in)) return fail (DRX_ERR_ARG, "drx_bvh_build: null input");
  if (n_in > (1 << 30)) return fail (DRX_ERR_ARG, "drx_bvh_build: %d boxes exceed the int32 tree index range", n_in);
  hipStream_t st = as_stream (stream);
  const float *aabbs = aabbs_in;
  const int32_t *prims = prims_in;
  const bool zero_case = (n_in == 0);
  const int n = n_in < 2 ? 2 : n_in;
  // dray pads 0 and 1 boxes to two (linear_bvh_builder.cpp:411-456)
  float hb[12] = { 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0 };
  int32_t hp[2] = { 0, 0 };
  if (n_in < 2)
  {
    if (n_in == 1)
    {
      DRX_HIP (hipMemcpyAsync (hb, aabbs_in, 6 * sizeof (float), hipMemcpyDeviceToHost, st));
      DRX_HIP (hipMemcpyAsync (hp, prims_in, sizeof (int32_t), hipMemcpyDeviceToHost, st));
      DRX_HIP (hipStreamSynchronize (st));
      hp[1] = hp[0];
    }
    float *pa = (float *)arena (SLOT_TREE2, 12 * sizeof (float) + 2 * sizeof (int32_t), st);
    if (!pa) return DRX_ERR_HIP;
    DRX_HIP (hipMemcpyAsync (pa, hb, 12 * sizeof (float), hipMemcpyHostToDevice, st));
    DRX_HIP (hipMemcpyAsync (pa + 12, hp, 2 * sizeof (int32_t), hipMemcpyHostToDevice, st));
    DRX_HIP (hipStreamSynchronize (st));
    aabbs = pa;
    prims = (const int32_t *)(pa + 12);
  }
  const int inner = n - 1;
  // scratch: keys/vals ping-pong, histogram, tree arrays
  const int n_tiles = (n + kSortTile - 1) / kSortTile;
  uint32_t *kA = (uint32_t *)arena (SLOT_SORT_KEYS_A, sizeof (uint32_t) * n, st);
  uint32_t *kB = (uint32_t *)arena (SLOT_SORT_KEYS_B, sizeof (uint32_t) * n, st);
  int32_t *vA = (int32_t *)arena (SLOT_SORT_VALS_A, sizeof (int32_t) * n, st);
  int32_t *vB = (int32_t *)arena (SLOT_SORT_VALS_B, sizeof (int32_t) * n, st);
  int32_t *hist = (int32_t *)arena (SLOT_SORT_HIST, sizeof (int32_t) * (2 * 256 * (size_t)n_tiles + 8), st);
  // tree slot: sorted aabbs (6n) | inner aabbs (6 inner) | left | right | parent (2n-1) | sorted prims | partial bounds
  const int red_blocks = std::min (1024, (int)nblk (n, kRedBlock));
  const size_t tree_floats = 6 * (size_t)n + 6 * (size_t)inner;
  const size_t tree_ints = (size_t)inner * 2 + (2 * (size_t)n - 1) + (size_t)n;
  char *tree = (char *)arena (SLOT_TREE, tree_floats * 4 + tree_ints * 4 + 6 * 4 * (size_t)red_blocks + 128, st);
  if (!kA || !kB || !vA || !vB || !hist || !tree) return DRX_ERR_HIP;
  float *s_aabbs = (float *)tree;
  float *i_aabbs = s_aabbs + 6 * (size_t)n;
  int32_t *left = (int32_t *)(i_aabbs + 6 * (size_t)inner);
  int32_t *right = left + inner;
  int32_t *parent = right + inner;
  int32_t *s_prims = parent + (2 * (size_t)n - 1);
  float *partial = (float *)(s_prims + n);

  // bounds, reduced on the device (no host round trip before the Morton codes); copied out at the end
  float *dbounds = partial + 6 * (size_t)red_blocks;   // bounds[6], Morton frame (mins, 1 / extent)[6]
  if (!zero_case)
  {
    hipLaunchKernelGGL (k_bounds_partial, dim3 (red_blocks), dim3 (kRedBlock), 0, st, aabbs, (int64_t)n, partial);
    DRX_LAUNCH_CHECK ();
  }
  hipLaunchKernelGGL (k_bounds_final, dim3 (1), dim3 (kRedBlock), 0, st, partial, zero_case ? 0 : red_blocks, dbounds);
  DRX_LAUNCH_CHECK ();
  hipLaunchKernelGGL (k_morton, dim3 (nblk (n, 256)), dim3 (256), 0, st, aabbs, n, dbounds + 6, kA, vA);
  DRX_LAUNCH_CHECK ();
  // stable radix sort, 8-bit digits over 30-bit keys
  uint32_t *ks = nullptr;
  int32_t *vs = nullptr;
  {
    int rc = radix_sort_pairs (kA, vA, kB, vB, hist, n, 30, st, &ks, &vs);
    if (rc != DRX_OK) return rc;
  }
  // ks = sorted codes, vs = ids
  DRX_HIP (hipMemcpyAsync (out_aabb_ids, vs, sizeof (int32_t) * n, hipMemcpyDeviceToDevice, st));
  if (out_mcodes) DRX_HIP (hipMemcpyAsync (out_mcodes, ks, sizeof (uint32_t) * n, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL (k_gather_boxes, dim3 (nblk (n, 256)), dim3 (256), 0, st, vs, aabbs, prims, n, s_aabbs, s_prims);
  DRX_LAUNCH_CHECK ();
  DRX_HIP (hipMemcpyAsync (out_leaf_nodes, s_prims, sizeof (int32_t) * n, hipMemcpyDeviceToDevice, st));
  // pyramid levels 1.. over the sorted leaf boxes (level 0), plus the node ranges
  size_t pyr_floats = 0;
  int n_levels = 1;
  for (int64_t m = n; m > 1; m = (m + 1) / 2)
  {
    pyr_floats += 6 * (size_t)((m + 1) / 2);
    ++n_levels;
  }
  if (n_levels > kMaxLevels) return fail (DRX_ERR_ARG, "drx_bvh_build: too many boxes for the refit pyramid");
  char *pyr_mem = (char *)arena (SLOT_PYRAMID, pyr_floats * sizeof (float) + 2 * sizeof (int32_t) * (size_t)inner + 64, st);
  if (!pyr_mem) return DRX_ERR_HIP;
  int32_t *range_lo = (int32_t *)pyr_mem;
  int32_t *range_hi = range_lo + inner;
  float *levels = (float *)(range_hi + inner);
  hipLaunchKernelGGL (k_build_tree, dim3 (nblk (inner, 256)), dim3 (256), 0, st, ks, inner, left, right, parent,
                      range_lo, range_hi);
  Pyramid pyr;
  pyr.level[0] = s_aabbs;
  {
    float *dst = levels;
    int l = 1;
    for (int64_t m = n; m > 1; m = (m + 1) / 2, ++l)
    {
      hipLaunchKernelGGL (k_pyramid_level, dim3 (nblk ((m + 1) / 2, 256)), dim3 (256), 0, st, pyr.level[l - 1], (int)m, dst);
      pyr.level[l] = dst;
      dst += 6 * (size_t)((m + 1) / 2);
    }
    for (; l < kMaxLevels; ++l) pyr.level[l] = nullptr;
  }
  hipLaunchKernelGGL (k_refit_ranges, dim3 (nblk (inner, 256)), dim3 (256), 0, st, pyr, range_lo, range_hi, inner, i_aabbs);
  hipLaunchKernelGGL (k_emit, dim3 (nblk (inner, 256)), dim3 (256), 0, st, left, right, s_aabbs, i_aabbs, inner,
                      (float4 *)out_inner);
  DRX_LAUNCH_CHECK ();
  DRX_HIP (hipMemcpyAsync (out_bounds, dbounds, 6 * sizeof (float), hipMemcpyDeviceToHost, st));
  DRX_HIP (hipStreamSynchronize (st));
  return DRX_OK;
}

int drx_bvh_pack_leaves (const drx_bvh *bvh, float *out_leaf_pack, void *stream)
{
  if (!bvh || !out_leaf_pack || !bvh->leaf_nodes || !bvh->aabb_ids || !bvh->subrefs)
    return fail (DRX_ERR_ARG, "drx_bvh_pack_leaves: null pointer");
  hipLaunchKernelGGL (k_pack_leaves, dim3 (nblk (bvh->n_leaves, 256)), dim3 (256), 0, as_stream (stream), bvh->leaf_nodes,
                      bvh->aabb_ids, bvh->subrefs, bvh->n_leaves, (float4 *)out_leaf_pack);
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

} // extern "C"

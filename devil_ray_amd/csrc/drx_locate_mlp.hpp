// Point location with the learned Newton initializer.  Same traversal and candidate order as
// locate_point (DeviceMesh::locate, device_mesh.hpp:247-354); at each candidate the wavefront runs the
// MLP for all its lanes at once (MFMA needs the whole wave), then each lane runs Newton from its
// learned guess with a small step budget and, if that run has not converged, the plain method's Newton
// from its own start (candidate_start), which also decides acceptance in that case.  The walk is the
// plain method's (next_candidate_leaf, over the compact nodes when the BVH has them).
#pragma once

#include "drx_element.hpp"
#include "drx_mlp.hpp"

namespace drx
{

template <int P, bool POLISH>
__device__ __forceinline__ bool try_element_learned (const LocateCfg &cfg, const float *g, const float pt[3],
                                                     float x[3], const float xc[3], int budget, LocateStats &st)
{
  const float t[3] = { pt[0] - g[0], pt[1] - g[1], pt[2] - g[2] };
  NewtonOut no;
  bool found = false;
  st.candidates += 1;
  // pass 0: from the learned guess within `budget` steps; pass 1 (only if pass 0 did not converge):
  // the plain method's Newton from its own start xc (candidate_start) with dray's 20-step budget
#pragma unroll 1
  for (int pass = 0; pass < 2; ++pass)
  {
    if (pass == 1)
    {
      x[0] = xc[0];
      x[1] = xc[1];
      x[2] = xc[2];
      st.fallbacks += 1;
    }
    found = newton_f32<P, false> (g, t, x, no, POLISH, pass == 0 ? budget : 20);
    st.newton += no.steps;
    if (no.conv) break;
  }
  if constexpr (POLISH)
  {
    if (found && !(no.cond <= cfg.cond_threshold))
    {
      const float3 r = polish_f64_call<P> (g, pt[0], pt[1], pt[2], x[0], x[1], x[2]);
      x[0] = r.x;
      x[1] = r.y;
      x[2] = r.z;
      st.polished += 1;
    }
  }
  return found;
}

template <bool POLISH>
__device__ __forceinline__ bool try_element_learned_mixed (const LocateCfg &cfg, int order, const float *g,
                                                           const float pt[3], float x[3], const float xc[3], int budget,
                                                           LocateStats &st)
{
  switch (order)
  {
  case 1: return try_element_learned<1, POLISH> (cfg, g, pt, x, xc, budget, st);
  case 2: return try_element_learned<2, POLISH> (cfg, g, pt, x, xc, budget, st);
  case 3: return try_element_learned<3, POLISH> (cfg, g, pt, x, xc, budget, st);
  default: return try_element_learned<4, POLISH> (cfg, g, pt, x, xc, budget, st);
  }
}

// One candidate per call, resumable like locate_point_step (start a point with cur = 0, sp = 0 and kBarrier
// in stack entry 0).  Every lane of the wavefront calls it, with have == false when it holds no point,
// because the MLP round is wavefront-wide.  Returns 1 if the candidate was accepted (ref, cell set), 0 if it
// was rejected (call again with the same point), -1 when the search is exhausted or the stack would
// overflow (overflow set) or the lane holds no point.
template <int P, bool POLISH>
__device__ __forceinline__ int locate_point_mlp_step (const LocateCfg &cfg, TravStack &todo, const float pt[3],
                                                      float ref[3], LocateStats &st, bool &overflow, bool have,
                                                      int &cur, int &sp, int &cell, const MlpNet &net, int budget,
                                                      __bf16 *lx, float *ly)
{
  int leaf = -1;
  float4 l0 = make_float4 (0.f, 0.f, 0.f, 0.f), l1 = l0, map[3] = { l0, l0, l0 };
  if (have)
  {
    leaf = next_candidate_leaf (cfg, todo, pt, cur, sp, l0, l1, map);
    if (leaf == kWalkOverflow) overflow = true;
  }
  const bool has = leaf >= 0;
  if (__ballot (has) == 0) return -1;
  float f[mlp::IN_PAD], base[3] = { 0.5f, 0.5f, 0.5f }, xc[3] = { 0.5f, 0.5f, 0.5f };
  const float *g = cfg.mesh;
  int el = -1, order = P;
  if (has)
  {
    el = __float_as_int (l0.x);
    candidate_start_rows<P> (cfg, el, map, pt, l0, l1, xc);   // the fallback start of the plain method
    if constexpr (P == 0)
    {
      g = cfg.mesh + cfg.offsets[el];
      order = cfg.orders[el];
      mlp_features_mixed (order, g, pt, f, base);
    }
    else
    {
      g = cfg.mesh + (int64_t)el * Ord<P>::GSTRIDE;
      mlp_features<P> (g, pt, f, base);
    }
  }
  else
  {
#pragma unroll
    for (int k = 0; k < mlp::IN_PAD; ++k) f[k] = 0.f;
  }
  float corr[3];
  mlp_wave_infer (net, f, has, corr, lx, ly);
  if (!has) return -1;
  float x[3] = { base[0] + corr[0], base[1] + corr[1], base[2] + corr[2] };
  bool found;
  if constexpr (P == 0)
    found = try_element_learned_mixed<POLISH> (cfg, order, g, pt, x, xc, budget, st);
  else
    found = try_element_learned<P, POLISH> (cfg, g, pt, x, xc, budget, st);
  if (found)
  {
    ref[0] = x[0];
    ref[1] = x[1];
    ref[2] = x[2];
    cell = el;
    return 1;
  }
  cur = todo.get (sp);
  --sp;
  return 0;
}

} // namespace drx

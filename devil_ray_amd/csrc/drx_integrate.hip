// Volume::integrate on gfx950: the per-ray march (locate + field / lit shading + LUT + blend, <= 5
// partials per ray) and partial compaction.
#include "drx_common.hpp"
#include "drx_element.hpp"
#include "drx_locate_mlp.hpp"
#include "drx_query_common.hpp"
#include "drx_wave.hpp"

#include <cmath>

namespace drx
{

// ---------------------------------------------------------------------------------------------
// Volume integration: integrate_partials' per-ray loop (rendering/volume.cpp:134-217).
// ---------------------------------------------------------------------------------------------
struct IntegrateArgs
{
  LocateCfg cfg;
  const float *field;
  const double *field64;       // FP64 field (SHADE 2)
  const int64_t *field_offsets;
  const Ray *rays;
  const int32_t *order;
  int n_rays;
  const float4 *lut;
  int lut_size;
  float fmin, fmax, inv_range;
  double fmin64, fmax64, inv_range64;   // the colour map's range for an FP64 field
  float sample_dist;
  Partial *slots;
  unsigned long long *stats;   // located, newton, candidates, polished
  int32_t *err;
  MlpNet net;                  // learned initial guess (MODE 3)
  int budget;                  // Newton steps from the learned guess before the plain fallback
  int n_lights;                // lit shading (LIT kernels)
  drx_point_light lights[DRX_MAX_LIGHTS];
};

template <int P, typename FT = float>
__device__ __forceinline__ FT field_at (const IntegrateArgs &a, const FT *field, int cell, const float ref[3])
{
  if constexpr (P == 0)
  {
    const FT *f = field + a.field_offsets[cell];
    switch (a.cfg.orders[cell])
    {
    case 1: return eval_scalar<1> (f, ref);
    case 2: return eval_scalar<2> (f, ref);
    case 3: return eval_scalar<3> (f, ref);
    default: return eval_scalar<4> (f, ref);
    }
  }
  else
    return eval_scalar<P> (field + (int64_t)cell * Ord<P>::FSTRIDE, ref);
}

// MODE: 0 binary per-lane traversal, 1 8-wide per-lane, 2 8-wide wave-coherent (identical results).
__device__ __forceinline__ float dot3 (const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ void normalize3 (float v[3])
{
  const float m = sqrtf (dot3 (v, v));
  v[0] /= m;
  v[1] /= m;
  v[2] /= m;
}
__device__ __forceinline__ float clamp01f (float v) { return fminf (1.f, fmaxf (0.f, v)); }

// VolumeShader::scalar_gradient + shaded_color (rendering/volume_shader.hpp:39-158) for order P:
// physical gradient = ref gradient x J^-1 (three LUP solves, like matrix_inverse), then per light
// ambient + diffuse + Blinn-Phong specular, gradient flipped toward the light and the half vector
// (flips persist across lights), channels clamped after each light; alpha <= 0.01 gives zero.
template <int P>
__device__ float4 lit_color_at (const IntegrateArgs &a, const float *g, const float *f, const float ref[3],
                                const Ray &ray)
{
  float y[3], J[3][3];
  eval_geom<P, float> (g, ref, y, J);
  float dref[3];
  const float scalar = eval_scalar_d<P> (f, ref, dref);
  float inv[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    float A[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) A[r][q] = J[r][q];
    const float e[3] = { c == 0 ? 1.f : 0.f, c == 1 ? 1.f : 0.f, c == 2 ? 1.f : 0.f };
    float x[3];
    lup_solve3<float> (A, e, x);
    inv[0][c] = x[0];
    inv[1][c] = x[1];
    inv[2][c] = x[2];
  }
  float grad[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) grad[c] = dref[0] * inv[0][c] + dref[1] * inv[1][c] + dref[2] * inv[2][c];
  const float4 col = color_lookup (a.lut, a.lut_size, a.fmin, a.fmax, a.inv_range, scalar);
  float4 acc = make_float4 (0.f, 0.f, 0.f, 0.f);
  if ((double)col.w > 0.01)
  {
    normalize3 (grad);
    const float wp[3] = { g[0] + y[0], g[1] + y[1], g[2] + y[2] };
    const float view[3] = { -ray.dir[0], -ray.dir[1], -ray.dir[2] };
    for (int l = 0; l < a.n_lights; ++l)
    {
      const drx_point_light &L = a.lights[l];
      float ld[3] = { L.pos[0] - wp[0], L.pos[1] - wp[1], L.pos[2] - wp[2] };
      normalize3 (ld);
      if (dot3 (ld, grad) < 0.f)
#pragma unroll
        for (int i = 0; i < 3; ++i) grad[i] = -grad[i];
      const float diffuse = clamp01f (dot3 (ld, grad));
      float sc[3] = { L.amb[0] * col.x, L.amb[1] * col.y, L.amb[2] * col.z };
      sc[0] += diffuse * L.diff[0] * col.x;
      sc[1] += diffuse * L.diff[1] * col.y;
      sc[2] += diffuse * L.diff[2] * col.z;
      float hv[3] = { 0.5f * (view[0] + ld[0]), 0.5f * (view[1] + ld[1]), 0.5f * (view[2] + ld[2]) };
      normalize3 (hv);
      if (dot3 (grad, hv) < 0.f)
#pragma unroll
        for (int i = 0; i < 3; ++i) grad[i] = -grad[i];
      const float doth = clamp01f (dot3 (grad, hv));
      const float intensity = powf (doth, L.spec_pow);
      sc[0] += intensity * L.spec[0] * col.x;
      sc[1] += intensity * L.spec[1] * col.y;
      sc[2] += intensity * L.spec[2] * col.z;
      acc.x = clamp01f (acc.x + sc[0]);
      acc.y = clamp01f (acc.y + sc[1]);
      acc.z = clamp01f (acc.z + sc[2]);
      acc.w += col.w;
    }
  }
  return acc;
}

template <int P>
__device__ __forceinline__ float4 lit_color (const IntegrateArgs &a, int cell, const float ref[3], const Ray &ray)
{
  if constexpr (P == 0)
  {
    const float *g = a.cfg.mesh + a.cfg.offsets[cell];
    const float *f = a.field + a.field_offsets[cell];
    switch (a.cfg.orders[cell])
    {
    case 1: return lit_color_at<1> (a, g, f, ref, ray);
    case 2: return lit_color_at<2> (a, g, f, ref, ray);
    case 3: return lit_color_at<3> (a, g, f, ref, ray);
    default: return lit_color_at<4> (a, g, f, ref, ray);
    }
  }
  else
    return lit_color_at<P> (a, a.cfg.mesh + (int64_t)cell * Ord<P>::GSTRIDE, a.field + (int64_t)cell * Ord<P>::FSTRIDE,
                            ref, ray);
}

#ifndef DRX_INTEGRATE_BLOCK
#define DRX_INTEGRATE_BLOCK 64
#endif
constexpr int kIntBlock = DRX_INTEGRATE_BLOCK;   // threads per workgroup of the integrate kernel
// DRX_TRAVERSE_AUTO in drx_volume_integrate: the wave walk for BVHs of at most this many leaves and frames
// of at least this many active rays per SIMD, else the binary walk (measured; see drx_volume_integrate).
constexpr int kWaveWalkMaxLeaves = 1 << 16;
constexpr int kWaveWalkMinRaysPerSimd = 48;
// DRX_LAUNCH_AUTO in drx_volume_integrate: the sample-parallel kernel below this many active rays per SIMD
// (measured on MI355X: 1.9-3.5x faster than one lane per ray at 6-150 rays per SIMD, 3% faster at 286,
// 10% slower at 374, the 1024^2 frame of a 64^3 order-3 mesh).
#ifndef DRX_SPLIT_MAX_RAYS_PER_SIMD
#define DRX_SPLIT_MAX_RAYS_PER_SIMD 300
#endif
constexpr int kSplitMaxRaysPerSimd = DRX_SPLIT_MAX_RAYS_PER_SIMD;
constexpr int kIntWaves = kIntBlock / 64;

// SHADE: 0 VolumeShader::color over an FP32 field, 1 shaded_color (lit), 2 color over an FP64 field.
template <int P, bool POLISH, int MODE, int SHADE>
__global__ void __launch_bounds__ (kIntBlock, waves_per_simd (P, POLISH, MODE)) k_integrate (IntegrateArgs a)
{
  constexpr bool LIT = SHADE == 1;
  constexpr bool WAVE = MODE == 2;
  constexpr bool LEARNED = MODE == 3;   // learned Newton start; the whole wavefront stays in the loop
  __shared__ int32_t stack_lds[WAVE ? 1 : kStackLds * kIntBlock];
  __shared__ int32_t w_code[WAVE ? kIntWaves * kWaveStack : 1];
  __shared__ uint64_t w_mask[WAVE ? kIntWaves * kWaveStack : 1];
  __shared__ int32_t w_cand[WAVE ? kIntWaves * kCandQ * 64 : 1];
  __shared__ __attribute__ ((aligned (16))) __bf16 mlp_x[LEARNED ? kIntWaves * 64 * mlp::IN_PAD : 8];
  __shared__ __attribute__ ((aligned (16))) float mlp_y[LEARNED ? kIntWaves * 64 * 4 : 4];
  // Element blocks staged in LDS for Newton: order 3 only.  Measured per order (binary walk): order 3
  // 15.5 -> 15.2 ms on the 64^3 frame; order 4 (two slots fit) 211 -> 234 ms and order 2 on a frame of one
  // wavefront per SIMD 4.3 -> 5.1 ms, where the copy and the wave synchronisation cost more than they save.
  constexpr bool STAGE = (MODE == 0 || MODE == 4) && P == 3;
  constexpr bool GUIDED = MODE == 4;    // binary walk with given starts and / or the record of every sample
  __shared__ __attribute__ ((aligned (16))) float stage_lds[STAGE ? kIntWaves * Stage<(P >= 1 ? P : 1)>::FLOATS : 4];
  WaveScratch ws;
  const int wv = threadIdx.x >> 6;
  {
    ws.st_code = w_code + (WAVE ? wv * kWaveStack : 0);
    ws.st_mask = w_mask + (WAVE ? wv * kWaveStack : 0);
    ws.cand = w_cand + (WAVE ? wv * kCandQ * 64 : 0);
  }
  TravStack todo;
  todo.lds = stack_lds;
  todo.bdim = blockDim.x;
  todo.tid = threadIdx.x;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool has_ray = t < a.n_rays;
  unsigned long long located = 0;
  LocateStats st { 0, 0, 0, 0 };
  bool overflow = false;
  if (LEARNED || has_ray)
  {
    const int ri = has_ray ? (a.order ? a.order[t] : t) : 0;
    Ray ray;
    if (has_ray)
      ray = a.rays[ri];
    else
    {
      ray.dir[0] = ray.dir[1] = ray.dir[2] = 0.f;
      ray.orig[0] = ray.orig[1] = ray.orig[2] = 0.f;
      ray.near_ = ray.far_ = 0.f;
      ray.pixel_id = -1;
    }
    Partial *out5 = a.slots + 5 * (int64_t)ri;
    if (has_ray)
#pragma unroll
      for (int s = 0; s < 5; ++s)
      {
        Partial p0;
        p0.pixel_id = -1;
        p0.depth = INFINITY;
        p0.rgba[0] = p0.rgba[1] = p0.rgba[2] = p0.rgba[3] = 0.f;
        out5[s] = p0;
      }
    // integrate_partials' per-ray loop as a state machine with one locate() per iteration:
    //   search: locate at `distance`; on a miss step on; on a hit open a segment and shade the sample;
    //   inside: locate at the advanced distance; keep shading while inside, before `far` and alpha < 0.95,
    //           else close the segment (<= 5 per ray); the next search re-locates at the same distance.
    float distance = ray.near_ + a.sample_dist;
    int segment = 0;
    bool inseg = false;
    float depth = 0.f;
    float rgba[4] = { 0.f, 0.f, 0.f, 0.f };
    bool alive = has_ray;
    // Binary and 8-wide walks: one candidate per iteration.  A lane whose candidate was rejected keeps its
    // depth-first state and resumes at the same sample point in the next iteration, while the other
    // lanes go on with their next samples, so the wavefront never runs a Newton round for a few lanes.
    int tcur = 0, tsp = 0;
    bool resume = false;
    int sample = 0;   // GUIDED: the number of locates this ray has finished
    Location given;   // GUIDED: the given start of the current sample
    given.cell = -1;
    for (;;)
    {
      if (alive && !inseg && !(distance < ray.far_)) alive = false;
      if constexpr (LEARNED)
      {
        if (__ballot (alive) == 0) break;
      }
      else if (!alive)
        break;
      const float pt[3] = { ray_point (ray.orig[0], distance, ray.dir[0]), ray_point (ray.orig[1], distance, ray.dir[1]),
                            ray_point (ray.orig[2], distance, ray.dir[2]) };
      float ref[3];
      int cell;
      if constexpr (LEARNED)
      {
        // one candidate per iteration as in the plain walks; every lane joins the wavefront's MLP round
        if (alive && !resume)
        {
          tcur = 0;
          tsp = 0;
          todo.set (0, kBarrier);
        }
        int c = -1;
        const int r = locate_point_mlp_step<P, POLISH> (a.cfg, todo, pt, ref, st, overflow, alive, tcur, tsp, c, a.net,
                                                        a.budget, mlp_x + wv * 64 * mlp::IN_PAD, mlp_y + wv * 64 * 4);
        if (!alive) continue;
        resume = r == 0;
        if (resume) continue;
        cell = r > 0 ? c : -1;
      }
      else if constexpr (WAVE)
        cell = locate_wave<P, POLISH> (a.cfg, ws, pt, ref, st, overflow);
      else if constexpr (MODE <= 1 || GUIDED)
      {
        if (!resume)
        {
          tcur = 0;
          tsp = 0;
          if constexpr (MODE != 1) todo.set (0, kBarrier);
          if constexpr (GUIDED)
          {
            given.cell = -1;
            if (a.cfg.guess && sample < a.cfg.samples_per_ray)
              given = a.cfg.guess[(int64_t)ri * a.cfg.samples_per_ray + sample];
          }
        }
        int c = -1;
        int r;
        if constexpr (MODE == 1)
          r = locate_point_wide_step<P, POLISH> (a.cfg, todo, pt, ref, st, overflow, tcur, tsp, c);
        else
          r = locate_point_step<P, POLISH> (a.cfg, todo, pt, ref, st, overflow, tcur, tsp, c,
                                            STAGE ? stage_lds + wv * Stage<(P >= 1 ? P : 1)>::FLOATS : nullptr,
                                            GUIDED ? &given : nullptr);
        resume = r == 0;
        if (resume) continue;
        cell = r > 0 ? c : -1;
        if constexpr (GUIDED)
        {
          if (a.cfg.record && sample < a.cfg.samples_per_ray && !overflow)
          {
            Location rec;
            rec.cell = cell;
            rec.ref[0] = cell >= 0 ? ref[0] : -1.f;
            rec.ref[1] = cell >= 0 ? ref[1] : -1.f;
            rec.ref[2] = cell >= 0 ? ref[2] : -1.f;
            a.cfg.record[(int64_t)ri * a.cfg.samples_per_ray + sample] = rec;
          }
          ++sample;
        }
      }
      else
        cell = locate_any<P, POLISH, MODE == 1> (a.cfg, todo, pt, ref, st, overflow);
      if (overflow)
      {
        alive = false;
        continue;
      }
      const bool found = cell != -1;
      if (found) ++located;
      bool shade = false;
      bool ray_done = false;
      if (!inseg)
      {
        if (found)
        {
          inseg = true;
          depth = distance;
          rgba[0] = rgba[1] = rgba[2] = rgba[3] = 0.f;
          shade = true;
        }
        else
          distance += a.sample_dist;
      }
      else if (distance < ray.far_ && found && rgba[3] < 0.95f)
        shade = true;
      else
      {
        Partial pp;
        pp.pixel_id = ray.pixel_id;
        pp.depth = depth;
        pp.rgba[0] = rgba[0];
        pp.rgba[1] = rgba[1];
        pp.rgba[2] = rgba[2];
        pp.rgba[3] = rgba[3];
        out5[segment] = pp;
        ++segment;
        if (distance >= ray.far_ || rgba[3] > 0.95f || segment == 5)
          ray_done = true;
        else
          inseg = false;
      }
      if (ray_done)
      {
        alive = false;
        continue;
      }
      if (shade)
      {
        if constexpr (LIT)
          blend_nc (rgba, lit_color<P> (a, cell, ref, ray));
        else if constexpr (SHADE == 2)
          blend_nc (rgba, color_lookup64 (a.lut, a.lut_size, a.fmin64, a.fmax64, a.inv_range64,
                                          field_at<P, double> (a, a.field64, cell, ref)));
        else
          blend_nc (rgba, color_lookup (a.lut, a.lut_size, a.fmin, a.fmax, a.inv_range, field_at<P> (a, a.field, cell, ref)));
        distance += a.sample_dist;
      }
    }
  }
  if (overflow) atomicOr (a.err, 1);
  // wave-aggregated statistics
  unsigned long long v[4] = { located, (unsigned long long)st.newton, (unsigned long long)st.candidates,
                              (unsigned long long)st.polished };
#pragma unroll
  for (int q = 0; q < 4; ++q)
  {
    unsigned long long x = v[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor (x, o, 64);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd (&a.stats[q], x);
  }
}

// ---- Sample-parallel form for frames with too few rays to fill the GPU at one lane per ray ----
// A block holds 64 rays (one per lane) and kSplitWaves wavefronts.  One round covers kSplitR consecutive
// samples of every ray: wave w locates and shades samples [base + w*T, base + (w+1)*T) of its lane's ray
// into LDS (a sample's location and colour depend on its point alone), then wave 0 advances each ray
// through k_integrate's march state machine over the round's samples, and the block stops when its rays
// are done.  The sample distances come from the same float additions as the march's, and every locate the
// march makes (including the repeated one after a segment closes) is replayed from its sample's record,
// so partials and statistics equal k_integrate's bit for bit; samples taken past a ray's end are wasted
// work (under kSplitR per ray), never results.
#ifndef DRX_SPLIT_WAVES
#define DRX_SPLIT_WAVES 4   // wavefronts per 64 rays
#endif
#ifndef DRX_SPLIT_T
#define DRX_SPLIT_T 2       // consecutive samples per wavefront and round
#endif
#ifndef DRX_SPLIT_WAVES_PER_SIMD
#define DRX_SPLIT_WAVES_PER_SIMD 4   // occupancy target of the split kernels (0: the integrate kernels' choice)
#endif
constexpr int kSplitWaves = DRX_SPLIT_WAVES, kSplitT = DRX_SPLIT_T, kSplitR = kSplitWaves * kSplitT;
__host__ __device__ constexpr int split_waves_per_simd (int P, bool polish)
{
  return DRX_SPLIT_WAVES_PER_SIMD > 0 ? DRX_SPLIT_WAVES_PER_SIMD : waves_per_simd (P, polish, 0);
}
constexpr int32_t kRecFound = 1 << 30, kRecPolished = 1 << 29, kRecOverflow = 1 << 28, kRecCand = (1 << 28) - 1;

template <int P, bool POLISH, int SHADE>
__global__ void __launch_bounds__ (64 * kSplitWaves, split_waves_per_simd (P, POLISH)) k_integrate_split (IntegrateArgs a)
{
  __shared__ int32_t stack_lds[kStackLds * 64 * kSplitWaves];
  __shared__ float4 rcol[kSplitR * 64];   // sample colours, [sample of the round][lane]
  __shared__ int2 rcnt[kSplitR * 64];     // Newton steps; candidates | found / polished / overflow bits
  __shared__ int32_t ray_live[64];
  __shared__ int32_t more;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  TravStack todo;
  todo.lds = stack_lds;
  todo.bdim = blockDim.x;
  todo.tid = threadIdx.x;
  const int i = blockIdx.x * 64 + lane;
  const bool has_ray = i < a.n_rays;
  const int ri = has_ray ? (a.order ? a.order[i] : i) : 0;
  Ray ray;
  if (has_ray)
    ray = a.rays[ri];
  else
  {
    ray.dir[0] = ray.dir[1] = ray.dir[2] = 0.f;
    ray.orig[0] = ray.orig[1] = ray.orig[2] = 0.f;
    ray.near_ = ray.far_ = 0.f;
    ray.pixel_id = -1;
  }
  Partial *out5 = a.slots + 5 * (int64_t)ri;
  if (wv == 0 && has_ray)
#pragma unroll
    for (int s = 0; s < 5; ++s)
    {
      Partial p0;
      p0.pixel_id = -1;
      p0.depth = INFINITY;
      p0.rgba[0] = p0.rgba[1] = p0.rgba[2] = p0.rgba[3] = 0.f;
      out5[s] = p0;
    }
  if (wv == 0) ray_live[lane] = has_ray;
  // march state (wave 0): the distance of sample k
  float distance = ray.near_ + a.sample_dist;
  int k = 0, segment = 0;
  bool inseg = false, alive = has_ray, overflow = false;
  float depth = 0.f;
  float rgba[4] = { 0.f, 0.f, 0.f, 0.f };
  unsigned long long v[4] = { 0, 0, 0, 0 };   // located, Newton steps, candidates, polished
  // sampling state: the distance of this wave's next sample and of the sample before it (the march only
  // locates sample k when sample k - 1 lies before `far`)
  float sd = ray.near_ + a.sample_dist, sprev = -INFINITY;
  for (int q = 0; q < wv * kSplitT; ++q)
  {
    sprev = sd;
    sd += a.sample_dist;
  }
  __syncthreads ();
  for (int base = 0;; base += kSplitR)
  {
    const bool live = ray_live[lane] != 0;
#pragma unroll 1
    for (int s = 0; s < kSplitT; ++s)
    {
      const int slot = (wv * kSplitT + s) * 64 + lane;
      int2 cn = { 0, 0 };
      float4 col = { 0.f, 0.f, 0.f, 0.f };
      if (live && sprev < ray.far_)
      {
        const float pt[3] = { ray_point (ray.orig[0], sd, ray.dir[0]), ray_point (ray.orig[1], sd, ray.dir[1]),
                              ray_point (ray.orig[2], sd, ray.dir[2]) };
        LocateStats st { 0, 0, 0, 0 };
        bool ovf = false;
        float ref[3];
        const int cell = locate_point<P, POLISH> (a.cfg, todo, pt, ref, st, ovf);
        cn.x = st.newton;
        cn.y = min (st.candidates, kRecCand) | (cell >= 0 ? kRecFound : 0) | (st.polished ? kRecPolished : 0) |
               (ovf ? kRecOverflow : 0);
        if (cell >= 0 && !ovf)
        {
          if constexpr (SHADE == 1)
            col = lit_color<P> (a, cell, ref, ray);
          else if constexpr (SHADE == 2)
            col = color_lookup64 (a.lut, a.lut_size, a.fmin64, a.fmax64, a.inv_range64,
                                  field_at<P, double> (a, a.field64, cell, ref));
          else
            col = color_lookup (a.lut, a.lut_size, a.fmin, a.fmax, a.inv_range, field_at<P> (a, a.field, cell, ref));
        }
      }
      rcol[slot] = col;
      rcnt[slot] = cn;
      sprev = sd;
      sd += a.sample_dist;
    }
    for (int q = 0; q < kSplitR - kSplitT; ++q)
    {
      sprev = sd;
      sd += a.sample_dist;
    }
    __syncthreads ();
    if (wv == 0)
    {
      while (alive)
      {
        if (!inseg && !(distance < ray.far_))
        {
          alive = false;
          break;
        }
        if (k >= base + kSplitR) break;   // the next round holds this sample
        const int slot = (k - base) * 64 + lane;
        const int2 cn = rcnt[slot];
        v[1] += (unsigned)cn.x;
        v[2] += (unsigned)(cn.y & kRecCand);
        v[3] += (cn.y & kRecPolished) ? 1 : 0;
        if (cn.y & kRecOverflow)
        {
          overflow = true;
          alive = false;
          break;
        }
        const bool found = (cn.y & kRecFound) != 0;
        if (found) ++v[0];
        bool shade = false;
        if (!inseg)
        {
          if (found)
          {
            inseg = true;
            depth = distance;
            rgba[0] = rgba[1] = rgba[2] = rgba[3] = 0.f;
            shade = true;
          }
          else
          {
            distance += a.sample_dist;
            ++k;
          }
        }
        else if (distance < ray.far_ && found && rgba[3] < 0.95f)
          shade = true;
        else
        {
          Partial pp;
          pp.pixel_id = ray.pixel_id;
          pp.depth = depth;
          pp.rgba[0] = rgba[0];
          pp.rgba[1] = rgba[1];
          pp.rgba[2] = rgba[2];
          pp.rgba[3] = rgba[3];
          out5[segment] = pp;
          ++segment;
          if (distance >= ray.far_ || rgba[3] > 0.95f || segment == 5)
          {
            alive = false;
            break;
          }
          inseg = false;
        }
        if (shade)
        {
          blend_nc (rgba, rcol[slot]);
          distance += a.sample_dist;
          ++k;
        }
      }
      ray_live[lane] = alive;
      const bool any = __ballot (alive) != 0;
      if (lane == 0) more = any;
    }
    __syncthreads ();
    if (!more) break;
  }
  if (overflow) atomicOr (a.err, 1);
#pragma unroll
  for (int q = 0; q < 4; ++q)
  {
    unsigned long long x = v[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor (x, o, 64);
    if (lane == 0 && x) atomicAdd (&a.stats[q], x);
  }
}

__global__ void k_flag_partials (const Partial *__restrict__ p, int64_t n, int32_t *__restrict__ flags)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = (p[i].depth == INFINITY) ? 0 : 1;
}

__global__ void k_compact_partials (const Partial *__restrict__ p, int64_t n, const int32_t *__restrict__ flags,
                                    const int32_t *__restrict__ pos, Partial *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flags[i]) out[pos[i]] = p[i];
}

} // namespace drx

using namespace drx;

extern "C" {

int drx_volume_integrate (const drx_mesh *mesh, const drx_mesh *field, const drx_bvh *bvh, const void *active_rays,
                          int32_t n_rays, const float *lut, int32_t lut_size, double field_min, double field_max,
                          float sample_dist, const int32_t *ray_order, const drx_locate_opts *opts,
                          const drx_shading *shading, void *out_partials, int32_t *out_count,
                          drx_integrate_stats *stats, void *stream)
{
  int rc = check_mesh (mesh, 3, "drx_volume_integrate(mesh)");
  if (rc) return rc;
  if ((rc = check_mesh (field, 1, "drx_volume_integrate(field)"))) return rc;
  if ((rc = check_bvh (bvh, "drx_volume_integrate"))) return rc;
  if ((mesh->elem_order == nullptr) != (field->elem_order == nullptr) ||
      (!mesh->elem_order && mesh->order != field->order))
    return fail (DRX_ERR_UNSUPPORTED, "drx_volume_integrate: mesh and field orders must match");
  if (!out_count) return fail (DRX_ERR_ARG, "drx_volume_integrate: null out_count");
  if (n_rays < 0 || lut_size < 1 || !lut) return fail (DRX_ERR_ARG, "drx_volume_integrate: bad rays/lut");
  if (!(field_min <= field_max)) return fail (DRX_ERR_ARG, "ColorMap scalar range never set");
  const bool learned = opts && opts->init_mode == DRX_INIT_MLP;
  const bool given = opts && opts->init_mode == DRX_INIT_GUESS;
  if (opts && opts->init_mode != DRX_INIT_AFFINE && opts->init_mode != DRX_INIT_SUBREF_CENTRE && !learned && !given)
    return fail (DRX_ERR_ARG, "drx_volume_integrate: unknown init_mode %d", opts->init_mode);
  const bool guided = given || (opts && opts->record);
  if (guided && (opts->samples_per_ray <= 0 || (given && !opts->guess) || learned))
    return fail (DRX_ERR_ARG, "drx_volume_integrate: guess / record need samples_per_ray > 0 (and opts->guess "
                 "with DRX_INIT_GUESS), and not the learned start");
  if (learned && !opts->mlp)
    return fail (DRX_ERR_ARG, "drx_volume_integrate: DRX_INIT_MLP needs opts->mlp (drx_mlp_pack)");
  const bool lit = shading && shading->use_lighting;
  const bool f64 = field->data64 != nullptr;
  if (f64 && (lit || learned))
    return fail (DRX_ERR_UNSUPPORTED, "drx_volume_integrate: an FP64 field is rendered unlit with the plain Newton start");
  if (lit && (shading->n_lights < 0 || shading->n_lights > DRX_MAX_LIGHTS || (shading->n_lights > 0 && !shading->lights)))
    return fail (DRX_ERR_ARG, "drx_volume_integrate: %d lights (at most %d, host array)", shading->n_lights,
                 DRX_MAX_LIGHTS);
  *out_count = 0;
  if (stats) *stats = drx_integrate_stats { 0, 0, 0, 0, 0.f, 0.f, 0 };
  if (n_rays == 0) return DRX_OK;
  if (!active_rays || !out_partials) return fail (DRX_ERR_ARG, "drx_volume_integrate: null rays/output");
  if ((int64_t)n_rays * 5 > (1ll << 31) - 1) return fail (DRX_ERR_ARG, "drx_volume_integrate: too many rays");
  hipStream_t st = as_stream (stream);
  const int64_t n_slots = 5 * (int64_t)n_rays;
  Partial *slots = (Partial *)arena (SLOT_PARTIALS, sizeof (Partial) * (size_t)n_slots, st);
  int32_t *flags = (int32_t *)arena (SLOT_FLAGS, sizeof (int32_t) * 2 * (size_t)n_slots + 64, st);
  unsigned long long *dstats = (unsigned long long *)arena (SLOT_STATS, 8 * sizeof (unsigned long long), st);
  if (!slots || !flags || !dstats) return DRX_ERR_HIP;
  int32_t *pos = flags + n_slots;
  int32_t *err = (int32_t *)(dstats + 4);
  hipLaunchKernelGGL (k_fill_i32q, dim3 (1), dim3 (64), 0, st, (int32_t *)dstats, (int64_t)16, 0);
  IntegrateArgs a;
  a.cfg = make_cfg (mesh, bvh, opts);
  a.cfg.polished_out = nullptr;   // per-point flags belong to batched locate
  a.field = field->data;
  a.field64 = field->data64;
  a.field_offsets = field->elem_offset;
  a.rays = (const Ray *)active_rays;
  a.order = ray_order;
  a.n_rays = n_rays;
  a.lut = (const float4 *)lut;
  a.lut_size = lut_size;
  a.fmin = (float)field_min;
  a.fmax = (float)field_max;
  {
    // rcp_safe(range length) on the host in the colour map's Float (device_color_map.hpp:45, math.hpp:213-217)
    const float len = a.fmax - a.fmin;
    a.inv_range = 1.0f / ((std::fabs (len) < 1e-8f) ? (std::signbit (len) ? -1e-8f : 1e-8f) : len);
    const double len64 = field_max - field_min;
    a.fmin64 = field_min;
    a.fmax64 = field_max;
    a.inv_range64 = 1.0 / ((std::fabs (len64) < 1e-8) ? (std::signbit (len64) ? -1e-8 : 1e-8) : len64);
  }
  a.sample_dist = sample_dist;
  a.slots = slots;
  a.stats = dstats;
  a.err = err;
  a.net = mlp_net (learned ? opts->mlp : nullptr);
  a.budget = (learned && opts->mlp_newton_steps > 0) ? opts->mlp_newton_steps : 4;
  a.n_lights = lit ? shading->n_lights : 0;
  for (int l = 0; l < a.n_lights; ++l) a.lights[l] = shading->lights[l];
  const dim3 g (nblk (n_rays, kIntBlock)), b (kIntBlock);
  const int P = mesh->elem_order ? 0 : mesh->order;
  const int trav = opts ? opts->traversal : DRX_TRAVERSE_AUTO;
  if (trav == DRX_TRAVERSE_WAVE && !a.cfg.wide)
    return fail (DRX_ERR_ARG, "drx_volume_integrate: wave traversal needs bvh->wide_nodes");
  // Sample-parallel launch (k_integrate_split): asked for, or picked for frames of fewer than
  // kSplitMaxRaysPerSimd active rays per SIMD under the automatic walk and launch.
  const int launch = opts ? opts->launch : DRX_LAUNCH_AUTO;
  const bool plain = !guided && !learned && P <= DRX_MAX_MIXED_ORDER;
  bool split = false;
  if (launch == DRX_LAUNCH_SAMPLES)
  {
    if (!plain || (trav != DRX_TRAVERSE_AUTO && trav != DRX_TRAVERSE_BINARY))
      return fail (DRX_ERR_UNSUPPORTED, "drx_volume_integrate: the sample-parallel launch runs the binary walk with "
                   "the plain start, orders up to %d", DRX_MAX_MIXED_ORDER);
    split = true;
  }
  else if (launch == DRX_LAUNCH_AUTO && trav == DRX_TRAVERSE_AUTO && plain)
  {
    int dev = 0;
    split = hipGetDevice (&dev) == hipSuccess &&
            (int64_t)n_rays < (int64_t)kSplitMaxRaysPerSimd * 4 * compute_units (dev);
  }
  int walk = split ? DRX_TRAVERSE_BINARY : trav;
  if (walk == DRX_TRAVERSE_AUTO && a.cfg.wide && !learned && P <= DRX_MAX_MIXED_ORDER)
  {
    // The wave walk (uniform node reads shared by a wavefront's coherent rays) pays on small BVHs with at
    // least ~48 active rays per SIMD: an 8^3 order-2 mesh (16,896 leaves), orthographic frames of
    // 224^2..384^2 rays run 1-13% faster than the binary walk, 192^2 (36 rays per SIMD) 11% slower.  On the
    // 64^3 order-3 mesh (8.65 M leaves) it is 2.6-4.9x slower at every frame size from 128^2 to 512^2.
    int dev = 0;
    if (bvh->n_leaves <= kWaveWalkMaxLeaves && hipGetDevice (&dev) == hipSuccess &&
        (int64_t)n_rays >= (int64_t)kWaveWalkMinRaysPerSimd * 4 * compute_units (dev))
      walk = DRX_TRAVERSE_WAVE;
  }
  const int mode = guided ? 4 : learned ? 3 : walk == DRX_TRAVERSE_WAVE ? 2 : (walk == DRX_TRAVERSE_LANE && a.cfg.wide) ? 1 : 0;
  if (P > DRX_MAX_MIXED_ORDER && mode != 0)
    return fail (DRX_ERR_UNSUPPORTED, "drx_volume_integrate: order %d runs the binary walk with the plain start "
                 "only (orders up to %d take the other forms)", P, DRX_MAX_MIXED_ORDER);
  hipEvent_t ev[3];
  for (int q = 0; q < 3; ++q) DRX_HIP (hipEventCreate (&ev[q]));
  DRX_HIP (hipEventRecord (ev[0], st));
#define DRX_INT2(PP, PO, LT)                                             \
  do                                                                     \
  {                                                                      \
    if (mode == 4)                                                       \
      hipLaunchKernelGGL ((k_integrate<PP, PO, 4, LT>), g, b, 0, st, a); \
    else if (mode == 3)                                                  \
      hipLaunchKernelGGL ((k_integrate<PP, PO, 3, LT>), g, b, 0, st, a); \
    else if (mode == 2)                                                  \
      hipLaunchKernelGGL ((k_integrate<PP, PO, 2, LT>), g, b, 0, st, a); \
    else if (mode == 1)                                                  \
      hipLaunchKernelGGL ((k_integrate<PP, PO, 1, LT>), g, b, 0, st, a); \
    else                                                                 \
      hipLaunchKernelGGL ((k_integrate<PP, PO, 0, LT>), g, b, 0, st, a); \
  } while (0)
#define DRX_INT64(PP, PO)                                                \
  do                                                                     \
  {                                                                      \
    if (mode == 4)                                                       \
      hipLaunchKernelGGL ((k_integrate<PP, PO, 4, 2>), g, b, 0, st, a);  \
    else if (mode == 2)                                                  \
      hipLaunchKernelGGL ((k_integrate<PP, PO, 2, 2>), g, b, 0, st, a);  \
    else if (mode == 1)                                                  \
      hipLaunchKernelGGL ((k_integrate<PP, PO, 1, 2>), g, b, 0, st, a);  \
    else                                                                 \
      hipLaunchKernelGGL ((k_integrate<PP, PO, 0, 2>), g, b, 0, st, a);  \
  } while (0)
#define DRX_INT(PP, PO)          \
  do                             \
  {                              \
    if (f64)                     \
      DRX_INT64 (PP, PO);        \
    else if (lit)                \
      DRX_INT2 (PP, PO, 1);      \
    else                         \
      DRX_INT2 (PP, PO, 0);      \
  } while (0)
#define DRX_SPLIT(PP, PO)                                                                  \
  do                                                                                       \
  {                                                                                        \
    const dim3 gs (nblk (n_rays, 64)), bs (64 * kSplitWaves);                              \
    if (f64)                                                                               \
      hipLaunchKernelGGL ((k_integrate_split<PP, PO, 2>), gs, bs, 0, st, a);               \
    else if (lit)                                                                          \
      hipLaunchKernelGGL ((k_integrate_split<PP, PO, 1>), gs, bs, 0, st, a);               \
    else                                                                                   \
      hipLaunchKernelGGL ((k_integrate_split<PP, PO, 0>), gs, bs, 0, st, a);               \
  } while (0)
#define DRX_INT_HI(PP, PO)                                                   \
  do                                                                         \
  {                                                                          \
    if (f64)                                                                 \
      hipLaunchKernelGGL ((k_integrate<PP, PO, 0, 2>), g, b, 0, st, a);      \
    else if (lit)                                                            \
      hipLaunchKernelGGL ((k_integrate<PP, PO, 0, 1>), g, b, 0, st, a);      \
    else                                                                     \
      hipLaunchKernelGGL ((k_integrate<PP, PO, 0, 0>), g, b, 0, st, a);      \
  } while (0)
  if (split)
  {
    if (a.cfg.polish)
      switch (P)
      {
      case 0: DRX_SPLIT (0, true); break;
      case 1: DRX_SPLIT (1, true); break;
      case 2: DRX_SPLIT (2, true); break;
      case 3: DRX_SPLIT (3, true); break;
      default: DRX_SPLIT (4, true); break;
      }
    else
      switch (P)
      {
      case 0: DRX_SPLIT (0, false); break;
      case 1: DRX_SPLIT (1, false); break;
      case 2: DRX_SPLIT (2, false); break;
      case 3: DRX_SPLIT (3, false); break;
      default: DRX_SPLIT (4, false); break;
      }
  }
  else if (a.cfg.polish)
    switch (P)
    {
    case 0: DRX_INT (0, true); break;
    case 1: DRX_INT (1, true); break;
    case 2: DRX_INT (2, true); break;
    case 3: DRX_INT (3, true); break;
    case 4: DRX_INT (4, true); break;
    case 5: DRX_INT_HI (5, true); break;
    case 6: DRX_INT_HI (6, true); break;
    case 7: DRX_INT_HI (7, true); break;
    default: DRX_INT_HI (8, true); break;
    }
  else
    switch (P)
    {
    case 0: DRX_INT (0, false); break;
    case 1: DRX_INT (1, false); break;
    case 2: DRX_INT (2, false); break;
    case 3: DRX_INT (3, false); break;
    case 4: DRX_INT (4, false); break;
    case 5: DRX_INT_HI (5, false); break;
    case 6: DRX_INT_HI (6, false); break;
    case 7: DRX_INT_HI (7, false); break;
    default: DRX_INT_HI (8, false); break;
    }
#undef DRX_INT_HI
#undef DRX_SPLIT
#undef DRX_INT
#undef DRX_INT64
#undef DRX_INT2
  DRX_LAUNCH_CHECK ();
  DRX_HIP (hipEventRecord (ev[1], st));
  DRX_LAUNCH_CHECK ();
  hipLaunchKernelGGL (k_flag_partials, dim3 (nblk (n_slots, 256)), dim3 (256), 0, st, slots, n_slots, flags);
  DRX_LAUNCH_CHECK ();
  int32_t total = 0;
  rc = exclusive_scan_i32 (flags, pos, (int32_t)n_slots, &total, st);
  if (rc) return rc;
  hipLaunchKernelGGL (k_compact_partials, dim3 (nblk (n_slots, 256)), dim3 (256), 0, st, slots, n_slots, flags, pos,
                      (Partial *)out_partials);
  DRX_LAUNCH_CHECK ();
  DRX_HIP (hipEventRecord (ev[2], st));
  unsigned long long hs[8];
  DRX_HIP (hipMemcpyAsync (hs, dstats, sizeof hs, hipMemcpyDeviceToHost, st));
  DRX_HIP (hipStreamSynchronize (st));
  float k_ms = 0.f, t_ms = 0.f;
  DRX_HIP (hipEventElapsedTime (&k_ms, ev[0], ev[1]));
  DRX_HIP (hipEventElapsedTime (&t_ms, ev[0], ev[2]));
  for (int q = 0; q < 3; ++q) (void)hipEventDestroy (ev[q]);
  if (((int32_t *)(hs + 4))[0])
    return fail (DRX_ERR_OVERFLOW, "drx_volume_integrate: BVH traversal stack (%d entries) overflowed", kStackDepth);
  *out_count = total;
  if (stats)
  {
    stats->located_samples = (int64_t)hs[0];
    stats->newton_steps = (int64_t)hs[1];
    stats->candidates = (int64_t)hs[2];
    stats->polished = (int64_t)hs[3];
    stats->kernel_ms = k_ms;
    stats->total_ms = t_ms;
    stats->walk = split ? 5 : mode;
  }
  return DRX_OK;
}

// detail::default_light (rendering/renderer.cpp:141-159): 0.1 x |look_at - pos| above the camera along its
// orthonormalised up vector; the other members keep PointLight's defaults.
int drx_default_light (const drx_camera *cam, drx_point_light *out)
{
  if (!cam || !out) return fail (DRX_ERR_ARG, "drx_default_light: null pointer");
  auto normalize = [] (float v[3]) {
    float s = 0.f;
    for (int i = 0; i < 3; ++i) s += v[i] * v[i];
    const float m = sqrtf (s);
    for (int i = 0; i < 3; ++i) v[i] /= m;
  };
  float up[3] = { cam->up[0], cam->up[1], cam->up[2] };
  normalize (up);
  float look[3];
  for (int i = 0; i < 3; ++i) look[i] = cam->look_at[i] - cam->pos[i];
  float m2 = 0.f;
  for (int i = 0; i < 3; ++i) m2 += look[i] * look[i];
  const float mag = sqrtf (m2);
  float right[3] = { look[1] * up[2] - look[2] * up[1], look[2] * up[0] - look[0] * up[2], look[0] * up[1] - look[1] * up[0] };
  normalize (right);
  float mu[3] = { right[1] * look[2] - right[2] * look[1], right[2] * look[0] - right[0] * look[2],
                  right[0] * look[1] - right[1] * look[0] };
  normalize (mu);
  for (int i = 0; i < 3; ++i)
  {
    out->pos[i] = cam->pos[i] + .1f * mag * mu[i];
    out->amb[i] = 0.4f;
    out->diff[i] = 0.75f;
    out->spec[i] = 0.8f;
  }
  out->spec_pow = 15.f;
  return DRX_OK;
}

} // extern "C"

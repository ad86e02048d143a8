// Point location, field evaluation, camera rays, ray set-up and compositing for gfx950.
#include "drx_common.hpp"
#include "drx_element.hpp"
#include "drx_locate_mlp.hpp"
#include "drx_query_common.hpp"
#include "drx_wave.hpp"

#include <cmath>
#include <mutex>
#include <vector>

namespace drx
{

// ---------------------------------------------------------------------------------------------
// Batched locate (UnstructuredMesh::locate, unstructured_mesh.cpp:123-148): one lane per point.
// ---------------------------------------------------------------------------------------------
template <int P, bool POLISH, int MODE>
__global__ void __launch_bounds__ (256, waves_per_simd (P, POLISH, MODE)) k_locate (LocateCfg cfg, const float *__restrict__ pts, int64_t n,
                                                  const int32_t *__restrict__ order, Location *__restrict__ out,
                                                  int32_t *__restrict__ steps_out, int32_t *__restrict__ cand_out,
                                                  int32_t *__restrict__ err)
{
  __shared__ int32_t stack_lds[MODE == 2 ? 1 : kStackLds * 256];
  __shared__ int32_t w_code[MODE == 2 ? 4 * kWaveStack : 1];
  __shared__ uint64_t w_mask[MODE == 2 ? 4 * kWaveStack : 1];
  __shared__ int32_t w_cand[MODE == 2 ? 4 * kCandQ * 64 : 1];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t i = order ? (int64_t)order[t] : t;
  TravStack todo;
  todo.lds = stack_lds;
  todo.bdim = blockDim.x;
  todo.tid = threadIdx.x;
  const float pt[3] = { pts[3 * i], pts[3 * i + 1], pts[3 * i + 2] };
  float ref[3] = { -1.f, -1.f, -1.f };
  LocateStats st { 0, 0, 0, 0 };
  bool overflow = false;
  int cell;
  if constexpr (MODE == 2)
  {
    const int wv = threadIdx.x >> 6;
    WaveScratch ws { w_code + wv * kWaveStack, w_mask + wv * kWaveStack, w_cand + wv * kCandQ * 64 };
    cell = locate_wave<P, POLISH> (cfg, ws, pt, ref, st, overflow);
  }
  else
    cell = locate_any<P, POLISH, MODE == 1> (cfg, todo, pt, ref, st, overflow);
  if (overflow) atomicOr (err, 1);
  Location loc;
  loc.cell = cell;
  loc.ref[0] = cell >= 0 ? ref[0] : -1.f;
  loc.ref[1] = cell >= 0 ? ref[1] : -1.f;
  loc.ref[2] = cell >= 0 ? ref[2] : -1.f;
  out[i] = loc;
  if (steps_out) steps_out[i] = st.newton;
  if (cand_out) cand_out[i] = st.candidates;
        if (cfg.polished_out) cfg.polished_out[i] = st.polished;
}

// Newton start map of every leaf (drx_bvh_leaf_starts): the inverse of its element's corner-affine
// approximation, x0 = M p + b in absolute coordinates (b NaN when singular).
__global__ void k_leaf_starts (const float *__restrict__ mesh, const int64_t *__restrict__ offsets,
                               const int32_t *__restrict__ orders, int stride, int order, const int32_t *__restrict__ leaf_el,
                               int n, float4 *__restrict__ out)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int el = leaf_el[i];
  const float *g = offsets ? mesh + offsets[el] : mesh + (int64_t)el * stride;
  const int p = orders ? orders[el] : order;
  const int n1 = p + 1, np = (n1 * n1 * n1 + 3) & ~3;
  float m[3], eu[3], ev[3], ew[3];
  for (int d = 0; d < 3; ++d)
  {
    const float *X = g + 4 + d * np;
    float c[8];
    for (int q = 0; q < 8; ++q) c[q] = X[(q & 1) * p + n1 * (((q >> 1) & 1) * p) + n1 * n1 * ((q >> 2) * p)];
    m[d] = (((c[0] + c[1]) + (c[2] + c[3])) + ((c[4] + c[5]) + (c[6] + c[7]))) * 0.125f + g[d];
    eu[d] = (((c[1] - c[0]) + (c[3] - c[2])) + ((c[5] - c[4]) + (c[7] - c[6]))) * 0.25f;
    ev[d] = (((c[2] - c[0]) + (c[3] - c[1])) + ((c[6] - c[4]) + (c[7] - c[5]))) * 0.25f;
    ew[d] = (((c[4] - c[0]) + (c[5] - c[1])) + ((c[6] - c[2]) + (c[7] - c[3]))) * 0.25f;
  }
  // M = [eu ev ew]^-1 = adjugate / det: row d of the inverse is (column vector cross products) / det
  const float r0[3] = { ev[1] * ew[2] - ev[2] * ew[1], ev[2] * ew[0] - ev[0] * ew[2], ev[0] * ew[1] - ev[1] * ew[0] };
  const float r1[3] = { ew[1] * eu[2] - ew[2] * eu[1], ew[2] * eu[0] - ew[0] * eu[2], ew[0] * eu[1] - ew[1] * eu[0] };
  const float r2[3] = { eu[1] * ev[2] - eu[2] * ev[1], eu[2] * ev[0] - eu[0] * ev[2], eu[0] * ev[1] - eu[1] * ev[0] };
  const float det = (eu[0] * r0[0] + eu[1] * r0[1]) + eu[2] * r0[2];
  const float inv = 1.0f / det;
  const bool ok = det != 0.f && fabsf (inv) < 1e30f;
  const float *rows[3] = { r0, r1, r2 };
  for (int d = 0; d < 3; ++d)
  {
    const float a = rows[d][0] * inv, b = rows[d][1] * inv, c = rows[d][2] * inv;
    const float off = ok ? 0.5f - ((a * m[0] + b * m[1]) + c * m[2]) : __builtin_nanf ("");
    out[3 * (int64_t)i + d] = make_float4 (ok ? a : 0.f, ok ? b : 0.f, ok ? c : 0.f, off);
  }
}

// Compact traversal nodes (drx_bvh_quantize): one thread per inner node quantizes both child boxes to
// grid cells and records the exact box of each child that is a leaf.
__global__ void k_quantize_nodes (const float4 *__restrict__ inner, int n_inner, float ox, float oy, float oz, float ix,
                                  float iy, float iz, uint4 *__restrict__ qn, float4 *__restrict__ leaf_boxes)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inner) return;
  const float4 a = inner[4 * (int64_t)i], b = inner[4 * (int64_t)i + 1], c = inner[4 * (int64_t)i + 2],
               ch = inner[4 * (int64_t)i + 3];
  const float lo[2][3] = { { a.x, a.y, a.z }, { b.z, b.w, c.x } };
  const float hi[2][3] = { { a.w, b.x, b.y }, { c.y, c.z, c.w } };
  const float org[3] = { ox, oy, oz }, inv[3] = { ix, iy, iz };
  uint32_t w[6];
  for (int k = 0; k < 2; ++k)
  {
    uint32_t ql[3], qh[3];
    for (int d = 0; d < 3; ++d)
    {
      ql[d] = grid_cell (lo[k][d], org[d], inv[d]);
      qh[d] = grid_cell (hi[k][d], org[d], inv[d]);
    }
    w[3 * k] = ql[0] | (ql[1] << 16);
    w[3 * k + 1] = ql[2] | (qh[0] << 16);
    w[3 * k + 2] = qh[1] | (qh[2] << 16);
  }
  const int child[2] = { __float_as_int (ch.x), __float_as_int (ch.y) };
  int q[2];
  for (int k = 0; k < 2; ++k)
  {
    q[k] = child[k] >= 0 ? child[k] / 2 : child[k];   // float4 offset 4 * index -> uint4 offset 2 * index
    if (child[k] < 0)
    {
      const int64_t leaf = -(int64_t)child[k] - 1;
      leaf_boxes[2 * leaf] = make_float4 (lo[k][0], lo[k][1], lo[k][2], hi[k][0]);
      leaf_boxes[2 * leaf + 1] = make_float4 (hi[k][1], hi[k][2], 0.f, 0.f);
    }
  }
  qn[2 * (int64_t)i] = make_uint4 (w[0], w[1], w[2], w[3]);
  qn[2 * (int64_t)i + 1] = make_uint4 (w[4], w[5], (uint32_t)q[0], (uint32_t)q[1]);
}

// Per-lane walks, streamed: each wavefront works through `chunk` consecutive entries of the processing
// order and every iteration advances each lane's point by one candidate (locate_point_step); a lane whose
// point is settled (accepted or exhausted) writes its Location and takes the next unclaimed entry.  The
// wavefront therefore never waits for the lanes whose points need several candidates.  Without a queue
// wavefront b owns chunk b; with one (persistent launch, a grid of about the resident wavefronts) each
// wavefront takes the next chunk from the queue counter whenever its current chunk is used up, so its
// lanes refill across chunks until the last chunk is gone.  Same candidate sequence and results as one
// lane per point, whatever the chunk size and launch form.
constexpr int kLocateChunk = 1024;   // entries per chunk unless tuned (drx_locate_autotune)

// Wave-uniform pool of unclaimed processing-order entries [claimed, end) for the streamed kernels; `more`:
// further chunks may come from the queue.
struct EntryPool
{
  int64_t claimed, end, n;
  int chunk;
  int32_t *queue;
  bool more;

  __device__ __forceinline__ EntryPool (int64_t n_, int chunk_, int32_t *queue_)
      : claimed (0), end (0), n (n_), chunk (chunk_), queue (queue_), more (queue_ != nullptr)
  {
    if (!queue)
    {
      claimed = (int64_t)blockIdx.x * chunk;
      end = claimed + chunk < n ? claimed + chunk : n;
    }
  }

  // The lanes in m (wave-uniform mask of lanes without a point) take the next entries in lane order,
  // fetching chunks from the queue while the pool is empty; a lane that gets entry k calls take(k).
  template <typename F> __device__ __forceinline__ void refill (uint64_t m, int lane, F &&take)
  {
    while (m)
    {
      if (claimed >= end)
      {
        if (!more) break;
        int c = 0;
        if (lane == 0) c = atomicAdd (queue, 1);
        c = __builtin_amdgcn_readfirstlane (c);
        claimed = (int64_t)c * chunk;
        if (claimed >= n)
        {
          more = false;
          break;
        }
        end = claimed + chunk < n ? claimed + chunk : n;
      }
      const int64_t avail = end - claimed;
      const int rank = __popcll (m & ((1ull << lane) - 1));
      const int cnt = __popcll (m) < avail ? __popcll (m) : (int)avail;
      if (((m >> lane) & 1ull) && rank < cnt) take (claimed + rank);
      claimed += cnt;
      for (int t = 0; t < cnt; ++t) m &= m - 1;   // drop the cnt lowest lanes of m
    }
  }
};

template <int P, bool POLISH, int MODE>
__global__ void __launch_bounds__ (64, locate_waves_per_simd (P, POLISH, MODE))
k_locate_stream (LocateCfg cfg, const float *__restrict__ pts, int64_t n, const int32_t *__restrict__ order,
                 Location *__restrict__ out, int32_t *__restrict__ steps_out, int32_t *__restrict__ cand_out,
                 int32_t *__restrict__ err, int chunk, int32_t *__restrict__ queue)
{
  __shared__ int32_t stack_lds[kStackLds * 64];
  const int lane = threadIdx.x;
  TravStack todo;
  todo.lds = stack_lds;
  todo.bdim = 64;
  todo.tid = lane;
  EntryPool pool (n, chunk, queue);
  int64_t k = 0;
  bool have = false;
  int64_t i = 0;
  float pt[3] = { 0.f, 0.f, 0.f };
  int cur = 0, sp = 0;
  LocateStats st { 0, 0, 0, 0 };
  Location given;   // MODE 4: the point's given start
  given.cell = -1;
  auto start = [&] () {
    i = order ? (int64_t)order[k] : k;
    pt[0] = pts[3 * i];
    pt[1] = pts[3 * i + 1];
    pt[2] = pts[3 * i + 2];
    cur = 0;
    sp = 0;
    if constexpr (MODE != 1) todo.set (0, kBarrier);
    if constexpr (MODE == 4) given = cfg.guess[i];
    st = LocateStats { 0, 0, 0, 0 };
  };
  bool done = true;   // every lane starts without a point
  for (;;)
  {
    // refill: lanes without a point take the next entries of the processing order
    pool.refill (__ballot (done), lane, [&] (int64_t e) {
      k = e;
      have = true;
      start ();
    });
    if (__ballot (have) == 0) break;
    done = false;
    if (have)
    {
      float ref[3];
      int cell = -1;
      bool overflow = false;
      int r;
      if constexpr (MODE == 1)
        r = locate_point_wide_step<P, POLISH> (cfg, todo, pt, ref, st, overflow, cur, sp, cell);
      else
        r = locate_point_step<P, POLISH> (cfg, todo, pt, ref, st, overflow, cur, sp, cell, nullptr,
                                          MODE == 4 ? &given : nullptr);
      if (r != 0)
      {
        done = true;
        have = false;
        if (overflow) atomicOr (err, 1);
        const bool hit = r > 0;
        Location loc;
        loc.cell = hit ? cell : -1;
        loc.ref[0] = hit ? ref[0] : -1.f;
        loc.ref[1] = hit ? ref[1] : -1.f;
        loc.ref[2] = hit ? ref[2] : -1.f;
        out[i] = loc;
        if (steps_out) steps_out[i] = st.newton;
        if (cand_out) cand_out[i] = st.candidates;
        if (cfg.polished_out) cfg.polished_out[i] = st.polished;
      }
    }
  }
}

// Mesh::locate with the learned initial guess, streamed like k_locate_stream: one wavefront per block, one
// candidate per lane per iteration with a wavefront-wide MLP round for the lanes that hold one, settled
// lanes refilled from the chunk (or the queue).  Same candidates, guesses and results as one lane per
// point: a lane's MLP output is its own MFMA column, independent of the other lanes in the round.
template <int P, bool POLISH>
__global__ void __launch_bounds__ (64, locate_waves_per_simd (P, POLISH, 3))
k_locate_mlp_stream (LocateCfg cfg, MlpNet net, int budget, const float *__restrict__ pts, int64_t n,
                     const int32_t *__restrict__ order, Location *__restrict__ out, int32_t *__restrict__ steps_out,
                     int32_t *__restrict__ cand_out, int32_t *__restrict__ err, int chunk, int32_t *__restrict__ queue)
{
  __shared__ int32_t stack_lds[kStackLds * 64];
  __shared__ __attribute__ ((aligned (16))) __bf16 lx[64 * mlp::IN_PAD];
  __shared__ __attribute__ ((aligned (16))) float ly[64 * 4];
  const int lane = threadIdx.x;
  TravStack todo;
  todo.lds = stack_lds;
  todo.bdim = 64;
  todo.tid = lane;
  EntryPool pool (n, chunk, queue);
  int64_t k = 0, i = 0;
  bool have = false, done = true;
  float pt[3] = { 0.f, 0.f, 0.f };
  int cur = 0, sp = 0;
  LocateStats st { 0, 0, 0, 0 };
  for (;;)
  {
    pool.refill (__ballot (done), lane, [&] (int64_t e) {
      k = e;
      have = true;
      i = order ? (int64_t)order[k] : k;
      pt[0] = pts[3 * i];
      pt[1] = pts[3 * i + 1];
      pt[2] = pts[3 * i + 2];
      cur = 0;
      sp = 0;
      todo.set (0, kBarrier);
      st = LocateStats { 0, 0, 0, 0 };
    });
    if (__ballot (have) == 0) break;
    done = false;
    float ref[3];
    int cell = -1;
    bool overflow = false;
    const int r = locate_point_mlp_step<P, POLISH> (cfg, todo, pt, ref, st, overflow, have, cur, sp, cell, net, budget,
                                                    lx, ly);
    if (have && r != 0)
    {
      done = true;
      have = false;
      if (overflow) atomicOr (err, 1);
      const bool hit = r > 0;
      Location loc;
      loc.cell = hit ? cell : -1;
      loc.ref[0] = hit ? ref[0] : -1.f;
      loc.ref[1] = hit ? ref[1] : -1.f;
      loc.ref[2] = hit ? ref[2] : -1.f;
      out[i] = loc;
      if (steps_out) steps_out[i] = st.newton;
      if (cand_out) cand_out[i] = st.candidates;
        if (cfg.polished_out) cfg.polished_out[i] = st.polished;
    }
  }
}


template <int P, typename FT>
__global__ void __launch_bounds__ (256) k_field_eval (const FT *__restrict__ field, const int64_t *__restrict__ offs,
                                                      const int32_t *__restrict__ orders, const Location *__restrict__ locs,
                                                      int64_t n, float *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Location l = locs[i];
  if (l.cell == -1) return;   // Field::eval leaves misses untouched
  if constexpr (P == 0)
  {
    const FT *f = field + offs[l.cell];
    switch (orders[l.cell])
    {
    case 1: out[i] = (float)eval_scalar<1> (f, l.ref); break;
    case 2: out[i] = (float)eval_scalar<2> (f, l.ref); break;
    case 3: out[i] = (float)eval_scalar<3> (f, l.ref); break;
    default: out[i] = (float)eval_scalar<4> (f, l.ref); break;
    }
  }
  else
    out[i] = (float)eval_scalar<P> (field + (int64_t)l.cell * Ord<P>::FSTRIDE, l.ref);
}

// ---------------------------------------------------------------------------------------------
// Camera::gen_perspective per-pixel part (rendering/camera.cpp:457-481)
// ---------------------------------------------------------------------------------------------
struct CamSetup
{
  float nlook[3], dx[3], dy[3], pos[3];
  int w, h;
};

__global__ void k_camera (CamSetup cs, Ray *__restrict__ out)
{
#pragma clang fp contract(off)
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= cs.w * cs.h) return;
  Ray r;
  const int i = idx % cs.w, j = idx / cs.w;
  r.orig[0] = cs.pos[0];
  r.orig[1] = cs.pos[1];
  r.orig[2] = cs.pos[2];
  r.near_ = 0.f;
  r.far_ = INFINITY;
  r.pixel_id = j * cs.w + i;
  const float a = (2.f * (float)i - (float)cs.w) / 2.0f;
  const float b = (2.f * (float)j - (float)cs.w) / 2.0f;   // dray uses w for both axes (camera.cpp:469-470)
  float s = 0.f;
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    const float t1 = cs.dx[d] * a;
    const float t2 = cs.dy[d] * b;
    float v = (cs.nlook[d] + t1) + t2;
    if (v == 0.f) v += 0.0000001f;
    r.dir[d] = v;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) s += r.dir[d] * r.dir[d];
  const float m = sqrtf (s);   // correctly rounded (HIP default); __fsqrt_rn is the native approximation
#pragma unroll
  for (int d = 0; d < 3; ++d) r.dir[d] = r.dir[d] / m;
  out[idx] = r;
}

static void vnormalize_host (float v[3])
{
  float s = 0.f;
  for (int i = 0; i < 3; ++i) s += v[i] * v[i];
  const float m = sqrtf (s);
  for (int i = 0; i < 3; ++i) v[i] /= m;
}

// Camera host set-up: set_fov (camera.cpp:132-163), gen_perspective (:402-440), create_rays (:219-243).
static CamSetup camera_setup (const drx_camera &c)
{
  const double pi = 3.14159265358979323846264338327950288;
  const int w = c.width, h = c.height;
  const float fov_y = c.fov_y;
  const float fov_y_rad = (fov_y * static_cast<float> (pi)) / 180.0f;
  const float vertical = std::tan (0.5f * fov_y_rad);
  const float aspect = float (w) / float (h);
  const float horizontal = aspect * vertical;
  const float fov_x_rad = 2.0f * std::atan (horizontal);
  const float fov_x = 180.0f * (fov_x_rad / static_cast<float> (pi));
  float up[3] = { c.up[0], c.up[1], c.up[2] };
  vnormalize_host (up);
  float look[3];
  for (int i = 0; i < 3; ++i) look[i] = c.look_at[i] - c.pos[i];
  const float thx = tanf ((fov_x * float (pi) / 180.f) * .5f);
  const float thy = tanf ((fov_y * float (pi) / 180.f) * .5f);
  float ru[3] = { look[1] * up[2] - look[2] * up[1], look[2] * up[0] - look[0] * up[2], look[0] * up[1] - look[1] * up[0] };
  const float ruf[3] = { ru[0], ru[1], ru[2] };
  vnormalize_host (ru);
  float rv[3] = { ruf[1] * look[2] - ruf[2] * look[1], ruf[2] * look[0] - ruf[0] * look[2],
                  ruf[0] * look[1] - ruf[1] * look[0] };
  vnormalize_host (rv);
  CamSetup cs;
  for (int i = 0; i < 3; ++i)
  {
    cs.dx[i] = ru[i] * (2 * thx / (float)w);
    cs.dy[i] = rv[i] * (2 * thy / (float)h);
    if (c.zoom > 0)
    {
      cs.dx[i] = cs.dx[i] / c.zoom;
      cs.dy[i] = cs.dy[i] / c.zoom;
    }
    cs.nlook[i] = look[i];
    cs.pos[i] = c.pos[i];
  }
  vnormalize_host (cs.nlook);
  cs.w = w;
  cs.h = h;
  return cs;
}

// ---------------------------------------------------------------------------------------------
// remove_missed_rays (ray.cpp:102-241): slab test against bounds scaled x1.001, compaction, near/far.
// ---------------------------------------------------------------------------------------------
struct Bounds6
{
  float b[6];
};

__device__ __forceinline__ float rcp_safe (float f)
{
  return 1.0f / ((fabsf (f) < 1e-8f) ? (signbit (f) ? -1e-8f : 1e-8f) : f);
}

__device__ __forceinline__ bool ray_slab (const Ray &ray, const Bounds6 &mb, float &near_o, float &far_o)
{
  const float idx = rcp_safe (ray.dir[0]), idy = rcp_safe (ray.dir[1]), idz = rcp_safe (ray.dir[2]);
  float ox, oy, oz;
  {
#pragma clang fp contract(off)
    ox = ray.orig[0] * idx;
    oy = ray.orig[1] * idy;
    oz = ray.orig[2] * idz;
  }
  const float xmin = mul_sub_nc (mb.b[0], idx, ox), ymin = mul_sub_nc (mb.b[2], idy, oy), zmin = mul_sub_nc (mb.b[4], idz, oz);
  const float xmax = mul_sub_nc (mb.b[1], idx, ox), ymax = mul_sub_nc (mb.b[3], idy, oy), zmax = mul_sub_nc (mb.b[5], idz, oz);
  // std::min/std::max semantics: (a < b) ? a : b
  auto mn = [] (float a, float b) { return (b < a) ? b : a; };
  auto mx = [] (float a, float b) { return (a < b) ? b : a; };
  const float mind = mx (mx (mx (mn (ymin, ymax), mn (xmin, xmax)), mn (zmin, zmax)), ray.near_);
  float maxd = mn (mn (mx (ymin, ymax), mx (xmin, xmax)), mx (zmin, zmax));
  maxd = mn (maxd, ray.far_);
  near_o = mind;
  far_o = maxd;
  return maxd > mind;
}

__global__ void k_mark_active (const Ray *__restrict__ rays, int n, Bounds6 mb, int32_t *__restrict__ flags)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float a, b;
  flags[i] = ray_slab (rays[i], mb, a, b) ? 1 : 0;
}

__global__ void k_gather_active (const Ray *__restrict__ rays, int n, Bounds6 mb, const int32_t *__restrict__ flags,
                                 const int32_t *__restrict__ pos, Ray *__restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flags[i]) return;
  Ray r = rays[i];
  float nr, fr;
  ray_slab (r, mb, nr, fr);
  r.near_ = nr;
  r.far_ = fr;
  out[pos[i]] = r;
}

// Ray tiles of 64 pixels, kTileW x kTileH (one wavefront each).
#ifndef DRX_TILE_LOG2_W
#define DRX_TILE_LOG2_W 3   // 8 x 8 (64^3 frame: 13.9 ms; 16 x 4: 14.5 ms; 4 x 16: 14.7 ms)
#endif
constexpr int kTileLW = DRX_TILE_LOG2_W, kTileW = 1 << kTileLW, kTileLH = 6 - kTileLW, kTileH = 1 << kTileLH;

// Tile-major rank of a pixel in a w x h image.
__device__ __forceinline__ int64_t tile_rank (int pid, int w, int h)
{
  const int i = pid % w, j = pid / w;
  const int tx = i >> kTileLW, ty = j >> kTileLH;
  const int tiles_x = (w + kTileW - 1) >> kTileLW;
  return ((int64_t)(ty * tiles_x + tx) << 6) + ((j & (kTileH - 1)) << kTileLW) + (i & (kTileW - 1));
}


__global__ void k_tile_scatter (const Ray *__restrict__ rays, int n, int w, int h, int32_t *__restrict__ slots)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int pid = rays[k].pixel_id;
  if (pid >= 0 && pid < w * h) slots[tile_rank (pid, w, h)] = k;
}

// Tile keys for longest-first processing: the longest [near, far] segment among the tile's rays, as a
// descending unsigned key (tiles without rays sort last).
__global__ void k_tile_keys (const Ray *__restrict__ rays, const int32_t *__restrict__ slots, int n_tiles,
                             uint32_t *__restrict__ keys, int32_t *__restrict__ vals)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tiles) return;
  float len = -1.f;
  for (int l = 0; l < 64; ++l)
  {
    const int r = slots[((int64_t)t << 6) + l];
    if (r >= 0) len = fmaxf (len, fmaxf (rays[r].far_ - rays[r].near_, 0.f));
  }
  keys[t] = len < 0.f ? 0xFFFFFFFFu : 0x7FFFFFFFu - __float_as_uint (len);
  vals[t] = t;
}

__global__ void k_gather_tiles (const int32_t *__restrict__ slots, const int32_t *__restrict__ perm, int64_t n,
                                int32_t *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = slots[((int64_t)perm[i >> 6] << 6) + (i & 63)];
}

__global__ void k_flag_nonneg (const int32_t *__restrict__ v, int64_t n, int32_t *__restrict__ flags)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = v[i] >= 0 ? 1 : 0;
}

__global__ void k_compact_i32 (const int32_t *__restrict__ v, int64_t n, const int32_t *__restrict__ pos,
                               int32_t *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && v[i] >= 0) out[pos[i]] = v[i];
}

// ---------------------------------------------------------------------------------------------
// Volume::save compositing (volume.cpp:369-416) and Framebuffer::composite_background
// (framebuffer.cpp:146-162).  One lane per run of equal pixel ids.
// ---------------------------------------------------------------------------------------------
// Keys for the cross-domain partial sort: depth as an order-preserving unsigned key, then pixel id.
__global__ void k_depth_keys (const Partial *__restrict__ p, int n, uint32_t *__restrict__ keys, int32_t *__restrict__ vals)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t b = __float_as_uint (p[i].depth);
  keys[i] = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  vals[i] = i;
}

__global__ void k_pixel_keys (const Partial *__restrict__ p, const int32_t *__restrict__ perm, int n,
                              uint32_t *__restrict__ keys, int32_t *__restrict__ vals)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  keys[i] = (uint32_t)p[j].pixel_id;
  vals[i] = j;
}

__global__ void k_gather_partials (const Partial *__restrict__ p, const int32_t *__restrict__ perm, int n,
                                   Partial *__restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = p[perm[i]];
}

__global__ void k_fill_f32 (float *p, int64_t n, float v)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_f4 (float4 *p, int64_t n, float4 v)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// One lane per run of equal pixel ids (partials grouped per pixel, front to back).  The run is blended
// with pre_mult_alpha_blend_host (colors.hpp:45-53) and written as detail::partials_to_framebuffer does
// (rendering/renderer.cpp:97-138): blend = 0 replaces the pixel; blend = 1 puts the run over what the
// framebuffer holds (a surface behind the volume): colour = run + (1 - run alpha) * framebuffer.  depth
// (nullable) takes the depth of the run's first (nearest) partial.
__global__ void k_composite_runs (const Partial *__restrict__ p, int n, int npix, float4 *__restrict__ img,
                                  float *__restrict__ depth, int blend)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int pid = p[i].pixel_id;
  if (i > 0 && p[i - 1].pixel_id == pid) return;
  float c[4] = { p[i].rgba[0], p[i].rgba[1], p[i].rgba[2], p[i].rgba[3] };
  for (int k = i + 1; k < n && p[k].pixel_id == pid; ++k)
  {
    const float a = 1.f - c[3];
    c[0] = c[0] + p[k].rgba[0] * a;
    c[1] = c[1] + p[k].rgba[1] * a;
    c[2] = c[2] + p[k].rgba[2] * a;
    c[3] = c[3] + p[k].rgba[3] * a;
  }
  if (pid < 0 || pid >= npix) return;
  if (blend)
  {
    const float4 b = img[pid];
    const float o = 1.f - c[3];
    c[0] = c[0] + o * b.x;
    c[1] = c[1] + o * b.y;
    c[2] = c[2] + o * b.z;
    c[3] = c[3] + o * b.w;
  }
  img[pid] = make_float4 (c[0], c[1], c[2], c[3]);
  if (depth) depth[pid] = p[i].depth;
}

// Framebuffer::composite_background (framebuffer.cpp:146-162) with blend_pre_alpha (colors.hpp:25-33):
// pixels with alpha < 1 are put over the background colour.
__global__ void k_background (float4 *__restrict__ img, int npix, float4 bg)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  float4 c = img[i];
  if (c.w < 1.f)
  {
    const float a = 1.f - c.w;
    c.x = c.x + bg.x * a;
    c.y = c.y + bg.y * a;
    c.z = c.z + bg.z * a;
    c.w = c.w + bg.w * a;
    img[i] = c;
  }
}

} // namespace drx

namespace drx
{
namespace
{
// Launch settings drx_locate_autotune picked, per device, order (0: mixed), FP64 polish and walk (0 binary,
// 1 8-wide, 2 learned start); chunk 0 = not tuned.  Written by the tuner, read by drx_locate on the same
// host thread.
struct LocateTune
{
  int launch;
  int chunk;
};
constexpr int kTuneDevices = 16;
LocateTune g_tune[kTuneDevices][DRX_MAX_ORDER + 1][2][3];

int stream_walk (const drx_bvh *bvh, const drx_locate_opts *opts)
{
  if (opts && opts->init_mode == DRX_INIT_MLP) return 2;
  const int trav = opts ? opts->traversal : DRX_TRAVERSE_AUTO;
  return (trav == DRX_TRAVERSE_LANE && bvh->wide_nodes) ? 1 : 0;
}
} // namespace
// ---- allocation-free locate ----------------------------------------------------------------------
// drx_locate_noalloc keeps its scratch in this statically sized device array instead of the arena: the
// two counters of locate_impl in the first 256 bytes, then the Morton-order buffers for up to
// DRX_LOCATE_NOALLOC_MAX_ORDERED points.  Every GPU holds its own copy (loaded with the code object,
// kept until the process ends).  Host code must not pass the array's name to a kernel (on the host it is
// a stand-in, not a device address): g_scratch_dev[dev] holds the device address hipGetSymbolAddress
// returned for each device.
constexpr int64_t kNaOrdered = DRX_LOCATE_NOALLOC_MAX_ORDERED;
constexpr int64_t kNaHeaderWords = 64;
constexpr int64_t kNaHistWords = 2 * 256 * ((kNaOrdered + 4095) / 4096) + 8;
constexpr int64_t kNaSumsWords = (256 * ((kNaOrdered + 4095) / 4096) + 2047) / 2048 + 1;
constexpr int64_t kNaWords = kNaHeaderWords + 4 * kNaOrdered + ((kNaHistWords + 63) & ~63) + ((kNaSumsWords + 63) & ~63);
static __device__ __align__ (256) int32_t scratch[kNaWords];

constexpr int kNaDevices = 16;
static int32_t *g_scratch_dev[kNaDevices];
static std::mutex g_scratch_mutex;

// Resolves the current device's copy of `scratch` (once per device).  Called by drx_init and by the first
// drx_locate_noalloc on a device drx_init did not set up.
int locate_noalloc_setup ()
{
  int dev = 0;
  DRX_HIP (hipGetDevice (&dev));
  if (dev < 0 || dev >= kNaDevices) return fail (DRX_ERR_STATE, "drx_locate_noalloc: device %d >= %d", dev, kNaDevices);
  std::lock_guard<std::mutex> lock (g_scratch_mutex);
  if (g_scratch_dev[dev]) return DRX_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress (&p, HIP_SYMBOL (scratch));
  if (e != hipSuccess || !p)
    return fail (DRX_ERR_HIP, "drx_locate_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                 hipGetErrorString (e));
  g_scratch_dev[dev] = (int32_t *)p;
  return DRX_OK;
}

static int32_t *scratch_of_current_device ()
{
  int dev = 0;
  if (hipGetDevice (&dev) != hipSuccess || dev < 0 || dev >= kNaDevices) return nullptr;
  std::lock_guard<std::mutex> lock (g_scratch_mutex);
  return g_scratch_dev[dev];
}

} // namespace drx

using namespace drx;

extern "C" {

int drx_bvh_leaf_starts (const drx_mesh *mesh, const drx_bvh *bvh, float *out_leaf_starts, void *stream)
{
  int rc = check_mesh (mesh, 3, "drx_bvh_leaf_starts");
  if (rc) return rc;
  if (!bvh || !bvh->leaf_nodes || bvh->n_leaves < 0 || !out_leaf_starts)
    return fail (DRX_ERR_ARG, "drx_bvh_leaf_starts: bvh needs leaf_nodes; output required");
  if (bvh->n_leaves == 0) return DRX_OK;
  hipLaunchKernelGGL (k_leaf_starts, dim3 (nblk (bvh->n_leaves, 256)), dim3 (256), 0, as_stream (stream), mesh->data,
                      mesh->elem_offset, mesh->elem_order, mesh->stride, mesh->order, bvh->leaf_nodes, bvh->n_leaves,
                      (float4 *)out_leaf_starts);
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

int drx_bvh_quantize (const drx_bvh *bvh, uint32_t *out_qnodes, float *out_leaf_boxes, float *out_grid, void *stream)
{
  if (!bvh || !bvh->inner_nodes || bvh->n_leaves < 2 || !out_qnodes || !out_leaf_boxes || !out_grid)
    return fail (DRX_ERR_ARG, "drx_bvh_quantize: bvh needs inner_nodes and >= 2 leaves; outputs required");
  // 65,000 cells over the bounds plus a 200-cell margin on each side
  for (int d = 0; d < 3; ++d)
  {
    const float lo = bvh->bounds[2 * d], hi = bvh->bounds[2 * d + 1];
    if (!(lo <= hi) || !std::isfinite (lo) || !std::isfinite (hi))
      return fail (DRX_ERR_ARG, "drx_bvh_quantize: bounds not finite");
    const double ext = (double)hi - (double)lo;
    const double cell = ext > 0.0 ? ext / 65000.0 : 1.0;
    out_grid[d] = (float)((double)lo - 200.0 * cell);
    out_grid[3 + d] = (float)(1.0 / cell);
  }
  const int n_inner = bvh->n_leaves - 1;
  hipLaunchKernelGGL (k_quantize_nodes, dim3 (nblk (n_inner, 256)), dim3 (256), 0, as_stream (stream),
                      (const float4 *)bvh->inner_nodes, n_inner, out_grid[0], out_grid[1], out_grid[2], out_grid[3],
                      out_grid[4], out_grid[5], (uint4 *)out_qnodes, (float4 *)out_leaf_boxes);
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

// The body of drx_locate and drx_locate_noalloc.  `counters` (nullptr: the arena's flag slot) holds two
// int32: the traversal-stack overflow flag and the persistent launch's chunk counter; `ws` (nullptr: the
// arena's sort slots) the buffers of the Morton order.  Arguments are checked by the caller.
static int locate_impl (const drx_mesh *mesh, const drx_bvh *bvh, const float *points, int64_t n,
                        void *out_locations, const drx_locate_opts *opts, void *stream, int32_t *counters,
                        const SortScratch *ws)
{
  const bool learned = opts && opts->init_mode == DRX_INIT_MLP;
  const bool given = opts && opts->init_mode == DRX_INIT_GUESS;
  if (opts && opts->init_mode != DRX_INIT_AFFINE && opts->init_mode != DRX_INIT_SUBREF_CENTRE && !learned && !given)
    return fail (DRX_ERR_ARG, "drx_locate: unknown init_mode %d", opts->init_mode);
  if (given && !opts->guess) return fail (DRX_ERR_ARG, "drx_locate: DRX_INIT_GUESS needs opts->guess");
  if (learned && !opts->mlp) return fail (DRX_ERR_ARG, "drx_locate: DRX_INIT_MLP needs opts->mlp (drx_mlp_pack)");
  int launch = opts ? opts->launch : DRX_LAUNCH_AUTO;
  int chunk = opts ? opts->chunk : 0;
  if (launch < DRX_LAUNCH_AUTO || launch > DRX_LAUNCH_PERSISTENT || chunk < 0)
    return fail (DRX_ERR_ARG, "drx_locate: launch %d / chunk %d", launch, chunk);
  hipStream_t st = as_stream (stream);
  int32_t *err = counters ? counters : (int32_t *)arena (SLOT_FLAGS, 64, st);
  if (!err) return DRX_ERR_HIP;
  int32_t *queue = err + 1;   // chunk counter of the persistent launch
  hipLaunchKernelGGL (k_fill_i32q, dim3 (1), dim3 (64), 0, st, err, (int64_t)2, 0);
  const LocateCfg cfg = make_cfg (mesh, bvh, opts);
  int32_t *steps = opts ? opts->newton_steps : nullptr;
  int32_t *cands = opts ? opts->candidates : nullptr;
  const int P = mesh->elem_order ? 0 : mesh->order;
  int dev = 0;
  DRX_HIP (hipGetDevice (&dev));
  const int walk = stream_walk (bvh, opts);
  const bool tunable = learned || !(opts && opts->traversal == DRX_TRAVERSE_WAVE);
  if (tunable && (launch == DRX_LAUNCH_AUTO || chunk == 0) && dev >= 0 && dev < kTuneDevices)
  {
    const LocateTune &t = g_tune[dev][P][cfg.polish ? 1 : 0][walk];
    if (t.chunk > 0)
    {
      if (launch == DRX_LAUNCH_AUTO) launch = t.launch;
      if (chunk == 0) chunk = t.chunk;
    }
  }
  if (launch == DRX_LAUNCH_AUTO) launch = DRX_LAUNCH_CHUNKED;
  if (chunk == 0) chunk = kLocateChunk;
  if (chunk < 64 || chunk > (1 << 20)) return fail (DRX_ERR_ARG, "drx_locate: chunk %d outside [64, 2^20]", chunk);
  const int64_t n_chunks = (n + chunk - 1) / chunk;
  if (n_chunks > (1ll << 31) - 1) return fail (DRX_ERR_ARG, "drx_locate: too many chunks");
  int64_t grid_s = n_chunks;
  if (launch == DRX_LAUNCH_PERSISTENT)
  {
    const int64_t resident = (int64_t)compute_units (dev) * 4 * locate_waves_per_simd (P, cfg.polish, learned ? 3 : walk);
    grid_s = n_chunks < resident ? n_chunks : resident;
  }
  else
    queue = nullptr;
  const dim3 g (nblk (n, 256)), b (256);
  const dim3 gs ((unsigned)grid_s), bs (64);
  Location *out = (Location *)out_locations;
  const int trav = opts ? opts->traversal : DRX_TRAVERSE_AUTO;
  if (trav == DRX_TRAVERSE_WAVE && !cfg.wide)
    return fail (DRX_ERR_ARG, "drx_locate: wave traversal needs bvh->wide_nodes");
  // given starts ride on the binary walk
  const int mode = given ? 4 : trav == DRX_TRAVERSE_WAVE ? 2 : (trav == DRX_TRAVERSE_LANE && cfg.wide) ? 1 : 0;
  const int32_t *order = nullptr;
  if (opts && opts->coherent_order)
  {
    if (n > (1ll << 31) - 1) return fail (DRX_ERR_ARG, "drx_locate: coherent_order supports < 2^31 points");
    int rc2 = morton_order_points (points, (int)n, bvh->bounds, st, &order, ws);
    if (rc2) return rc2;
  }
  const MlpNet net = mlp_net (learned ? opts->mlp : nullptr);
  const int budget = (learned && opts->mlp_newton_steps > 0) ? opts->mlp_newton_steps : 4;
  if (P > DRX_MAX_MIXED_ORDER && (learned || mode != 0))
    return fail (DRX_ERR_UNSUPPORTED, "drx_locate: order %d runs the binary walk with the plain start only "
                 "(orders up to %d take the wide, wave, learned and given-start forms)", P, DRX_MAX_MIXED_ORDER);
#define DRX_LOC_HI(PP, PO)                                                                          \
  hipLaunchKernelGGL ((k_locate_stream<PP, PO, 0>), gs, bs, 0, st, cfg, points, n, order, out, steps, cands, err, \
                      chunk, queue)
#define DRX_LOC(PP, PO)                                                                             \
  do                                                                                                \
  {                                                                                                 \
    if (learned)                                                                                    \
      hipLaunchKernelGGL ((k_locate_mlp_stream<PP, PO>), gs, bs, 0, st, cfg, net, budget, points, n, order, out, \
                          steps, cands, err, chunk, queue);                                         \
    else if (mode == 2)                                                                             \
      hipLaunchKernelGGL ((k_locate<PP, PO, 2>), g, b, 0, st, cfg, points, n, order, out, steps, cands, err); \
    else if (mode == 1)                                                                             \
      hipLaunchKernelGGL ((k_locate_stream<PP, PO, 1>), gs, bs, 0, st, cfg, points, n, order, out, steps, cands, err, \
                          chunk, queue);                                                            \
    else if (mode == 4)                                                                             \
      hipLaunchKernelGGL ((k_locate_stream<PP, PO, 4>), gs, bs, 0, st, cfg, points, n, order, out, steps, cands, err, \
                          chunk, queue);                                                            \
    else                                                                                            \
      hipLaunchKernelGGL ((k_locate_stream<PP, PO, 0>), gs, bs, 0, st, cfg, points, n, order, out, steps, cands, err, \
                          chunk, queue);                                                            \
  } while (0)
  if (cfg.polish)
    switch (P)
    {
    case 0: DRX_LOC (0, true); break;
    case 1: DRX_LOC (1, true); break;
    case 2: DRX_LOC (2, true); break;
    case 3: DRX_LOC (3, true); break;
    case 4: DRX_LOC (4, true); break;
    case 5: DRX_LOC_HI (5, true); break;
    case 6: DRX_LOC_HI (6, true); break;
    case 7: DRX_LOC_HI (7, true); break;
    default: DRX_LOC_HI (8, true); break;
    }
  else
    switch (P)
    {
    case 0: DRX_LOC (0, false); break;
    case 1: DRX_LOC (1, false); break;
    case 2: DRX_LOC (2, false); break;
    case 3: DRX_LOC (3, false); break;
    case 4: DRX_LOC (4, false); break;
    case 5: DRX_LOC_HI (5, false); break;
    case 6: DRX_LOC_HI (6, false); break;
    case 7: DRX_LOC_HI (7, false); break;
    default: DRX_LOC_HI (8, false); break;
    }
#undef DRX_LOC
#undef DRX_LOC_HI
  DRX_LAUNCH_CHECK ();
  int32_t herr = 0;
  DRX_HIP (hipMemcpyAsync (&herr, err, sizeof (int32_t), hipMemcpyDeviceToHost, st));
  DRX_HIP (hipStreamSynchronize (st));
  if (herr) return fail (DRX_ERR_OVERFLOW, "drx_locate: BVH traversal stack (%d entries) overflowed", kStackDepth);
  return DRX_OK;
}

static int locate_check (const drx_mesh *mesh, const drx_bvh *bvh, const float *points, int64_t n,
                         void *out_locations, const char *who)
{
  int rc = check_mesh (mesh, 3, who);
  if (rc) return rc;
  if ((rc = check_bvh (bvh, who))) return rc;
  if (n < 0) return fail (DRX_ERR_ARG, "%s: n < 0", who);
  if (n > 0 && (!points || !out_locations)) return fail (DRX_ERR_ARG, "%s: null points/output", who);
  return DRX_OK;
}


int drx_locate (const drx_mesh *mesh, const drx_bvh *bvh, const float *points, int64_t n, void *out_locations,
                const drx_locate_opts *opts, void *stream)
{
  int rc = locate_check (mesh, bvh, points, n, out_locations, "drx_locate");
  if (rc || n == 0) return rc;
  return locate_impl (mesh, bvh, points, n, out_locations, opts, stream, nullptr, nullptr);
}

int drx_locate_noalloc (const drx_mesh *mesh, const drx_bvh *bvh, const float *points, int64_t n,
                        void *out_locations, const drx_locate_opts *opts, void *stream)
{
  int rc = locate_check (mesh, bvh, points, n, out_locations, "drx_locate_noalloc");
  if (rc || n == 0) return rc;
  if (opts && opts->coherent_order && n > kNaOrdered)
    return fail (DRX_ERR_ARG, "drx_locate_noalloc: coherent_order over %lld points exceeds the scratch limit of %lld "
                 "(DRX_LOCATE_NOALLOC_MAX_ORDERED); split the batch or use drx_locate",
                 (long long)n, (long long)kNaOrdered);
  const int64_t hist_need = radix_hist_len ((int)(n < kNaOrdered ? n : kNaOrdered));
  const int64_t sums_need = scan_sums_len (256 * radix_tiles ((int)(n < kNaOrdered ? n : kNaOrdered)));
  if (hist_need > kNaHistWords || sums_need > kNaSumsWords)
    return fail (DRX_ERR_STATE, "drx_locate_noalloc: scratch layout too small (%lld/%lld hist, %lld/%lld sums)",
                 (long long)hist_need, (long long)kNaHistWords, (long long)sums_need, (long long)kNaSumsWords);
  int32_t *base = scratch_of_current_device ();
  if (!base)
  {
    if ((rc = locate_noalloc_setup ())) return rc;
    base = scratch_of_current_device ();
  }
  SortScratch ws;
  ws.kA = (uint32_t *)(base + kNaHeaderWords);
  ws.kB = ws.kA + kNaOrdered;
  ws.vA = (int32_t *)(ws.kB + kNaOrdered);
  ws.vB = ws.vA + kNaOrdered;
  ws.hist = ws.vB + kNaOrdered;
  ws.scan_sums = ws.hist + ((kNaHistWords + 63) & ~63);
  return locate_impl (mesh, bvh, points, n, out_locations, opts, stream, base, &ws);
}

int drx_locate_autotune (const drx_mesh *mesh, const drx_bvh *bvh, const float *points, int64_t n,
                         void *out_locations, const drx_locate_opts *opts, int32_t *out_launch, int32_t *out_chunk,
                         float *out_ms, void *stream)
{
  if (!out_launch || !out_chunk) return fail (DRX_ERR_ARG, "drx_locate_autotune: null output");
  if (n <= 0) return fail (DRX_ERR_ARG, "drx_locate_autotune: no points");
  if (opts && opts->init_mode != DRX_INIT_MLP && opts->traversal == DRX_TRAVERSE_WAVE)
    return fail (DRX_ERR_ARG, "drx_locate_autotune: the wave walk is not streamed");
  int rc = check_mesh (mesh, 3, "drx_locate_autotune");
  if (rc) return rc;
  if ((rc = check_bvh (bvh, "drx_locate_autotune"))) return rc;
  drx_locate_opts o {};
  if (opts)
    o = *opts;
  else
  {
    o.cond_threshold = 1e4f;
    o.traversal = DRX_TRAVERSE_AUTO;
  }
  int dev = 0;
  DRX_HIP (hipGetDevice (&dev));
  hipStream_t st = as_stream (stream);
  static const int kChunks[5] = { 256, 512, 1024, 2048, 4096 };
  hipEvent_t e0, e1;
  DRX_HIP (hipEventCreate (&e0));
  if (hipEventCreate (&e1) != hipSuccess)
  {
    (void)hipEventDestroy (e0);
    return fail (DRX_ERR_HIP, "drx_locate_autotune: hipEventCreate");
  }
  float best = INFINITY;
  int best_launch = DRX_LAUNCH_CHUNKED, best_chunk = kLocateChunk;
  for (int c = 0; c < 5 && rc == DRX_OK; ++c)
    for (int l = 0; l < 2 && rc == DRX_OK; ++l)
    {
      o.launch = l ? DRX_LAUNCH_PERSISTENT : DRX_LAUNCH_CHUNKED;
      o.chunk = kChunks[c];
      rc = drx_locate (mesh, bvh, points, n, out_locations, &o, stream);   // warm-up
      if (rc) break;
      if (hipEventRecord (e0, st) != hipSuccess) rc = fail (DRX_ERR_HIP, "drx_locate_autotune: hipEventRecord");
      for (int r = 0; r < 2 && rc == DRX_OK; ++r) rc = drx_locate (mesh, bvh, points, n, out_locations, &o, stream);
      float ms = 0.f;
      if (rc == DRX_OK && (hipEventRecord (e1, st) != hipSuccess || hipEventSynchronize (e1) != hipSuccess ||
                           hipEventElapsedTime (&ms, e0, e1) != hipSuccess))
        rc = fail (DRX_ERR_HIP, "drx_locate_autotune: event timing");
      if (rc) break;
      ms *= 0.5f;
      if (out_ms) out_ms[2 * c + l] = ms;
      if (ms < best)
      {
        best = ms;
        best_launch = o.launch;
        best_chunk = o.chunk;
      }
    }
  (void)hipEventDestroy (e0);
  (void)hipEventDestroy (e1);
  if (rc) return rc;
  const int P = mesh->elem_order ? 0 : mesh->order;
  const bool polish = o.fp64_polish != 0;
  if (dev >= 0 && dev < kTuneDevices) g_tune[dev][P][polish ? 1 : 0][stream_walk (bvh, &o)] = LocateTune { best_launch, best_chunk };
  *out_launch = best_launch;
  *out_chunk = best_chunk;
  return DRX_OK;
}

int drx_field_eval (const drx_mesh *field, const void *locations, int64_t n, float *out_values, void *stream)
{
  int rc = check_mesh (field, 1, "drx_field_eval");
  if (rc) return rc;
  if (n < 0) return fail (DRX_ERR_ARG, "drx_field_eval: n < 0");
  if (n == 0) return DRX_OK;
  if (!locations || !out_values) return fail (DRX_ERR_ARG, "drx_field_eval: null pointer");
  hipStream_t st = as_stream (stream);
  const Location *l = (const Location *)locations;
  const dim3 g (nblk (n, 256)), b (256);
#define DRX_FE(FT, PTR)                                                                                               \
  switch (field->elem_order ? 0 : field->order)                                                                       \
  {                                                                                                                   \
  case 0: hipLaunchKernelGGL ((k_field_eval<0, FT>), g, b, 0, st, PTR, field->elem_offset, field->elem_order, l, n, out_values); break; \
  case 1: hipLaunchKernelGGL ((k_field_eval<1, FT>), g, b, 0, st, PTR, nullptr, nullptr, l, n, out_values); break;   \
  case 2: hipLaunchKernelGGL ((k_field_eval<2, FT>), g, b, 0, st, PTR, nullptr, nullptr, l, n, out_values); break;   \
  case 3: hipLaunchKernelGGL ((k_field_eval<3, FT>), g, b, 0, st, PTR, nullptr, nullptr, l, n, out_values); break;   \
  case 4: hipLaunchKernelGGL ((k_field_eval<4, FT>), g, b, 0, st, PTR, nullptr, nullptr, l, n, out_values); break;   \
  case 5: hipLaunchKernelGGL ((k_field_eval<5, FT>), g, b, 0, st, PTR, nullptr, nullptr, l, n, out_values); break;   \
  case 6: hipLaunchKernelGGL ((k_field_eval<6, FT>), g, b, 0, st, PTR, nullptr, nullptr, l, n, out_values); break;   \
  case 7: hipLaunchKernelGGL ((k_field_eval<7, FT>), g, b, 0, st, PTR, nullptr, nullptr, l, n, out_values); break;   \
  default: hipLaunchKernelGGL ((k_field_eval<8, FT>), g, b, 0, st, PTR, nullptr, nullptr, l, n, out_values); break;  \
  }
  if (field->data64)
  {
    DRX_FE (double, field->data64)
  }
  else
  {
    DRX_FE (float, field->data)
  }
#undef DRX_FE
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

int drx_camera_rays (const drx_camera *cam, void *out_rays, void *stream)
{
  if (!cam || !out_rays) return fail (DRX_ERR_ARG, "drx_camera_rays: null pointer");
  if (cam->width <= 0 || cam->height <= 0) return fail (DRX_ERR_ARG, "Camera width/height must be greater than zero.");
  if (cam->fov_y <= 0 || cam->fov_y > 180) return fail (DRX_ERR_ARG, "Camera field of view must be in (0, 180].");
  const CamSetup cs = camera_setup (*cam);
  const int64_t npix = (int64_t)cam->width * cam->height;
  hipLaunchKernelGGL (k_camera, dim3 (nblk (npix, 256)), dim3 (256), 0, as_stream (stream), cs, (Ray *)out_rays);
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

int drx_remove_missed_rays (const void *rays, int32_t n, const float *bounds6, void *out_rays, int32_t *out_count,
                            void *stream)
{
  if (n < 0 || !bounds6 || !out_count) return fail (DRX_ERR_ARG, "drx_remove_missed_rays: bad arguments");
  *out_count = 0;
  if (n == 0) return DRX_OK;
  if (!rays || !out_rays) return fail (DRX_ERR_ARG, "drx_remove_missed_rays: null rays");
  // an empty bounds box yields no rays (ray.cpp:105)
  bool empty = true;
  for (int d = 0; d < 3; ++d) empty = empty && (bounds6[2 * d] > bounds6[2 * d + 1]);
  if (empty) return DRX_OK;
  // mesh_bounds.scale(1.001f) on the host, float32 like Range::scale
  Bounds6 mb;
  for (int d = 0; d < 3; ++d)
  {
    float lo = bounds6[2 * d], hi = bounds6[2 * d + 1];
    if (!(lo > hi))
    {
      volatile float c = 0.5f * (lo + hi);
      volatile float len = hi - lo;
      volatile float half = 1.001f * 0.5f;
      volatile float delta = half * len;
      volatile float v0 = c - delta, v1 = c + delta;
      lo = std::fmin (lo, (float)v0);
      hi = std::fmax (hi, (float)v0);
      lo = std::fmin (lo, (float)v1);
      hi = std::fmax (hi, (float)v1);
    }
    mb.b[2 * d] = lo;
    mb.b[2 * d + 1] = hi;
  }
  hipStream_t st = as_stream (stream);
  int32_t *flags = (int32_t *)arena (SLOT_FLAGS, sizeof (int32_t) * 2 * (size_t)n + 64, st);
  if (!flags) return DRX_ERR_HIP;
  int32_t *pos = flags + n;
  hipLaunchKernelGGL (k_mark_active, dim3 (nblk (n, 256)), dim3 (256), 0, st, (const Ray *)rays, n, mb, flags);
  DRX_LAUNCH_CHECK ();
  int32_t total = 0;
  int rc = exclusive_scan_i32 (flags, pos, n, &total, st);
  if (rc) return rc;
  hipLaunchKernelGGL (k_gather_active, dim3 (nblk (n, 256)), dim3 (256), 0, st, (const Ray *)rays, n, mb, flags, pos,
                      (Ray *)out_rays);
  DRX_LAUNCH_CHECK ();
  DRX_HIP (hipStreamSynchronize (st));
  *out_count = total;
  return DRX_OK;
}

int drx_ray_tile_order (const void *rays, int32_t n, int32_t width, int32_t height, int32_t *out_order, void *stream)
{
  if (n < 0 || width <= 0 || height <= 0) return fail (DRX_ERR_ARG, "drx_ray_tile_order: bad sizes");
  if (n == 0) return DRX_OK;
  if (!rays || !out_order) return fail (DRX_ERR_ARG, "drx_ray_tile_order: null pointer");
  hipStream_t st = as_stream (stream);
  const int tiles_x = (width + kTileW - 1) / kTileW, tiles_y = (height + kTileH - 1) / kTileH;
  const int64_t slots_n = (int64_t)tiles_x * tiles_y * 64;
  if (slots_n > (1ll << 30)) return fail (DRX_ERR_ARG, "drx_ray_tile_order: image too large");
  int32_t *slots0 = (int32_t *)arena (SLOT_TREE2, sizeof (int32_t) * 4 * (size_t)slots_n + 64, st);
  if (!slots0) return DRX_ERR_HIP;
  int32_t *slots = slots0 + slots_n, *flags = slots + slots_n, *pos = flags + slots_n;
  hipLaunchKernelGGL (k_fill_i32q, dim3 (nblk (slots_n, 256)), dim3 (256), 0, st, slots0, slots_n, -1);
  hipLaunchKernelGGL (k_tile_scatter, dim3 (nblk (n, 256)), dim3 (256), 0, st, (const Ray *)rays, n, width, height, slots0);
  DRX_LAUNCH_CHECK ();
  // Tiles longest-first (stable): a frame needs more wavefronts than the GPU holds at once, and the ones
  // dispatched last run on an emptying GPU, so they should be the short ones.
  const int n_img_tiles = tiles_x * tiles_y;
  const int n_rt = radix_tiles (n_img_tiles);
  uint32_t *kA = (uint32_t *)arena (SLOT_SORT_KEYS_A, sizeof (uint32_t) * (size_t)n_img_tiles, st);
  uint32_t *kB = (uint32_t *)arena (SLOT_SORT_KEYS_B, sizeof (uint32_t) * (size_t)n_img_tiles, st);
  int32_t *vA = (int32_t *)arena (SLOT_SORT_VALS_A, sizeof (int32_t) * (size_t)n_img_tiles, st);
  int32_t *vB = (int32_t *)arena (SLOT_SORT_VALS_B, sizeof (int32_t) * (size_t)n_img_tiles, st);
  int32_t *hist = (int32_t *)arena (SLOT_SORT_HIST, sizeof (int32_t) * (2 * 256 * (size_t)n_rt + 8), st);
  if (!kA || !kB || !vA || !vB || !hist) return DRX_ERR_HIP;
  hipLaunchKernelGGL (k_tile_keys, dim3 (nblk (n_img_tiles, 256)), dim3 (256), 0, st, (const Ray *)rays, slots0,
                      n_img_tiles, kA, vA);
  DRX_LAUNCH_CHECK ();
  uint32_t *ks;
  int32_t *vs;
  int rc0 = radix_sort_pairs (kA, vA, kB, vB, hist, n_img_tiles, 32, st, &ks, &vs);
  if (rc0) return rc0;
  hipLaunchKernelGGL (k_gather_tiles, dim3 (nblk (slots_n, 256)), dim3 (256), 0, st, slots0, vs, slots_n, slots);
  hipLaunchKernelGGL (k_flag_nonneg, dim3 (nblk (slots_n, 256)), dim3 (256), 0, st, slots, slots_n, flags);
  DRX_LAUNCH_CHECK ();
  int32_t total = 0;
  int rc = exclusive_scan_i32 (flags, pos, (int32_t)slots_n, &total, st);
  if (rc) return rc;
  if (total != n)
    return fail (DRX_ERR_ARG, "drx_ray_tile_order: rays are not distinct pixels of a %dx%d image", width, height);
  hipLaunchKernelGGL (k_compact_i32, dim3 (nblk (slots_n, 256)), dim3 (256), 0, st, slots, slots_n, pos, out_order);
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

// Stable (pixel id, depth) order of partials from any number of domains, ties kept in input order: an
// LSD radix sort by depth (order-preserving unsigned key), then stably by pixel id.
static int sort_partials_pixel_depth (const Partial *in, int32_t n, int64_t npix, hipStream_t st, Partial **out)
{
  const int n_tiles = radix_tiles (n);
  uint32_t *kA = (uint32_t *)arena (SLOT_SORT_KEYS_A, sizeof (uint32_t) * (size_t)n, st);
  uint32_t *kB = (uint32_t *)arena (SLOT_SORT_KEYS_B, sizeof (uint32_t) * (size_t)n, st);
  int32_t *vA = (int32_t *)arena (SLOT_SORT_VALS_A, sizeof (int32_t) * (size_t)n, st);
  int32_t *vB = (int32_t *)arena (SLOT_SORT_VALS_B, sizeof (int32_t) * (size_t)n, st);
  int32_t *hist = (int32_t *)arena (SLOT_SORT_HIST, sizeof (int32_t) * (2 * 256 * (size_t)n_tiles + 8), st);
  Partial *sorted = (Partial *)arena (SLOT_PARTIALS, sizeof (Partial) * (size_t)n, st);
  if (!kA || !kB || !vA || !vB || !hist || !sorted) return DRX_ERR_HIP;
  hipLaunchKernelGGL (k_depth_keys, dim3 (nblk (n, 256)), dim3 (256), 0, st, in, n, kA, vA);
  DRX_LAUNCH_CHECK ();
  uint32_t *ks;
  int32_t *vs;
  int rc = radix_sort_pairs (kA, vA, kB, vB, hist, n, 32, st, &ks, &vs);
  if (rc) return rc;
  uint32_t *k2 = (ks == kA) ? kB : kA;
  int32_t *v2 = (vs == vA) ? vB : vA;
  hipLaunchKernelGGL (k_pixel_keys, dim3 (nblk (n, 256)), dim3 (256), 0, st, in, vs, n, k2, v2);
  DRX_LAUNCH_CHECK ();
  int pix_bits = 1;
  while (pix_bits < 32 && (int64_t (1) << pix_bits) < npix) ++pix_bits;
  rc = radix_sort_pairs (k2, v2, ks, vs, hist, n, pix_bits, st, &ks, &vs);
  if (rc) return rc;
  hipLaunchKernelGGL (k_gather_partials, dim3 (nblk (n, 256)), dim3 (256), 0, st, in, vs, n, sorted);
  DRX_LAUNCH_CHECK ();
  *out = sorted;
  return DRX_OK;
}

static bool image_ok (int32_t width, int32_t height)
{
  return width > 0 && height > 0 && (int64_t)width * height <= (int64_t (1) << 30);
}

int drx_framebuffer_clear (float *colors, float *depths, int32_t n_pixels, const float color[4], void *stream)
{
  if (n_pixels < 0 || !color || (n_pixels > 0 && !colors))
    return fail (DRX_ERR_ARG, "drx_framebuffer_clear: bad arguments");
  if (n_pixels == 0) return DRX_OK;
  hipStream_t st = as_stream (stream);
  hipLaunchKernelGGL (k_fill_f4, dim3 (nblk (n_pixels, 256)), dim3 (256), 0, st, (float4 *)colors, (int64_t)n_pixels,
                      make_float4 (color[0], color[1], color[2], color[3]));
  if (depths)
    hipLaunchKernelGGL (k_fill_f32, dim3 (nblk (n_pixels, 256)), dim3 (256), 0, st, depths, (int64_t)n_pixels,
                        __builtin_huge_valf ());
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

int drx_partials_to_framebuffer (const void *partials, int32_t n, int32_t width, int32_t height, float *colors,
                                 float *depths, int32_t blend, void *stream)
{
  if (n < 0 || !image_ok (width, height) || !colors || (n > 0 && !partials) || (blend != 0 && blend != 1))
    return fail (DRX_ERR_ARG, "drx_partials_to_framebuffer: bad arguments");
  if (n == 0) return DRX_OK;
  hipStream_t st = as_stream (stream);
  Partial *sorted = nullptr;
  int rc = sort_partials_pixel_depth ((const Partial *)partials, n, (int64_t)width * height, st, &sorted);
  if (rc) return rc;
  hipLaunchKernelGGL (k_composite_runs, dim3 (nblk (n, 256)), dim3 (256), 0, st, sorted, n, width * height,
                      (float4 *)colors, depths, (int)blend);
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

int drx_framebuffer_composite_background (float *colors, int32_t n_pixels, const float bg[4], void *stream)
{
  if (n_pixels < 0 || !bg || (n_pixels > 0 && !colors))
    return fail (DRX_ERR_ARG, "drx_framebuffer_composite_background: bad arguments");
  if (n_pixels == 0) return DRX_OK;
  hipLaunchKernelGGL (k_background, dim3 (nblk (n_pixels, 256)), dim3 (256), 0, as_stream (stream), (float4 *)colors,
                      (int)n_pixels, make_float4 (bg[0], bg[1], bg[2], bg[3]));
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

int drx_composite_domains (const void *partials, int32_t n, int32_t width, int32_t height, float *out_rgba,
                           void *stream)
{
  if (n < 0 || !image_ok (width, height) || !out_rgba || (n > 0 && !partials))
    return fail (DRX_ERR_ARG, "drx_composite_domains: bad arguments");
  const float zero[4] = { 0.f, 0.f, 0.f, 0.f }, white[4] = { 1.f, 1.f, 1.f, 1.f };
  const int32_t npix = width * height;
  int rc = drx_framebuffer_clear (out_rgba, nullptr, npix, zero, stream);
  if (!rc) rc = drx_partials_to_framebuffer (partials, n, width, height, out_rgba, nullptr, 0, stream);
  if (!rc) rc = drx_framebuffer_composite_background (out_rgba, npix, white, stream);
  return rc;
}

int drx_composite (const void *partials, int32_t n, int32_t width, int32_t height, float *out_rgba, void *stream)
{
  if (n < 0 || !image_ok (width, height) || !out_rgba || (n > 0 && !partials))
    return fail (DRX_ERR_ARG, "drx_composite: bad arguments");
  hipStream_t st = as_stream (stream);
  const int npix = width * height;
  hipLaunchKernelGGL (k_fill_f32, dim3 (nblk (4 * (int64_t)npix, 256)), dim3 (256), 0, st, out_rgba, 4 * (int64_t)npix, 0.f);
  if (n > 0)
    hipLaunchKernelGGL (k_composite_runs, dim3 (nblk (n, 256)), dim3 (256), 0, st, (const Partial *)partials, n, npix,
                        (float4 *)out_rgba, (float *)nullptr, 0);
  hipLaunchKernelGGL (k_background, dim3 (nblk (npix, 256)), dim3 (256), 0, st, (float4 *)out_rgba, npix,
                      make_float4 (1.f, 1.f, 1.f, 1.f));
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

} // extern "C"

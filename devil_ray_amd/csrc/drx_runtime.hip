// libdrx runtime: error state, device arena, init, scans, mesh packing, colour tables.
#include "drx_common.hpp"
#include "drx_wave.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "drx_color_presets.inc"

namespace drx
{

static thread_local std::string g_last_error;

void set_error (const char *fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start (ap, fmt);
  vsnprintf (buf, sizeof buf, fmt, ap);
  va_end (ap);
  g_last_error = buf;
}

int fail (int status, const char *fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start (ap, fmt);
  vsnprintf (buf, sizeof buf, fmt, ap);
  va_end (ap);
  g_last_error = buf;
  return status;
}

// ---- arena ---------------------------------------------------------------------------------------
// One set of growable scratch slots per (device, stream): work on different streams never shares scratch,
// and work on one stream is ordered by the stream, so a slot may be reused by the next call on its stream
// while the previous call's kernels are still queued.  Growing a slot frees the old buffer and allocates
// the new one in stream order (hipFreeAsync / hipMallocAsync on the slot's stream): no device-wide wait.
namespace
{
struct Slot
{
  void *ptr = nullptr;
  size_t bytes = 0;
};
struct SlotSet
{
  int dev;
  hipStream_t stream;
  Slot slots[SLOT_COUNT];
};
std::vector<SlotSet> g_sets;
std::mutex g_arena_mutex;
} // namespace

void *arena (int slot, size_t bytes, hipStream_t stream)
{
  int dev = 0;
  if (hipGetDevice (&dev) != hipSuccess || dev < 0)
  {
    set_error ("arena: no current device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock (g_arena_mutex);
  SlotSet *set = nullptr;
  for (auto &x : g_sets)
    if (x.dev == dev && x.stream == stream) set = &x;
  if (!set)
  {
    g_sets.push_back (SlotSet { dev, stream, {} });
    set = &g_sets.back ();
  }
  Slot &s = set->slots[slot];
  if (bytes == 0) bytes = 256;
  if (s.bytes >= bytes) return s.ptr;
  if (s.ptr)
  {
    // only work queued on `stream` can still read the old buffer: free it in stream order
    const hipError_t ef = hipFreeAsync (s.ptr, stream);
    s.ptr = nullptr;
    s.bytes = 0;
    if (ef != hipSuccess)
    {
      set_error ("arena: growing slot %d: hipFreeAsync: %s", slot, hipGetErrorString (ef));
      return nullptr;
    }
  }
  const size_t want = bytes + bytes / 8;   // grow with headroom
  hipError_t e = hipMallocAsync (&s.ptr, want, stream);
  if (e != hipSuccess)
  {
    s.ptr = nullptr;
    set_error ("arena: hipMallocAsync(%zu) failed: %s", want, hipGetErrorString (e));
    return nullptr;
  }
  s.bytes = want;
  return s.ptr;
}

int arena_release ()
{
  int dev = 0;
  DRX_HIP (hipGetDevice (&dev));
  DRX_HIP (hipDeviceSynchronize ());
  std::lock_guard<std::mutex> lock (g_arena_mutex);
  hipError_t first = hipSuccess;
  for (auto it = g_sets.begin (); it != g_sets.end ();)
  {
    if (it->dev != dev)
    {
      ++it;
      continue;
    }
    for (int i = 0; i < SLOT_COUNT; ++i)
      if (it->slots[i].ptr)
      {
        const hipError_t e = hipFree (it->slots[i].ptr);
        if (e != hipSuccess && first == hipSuccess) first = e;
      }
    it = g_sets.erase (it);
  }
  if (first != hipSuccess) return fail (DRX_ERR_HIP, "drx_release: hipFree: %s", hipGetErrorString (first));
  return DRX_OK;
}

// ---- exclusive scan (int32) ---------------------------------------------------------------------
// Three passes: per-block totals, one-block scan of the totals, per-block rescan with its offset.
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ void __launch_bounds__ (kScanBlock) k_scan_reduce (const int32_t *in, int32_t n, int32_t *block_sums)
{
  __shared__ int32_t wsum[kScanBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int32_t s = 0;
  for (int k = 0; k < kScanItems; ++k)
  {
    const int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  s = wave_reduce_add (s);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads ();
  if (threadIdx.x == 0)
  {
    int32_t t = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) t += wsum[w];
    block_sums[blockIdx.x] = t;
  }
}

// Single block: exclusive scan of block_sums in place; total -> block_sums[nb].
__global__ void __launch_bounds__ (kScanBlock) k_scan_sums (int32_t *block_sums, int32_t nb)
{
  __shared__ int32_t wsum[kScanBlock / 64];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads ();
  for (int32_t base = 0; base < nb; base += kScanBlock)
  {
    const int32_t i = base + threadIdx.x;
    const int32_t v = (i < nb) ? block_sums[i] : 0;
    const int32_t incl = wave_scan_incl_add (v);
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = incl;
    __syncthreads ();
    int32_t woff = 0;
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) woff += wsum[w];
    const int32_t c = carry;
    if (i < nb) block_sums[i] = c + woff + incl - v;
    __syncthreads ();
    if (threadIdx.x == kScanBlock - 1) carry = c + woff + incl;
    __syncthreads ();
  }
  if (threadIdx.x == 0) block_sums[nb] = carry;
}

__global__ void __launch_bounds__ (kScanBlock) k_scan_down (const int32_t *in, int32_t *out, int32_t n,
                                                             const int32_t *block_sums)
{
  __shared__ int32_t wsum[kScanBlock / 64];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = block_sums[blockIdx.x];
  __syncthreads ();
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  for (int k = 0; k < kScanItems; ++k)
  {
    const int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
    const int32_t v = (i < n) ? in[i] : 0;
    const int32_t incl = wave_scan_incl_add (v);
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = incl;
    __syncthreads ();
    int32_t woff = 0;
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) woff += wsum[w];
    const int32_t c = carry;
    if (i < n) out[i] = c + woff + incl - v;
    __syncthreads ();
    if (threadIdx.x == kScanBlock - 1) carry = c + woff + incl;
    __syncthreads ();
  }
}

int32_t scan_sums_len (int32_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

int exclusive_scan_i32 (const int32_t *in, int32_t *out, int32_t n, int32_t *total_host, hipStream_t st, int32_t *sums)
{
  if (n <= 0)
  {
    if (total_host) *total_host = 0;
    return DRX_OK;
  }
  const int32_t nb = (n + kScanTile - 1) / kScanTile;
  if (!sums) sums = (int32_t *)arena (SLOT_SCAN, sizeof (int32_t) * (nb + 1), st);
  if (!sums) return DRX_ERR_HIP;
  hipLaunchKernelGGL (k_scan_reduce, dim3 (nb), dim3 (kScanBlock), 0, st, in, n, sums);
  hipLaunchKernelGGL (k_scan_sums, dim3 (1), dim3 (kScanBlock), 0, st, sums, nb);
  hipLaunchKernelGGL (k_scan_down, dim3 (nb), dim3 (kScanBlock), 0, st, in, out, n, sums);
  DRX_LAUNCH_CHECK ();
  if (total_host)
  {
    DRX_HIP (hipMemcpyAsync (total_host, sums + nb, sizeof (int32_t), hipMemcpyDeviceToHost, st));
    DRX_HIP (hipStreamSynchronize (st));
  }
  return DRX_OK;
}

// ---- mesh packing -------------------------------------------------------------------------------
// One wavefront per element: lane d handles dofs d, d+64, ...; the element origin is the centre of the
// control-net box, reduced across the wave.
__global__ void __launch_bounds__ (64) k_mesh_pack (const int32_t *__restrict__ ctrl_idx, const float *__restrict__ values,
                                                     int ncomp, const int32_t *__restrict__ orders,
                                                     const int64_t *__restrict__ offsets,
                                                     const int64_t *__restrict__ out_offsets, int order, int n_el,
                                                     float *__restrict__ out)
{
  const int e = blockIdx.x;
  if (e >= n_el) return;
  const int p = orders ? orders[e] : order;
  const int npe = (p + 1) * (p + 1) * (p + 1);
  const int npe_pad = (npe + 3) & ~3;
  const int64_t in_off = offsets ? offsets[e] : (int64_t)e * npe;
  const int stride = (ncomp == 3) ? 4 + 3 * npe_pad : npe_pad;
  const int64_t o = out_offsets ? out_offsets[e] : (int64_t)e * stride;
  const int lane = threadIdx.x;
  if (ncomp == 1)
  {
    for (int d = lane; d < npe_pad; d += 64)
      out[o + d] = (d < npe) ? values[ctrl_idx[in_off + d]] : 0.f;
    return;
  }
  float mn[3] = { INFINITY, INFINITY, INFINITY }, mx[3] = { -INFINITY, -INFINITY, -INFINITY };
  for (int d = lane; d < npe; d += 64)
  {
    const int64_t g = (int64_t)ctrl_idx[in_off + d] * 3;
    for (int c = 0; c < 3; ++c)
    {
      const float v = values[g + c];
      mn[c] = fminf (mn[c], v);
      mx[c] = fmaxf (mx[c], v);
    }
  }
  float org[3];
  for (int c = 0; c < 3; ++c)
  {
    const float a = wave_reduce_min (mn[c]), b = wave_reduce_max (mx[c]);
    org[c] = (float)(0.5 * ((double)a + (double)b));
  }
  if (lane < 4) out[o + lane] = (lane < 3) ? org[lane] : 0.f;
  for (int d = lane; d < npe_pad; d += 64)
  {
    if (d < npe)
    {
      const int64_t g = (int64_t)ctrl_idx[in_off + d] * 3;
      for (int c = 0; c < 3; ++c) out[o + 4 + c * npe_pad + d] = (float)((double)values[g + c] - (double)org[c]);
    }
    else
      for (int c = 0; c < 3; ++c) out[o + 4 + c * npe_pad + d] = 0.f;
  }
}

// FP64 scalar field: one wavefront per element, lane d copies dofs d, d+64, ... (zero padding up to npe_pad).
__global__ void __launch_bounds__ (64) k_mesh_pack_f64 (const int32_t *__restrict__ ctrl_idx,
                                                         const double *__restrict__ values,
                                                         const int32_t *__restrict__ orders,
                                                         const int64_t *__restrict__ offsets,
                                                         const int64_t *__restrict__ out_offsets, int order, int n_el,
                                                         double *__restrict__ out)
{
  const int e = blockIdx.x;
  if (e >= n_el) return;
  const int p = orders ? orders[e] : order;
  const int npe = (p + 1) * (p + 1) * (p + 1);
  const int npe_pad = (npe + 3) & ~3;
  const int64_t in_off = offsets ? offsets[e] : (int64_t)e * npe;
  const int64_t o = out_offsets ? out_offsets[e] : (int64_t)e * npe_pad;
  for (int d = threadIdx.x; d < npe_pad; d += 64) out[o + d] = (d < npe) ? values[ctrl_idx[in_off + d]] : 0.0;
}

} // namespace drx

using namespace drx;

extern "C" {

int drx_init (int device)
{
  int n = 0;
  DRX_HIP (hipGetDeviceCount (&n));
  if (device < 0 || device >= n) return fail (DRX_ERR_ARG, "drx_init: device %d out of range (%d devices)", device, n);
  DRX_HIP (hipSetDevice (device));
  hipDeviceProp_t prop;
  DRX_HIP (hipGetDeviceProperties (&prop, device));
  if (std::strncmp (prop.gcnArchName, "gfx950", 6) != 0)
    return fail (DRX_ERR_UNSUPPORTED, "drx_init: device %d is %s; this library is built for gfx950 only", device,
                 prop.gcnArchName);
  return locate_noalloc_setup ();
}

const char *drx_last_error (void) { return g_last_error.c_str (); }

const char *drx_version (void) { return "drx 0.1 (gfx950)"; }

int drx_release (void) { return arena_release (); }

int32_t drx_mesh_stride (int32_t order, int32_t ncomp)
{
  if (order < 1) return 0;
  const int32_t npe_pad = npe_pad_of (order);
  return ncomp == 3 ? 4 + 3 * npe_pad : npe_pad;
}

int64_t drx_mesh_pack_layout (const int32_t *h_orders, int32_t n_el, int32_t ncomp, int64_t *h_offsets)
{
  int64_t total = 0;
  for (int32_t e = 0; e < n_el; ++e)
  {
    if (h_orders[e] < 1 || h_orders[e] > DRX_MAX_MIXED_ORDER)
    {
      fail (DRX_ERR_ARG, "drx_mesh_pack_layout: element %d has order %d (mixed-order meshes take 1..%d)", e,
            h_orders[e], DRX_MAX_MIXED_ORDER);
      return -1;
    }
    h_offsets[e] = total;
    total += drx_mesh_stride (h_orders[e], ncomp);
  }
  return total;
}

int drx_mesh_pack (const int32_t *ctrl_idx, const float *values, int32_t ncomp, const int32_t *orders,
                   const int64_t *offsets, const int64_t *out_offsets, int32_t order, int32_t n_el, float *out,
                   void *stream)
{
  if (n_el < 0 || (ncomp != 1 && ncomp != 3)) return fail (DRX_ERR_ARG, "drx_mesh_pack: bad n_el/ncomp");
  if (n_el == 0) return DRX_OK;
  if (!ctrl_idx || !values || !out) return fail (DRX_ERR_ARG, "drx_mesh_pack: null pointer");
  if (!orders && (order < 1 || order > DRX_MAX_ORDER))
    return fail (DRX_ERR_ARG, "drx_mesh_pack: order %d outside 1..%d", order, DRX_MAX_ORDER);
  if ((orders == nullptr) != (out_offsets == nullptr) || (orders == nullptr) != (offsets == nullptr))
    return fail (DRX_ERR_ARG, "drx_mesh_pack: mixed-order meshes need orders, offsets and out_offsets");
  hipLaunchKernelGGL (k_mesh_pack, dim3 (n_el), dim3 (64), 0, as_stream (stream), ctrl_idx, values, ncomp, orders,
                      offsets, out_offsets, order, n_el, out);
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

int drx_mesh_pack_f64 (const int32_t *ctrl_idx, const double *values, const int32_t *orders, const int64_t *offsets,
                       const int64_t *out_offsets, int32_t order, int32_t n_el, double *out, void *stream)
{
  if (n_el < 0) return fail (DRX_ERR_ARG, "drx_mesh_pack_f64: n_el < 0");
  if (n_el == 0) return DRX_OK;
  if (!ctrl_idx || !values || !out) return fail (DRX_ERR_ARG, "drx_mesh_pack_f64: null pointer");
  if (!orders && (order < 1 || order > DRX_MAX_ORDER))
    return fail (DRX_ERR_ARG, "drx_mesh_pack_f64: order %d outside 1..%d", order, DRX_MAX_ORDER);
  if ((orders == nullptr) != (out_offsets == nullptr) || (orders == nullptr) != (offsets == nullptr))
    return fail (DRX_ERR_ARG, "drx_mesh_pack_f64: mixed-order meshes need orders, offsets and out_offsets");
  hipLaunchKernelGGL (k_mesh_pack_f64, dim3 (n_el), dim3 (64), 0, as_stream (stream), ctrl_idx, values, orders, offsets,
                      out_offsets, order, n_el, out);
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

// ColorTable::correct_opacity + ColorTable::sample (color_table.cpp:152-324), host side as in dray.
int drx_color_table_sample (const float *rgb, int32_t n_rgb, const float *alpha, int32_t n_alpha, int32_t smooth,
                            float factor, int32_t num_samples, float alpha_scale, float *out)
{
  if (num_samples < 2 || !out || (n_rgb > 0 && !rgb) || (n_alpha > 0 && !alpha))
    return fail (DRX_ERR_ARG, "drx_color_table_sample: bad arguments");
  std::vector<float> ap (alpha, alpha + 2 * n_alpha);
  if (factor > 0.f)
    for (int i = 0; i < n_alpha; ++i) ap[2 * i + 1] = 1.f - std::pow ((1.f - ap[2 * i + 1]), factor);
  auto lerp_pick = [smooth] (float c0, float c1, float a) {
    return smooth ? (c0 * (1.0f - a) + c1 * a) : (a < .5 ? c0 : c1);
  };
  for (int i = 0; i < num_samples; ++i)
  {
    const float s = static_cast<float> (i) / static_cast<float> (num_samples - 1);
    float c[4];
    if (n_rgb == 0)
    {
      c[0] = c[1] = c[2] = 0.5f;
      c[3] = 1.f;
    }
    else if (n_rgb == 1 || s <= rgb[0])
      for (int k = 0; k < 4; ++k) c[k] = rgb[1 + k];
    else if (s >= rgb[5 * (n_rgb - 1)])
      for (int k = 0; k < 4; ++k) c[k] = rgb[5 * (n_rgb - 1) + 1 + k];
    else
    {
      int sec = 1;
      for (; sec < n_rgb - 1; ++sec)
        if (s < rgb[5 * sec]) break;
      const int fst = sec - 1;
      const float seg = rgb[5 * sec] - rgb[5 * fst];
      const float a = (seg == 0.f) ? .5f : (s - rgb[5 * fst]) / seg;
      for (int k = 0; k < 4; ++k) c[k] = lerp_pick (rgb[5 * fst + 1 + k], rgb[5 * sec + 1 + k], a);
    }
    float al;
    if (n_alpha == 0)
      al = 1.f;
    else if (n_alpha == 1 || s <= ap[0])
      al = ap[1];
    else if (s >= ap[2 * (n_alpha - 1)])
      al = ap[2 * (n_alpha - 1) + 1];
    else
    {
      int sec = 1;
      for (; sec < n_alpha - 1; ++sec)
        if (s < ap[2 * sec]) break;
      const int fst = sec - 1;
      const float seg = ap[2 * sec] - ap[2 * fst];
      const float a = (seg == 0.f) ? .5f : (s - ap[2 * fst]) / seg;
      al = smooth ? (ap[2 * fst + 1] * (1.f - a) + ap[2 * sec + 1] * a) : (a < .5 ? ap[2 * fst + 1] : ap[2 * sec + 1]);
    }
    c[3] = al * alpha_scale;
    for (int k = 0; k < 4; ++k) out[4 * i + k] = c[k];
  }
  return DRX_OK;
}

// ColorTable (const std::string &name) (color_table.cpp:378-1330): the named table's rgb control points
// (alpha 1) as n x (pos, r, g, b, a).  "" and "default" name cool2warm; named tables are smooth except
// "sharp"; an unknown name prints a notice and takes YlOrRd, as dray does.  With rgb_pts null (or capacity
// below the count) only *n_pts and *smooth are written.
int drx_color_table_preset (const char *name, float *rgb_pts, int32_t capacity, int32_t *n_pts, int32_t *smooth)
{
  if (!name || !n_pts || !smooth) return fail (DRX_ERR_ARG, "drx_color_table_preset: null argument");
  std::string nm (name);
  if (nm.empty () || nm == "default") nm = "cool2warm";
  const PresetTable *t = nullptr, *fallback = nullptr;
  for (const PresetTable &e : kPresetTables)
  {
    if (nm == e.name) t = &e;
    if (std::string (e.name) == "YlOrRd") fallback = &e;
  }
  if (!t)
  {
    std::fprintf (stderr, "Unknown Color Table: '%s' defaulting\n", name);
    t = fallback;
  }
  *n_pts = t->count;
  *smooth = t->smooth;
  if (rgb_pts && capacity >= t->count)
    for (int i = 0; i < t->count; ++i)
    {
      const float *q = kPresetPoints[t->first + i];
      rgb_pts[5 * i + 0] = q[0];
      rgb_pts[5 * i + 1] = q[1];
      rgb_pts[5 * i + 2] = q[2];
      rgb_pts[5 * i + 3] = q[3];
      rgb_pts[5 * i + 4] = 1.f;
    }
  return DRX_OK;
}

} // extern "C"

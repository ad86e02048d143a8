// Host C++ mirror of dray's API for the volume-render / point-location path, over the libdrx C ABI.
//
// Class and method names, argument meaning and error behaviour follow LLNL/devil_ray (src/dray/):
//   Array<T>                 array.hpp:19-42 (host/device mirror, lazy copies, shallow shared copies)
//   LinearBVHBuilder         linear_bvh_builder.hpp:16-22
//   UnstructuredMesh         data_model/unstructured_mesh.hpp:117-135 (locate :84, get_bvh, bounds)
//   UnstructuredField        data_model/unstructured_field.hpp (eval, range)
//   ColorTable               color_table.hpp (add_point, add_alpha, set_smooth, correct_opacity, sample)
//   Camera                   rendering/camera.hpp (set_width/height/fov/pos/look_at/up, create_rays)
//   Volume                   rendering/volume.hpp:25-60 (field, samples, use_lighting, color_map, integrate, save)
// Errors throw dray::DRayError (error.hpp) carrying drx_last_error().  All compute runs in libdrx's
// gfx950 kernels; there is no host fallback.
#pragma once

#include "../drx.h"

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace dray
{

typedef int32_t int32;
typedef float float32;
typedef double float64;
typedef float Float;

class DRayError : public std::runtime_error
{
public:
  explicit DRayError (const std::string &m) : std::runtime_error (m) {}
};

void drx_check (int status, const char *what);

// ---- Array: host/device mirror (dray's Array<T> semantics: dirty flags, shallow copies) ------------
namespace detail
{
struct Buffer
{
  void *host = nullptr;
  void *device = nullptr;
  size_t bytes = 0;
  bool host_dirty = false;     // device copy is newer
  bool device_dirty = false;   // host copy is newer
  ~Buffer ();
  void resize (size_t b);
  void *get_host ();
  void *get_device ();
};
} // namespace detail

template <typename T> class Array
{
public:
  Array () : m_buf (std::make_shared<detail::Buffer> ()) {}
  explicit Array (size_t n) : Array () { resize (n); }
  void resize (size_t n)
  {
    m_size = n;
    m_buf->resize (n * sizeof (T));
  }
  size_t size () const { return m_size; }
  T *get_host_ptr ()
  {
    T *p = (T *)m_buf->get_host ();
    m_buf->device_dirty = true;
    return p;
  }
  const T *get_host_ptr_const () const { return (const T *)m_buf->get_host (); }
  T *get_device_ptr ()
  {
    T *p = (T *)m_buf->get_device ();
    m_buf->host_dirty = true;
    return p;
  }
  const T *get_device_ptr_const () const { return (const T *)m_buf->get_device (); }

private:
  std::shared_ptr<detail::Buffer> m_buf;
  size_t m_size = 0;
};

// ---- small POD types (vec.hpp, range.hpp, aabb.hpp, location.hpp, ray.hpp, volume_partial.hpp) ----
template <typename T, int S> struct Vec
{
  T m_data[S];
  T &operator[] (int i) { return m_data[i]; }
  const T &operator[] (int i) const { return m_data[i]; }
};
typedef Vec<float32, 3> Vec3f;
typedef Vec<float32, 4> Vec4f;

struct Range
{
  float32 m_min, m_max;
  float32 min () const { return m_min; }
  float32 max () const { return m_max; }
};
template <int dim = 3> struct AABB
{
  Range m_ranges[dim];   // [xmin, xmax, ymin, ymax, zmin, zmax]
};

struct Location
{
  int32 m_cell_id;
  Vec<Float, 3> m_ref_pt;
};
struct Ray
{
  Vec<Float, 3> m_dir;
  Float m_near;
  Vec<Float, 3> m_orig;
  Float m_far;
  int32 m_pixel_id;
};
struct VolumePartial
{
  int32 m_pixel_id;
  float32 m_depth;
  Vec<float32, 4> m_color;
};
static_assert (sizeof (Location) == 16 && sizeof (Ray) == 36 && sizeof (VolumePartial) == 24, "dray record layouts");

// ---- BVH + builder (bvh.hpp, linear_bvh_builder.hpp) ------------------------------------------------
struct BVH
{
  Array<Vec<float32, 4>> m_inner_nodes;
  Array<int32> m_leaf_nodes;
  AABB<> m_bounds;
  Array<int32> m_aabb_ids;
  // traversal copies (libdrx): fused leaf records and the 8-wide collapse
  Array<float32> m_leaf_pack;
  Array<float32> m_wide_nodes;
};

class LinearBVHBuilder
{
public:
  BVH construct (Array<AABB<>> aabbs);
  BVH construct (Array<AABB<>> aabbs, Array<int32> prim_ids);
};

// ---- mesh / field --------------------------------------------------------------------------------------
template <int32 NC> struct GridFunction
{
  Array<int32> m_ctrl_idx;          // (p+1)^3 per element, dof i + (p+1)(j + (p+1)k)
  Array<Vec<Float, NC>> m_values;
  int32 m_el_dofs = 0;
  int32 m_size_el = 0;
};

class UnstructuredMesh
{
public:
  UnstructuredMesh (const GridFunction<3> &dof_data, int32 poly_order);
  int32 cells () const { return m_dof_data.m_size_el; }
  int32 order () const { return m_poly_order; }
  BVH &get_bvh ();                                   // built on first use (unstructured_mesh.cpp:23-31)
  AABB<> bounds ();                                  // = get_bvh().m_bounds
  Array<Location> locate (Array<Vec<Float, 3>> &wpoints);
  Array<AABB<>> &ref_aabbs () { return m_ref_aabbs; }
  // libdrx views
  drx_mesh device_view ();
  drx_bvh bvh_view ();

private:
  GridFunction<3> m_dof_data;
  int32 m_poly_order;
  bool m_is_constructed = false;
  BVH m_bvh;
  Array<AABB<>> m_ref_aabbs;   // SubRef boxes (lo xyz, hi xyz), 6 floats like AABB
  Array<float32> m_packed;
};

class UnstructuredField
{
public:
  UnstructuredField (const GridFunction<1> &dof_data, int32 poly_order, const std::string &name);
  const std::string &name () const { return m_name; }
  Range range () const;                              // min/max of the control values
  void eval (const Array<Location> locs, Array<Float> &values);
  drx_mesh device_view ();

private:
  GridFunction<1> m_dof_data;
  int32 m_poly_order;
  std::string m_name;
  Array<float32> m_packed;
};

// ---- colour table / map ---------------------------------------------------------------------------------
class ColorTable
{
public:
  ColorTable () {}
  void add_point (float32 position, const Vec3f &color);
  void add_point (float32 position, const Vec4f &color);
  void add_alpha (float32 position, float32 alpha);
  void set_smooth (bool smooth) { m_smooth = smooth; }
  bool get_smooth () const { return m_smooth; }
  ColorTable correct_opacity (const float32 &factor) const;
  void sample (int num_samples, Array<Vec<float32, 4>> &colors, float32 alpha_scale = 1.f) const;

private:
  std::vector<float32> m_rgb;     // (pos, r, g, b, a)
  std::vector<float32> m_alpha;   // (pos, alpha)
  bool m_smooth = false;          // dray's default (color_table.cpp:67-71)
};

// ---- lights (rendering/point_light.hpp:15-33) ---------------------------------------------------------------
typedef drx_point_light PointLight;   // {pos, amb, diff, spec, spec_pow}
PointLight make_point_light (const Vec3f &pos);   // dray's member defaults (amb 0.4, diff 0.75, spec 0.8, pow 15)

// ---- camera ----------------------------------------------------------------------------------------------
class Camera
{
public:
  void set_width (const int32 &w);
  void set_height (const int32 &h);
  void set_fov (const float32 &degrees);
  void set_pos (const Vec3f &p) { m_pos = p; }
  void set_look_at (const Vec3f &p) { m_look_at = p; }
  void set_up (const Vec3f &u) { m_up = u; }
  void set_zoom (float32 z) { m_zoom = z; }
  int32 get_width () const { return m_width; }
  int32 get_height () const { return m_height; }
  Vec3f get_pos () const { return m_pos; }
  Vec3f get_look_at () const { return m_look_at; }
  Vec3f get_up () const { return m_up; }
  void create_rays (Array<Ray> &rays);
  drx_camera view () const;

private:
  int32 m_width = 500, m_height = 500;
  float32 m_fov_y = 30.f, m_zoom = 1.f;
  Vec3f m_pos = { { 0.f, 0.f, 0.f } }, m_look_at = { { 0.f, 0.f, -1.f } }, m_up = { { 0.f, 1.f, 0.f } };
};

// ---- volume ----------------------------------------------------------------------------------------------
class Volume
{
public:
  Volume (UnstructuredMesh &mesh, UnstructuredField &field);
  void samples (int32 num_samples) { m_samples = num_samples; }
  void use_lighting (bool do_it) { m_use_lighting = do_it; }
  ColorTable &color_table () { return m_table; }
  void field_range (float32 lo, float32 hi) { m_range = { lo, hi }; m_range_set = true; }
  // libdrx extra: rays came from a width x height camera; process them in 8x8 pixel tiles (same results)
  void tile_order (int32 width, int32 height) { m_tile_w = width; m_tile_h = height; }
  // Volume::integrate (volume.cpp:323-351): returns compacted partials (<= 5 per active ray).  With
  // use_lighting (dray's default) every light shades each sample (VolumeShader::shaded_color); without
  // lights a lit sample contributes nothing, as in dray.
  Array<VolumePartial> integrate (Array<Ray> &rays, Array<PointLight> &lights);
  Array<VolumePartial> integrate (Array<Ray> &rays);   // no lights
  // Volume::save's compositing + white background (volume.cpp:369-416), RGBA per pixel (no PNG)
  Array<Vec<float32, 4>> composite (Array<VolumePartial> &partials, int32 width, int32 height);
  drx_integrate_stats last_stats () const { return m_stats; }

private:
  UnstructuredMesh &m_mesh;
  UnstructuredField &m_field;
  int32 m_samples = 100;          // volume.cpp:275
  bool m_use_lighting = true;     // volume.cpp:277
  ColorTable m_table;
  Range m_range = { 0.f, 0.f };
  bool m_range_set = false;
  int32 m_tile_w = 0, m_tile_h = 0;
  drx_integrate_stats m_stats {};
};

// detail::default_light (rendering/renderer.cpp:141-159): the light the Renderer adds when none is given.
PointLight default_light (Camera &camera);

} // namespace dray

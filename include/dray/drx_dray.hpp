// Host C++ mirror of dray's API for the volume-render / point-location path, over the libdrx C ABI.
//
// Class and method names, argument meaning and error behaviour follow LLNL/devil_ray (src/dray/):
//   Array<T>                 array.hpp:19-42 (host/device mirror, lazy copies, shallow shared copies)
//   LinearBVHBuilder         linear_bvh_builder.hpp:16-22
//   Mesh, Field              data_model/mesh.hpp:18-36, data_model/field.hpp:19-54 (abstract bases)
//   UnstructuredMesh<E>      data_model/unstructured_mesh.hpp:84-135 (locate, get_bvh, bounds; HexMesh[_P1|_P2])
//   UnstructuredField<E>     data_model/unstructured_field.hpp:30-93 (eval, range; HexScalar[_P1|_P2])
//   dispatch_3d              dispatcher.hpp:208-220 (Mesh* / Field* -> derived types)
//   ColorTable               color_table.hpp (add_point, add_alpha, set_smooth, correct_opacity, sample)
//   Camera                   rendering/camera.hpp (set_width/height/fov/pos/look_at/up, create_rays,
//                            reset_to_bounds, elevate, azimuth)
//   DataSet, Collection      data_model/data_set.hpp:20-59, data_model/collection.hpp:16-55
//   ColorMap                 color_map.hpp (color_table, scalar_range)
//   Volume                   rendering/volume.hpp:25-60 (Volume(Collection&), field, samples, use_lighting,
//                            color_map, active_domain, integrate)
//   PointLocation            queries/point_location.hpp:14-33 (add_var, empty_val, execute)
//   Framebuffer              rendering/framebuffer.hpp:19-52 (clear, composite_background, save, colors)
//   Renderer                 rendering/renderer.hpp:22-53 (volume, add_light, clear_lights, render)
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
typedef uint32_t uint32;
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
  float32 m_min = 3.402823466e+38f, m_max = -3.402823466e+38f;   // empty (range.hpp)
  float32 min () const { return m_min; }
  float32 max () const { return m_max; }
  bool is_empty () const { return m_min > m_max; }
  float32 length () const { return m_max - m_min; }
  float32 center () const { return 0.5f * (m_min + m_max); }
  void include (float32 v)
  {
    m_min = m_min < v ? m_min : v;
    m_max = m_max > v ? m_max : v;
  }
  void include (const Range &r)
  {
    if (r.is_empty ()) return;
    include (r.m_min);
    include (r.m_max);
  }
};
template <int dim = 3> struct AABB
{
  Range m_ranges[dim];   // [xmin, xmax, ymin, ymax, zmin, zmax]
  void include (const AABB &o)
  {
    for (int d = 0; d < dim; ++d) m_ranges[d].include (o.m_ranges[d]);
  }
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
  Array<float32> m_leaf_starts;   // Newton start map per leaf (drx_bvh_leaf_starts)
  Array<uint32> m_qnodes;         // compact traversal nodes (drx_bvh_quantize)
  Array<float32> m_leaf_boxes;    // exact leaf boxes for the compact walk
  float32 m_qgrid[6] = { 0, 0, 0, 0, 0, 0 };
  Array<float32> m_wide_nodes;
};

class LinearBVHBuilder
{
public:
  BVH construct (Array<AABB<>> aabbs);
  BVH construct (Array<AABB<>> aabbs, Array<int32> prim_ids);
};

// ---- elements (data_model/element.hpp, elem_attr.hpp): the type tags meshes and fields are templated on ----
enum class ElemType
{
  Tensor,
  Simplex
};
namespace Order
{
enum : int32
{
  General = -1,   // any order, given at run time (dray's Order::General)
  Constant = 0,
  Linear = 1,
  Quadratic = 2
};
}
// Element<topological dims, components, shape, order>: dray's element type, here a tag only (the GPU
// kernels evaluate every order 1..DRX_MAX_ORDER with one code path per order).
template <int32 dim, int32 ncomp, ElemType etype, int32 P> struct Element
{
  static constexpr int32 get_dim () { return dim; }
  static constexpr int32 get_ncomp () { return ncomp; }
  static constexpr ElemType get_etype () { return etype; }
  static constexpr int32 get_P () { return P; }   // Order::General (-1) or the fixed order
};
using Hex3 = Element<3, 3, ElemType::Tensor, Order::General>;          // unstructured_mesh.hpp:117-119
using Hex_P1 = Element<3, 3, ElemType::Tensor, Order::Linear>;
using Hex_P2 = Element<3, 3, ElemType::Tensor, Order::Quadratic>;
using HexScalar = Element<3, 1, ElemType::Tensor, Order::General>;     // unstructured_field.hpp:90-93
using HexScalar_P1 = Element<3, 1, ElemType::Tensor, Order::Linear>;
using HexScalar_P2 = Element<3, 1, ElemType::Tensor, Order::Quadratic>;
template <typename ElemT> std::string element_name ();   // "Hex_General", "Hex_P1", ... (dispatcher.hpp)

// ---- grid function (data_model/grid_function.hpp) ------------------------------------------------------------
template <int32 NC> struct GridFunction
{
  Array<int32> m_ctrl_idx;          // (p+1)^3 per element, dof i + (p+1)(j + (p+1)k)
  Array<Vec<Float, NC>> m_values;
  int32 m_el_dofs = 0;
  int32 m_size_el = 0;
};

// ---- Mesh and Field: dray's abstract bases (data_model/mesh.hpp:18-36, data_model/field.hpp:19-54) -----------
// The Conduit export (to_node) is left out: this build has no Conduit (I/O is outside the path).
class Mesh
{
protected:
  std::string m_name;

public:
  virtual ~Mesh () {}
  std::string name () const { return m_name; }
  void name (const std::string &name) { m_name = name; }
  virtual std::string type_name () const = 0;
  virtual int32 cells () const = 0;
  virtual int32 order () const = 0;
  virtual int32 dims () const = 0;
  virtual AABB<3> bounds () = 0;
  virtual Array<Location> locate (Array<Vec<Float, 3>> &wpoints) = 0;
};

class Field
{
protected:
  std::string m_name;
  std::string m_mesh;   // each field is associated with one mesh

public:
  virtual ~Field () {}
  std::string name () const { return m_name; }
  void name (const std::string &name) { m_name = name; }
  void mesh_name (const std::string &name) { m_mesh = name; }
  std::string mesh_name () const { return m_mesh; }
  virtual std::vector<Range> range () const = 0;
  virtual int32 order () const = 0;
  virtual int32 components () const = 0;
  virtual std::string type_name () const = 0;
  virtual void eval (const Array<Location> locs, Array<Float> &values) = 0;
};

// ---- UnstructuredMesh<Element> (data_model/unstructured_mesh.hpp:84-135) ------------------------------------
// Instantiated for Hex3 (any order 1..DRX_MAX_ORDER), Hex_P1 and Hex_P2 (the order must match).
template <class ElemT> class UnstructuredMesh : public Mesh
{
public:
  static constexpr int32 dim = ElemT::get_dim ();
  using ElementType = ElemT;

  UnstructuredMesh () = delete;
  UnstructuredMesh (const GridFunction<3> &dof_data, int32 poly_order);
  std::string type_name () const override;
  int32 cells () const override { return m_dof_data.m_size_el; }
  int32 order () const override { return m_poly_order; }
  int32 dims () const override { return dim; }
  AABB<3> bounds () override;                        // = get_bvh().m_bounds
  Array<Location> locate (Array<Vec<Float, 3>> &wpoints) override;
  BVH &get_bvh ();                                   // built on first use (unstructured_mesh.cpp:23-31)
  GridFunction<3> get_dof_data () { return m_dof_data; }
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
using HexMesh = UnstructuredMesh<Hex3>;
using HexMesh_P1 = UnstructuredMesh<Hex_P1>;
using HexMesh_P2 = UnstructuredMesh<Hex_P2>;

// ---- UnstructuredField<Element> (data_model/unstructured_field.hpp:30-88) -----------------------------------
template <class ElemT> class UnstructuredField : public Field
{
public:
  using ElementType = ElemT;

  UnstructuredField () = delete;
  UnstructuredField (const GridFunction<ElemT::get_ncomp ()> &dof_data, int32 poly_order, const std::string &name = "");
  std::vector<Range> range () const override;        // min/max of the control values per component
  // Fix the range to report (not in dray): a field whose coefficients were degree-elevated on import keeps
  // the range of the coefficients it was read with, as dray's range of the unelevated field.
  void range_override (const std::vector<Range> &r) { m_range_override = r; }
  int32 order () const override { return m_poly_order; }
  int32 components () const override { return ElemT::get_ncomp (); }
  std::string type_name () const override;
  void eval (const Array<Location> locs, Array<Float> &values) override;
  GridFunction<ElemT::get_ncomp ()> get_dof_data () { return m_dof_data; }
  drx_mesh device_view ();

private:
  GridFunction<ElemT::get_ncomp ()> m_dof_data;
  int32 m_poly_order;
  Array<float32> m_packed;
  std::vector<Range> m_range_override;
};

// ---- dispatch_3d (dispatcher.hpp:208-220): Mesh* / Field* -> the derived types, then func(mesh, field) -------
namespace detail
{
[[noreturn]] void cast_mesh_failed (Mesh *mesh, const char *file, unsigned long long line);
[[noreturn]] void cast_field_failed (Field *field, const char *file, unsigned long long line);

template <typename DerivedMeshT, typename Functor> void dispatch_scalar_field (Field *field, DerivedMeshT *mesh, Functor &func)
{
  using MElemT = typename DerivedMeshT::ElementType;
  using S = Element<MElemT::get_dim (), 1, MElemT::get_etype (), Order::General>;
  using S1 = Element<MElemT::get_dim (), 1, MElemT::get_etype (), Order::Linear>;
  using S2 = Element<MElemT::get_dim (), 1, MElemT::get_etype (), Order::Quadratic>;
  if (auto *f = dynamic_cast<UnstructuredField<S> *> (field))
    func (*mesh, *f);
  else if (auto *f1 = dynamic_cast<UnstructuredField<S1> *> (field))
    func (*mesh, *f1);
  else if (auto *f2 = dynamic_cast<UnstructuredField<S2> *> (field))
    func (*mesh, *f2);
  else
    cast_field_failed (field, __FILE__, __LINE__);
}

template <typename MeshGuessT, typename Functor> bool dispatch_mesh_field (MeshGuessT *, Mesh *mesh, Field *field, Functor &func)
{
  MeshGuessT *derived = dynamic_cast<MeshGuessT *> (mesh);
  if (derived) dispatch_scalar_field (field, derived, func);
  return derived != nullptr;
}

template <typename MeshGuessT, typename Functor> bool dispatch_mesh_only (MeshGuessT *, Mesh *mesh, Functor &func)
{
  MeshGuessT *derived = dynamic_cast<MeshGuessT *> (mesh);
  if (derived) func (*derived);
  return derived != nullptr;
}
} // namespace detail

// Hexahedral meshes of any order (HexMesh, HexMesh_P1, HexMesh_P2) with a scalar field; dray's tet meshes
// are outside this path (its GPU kernels are hex-only), so a tet mesh reaches cast_mesh_failed.
template <typename Functor> bool dispatch_3d (Mesh *mesh, Field *field, Functor &func)
{
  if (!detail::dispatch_mesh_field ((HexMesh *)0, mesh, field, func) &&
      !detail::dispatch_mesh_field ((HexMesh_P1 *)0, mesh, field, func) &&
      !detail::dispatch_mesh_field ((HexMesh_P2 *)0, mesh, field, func))
    detail::cast_mesh_failed (mesh, __FILE__, __LINE__);
  return true;
}

template <typename Functor> bool dispatch_3d (Mesh *mesh, Functor &func)
{
  if (!detail::dispatch_mesh_only ((HexMesh *)0, mesh, func) && !detail::dispatch_mesh_only ((HexMesh_P1 *)0, mesh, func) &&
      !detail::dispatch_mesh_only ((HexMesh_P2 *)0, mesh, func))
    detail::cast_mesh_failed (mesh, __FILE__, __LINE__);
  return true;
}

// ---- colour table / map ---------------------------------------------------------------------------------
class ColorTable
{
public:
  ColorTable () {}
  // ColorTable (name) (color_table.cpp:378-1330): the named table's rgb points, smooth except "sharp";
  // "" and "default" name cool2warm; an unknown name prints a notice and takes YlOrRd.
  explicit ColorTable (const std::string &name);
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
  float32 get_fov () const { return m_fov_y; }
  // Camera::reset_to_bounds (camera.cpp:485-514): look at the padded bounds' centre from one diagonal
  // length away along the current viewing direction; field of view 60 degrees.
  void reset_to_bounds (const AABB<3> bounds, const float64 xpad = 0., const float64 ypad = 0.,
                        const float64 zpad = 0.);
  // Camera::elevate / azimuth (camera.cpp:516-552): orbit the look-at point about the camera's right /
  // up axis by the given degrees (the up vector is kept).
  void elevate (const float32 degrees);
  void azimuth (const float32 degrees);
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

// ---- data sets and collections (data_model/data_set.hpp, data_model/collection.hpp) ----------------------
class DataSet
{
public:
  DataSet () {}
  explicit DataSet (std::shared_ptr<Mesh> mesh) { add_mesh (std::move (mesh)); }
  void domain_id (const int32 id) { m_domain_id = id; }
  int32 domain_id () const { return m_domain_id; }
  void clear_meshes () { m_meshes.clear (); }
  int32 number_of_meshes () const { return (int32)m_meshes.size (); }
  void add_mesh (std::shared_ptr<Mesh> mesh) { m_meshes.push_back (std::move (mesh)); }
  bool has_mesh (const std::string &name) const;
  Mesh *mesh (const int32 index = 0);                   // throws DRayError when out of range
  Mesh *mesh (const std::string name);                  // throws DRayError when missing
  std::vector<std::string> meshes () const;
  int32 number_of_fields () const { return (int32)m_fields.size (); }
  void clear_fields () { m_fields.clear (); }
  bool has_field (const std::string &name) const;
  std::vector<std::string> fields () const;
  Field *field (const std::string &name);               // throws DRayError when missing
  Field *field (const int &index);
  std::shared_ptr<Field> field_shared (const int &index);
  std::shared_ptr<Field> field_shared (const std::string &name);
  void add_field (std::shared_ptr<Field> field) { m_fields.push_back (std::move (field)); }
  std::string field_info ();

private:
  std::vector<std::shared_ptr<Mesh>> m_meshes;
  std::vector<std::shared_ptr<Field>> m_fields;
  int32 m_domain_id = 0;
};

class Collection
{
public:
  void add_domain (const DataSet &domain) { m_domains.push_back (domain); }
  DataSet domain (int32 index);
  int32 size () const { return (int32)m_domains.size (); }         // one rank: all domains are local
  int32 local_size () const { return (int32)m_domains.size (); }
  bool has_field (const std::string &name);                         // every domain has it
  Range range (const std::string &name);                            // union of the domains' field ranges
  AABB<3> bounds ();                                                // union of the domains' mesh bounds
  std::string field_list ();

private:
  std::vector<DataSet> m_domains;
};

// ---- Conduit-Blueprint reader (io/blueprint_reader.cpp:280-440) ---------------------------------------------
// HDF5 files as Conduit's relay writes them (superblock 2/3, compact groups, contiguous or compact datasets);
// one unstructured hex topology per domain with an MFEM H1 node field of order <= 2 and scalar H1 (order <= 2)
// or L2 fields.  Data are converted to Bernstein form on the host (MFEM dof numbering, nodal-to-Bernstein
// interpolation) and each domain's mesh and fields are degree-elevated to one order, the GPU kernels' form.
// Unsupported content throws DRayError.
class BlueprintReader
{
public:
  static Collection load (const std::string &root_file);
  static DataSet load_domain (const std::string &domain_file);
};

// ---- colour map (color_map.hpp) ----------------------------------------------------------------------------
class ColorMap
{
public:
  ColorTable &color_table () { return m_table; }
  void color_table (const ColorTable &t) { m_table = t; }
  void scalar_range (const Range &r)
  {
    m_range = r;
    m_range_set = true;
  }
  Range scalar_range () const { return m_range; }
  bool range_set () const { return m_range_set; }

private:
  ColorTable m_table { "cool2warm" };   // ColorMap () (color_map.cpp:11-18)
  Range m_range;
  bool m_range_set = false;
};

// ---- volume ----------------------------------------------------------------------------------------------
class Volume
{
public:
  explicit Volume (Collection &collection);   // adds dray's default alpha ramp (0.1, 0) - (1.0, 0.7)
  void input (Collection &collection);
  void field (const std::string &name);       // also takes the collection's range of the field
  std::string field () const { return m_field; }
  void samples (int32 num_samples) { m_samples = num_samples; }
  void use_lighting (bool do_it) { m_use_lighting = do_it; }
  ColorMap &color_map () { return m_color_map; }
  void active_domain (int32 index) { m_active_domain = index; }
  int32 active_domain () const { return m_active_domain; }
  int32 num_domains () { return m_collection.size (); }
  // libdrx extra: rays came from a width x height camera; process them in 8x8 pixel tiles (same results)
  void tile_order (int32 width, int32 height)
  {
    m_tile_w = width;
    m_tile_h = height;
  }
  // Volume::integrate (volume.cpp:323-351) on the active domain: compacted partials (<= 5 per active
  // ray).  With use_lighting (dray's default) every light shades each sample
  // (VolumeShader::shaded_color); without lights a lit sample contributes nothing, as in dray.
  Array<VolumePartial> integrate (Array<Ray> &rays, Array<PointLight> &lights);
  // Volume::save's compositing + white background (volume.cpp:369-416), RGBA per pixel (no PNG)
  Array<Vec<float32, 4>> composite (Array<VolumePartial> &partials, int32 width, int32 height);
  // Volume::save (volume.cpp:368-413): the compositing above written as name + ".png"
  void save (const std::string name, Array<VolumePartial> partials, const int32 width, const int32 height);
  drx_integrate_stats last_stats () const { return m_stats; }

private:
  int32 m_samples = 100;          // volume.cpp:275
  ColorMap m_color_map;
  Collection m_collection;        // shallow copy (its domains share meshes and fields)
  std::string m_field;
  AABB<3> m_bounds;
  bool m_use_lighting = true;     // volume.cpp:277
  int32 m_active_domain = 0;
  Range m_field_range;
  int32 m_tile_w = 0, m_tile_h = 0;
  drx_integrate_stats m_stats {};
};

// ---- point location query (queries/point_location.hpp) -----------------------------------------------------
class PointLocation
{
public:
  struct Result
  {
    Array<Vec<Float, 3>> m_points;
    std::vector<Array<Float>> m_values;
    std::vector<std::string> m_vars;
    Float m_empty_val;
  };
  void empty_val (const Float val) { m_empty_val = val; }
  void add_var (const std::string &var) { m_vars.push_back (var); }
  // every domain locates every point; a located point takes that domain's field values (later domains
  // overwrite earlier ones), others keep empty_val (point_location.cpp:176-287, one rank)
  Result execute (Collection &collection, Array<Vec<Float, 3>> &points);

private:
  Float m_empty_val = 0;
  std::vector<std::string> m_vars;
};

// detail::default_light (rendering/renderer.cpp:141-159): the light the Renderer adds when none is given.
PointLight default_light (Camera &camera);

// ---- framebuffer (rendering/framebuffer.hpp) ------------------------------------------------------------
// RGBA colours and depths per pixel (pixel id = x + y * width, row 0 at the bottom of the picture).
class Framebuffer
{
public:
  Framebuffer ();                                     // 1024 x 1024
  Framebuffer (const int32 width, const int32 height);
  int32 width () const { return m_width; }
  int32 height () const { return m_height; }
  void clear ();                                      // transparent black, depths +infinity
  void clear (const Vec<float32, 4> &color);
  void composite_background ();                       // pixels with alpha < 1 over the background colour
  void save (const std::string name);                 // name + ".png", 8-bit RGBA
  void background_color (const Vec<float32, 4> &color) { m_bg_color = color; }
  void foreground_color (const Vec<float32, 4> &color) { m_fg_color = color; }
  Vec<float32, 4> background_color () const { return m_bg_color; }
  Vec<float32, 4> foreground_color () const { return m_fg_color; }
  Array<Vec<float32, 4>> colors () const { return m_colors; }   // shallow, like dray's
  Array<float32> depths () const { return m_depths; }

private:
  int32 m_width, m_height;
  Vec<float32, 4> m_bg_color = { { 1.f, 1.f, 1.f, 1.f } };
  Vec<float32, 4> m_fg_color = { { 0.f, 0.f, 0.f, 1.f } };
  Array<Vec<float32, 4>> m_colors;
  Array<float32> m_depths;
};

// ---- renderer (rendering/renderer.hpp), volume branch ------------------------------------------------------
// Renderer::render (renderer.cpp:229-337) with a volume and no surface traceables: camera rays, a cleared
// framebuffer, the added lights (or default_light(camera)), Volume::integrate on every domain, then the
// partials of all domains composited by (pixel, depth) into the framebuffer.  Annotations (colour bars,
// triad, world annotations) are not drawn.
class Renderer
{
public:
  void volume (std::shared_ptr<Volume> volume) { m_volume = std::move (volume); }
  void add_light (const PointLight &light) { m_lights.push_back (light); }
  void clear_lights () { m_lights.clear (); }
  void use_lighting (bool use_it) { m_use_lighting = use_it; }   // surfaces only in dray; kept for the API
  // dray draws a colour bar (and, optionally, world annotations) over the composited image
  // (renderer.cpp:178-186,284-357); this build keeps the switches and draws neither.
  void color_bar (bool on) { m_color_bar = on; }
  bool color_bar () const { return m_color_bar; }
  void world_annotations (bool on) { m_world_annotations = on; }
  bool world_annotations () const { return m_world_annotations; }
  Framebuffer render (Camera &camera);

private:
  std::shared_ptr<Volume> m_volume;
  std::vector<PointLight> m_lights;
  bool m_use_lighting = true;
  bool m_color_bar = true, m_world_annotations = false;   // renderer.cpp:166-167
};

// Renderer::render's cross-domain volume compositing (rendering/renderer.cpp:307-330): the partials of
// every domain, sorted by pixel and depth, blended front to back over a white background -> RGBA.
Array<Vec<float32, 4>> composite_domains (std::vector<Array<VolumePartial>> &domain_partials, int32 width,
                                          int32 height);

} // namespace dray

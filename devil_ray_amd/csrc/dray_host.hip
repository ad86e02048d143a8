// Host C++ layer with dray's names over the libdrx C ABI (declared in include/dray/drx_dray.hpp).
#include "../../include/dray/drx_dray.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

namespace dray
{

void drx_check (int status, const char *what)
{
  if (status != DRX_OK) throw DRayError (std::string (what) + ": " + drx_last_error ());
}

static void hip_check (hipError_t e, const char *what)
{
  if (e != hipSuccess) throw DRayError (std::string (what) + ": " + hipGetErrorString (e));
}

namespace detail
{
Buffer::~Buffer ()
{
  if (device) (void)hipFree (device);
  if (host) std::free (host);
}

void Buffer::resize (size_t b)
{
  if (b == bytes) return;
  if (device) (void)hipFree (device);
  if (host) std::free (host);
  device = nullptr;
  host = nullptr;
  bytes = b;
  host_dirty = device_dirty = false;
}

void *Buffer::get_host ()
{
  if (bytes == 0) return nullptr;
  if (!host)
  {
    host = std::calloc (1, bytes);
    if (!host) throw DRayError ("Array: host allocation failed");
  }
  if (host_dirty && device)
    hip_check (hipMemcpy (host, device, bytes, hipMemcpyDeviceToHost), "Array: device->host copy");
  host_dirty = false;
  return host;
}

void *Buffer::get_device ()
{
  if (bytes == 0) return nullptr;
  if (!device)
  {
    hip_check (hipMalloc (&device, bytes), "Array: hipMalloc");
    if (!host) hip_check (hipMemset (device, 0, bytes), "Array: hipMemset");
    else device_dirty = true;
  }
  if (device_dirty && host)
    hip_check (hipMemcpy (device, host, bytes, hipMemcpyHostToDevice), "Array: host->device copy");
  device_dirty = false;
  return device;
}
} // namespace detail

// ---- LinearBVHBuilder ------------------------------------------------------------------------------------
BVH LinearBVHBuilder::construct (Array<AABB<>> aabbs)
{
  Array<int32> ids (aabbs.size ());
  int32 *p = ids.get_host_ptr ();
  for (size_t i = 0; i < aabbs.size (); ++i) p[i] = (int32)i;
  return construct (aabbs, ids);
}

BVH LinearBVHBuilder::construct (Array<AABB<>> aabbs, Array<int32> prim_ids)
{
  const int32 n = (int32)aabbs.size ();
  const int32 m = n < 2 ? 2 : n;
  BVH bvh;
  bvh.m_inner_nodes.resize (4 * (size_t)(m - 1));
  bvh.m_leaf_nodes.resize (m);
  bvh.m_aabb_ids.resize (m);
  float bounds[6];
  drx_check (drx_bvh_build (n ? (const float *)aabbs.get_device_ptr_const () : nullptr,
                            n ? prim_ids.get_device_ptr_const () : nullptr, n,
                            (float *)bvh.m_inner_nodes.get_device_ptr (), bvh.m_leaf_nodes.get_device_ptr (),
                            bvh.m_aabb_ids.get_device_ptr (), bounds, nullptr, nullptr),
             "LinearBVHBuilder::construct");
  for (int d = 0; d < 3; ++d) bvh.m_bounds.m_ranges[d] = { bounds[2 * d], bounds[2 * d + 1] };
  return bvh;
}

// ---- element names, dispatch failures (elem_attr.hpp:394-427, dispatcher.cpp) --------------------------
template <typename ElemT> std::string element_name ()
{
  std::string s = ElemT::get_dim () == 3 ? "3D_" : ElemT::get_dim () == 2 ? "2D_" : "";
  s += ElemT::get_etype () == ElemType::Tensor ? "Tensor" : "Simplex";
  return s + "_C" + std::to_string (ElemT::get_ncomp ()) + "_P" + std::to_string (ElemT::get_P ());
}
template std::string element_name<Hex3> ();
template std::string element_name<Hex_P1> ();
template std::string element_name<Hex_P2> ();
template std::string element_name<HexScalar> ();
template std::string element_name<HexScalar_P1> ();
template std::string element_name<HexScalar_P2> ();

namespace detail
{
void cast_mesh_failed (Mesh *mesh, const char *file, unsigned long long line)
{
  throw DRayError (std::string ("Cast of mesh '") + (mesh ? mesh->type_name () : std::string ("null")) +
                   "' failed (" + file + ":" + std::to_string (line) + ")");
}
void cast_field_failed (Field *field, const char *file, unsigned long long line)
{
  throw DRayError (std::string ("Cast of field '") + (field ? field->type_name () : std::string ("null")) +
                   "' failed (" + file + ":" + std::to_string (line) + ")");
}
} // namespace detail

static void check_fixed_order (int32 P, int32 poly_order, const char *what)
{
  if (P >= 1 && poly_order != P)
    throw DRayError (std::string (what) + ": order " + std::to_string (poly_order) + " given to an order-" +
                     std::to_string (P) + " element type");
}

// ---- UnstructuredMesh<ElemT> ------------------------------------------------------------------------------
template <class ElemT>
UnstructuredMesh<ElemT>::UnstructuredMesh (const GridFunction<3> &dof_data, int32 poly_order)
: m_dof_data (dof_data), m_poly_order (poly_order)
{
  check_fixed_order (ElemT::get_P (), poly_order, "UnstructuredMesh");
  if (poly_order < 1 || poly_order > DRX_MAX_ORDER)
    throw DRayError ("UnstructuredMesh: order " + std::to_string (poly_order) + " not supported on the GPU path");
  const int32 npe = (poly_order + 1) * (poly_order + 1) * (poly_order + 1);
  if (m_dof_data.m_el_dofs == 0) m_dof_data.m_el_dofs = npe;
  if (m_dof_data.m_el_dofs != npe) throw DRayError ("UnstructuredMesh: element dof count does not match the order");
  if (m_dof_data.m_size_el == 0) m_dof_data.m_size_el = (int32)(m_dof_data.m_ctrl_idx.size () / npe);
}

template <class ElemT> std::string UnstructuredMesh<ElemT>::type_name () const { return element_name<ElemT> (); }

template <class ElemT> drx_mesh UnstructuredMesh<ElemT>::device_view ()
{
  const int32 stride = drx_mesh_stride (m_poly_order, 3);
  if (m_packed.size () == 0)
  {
    m_packed.resize ((size_t)stride * cells ());
    drx_check (drx_mesh_pack (m_dof_data.m_ctrl_idx.get_device_ptr_const (),
                              (const float *)m_dof_data.m_values.get_device_ptr_const (), 3, nullptr, nullptr,
                              nullptr, m_poly_order, cells (), m_packed.get_device_ptr (), nullptr),
               "UnstructuredMesh: drx_mesh_pack");
  }
  drx_mesh v {};
  v.data = m_packed.get_device_ptr_const ();
  v.elem_offset = nullptr;
  v.elem_order = nullptr;
  v.n_el = cells ();
  v.order = m_poly_order;
  v.ncomp = 3;
  v.stride = stride;
  return v;
}

template <class ElemT> BVH &UnstructuredMesh<ElemT>::get_bvh ()
{
  if (!m_is_constructed)
  {
    // construct_bvh (data_model/mesh_utils.cpp:514-677)
    const int32 n = cells () * DRX_SPLITS_PER_ELEMENT;
    Array<AABB<>> aabbs (n);
    Array<int32> prims (n);
    m_ref_aabbs.resize (n);
    drx_check (drx_bvh_split_elements (m_dof_data.m_ctrl_idx.get_device_ptr_const (),
                                       (const float *)m_dof_data.m_values.get_device_ptr_const (), nullptr, nullptr,
                                       m_poly_order, cells (), (float *)aabbs.get_device_ptr (),
                                       prims.get_device_ptr (), (float *)m_ref_aabbs.get_device_ptr (), nullptr),
               "construct_bvh: drx_bvh_split_elements");
    m_bvh = LinearBVHBuilder ().construct (aabbs, prims);
    drx_bvh v = bvh_view ();
    m_bvh.m_leaf_pack.resize (8 * (size_t)m_bvh.m_leaf_nodes.size ());
    drx_check (drx_bvh_pack_leaves (&v, m_bvh.m_leaf_pack.get_device_ptr (), nullptr), "drx_bvh_pack_leaves");
    m_bvh.m_leaf_starts.resize (12 * (size_t)m_bvh.m_leaf_nodes.size ());
    const drx_mesh mv = device_view ();
    drx_check (drx_bvh_leaf_starts (&mv, &v, m_bvh.m_leaf_starts.get_device_ptr (), nullptr), "drx_bvh_leaf_starts");
    m_bvh.m_qnodes.resize (8 * (size_t)(m_bvh.m_leaf_nodes.size () - 1));
    m_bvh.m_leaf_boxes.resize (8 * (size_t)m_bvh.m_leaf_nodes.size ());
    drx_check (drx_bvh_quantize (&v, m_bvh.m_qnodes.get_device_ptr (), m_bvh.m_leaf_boxes.get_device_ptr (),
                                 m_bvh.m_qgrid, nullptr),
               "drx_bvh_quantize");
    // the 8-wide copy, read only by the wave walk that the automatic choice takes for small BVHs
    if (m_bvh.m_leaf_nodes.size () <= (size_t (1) << 16))
    {
      int64_t cnt = 0;
      v = bvh_view ();
      drx_check (drx_bvh_widen (&v, nullptr, 0, &cnt, nullptr), "drx_bvh_widen");
      m_bvh.m_wide_nodes.resize (64 * (size_t)cnt);
      drx_check (drx_bvh_widen (&v, m_bvh.m_wide_nodes.get_device_ptr (), cnt, &cnt, nullptr), "drx_bvh_widen");
    }
    m_is_constructed = true;
  }
  return m_bvh;
}

template <class ElemT> drx_bvh UnstructuredMesh<ElemT>::bvh_view ()
{
  drx_bvh v {};
  v.inner_nodes = (const float *)m_bvh.m_inner_nodes.get_device_ptr_const ();
  v.leaf_nodes = m_bvh.m_leaf_nodes.get_device_ptr_const ();
  v.aabb_ids = m_bvh.m_aabb_ids.get_device_ptr_const ();
  v.subrefs = (const float *)m_ref_aabbs.get_device_ptr_const ();
  v.leaf_pack = m_bvh.m_leaf_pack.size () ? m_bvh.m_leaf_pack.get_device_ptr_const () : nullptr;
  v.wide_nodes = m_bvh.m_wide_nodes.size () ? m_bvh.m_wide_nodes.get_device_ptr_const () : nullptr;
  v.n_leaves = (int32)m_bvh.m_leaf_nodes.size ();
  for (int d = 0; d < 3; ++d)
  {
    v.bounds[2 * d] = m_bvh.m_bounds.m_ranges[d].m_min;
    v.bounds[2 * d + 1] = m_bvh.m_bounds.m_ranges[d].m_max;
  }
  v.leaf_starts = m_bvh.m_leaf_starts.size () ? m_bvh.m_leaf_starts.get_device_ptr_const () : nullptr;
  if (m_bvh.m_qnodes.size ())
  {
    v.qnodes = m_bvh.m_qnodes.get_device_ptr_const ();
    v.leaf_boxes = m_bvh.m_leaf_boxes.get_device_ptr_const ();
    for (int q = 0; q < 6; ++q) v.qgrid[q] = m_bvh.m_qgrid[q];
  }
  return v;
}

template <class ElemT> AABB<> UnstructuredMesh<ElemT>::bounds () { return get_bvh ().m_bounds; }

template <class ElemT> Array<Location> UnstructuredMesh<ElemT>::locate (Array<Vec<Float, 3>> &wpoints)
{
  get_bvh ();
  drx_mesh mv = device_view ();
  drx_bvh bv = bvh_view ();
  Array<Location> locs (wpoints.size ());
  drx_locate_opts o {};
  o.init_mode = DRX_INIT_AFFINE;
  o.cond_threshold = 1e4f;
  o.traversal = DRX_TRAVERSE_AUTO;
  if (wpoints.size ())
    drx_check (drx_locate (&mv, &bv, (const float *)wpoints.get_device_ptr_const (), (int64_t)wpoints.size (),
                           locs.get_device_ptr (), &o, nullptr),
               "UnstructuredMesh::locate");
  return locs;
}


template class UnstructuredMesh<Hex3>;
template class UnstructuredMesh<Hex_P1>;
template class UnstructuredMesh<Hex_P2>;

// ---- UnstructuredField<ElemT> ------------------------------------------------------------------------------
template <class ElemT>
UnstructuredField<ElemT>::UnstructuredField (const GridFunction<ElemT::get_ncomp ()> &dof_data, int32 poly_order,
                                             const std::string &name)
: m_dof_data (dof_data), m_poly_order (poly_order)
{
  static_assert (ElemT::get_ncomp () == 1, "the GPU path evaluates scalar fields");
  check_fixed_order (ElemT::get_P (), poly_order, "UnstructuredField");
  m_name = name;
  const int32 npe = (poly_order + 1) * (poly_order + 1) * (poly_order + 1);
  if (m_dof_data.m_el_dofs == 0) m_dof_data.m_el_dofs = npe;
  if (m_dof_data.m_size_el == 0) m_dof_data.m_size_el = (int32)(m_dof_data.m_ctrl_idx.size () / npe);
}

template <class ElemT> std::string UnstructuredField<ElemT>::type_name () const { return element_name<ElemT> (); }

template <class ElemT> std::vector<Range> UnstructuredField<ElemT>::range () const
{
  if (!m_range_override.empty ()) return m_range_override;
  const Vec<Float, 1> *v = m_dof_data.m_values.get_host_ptr_const ();
  Range r = { INFINITY, -INFINITY };
  for (size_t i = 0; i < m_dof_data.m_values.size (); ++i)
  {
    r.m_min = std::min (r.m_min, v[i][0]);
    r.m_max = std::max (r.m_max, v[i][0]);
  }
  return std::vector<Range> { r };
}

template <class ElemT> drx_mesh UnstructuredField<ElemT>::device_view ()
{
  const int32 stride = drx_mesh_stride (m_poly_order, 1);
  if (m_packed.size () == 0)
  {
    m_packed.resize ((size_t)stride * m_dof_data.m_size_el);
    drx_check (drx_mesh_pack (m_dof_data.m_ctrl_idx.get_device_ptr_const (),
                              (const float *)m_dof_data.m_values.get_device_ptr_const (), 1, nullptr, nullptr,
                              nullptr, m_poly_order, m_dof_data.m_size_el, m_packed.get_device_ptr (), nullptr),
               "UnstructuredField: drx_mesh_pack");
  }
  drx_mesh v {};
  v.data = m_packed.get_device_ptr_const ();
  v.elem_offset = nullptr;
  v.elem_order = nullptr;
  v.n_el = m_dof_data.m_size_el;
  v.order = m_poly_order;
  v.ncomp = 1;
  v.stride = stride;
  return v;
}

template <class ElemT> void UnstructuredField<ElemT>::eval (const Array<Location> locs, Array<Float> &values)
{
  if (values.size () != locs.size ()) values.resize (locs.size ());
  drx_mesh fv = device_view ();
  if (locs.size ())
    drx_check (drx_field_eval (&fv, locs.get_device_ptr_const (), (int64_t)locs.size (), values.get_device_ptr (), nullptr),
               "Field::eval");
}

template class UnstructuredField<HexScalar>;
template class UnstructuredField<HexScalar_P1>;
template class UnstructuredField<HexScalar_P2>;

// ---- ColorTable ------------------------------------------------------------------------------------------
void ColorTable::add_point (float32 position, const Vec3f &c)
{
  m_rgb.insert (m_rgb.end (), { position, c[0], c[1], c[2], 1.f });
}
void ColorTable::add_point (float32 position, const Vec4f &c)
{
  m_rgb.insert (m_rgb.end (), { position, c[0], c[1], c[2], c[3] });
}
ColorTable::ColorTable (const std::string &name)
{
  int32 n = 0, smooth = 0;
  drx_check (drx_color_table_preset (name.c_str (), nullptr, 0, &n, &smooth), "ColorTable");
  m_rgb.resize ((size_t)n * 5);
  drx_check (drx_color_table_preset (name.c_str (), m_rgb.data (), n, &n, &smooth), "ColorTable");
  m_smooth = smooth != 0;
}
void ColorTable::add_alpha (float32 position, float32 alpha) { m_alpha.insert (m_alpha.end (), { position, alpha }); }

ColorTable ColorTable::correct_opacity (const float32 &factor) const
{
  ColorTable c (*this);
  for (size_t i = 1; i < c.m_alpha.size (); i += 2) c.m_alpha[i] = 1.f - std::pow ((1.f - c.m_alpha[i]), factor);
  return c;
}

void ColorTable::sample (int num_samples, Array<Vec<float32, 4>> &colors, float32 alpha_scale) const
{
  colors.resize (num_samples);
  drx_check (drx_color_table_sample (m_rgb.data (), (int32)(m_rgb.size () / 5), m_alpha.data (),
                                     (int32)(m_alpha.size () / 2), m_smooth ? 1 : 0, -1.f, num_samples, alpha_scale,
                                     (float *)colors.get_host_ptr ()),
             "ColorTable::sample");
}

// ---- Camera ----------------------------------------------------------------------------------------------
void Camera::set_width (const int32 &w)
{
  if (w <= 0) throw DRayError ("Camera width must be greater than zero.");
  m_width = w;
}
void Camera::set_height (const int32 &h)
{
  if (h <= 0) throw DRayError ("Camera height must be greater than zero.");
  m_height = h;
}
void Camera::set_fov (const float32 &degrees)
{
  if (degrees <= 0) throw DRayError ("Camera feild of view must be greater than zero.");
  if (degrees > 180) throw DRayError ("Camera feild of view must be less than 180.");
  m_fov_y = degrees;
}
drx_camera Camera::view () const
{
  drx_camera c;
  c.width = m_width;
  c.height = m_height;
  c.fov_y = m_fov_y;
  c.zoom = m_zoom;
  for (int i = 0; i < 3; ++i)
  {
    c.pos[i] = m_pos[i];
    c.look_at[i] = m_look_at[i];
    c.up[i] = m_up[i];
  }
  return c;
}

PointLight make_point_light (const Vec3f &pos)
{
  PointLight l;
  for (int i = 0; i < 3; ++i)
  {
    l.pos[i] = pos[i];
    l.amb[i] = 0.4f;
    l.diff[i] = 0.75f;
    l.spec[i] = 0.8f;
  }
  l.spec_pow = 15.f;
  return l;
}

PointLight default_light (Camera &camera)
{
  const drx_camera c = camera.view ();
  PointLight l;
  drx_check (drx_default_light (&c, &l), "default_light");
  return l;
}

void Camera::create_rays (Array<Ray> &rays)
{
  rays.resize ((size_t)m_width * m_height);
  drx_camera c;
  c.width = m_width;
  c.height = m_height;
  c.fov_y = m_fov_y;
  c.zoom = m_zoom;
  for (int i = 0; i < 3; ++i)
  {
    c.pos[i] = m_pos[i];
    c.look_at[i] = m_look_at[i];
    c.up[i] = m_up[i];
  }
  drx_check (drx_camera_rays (&c, rays.get_device_ptr (), nullptr), "Camera::create_rays");
}

// Orbit helpers restated from dray's Vec / Matrix / rotate (vec.hpp, matrix.hpp, transform_3d.hpp:126-157)
// with the same float operation order, so a rotated camera gives the same rays as dray's.
namespace
{
Vec3f v_sub (const Vec3f &a, const Vec3f &b) { return { { a[0] - b[0], a[1] - b[1], a[2] - b[2] } }; }
Vec3f v_add (const Vec3f &a, const Vec3f &b) { return { { a[0] + b[0], a[1] + b[1], a[2] + b[2] } }; }
Vec3f v_scale (const Vec3f &a, float32 s) { return { { a[0] * s, a[1] * s, a[2] * s } }; }
float32 v_mag (const Vec3f &a)
{
#pragma clang fp contract(off)
  float32 sum = 0.f;
  for (int i = 0; i < 3; ++i) sum += a[i] * a[i];
  return sqrtf (sum);
}
void v_normalize (Vec3f &a)
{
  const float32 m = v_mag (a);
  for (int i = 0; i < 3; ++i) a[i] /= m;
}
Vec3f v_cross (const Vec3f &a, const Vec3f &b)
{
#pragma clang fp contract(off)
  return { { a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0] } };
}
// rotate(degrees, axis) applied to a direction (transform_vector: w = 0)
Vec3f rotate_vector (float32 degrees, Vec3f axis, const Vec3f &v)
{
#pragma clang fp contract(off)
  const float32 rad = 0.01745329251994329547437168059786927f * degrees;
  v_normalize (axis);
  const float32 sn = sinf (rad), cs = cosf (rad);
  float32 m[3][3];
  m[0][0] = axis[0] * axis[0] * (1 - cs) + cs;
  m[0][1] = axis[0] * axis[1] * (1 - cs) - axis[2] * sn;
  m[0][2] = axis[0] * axis[2] * (1 - cs) + axis[1] * sn;
  m[1][0] = axis[1] * axis[0] * (1 - cs) + axis[2] * sn;
  m[1][1] = axis[1] * axis[1] * (1 - cs) + cs;
  m[1][2] = axis[1] * axis[2] * (1 - cs) - axis[0] * sn;
  m[2][0] = axis[2] * axis[0] * (1 - cs) - axis[1] * sn;
  m[2][1] = axis[2] * axis[1] * (1 - cs) + axis[0] * sn;
  m[2][2] = axis[2] * axis[2] * (1 - cs) + cs;
  Vec3f r;
  for (int i = 0; i < 3; ++i)
  {
    float32 d = 0.f;
    for (int j = 0; j < 3; ++j) d += m[i][j] * v[j];
    d += 0.f * 0.f;   // the homogeneous column (w = 0)
    r[i] = d;
  }
  return r;
}
} // namespace

void Camera::reset_to_bounds (const AABB<3> bounds, const float64 xpad, const float64 ypad, const float64 zpad)
{
#pragma clang fp contract(off)
  AABB<3> db;
  const float64 pads[3] = { xpad, ypad, zpad };
  for (int d = 0; d < 3; ++d)
  {
    const float64 pad = pads[d] * (bounds.m_ranges[d].max () - bounds.m_ranges[d].min ());
    db.m_ranges[d].include ((float32)(bounds.m_ranges[d].max () + pad));
    db.m_ranges[d].include ((float32)(bounds.m_ranges[d].min () - pad));
  }
  Vec3f proj_dir = v_sub (m_pos, m_look_at);
  v_normalize (proj_dir);
  const Vec3f center = { { db.m_ranges[0].center (), db.m_ranges[1].center (), db.m_ranges[2].center () } };
  m_look_at = center;
  const Vec3f extent = { { db.m_ranges[0].length (), db.m_ranges[1].length (), db.m_ranges[2].length () } };
  const float32 diagonal = v_mag (extent);
  m_pos = v_add (center, v_scale (v_scale (proj_dir, diagonal), 1.0f));
  set_fov (60.0f);
}

void Camera::elevate (const float32 degrees)
{
#pragma clang fp contract(off)
  Vec3f look = v_sub (m_look_at, m_pos);
  const float32 distance = v_mag (look);
  Vec3f right = v_cross (look, m_up);
  Vec3f up = v_cross (right, look);
  v_normalize (look);
  v_normalize (right);
  v_normalize (up);
  look = rotate_vector (-degrees, right, look);
  m_pos = v_sub (m_look_at, v_scale (look, distance));
}

void Camera::azimuth (const float32 degrees)
{
#pragma clang fp contract(off)
  Vec3f look = v_sub (m_look_at, m_pos);
  const float32 distance = v_mag (look);
  Vec3f right = v_cross (look, m_up);
  Vec3f up = v_cross (right, look);
  v_normalize (look);
  v_normalize (right);
  v_normalize (up);
  look = rotate_vector (degrees, up, look);
  m_pos = v_sub (m_look_at, v_scale (look, distance));
}

} // namespace dray

// The camera orbits as C entry points (drx_camera in, moved drx_camera out), for callers without the C++ layer.
namespace drx
{
int fail (int status, const char *fmt, ...);
}
namespace
{
template <typename Op> int camera_op (drx_camera *cam, const char *what, Op op)
{
  if (!cam) return drx::fail (DRX_ERR_ARG, "%s: null camera", what);
  try
  {
    dray::Camera k;
    k.set_width (cam->width);
    k.set_height (cam->height);
    k.set_fov (cam->fov_y);
    k.set_zoom (cam->zoom);
    k.set_pos ({ { cam->pos[0], cam->pos[1], cam->pos[2] } });
    k.set_look_at ({ { cam->look_at[0], cam->look_at[1], cam->look_at[2] } });
    k.set_up ({ { cam->up[0], cam->up[1], cam->up[2] } });
    op (k);
    *cam = k.view ();
    return DRX_OK;
  }
  catch (const std::exception &e)
  {
    return drx::fail (DRX_ERR_ARG, "%s: %s", what, e.what ());
  }
}
} // namespace

extern "C" {
int drx_camera_reset_to_bounds (drx_camera *cam, const float *bounds6, const double *pads)
{
  if (!bounds6) return drx::fail (DRX_ERR_ARG, "drx_camera_reset_to_bounds: null bounds");
  return camera_op (cam, "drx_camera_reset_to_bounds", [&] (dray::Camera &k) {
    dray::AABB<3> b;
    for (int d = 0; d < 3; ++d)
    {
      b.m_ranges[d].include (bounds6[2 * d]);
      b.m_ranges[d].include (bounds6[2 * d + 1]);
    }
    k.reset_to_bounds (b, pads ? pads[0] : 0., pads ? pads[1] : 0., pads ? pads[2] : 0.);
  });
}
int drx_camera_elevate (drx_camera *cam, float degrees)
{
  return camera_op (cam, "drx_camera_elevate", [&] (dray::Camera &k) { k.elevate (degrees); });
}
int drx_camera_azimuth (drx_camera *cam, float degrees)
{
  return camera_op (cam, "drx_camera_azimuth", [&] (dray::Camera &k) { k.azimuth (degrees); });
}
} // extern "C"

namespace dray
{

// ---- Volume ----------------------------------------------------------------------------------------------
// |bounds diagonal| / samples in float32, unfused like dray's (volume.cpp:89-91)
static float sample_distance (const float *bounds6, int32 samples)
{
#pragma clang fp contract(off)
  float s = 0.f;
  for (int d = 0; d < 3; ++d)
  {
    const float e = bounds6[2 * d + 1] - bounds6[2 * d];
    s += e * e;
  }
  return std::sqrt (s) / float (samples);
}

// ---- DataSet (data_model/data_set.cpp) ----------------------------------------------------------------------
bool DataSet::has_mesh (const std::string &name) const
{
  for (const auto &m : m_meshes)
    if (m->name () == name) return true;
  return false;
}

Mesh *DataSet::mesh (const int32 index)
{
  if (index < 0 || index >= number_of_meshes ())
    throw DRayError ("DataSet: no mesh at index " + std::to_string (index) + " (" + std::to_string (number_of_meshes ()) +
                     " meshes)");
  return m_meshes[index].get ();
}

Mesh *DataSet::mesh (const std::string name)
{
  for (auto &m : m_meshes)
    if (m->name () == name) return m.get ();
  throw DRayError ("DataSet: no mesh named '" + name + "'");
}

std::vector<std::string> DataSet::meshes () const
{
  std::vector<std::string> out;
  for (const auto &m : m_meshes) out.push_back (m->name ());
  return out;
}

bool DataSet::has_field (const std::string &name) const
{
  for (const auto &f : m_fields)
    if (f->name () == name) return true;
  return false;
}

Field *DataSet::field (const std::string &name) { return field_shared (name).get (); }

Field *DataSet::field (const int &index) { return field_shared (index).get (); }

std::shared_ptr<Field> DataSet::field_shared (const std::string &name)
{
  for (auto &f : m_fields)
    if (f->name () == name) return f;
  throw DRayError ("No field named '" + name + "'");
}

std::shared_ptr<Field> DataSet::field_shared (const int &index)
{
  if (index < 0 || index >= number_of_fields ()) throw DRayError ("DataSet: bad field index " + std::to_string (index));
  return m_fields[index];
}

std::vector<std::string> DataSet::fields () const
{
  std::vector<std::string> out;
  for (const auto &f : m_fields) out.push_back (f->name ());
  return out;
}

std::string DataSet::field_info ()
{
  std::string s;
  for (auto &f : m_fields) s += f->name () + " " + f->type_name () + "\n";
  return s;
}

// ---- Collection (data_model/collection.cpp) ----------------------------------------------------------------
DataSet Collection::domain (int32 index)
{
  if (index < 0 || index >= size ()) throw DRayError ("Collection: bad domain index " + std::to_string (index));
  return m_domains[index];
}

bool Collection::has_field (const std::string &name)
{
  bool res = true;
  for (auto &d : m_domains) res = res && d.has_field (name);
  return res;
}

Range Collection::range (const std::string &name)
{
  Range r;
  for (auto &d : m_domains) r.include (d.field (name)->range ()[0]);
  return r;
}

AABB<3> Collection::bounds ()
{
  AABB<3> b;
  for (auto &d : m_domains) b.include (d.mesh ()->bounds ());
  return b;
}

std::string Collection::field_list ()
{
  std::string s = "[";
  if (!m_domains.empty ())
    for (const auto &n : m_domains[0].fields ()) s += " " + n;
  return s + " ]";
}

// ---- Volume --------------------------------------------------------------------------------------------------
Volume::Volume (Collection &collection) : m_collection (collection)
{
  // dray's constructor adds a default alpha ramp to the colour map's table (volume.cpp:274-285)
  m_color_map.color_table ().add_alpha (0.1000f, .0f);
  m_color_map.color_table ().add_alpha (1.0000f, .7f);
  m_bounds = m_collection.bounds ();
}

void Volume::input (Collection &collection)
{
  m_collection = collection;
  m_active_domain = 0;
  m_bounds = m_collection.bounds ();
}

void Volume::field (const std::string &name)
{
  m_field = name;
  m_field_range = m_collection.range (name);
}

Array<VolumePartial> Volume::integrate (Array<Ray> &rays, Array<PointLight> &lights)
{
  if (m_field.empty ()) throw DRayError ("Field never set");
  if (!m_color_map.range_set ()) m_color_map.scalar_range (m_field_range);
  if (m_use_lighting && lights.size () > DRX_MAX_LIGHTS)
    throw DRayError ("Volume::integrate: at most " + std::to_string (DRX_MAX_LIGHTS) + " lights");
  DataSet ds = m_collection.domain (m_active_domain);
  // the domain's mesh and field as their derived types (dray's IntegratePartialsFunctor through dispatch_3d)
  drx_mesh mv {}, fv {};
  drx_bvh bv {};
  AABB<> mb;
  auto views = [&] (auto &mesh, auto &field) {
    mb = mesh.bounds ();   // builds the BVH on first use
    mv = mesh.device_view ();
    bv = mesh.bvh_view ();
    fv = field.device_view ();
  };
  dispatch_3d (ds.mesh (), ds.field (m_field), views);
  drx_shading shading;
  shading.use_lighting = m_use_lighting ? 1 : 0;
  shading.n_lights = (int32)lights.size ();
  shading.lights = lights.size () ? lights.get_host_ptr_const () : nullptr;
  const Range range = m_color_map.scalar_range ();
  const float mesh6[6] = { mb.m_ranges[0].m_min, mb.m_ranges[0].m_max, mb.m_ranges[1].m_min,
                           mb.m_ranges[1].m_max, mb.m_ranges[2].m_min, mb.m_ranges[2].m_max };
  const float bounds6[6] = { m_bounds.m_ranges[0].m_min, m_bounds.m_ranges[0].m_max, m_bounds.m_ranges[1].m_min,
                             m_bounds.m_ranges[1].m_max, m_bounds.m_ranges[2].m_min, m_bounds.m_ranges[2].m_max };
  // opacity-corrected table sampled to 1024 entries (volume.cpp:81-87, color_map.cpp:13)
  Array<Vec<float32, 4>> lut;
  m_color_map.color_table ().correct_opacity (10.f / m_samples).sample (1024, lut);
  // sample distance from the collection's bounds (volume.cpp:89-91); rays culled against the domain's
  // mesh bounds (volume.cpp:106)
  const float sample_dist = sample_distance (bounds6, m_samples);
  Array<Ray> active (std::max<size_t> (rays.size (), 1));
  int32 n_active = 0;
  if (rays.size ())
    drx_check (drx_remove_missed_rays (rays.get_device_ptr_const (), (int32)rays.size (), mesh6,
                                       active.get_device_ptr (), &n_active, nullptr),
               "remove_missed_rays");
  Array<int32> order;
  if (m_tile_w > 0 && m_tile_h > 0 && n_active > 0)
  {
    order.resize ((size_t)n_active);
    drx_check (drx_ray_tile_order (active.get_device_ptr_const (), n_active, m_tile_w, m_tile_h,
                                   order.get_device_ptr (), nullptr),
               "drx_ray_tile_order");
  }
  Array<VolumePartial> partials ((size_t)std::max (1, 5 * n_active));
  int32 count = 0;
  drx_locate_opts o {};
  o.cond_threshold = 1e4f;
  drx_check (drx_volume_integrate (&mv, &fv, &bv, active.get_device_ptr_const (), n_active,
                                   (const float *)lut.get_device_ptr_const (), 1024, range.m_min, range.m_max,
                                   sample_dist, order.size () ? order.get_device_ptr_const () : nullptr, &o,
                                   &shading, partials.get_device_ptr (), &count, &m_stats, nullptr),
             "Volume::integrate");
  // keep the first `count` partials (compacted, ray order)
  Array<VolumePartial> out ((size_t)count);
  if (count)
    hip_check (hipMemcpy (out.get_device_ptr (), partials.get_device_ptr_const (), sizeof (VolumePartial) * (size_t)count,
                          hipMemcpyDeviceToDevice),
               "Volume::integrate: partials copy");
  return out;
}

Array<Vec<float32, 4>> composite_domains (std::vector<Array<VolumePartial>> &domain_partials, int32 width,
                                          int32 height)
{
  size_t total = 0;
  for (auto &p : domain_partials) total += p.size ();
  Array<VolumePartial> all (std::max<size_t> (total, 1));
  size_t off = 0;
  for (auto &p : domain_partials)
  {
    if (p.size ())
      hip_check (hipMemcpy (all.get_device_ptr () + off, p.get_device_ptr_const (), sizeof (VolumePartial) * p.size (),
                            hipMemcpyDeviceToDevice),
                 "composite_domains: gather");
    off += p.size ();
  }
  Array<Vec<float32, 4>> img ((size_t)width * height);
  drx_check (drx_composite_domains (all.get_device_ptr_const (), (int32)total, width, height,
                                    (float *)img.get_device_ptr (), nullptr),
             "composite_domains");
  return img;
}

// ---- Framebuffer / Renderer ------------------------------------------------------------------------------
Framebuffer::Framebuffer () : Framebuffer (1024, 1024) {}

Framebuffer::Framebuffer (const int32 width, const int32 height) : m_width (width), m_height (height)
{
  if (width <= 0 || height <= 0) throw DRayError ("Framebuffer: width and height must be positive");
  m_colors.resize ((size_t)width * height);
  m_depths.resize ((size_t)width * height);
  clear ();
}

void Framebuffer::clear () { clear ({ { 0.f, 0.f, 0.f, 0.f } }); }

void Framebuffer::clear (const Vec<float32, 4> &color)
{
  drx_check (drx_framebuffer_clear ((float *)m_colors.get_device_ptr (), m_depths.get_device_ptr (),
                                    m_width * m_height, color.m_data, nullptr),
             "Framebuffer::clear");
}

void Framebuffer::composite_background ()
{
  drx_check (drx_framebuffer_composite_background ((float *)m_colors.get_device_ptr (), m_width * m_height,
                                                   m_bg_color.m_data, nullptr),
             "Framebuffer::composite_background");
}

void Framebuffer::save (const std::string name)
{
  const std::string path = name + ".png";
  drx_check (drx_save_png ((const float *)m_colors.get_host_ptr_const (), m_width, m_height, path.c_str ()),
             "Framebuffer::save");
}

Framebuffer Renderer::render (Camera &camera)
{
  Array<Ray> rays;
  camera.create_rays (rays);
  Framebuffer framebuffer (camera.get_width (), camera.get_height ());
  if (!m_volume) return framebuffer;
  Array<PointLight> lights;
  if (!m_lights.empty ())
  {
    lights.resize (m_lights.size ());
    for (size_t i = 0; i < m_lights.size (); ++i) lights.get_host_ptr ()[i] = m_lights[i];
  }
  else
  {
    lights.resize (1);
    lights.get_host_ptr ()[0] = default_light (camera);
  }
  m_volume->tile_order (camera.get_width (), camera.get_height ());   // 8x8 tile order, same results
  std::vector<Array<VolumePartial>> domain_partials;
  for (int32 d = 0; d < m_volume->num_domains (); ++d)
  {
    m_volume->active_domain (d);
    domain_partials.push_back (m_volume->integrate (rays, lights));
  }
  size_t total = 0;
  for (auto &p : domain_partials) total += p.size ();
  if (total == 0) return framebuffer;
  Array<VolumePartial> all (total);
  size_t off = 0;
  for (auto &p : domain_partials)
  {
    if (p.size ())
      hip_check (hipMemcpy (all.get_device_ptr () + off, p.get_device_ptr_const (), sizeof (VolumePartial) * p.size (),
                            hipMemcpyDeviceToDevice),
                 "Renderer::render: gather partials");
    off += p.size ();
  }
  Array<Vec<float32, 4>> colors = framebuffer.colors ();
  Array<float32> depths = framebuffer.depths ();
  // no surface traceables: the volume replaces the cleared pixels (partials_to_framebuffer, blend false)
  drx_check (drx_partials_to_framebuffer (all.get_device_ptr_const (), (int32)total, camera.get_width (),
                                          camera.get_height (), (float *)colors.get_device_ptr (),
                                          depths.get_device_ptr (), 0, nullptr),
             "Renderer::render");
  return framebuffer;
}

// ---- PointLocation -------------------------------------------------------------------------------------------
static __global__ void k_fill_value (Float *p, int64_t n, Float v)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

PointLocation::Result PointLocation::execute (Collection &collection, Array<Vec<Float, 3>> &points)
{
  if (m_vars.empty ()) throw DRayError ("PointLocation: must specify at least 1 variables.");
  if (points.size () == 0) throw DRayError ("PointLocation: must specify at least 1 point.");
  std::vector<std::string> valid;
  for (const auto &v : m_vars)
    if (collection.has_field (v)) valid.push_back (v);
  if (valid.empty ())
    throw DRayError ("PointLocation: must specify at least one valid field. Known fields " + collection.field_list ());
  const int64_t n = (int64_t)points.size ();
  std::vector<Array<Float>> values (valid.size ());
  for (auto &v : values)
  {
    v.resize ((size_t)n);
    hipLaunchKernelGGL (k_fill_value, dim3 ((unsigned)((n + 255) / 256)), dim3 (256), 0, 0, v.get_device_ptr (), n,
                        m_empty_val);
    hip_check (hipGetLastError (), "PointLocation: fill");
  }
  for (int32 d = 0; d < collection.local_size (); ++d)
  {
    DataSet ds = collection.domain (d);
    Array<Location> locs = ds.mesh ()->locate (points);
    // has_data: any point located in this domain
    const Location *lh = locs.get_host_ptr_const ();
    bool has = false;
    for (int64_t i = 0; i < n && !has; ++i) has = lh[i].m_cell_id != -1;
    if (!has) continue;
    for (size_t f = 0; f < valid.size (); ++f) ds.field (valid[f])->eval (locs, values[f]);
  }
  Result r;
  r.m_points = points;
  r.m_values = values;
  r.m_vars = m_vars;
  r.m_empty_val = m_empty_val;
  return r;
}

Array<Vec<float32, 4>> Volume::composite (Array<VolumePartial> &partials, int32 width, int32 height)
{
  Array<Vec<float32, 4>> img ((size_t)width * height);
  drx_check (drx_composite (partials.size () ? (const void *)partials.get_device_ptr_const () : nullptr,
                            (int32)partials.size (), width, height, (float *)img.get_device_ptr (), nullptr),
             "Volume::save");
  return img;
}

void Volume::save (const std::string name, Array<VolumePartial> partials, const int32 width, const int32 height)
{
  Array<Vec<float32, 4>> img = composite (partials, width, height);
  const std::string path = name + ".png";
  drx_check (drx_save_png ((const float *)img.get_host_ptr_const (), width, height, path.c_str ()), "Volume::save");
}

} // namespace dray

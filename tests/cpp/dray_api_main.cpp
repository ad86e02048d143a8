// Drives the whole volume-render path through dray's class names (include/dray/drx_dray.hpp), the way a
// dray application does (dray's tests/t_dray_volume_render.cpp: mesh -> Camera -> Volume::integrate ->
// composite), reading a mesh from raw files and writing every intermediate result back as raw files for
// the Python test to compare.
//
//   dray_api_test <in_dir> <out_dir>
//   dray_api_test --camera <out_dir>      (host only, no GPU: Camera::reset_to_bounds / azimuth / elevate)
// in_dir:  params.txt  "order n_el n_pts width height samples alpha0 fov px py pz lx ly lz"
//          ctrl.bin (int32), coords.bin (float32 xyz), field.bin (float32), points.bin (float32 xyz),
//          pl_points.bin (float32 xyz, PointLocation queries)
// out_dir: inner.bin leaf.bin ids.bin bounds.bin loc.bin vals.bin rays.bin partials.bin image.bin
//          lit_image.bin light.bin bounds2.bin range2.bin pl_values.bin two_image.bin two_parts0.bin
//          two_parts1.bin fb_colors.bin fb_depths.bin fb_final.bin render.png vol_save.png stats.txt, and "errors ok" on
//          stdout once the error paths have thrown as expected.  --camera writes camera.bin: pos, look_at,
//          fov after reset_to_bounds, then pos after azimuth(-60), then pos after elevate(25).
#include "dray/drx_dray.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

using namespace dray;

template <typename T> static std::vector<T> read_raw (const std::string &path)
{
  std::ifstream f (path, std::ios::binary | std::ios::ate);
  if (!f) throw DRayError ("cannot open " + path);
  const size_t bytes = (size_t)f.tellg ();
  std::vector<T> v (bytes / sizeof (T));
  f.seekg (0);
  f.read ((char *)v.data (), (std::streamsize)bytes);
  return v;
}

template <typename T> static void write_raw (const std::string &path, const T *p, size_t n)
{
  std::ofstream f (path, std::ios::binary);
  f.write ((const char *)p, (std::streamsize)(n * sizeof (T)));
  if (!f) throw DRayError ("cannot write " + path);
}

template <typename T> static void write_array (const std::string &path, const Array<T> &a)
{
  write_raw (path, a.get_host_ptr_const (), a.size ());
}

template <typename T> static Array<T> to_array (const std::vector<T> &v)
{
  Array<T> a (v.size ());
  T *p = a.get_host_ptr ();
  for (size_t i = 0; i < v.size (); ++i) p[i] = v[i];
  return a;
}

template <typename F> static bool throws (F f)
{
  try
  {
    f ();
  }
  catch (const DRayError &)
  {
    return true;
  }
  return false;
}

// Camera orbits as dray's triple-point test drives them (t_dray_volume_render.cpp:82-83), host only.
static int camera_only (const std::string &out)
{
  Camera c;
  c.set_pos ({ { 0.3f, 0.7f, 4.f } });
  c.set_look_at ({ { 0.5f, 0.5f, 0.5f } });
  c.set_up ({ { 0.f, 1.f, 0.f } });
  AABB<> b;
  const float lo[3] = { 0.f, 0.f, -1.f }, hi[3] = { 1.f, 2.f, 1.f };
  for (int d = 0; d < 3; ++d)
  {
    b.m_ranges[d].include (lo[d]);
    b.m_ranges[d].include (hi[d]);
  }
  std::vector<float> v;
  c.reset_to_bounds (b, 0.1, 0.2, 0.3);
  for (int d = 0; d < 3; ++d) v.push_back (c.get_pos ()[d]);
  for (int d = 0; d < 3; ++d) v.push_back (c.get_look_at ()[d]);
  v.push_back (c.get_fov ());
  c.azimuth (-60.f);
  for (int d = 0; d < 3; ++d) v.push_back (c.get_pos ()[d]);
  c.elevate (25.f);
  for (int d = 0; d < 3; ++d) v.push_back (c.get_pos ()[d]);
  write_raw (out + "/camera.bin", v.data (), v.size ());
  return 0;
}

int main (int argc, char **argv)
{
  if (argc == 3 && std::string (argv[1]) == "--camera") return camera_only (argv[2]);
  if (argc != 3)
  {
    std::fprintf (stderr, "usage: %s <in_dir> <out_dir>\n", argv[0]);
    return 2;
  }
  const std::string in = argv[1], out = argv[2];
  try
  {
    drx_check (drx_init (0), "drx_init");
    int order, n_el, n_pts, width, height, samples;
    float alpha0, fov, pos[3], look[3];
    {
      std::ifstream pf (in + "/params.txt");
      pf >> order >> n_el >> n_pts >> width >> height >> samples >> alpha0 >> fov >> pos[0] >> pos[1] >> pos[2] >>
      look[0] >> look[1] >> look[2];
      if (!pf) throw DRayError ("bad params.txt");
    }
    const std::vector<int32> ctrl = read_raw<int32> (in + "/ctrl.bin");
    const std::vector<Vec<Float, 3>> coords = read_raw<Vec<Float, 3>> (in + "/coords.bin");
    const std::vector<Vec<Float, 1>> field = read_raw<Vec<Float, 1>> (in + "/field.bin");
    const std::vector<Vec<Float, 3>> points = read_raw<Vec<Float, 3>> (in + "/points.bin");

    GridFunction<3> gf;
    gf.m_ctrl_idx = to_array (ctrl);
    gf.m_values = to_array (coords);
    gf.m_size_el = n_el;
    GridFunction<1> ff;
    ff.m_ctrl_idx = gf.m_ctrl_idx;   // shallow: same connectivity as dray's field on the mesh's dofs
    ff.m_values = to_array (field);
    ff.m_size_el = n_el;

    // a data set as dray's readers build one: the mesh and the field behind their abstract bases
    std::shared_ptr<Mesh> mesh_p = std::make_shared<HexMesh> (gf, order);
    mesh_p->name ("mesh");
    std::shared_ptr<Field> fld_p = std::make_shared<UnstructuredField<HexScalar>> (ff, order, "f");
    fld_p->mesh_name ("mesh");
    DataSet ds;
    ds.add_mesh (mesh_p);
    ds.add_field (fld_p);
    Collection col;
    col.add_domain (ds);
    Mesh *mesh = ds.mesh ();
    Field *fld = ds.field ("f");
    if (ds.number_of_meshes () != 1 || ds.mesh ("mesh") != mesh || mesh->cells () != n_el || mesh->order () != order ||
        mesh->dims () != 3 || fld->order () != order || fld->components () != 1 || ds.field (0) != fld)
      throw DRayError ("DataSet / Mesh / Field accessors");
    {
      std::ofstream tf (out + "/types.txt");
      tf << mesh->type_name () << " " << fld->type_name () << "\n";
    }

    // BVH (LinearBVHBuilder through construct_bvh), reached from Mesh* through dispatch_3d
    BVH *bvh_p = nullptr;
    auto grab_bvh = [&bvh_p] (auto &m) { bvh_p = &m.get_bvh (); };
    dispatch_3d (mesh, grab_bvh);
    BVH &bvh = *bvh_p;
    write_array (out + "/inner.bin", bvh.m_inner_nodes);
    write_array (out + "/leaf.bin", bvh.m_leaf_nodes);
    write_array (out + "/ids.bin", bvh.m_aabb_ids);
    AABB<> b = mesh->bounds ();
    write_raw (out + "/bounds.bin", &b.m_ranges[0].m_min, 6);

    // Mesh::locate + Field::eval (virtual calls)
    Array<Vec<Float, 3>> wp = to_array (points);
    Array<Location> locs = mesh->locate (wp);
    Array<Float> vals;
    fld->eval (locs, vals);
    write_array (out + "/loc.bin", locs);
    write_array (out + "/vals.bin", vals);
    // the fixed-order mesh type of the same order (HexMesh_P1 / _P2) locates the same points identically
    if (order == 1 || order == 2)
    {
      std::shared_ptr<Mesh> fixed;
      if (order == 1)
        fixed = std::make_shared<HexMesh_P1> (gf, order);
      else
        fixed = std::make_shared<HexMesh_P2> (gf, order);
      Array<Location> l2 = fixed->locate (wp);
      const Location *a = locs.get_host_ptr_const (), *c = l2.get_host_ptr_const ();
      for (size_t i = 0; i < locs.size (); ++i)
        if (a[i].m_cell_id != c[i].m_cell_id || a[i].m_ref_pt[0] != c[i].m_ref_pt[0] ||
            a[i].m_ref_pt[1] != c[i].m_ref_pt[1] || a[i].m_ref_pt[2] != c[i].m_ref_pt[2])
          throw DRayError ("fixed-order mesh type locates differently");
    }

    // Camera -> Volume::integrate -> composite
    Camera cam;
    cam.set_width (width);
    cam.set_height (height);
    cam.set_fov (fov);
    cam.set_pos ({ { pos[0], pos[1], pos[2] } });
    cam.set_look_at ({ { look[0], look[1], look[2] } });
    Array<Ray> rays;
    cam.create_rays (rays);
    write_array (out + "/rays.bin", rays);

    Volume vol (col);
    vol.field ("f");
    vol.use_lighting (false);
    vol.samples (samples);
    vol.tile_order (width, height);
    ColorTable tab;
    tab.add_point (0.f, Vec3f { { 0.5f, 0.5f, 0.5f } });
    tab.add_point (1.f, Vec3f { { 1.f, 0.f, 0.f } });
    tab.add_alpha (0.f, alpha0);
    tab.set_smooth (true);
    vol.color_map ().color_table (tab);
    Array<PointLight> no_lights;
    Array<VolumePartial> parts = vol.integrate (rays, no_lights);
    Array<Vec<float32, 4>> img = vol.composite (parts, width, height);
    write_array (out + "/partials.bin", parts);
    write_array (out + "/image.bin", img);
    vol.save (out + "/vol_save", parts, width, height);   // Volume::save -> vol_save.png
    // lit render with the Renderer's default light (dray's default shading)
    {
      Volume lit (col);
      lit.field ("f");
      lit.samples (samples);
      lit.tile_order (width, height);
      lit.color_map ().color_table (tab);
      Array<PointLight> lights (1);
      lights.get_host_ptr ()[0] = default_light (cam);
      Array<VolumePartial> lp = lit.integrate (rays, lights);
      Array<Vec<float32, 4>> limg = lit.composite (lp, width, height);
      write_array (out + "/lit_image.bin", limg);
      write_raw (out + "/light.bin", &lights.get_host_ptr_const ()[0].pos[0], 13);
    }
    const drx_integrate_stats st = vol.last_stats ();

    // PointLocation over two domains: the mesh and a copy shifted by +1 in x whose field is f + 10
    {
      std::vector<Vec<Float, 3>> shifted = coords;
      for (auto &v : shifted) v[0] = v[0] + 1.0f;
      std::vector<Vec<Float, 1>> f10 = field;
      for (auto &v : f10) v[0] = v[0] + 10.0f;
      GridFunction<3> gf2 = gf;
      gf2.m_values = to_array (shifted);
      GridFunction<1> ff2 = ff;
      ff2.m_values = to_array (f10);
      DataSet ds2 (std::make_shared<HexMesh> (gf2, order));
      ds2.add_field (std::make_shared<UnstructuredField<HexScalar>> (ff2, order, "f"));
      Collection two;
      two.add_domain (ds);
      two.add_domain (ds2);
      const AABB<> b2 = two.bounds ();
      write_raw (out + "/bounds2.bin", &b2.m_ranges[0].m_min, 6);
      const Range r2 = two.range ("f");
      write_raw (out + "/range2.bin", &r2.m_min, 2);
      std::vector<Vec<Float, 3>> qp = read_raw<Vec<Float, 3>> (in + "/pl_points.bin");
      Array<Vec<Float, 3>> qa = to_array (qp);
      PointLocation pl;
      pl.add_var ("f");
      pl.add_var ("missing");
      pl.empty_val (-7.f);
      PointLocation::Result res = pl.execute (two, qa);
      write_array (out + "/pl_values.bin", res.m_values[0]);

      // Renderer-style volume render of the two-domain collection: every domain, then cross-domain compositing
      Volume v2 (two);
      v2.field ("f");
      v2.use_lighting (false);
      v2.samples (samples);
      v2.color_map ().color_table (tab);
      std::vector<Array<VolumePartial>> dp;
      for (int d = 0; d < v2.num_domains (); ++d)
      {
        v2.active_domain (d);
        dp.push_back (v2.integrate (rays, no_lights));
      }
      Array<Vec<float32, 4>> timg = composite_domains (dp, width, height);
      write_array (out + "/two_image.bin", timg);
      write_array (out + "/two_parts0.bin", dp[0]);
      write_array (out + "/two_parts1.bin", dp[1]);

      // the same through Renderer::render -> Framebuffer::composite_background -> Framebuffer::save
      auto vr = std::make_shared<Volume> (two);
      vr->field ("f");
      vr->use_lighting (false);
      vr->samples (samples);
      vr->color_map ().color_table (tab);
      Renderer renderer;
      renderer.volume (vr);
      Framebuffer fb = renderer.render (cam);
      write_array (out + "/fb_colors.bin", fb.colors ());
      write_array (out + "/fb_depths.bin", fb.depths ());
      fb.composite_background ();
      write_array (out + "/fb_final.bin", fb.colors ());
      fb.save (out + "/render");
    }
    {
      std::ofstream sf (out + "/stats.txt");
      sf << st.located_samples << " " << st.newton_steps << " " << st.candidates << " " << parts.size () << "\n";
    }

    // error behaviour mirrors dray's DRAY_ERROR checks
    bool ok = true;
    ok &= throws ([] { Camera c; c.set_width (0); });
    ok &= throws ([] { Camera c; c.set_fov (181.f); });
    ok &= throws ([] { Framebuffer f (0, 4); });
    ok &= throws ([&] { HexMesh m (gf, DRX_MAX_ORDER + 1); });   // above the highest order
    ok &= throws ([&] { HexMesh_P1 m (gf, 2); });              // fixed-order type, other order
    ok &= throws ([&] { DataSet d; d.mesh (0); });             // no mesh
    ok &= throws ([&] { ds.mesh ("nope"); });
    ok &= throws ([&] {   // a field of no scalar element type dispatch_3d knows
      struct OtherField : Field
      {
        std::vector<Range> range () const override { return {}; }
        int32 order () const override { return 1; }
        int32 components () const override { return 1; }
        std::string type_name () const override { return "other"; }
        void eval (const Array<Location>, Array<Float> &) override {}
      } other;
      auto noop = [] (auto &, auto &) {};
      dispatch_3d (mesh, &other, noop);
    });
    ok &= throws ([&] {   // more lights than the library takes
      Volume v (col);
      v.field ("f");
      Array<PointLight> many (DRX_MAX_LIGHTS + 1);
      v.integrate (rays, many);
    });
    ok &= throws ([&] {   // Volume::integrate before field() (volume.cpp:331-334)
      Volume v (col);
      v.integrate (rays, no_lights);
    });
    ok &= throws ([&] {   // PointLocation without a valid variable
      PointLocation pl;
      pl.add_var ("nope");
      Array<Vec<Float, 3>> one = to_array (std::vector<Vec<Float, 3>> (1));
      pl.execute (col, one);
    });
    std::printf ("%s\n", ok ? "errors ok" : "errors FAILED");
    return ok ? 0 : 1;
  }
  catch (const std::exception &e)
  {
    std::fprintf (stderr, "dray_api_test: %s\n", e.what ());
    return 1;
  }
}

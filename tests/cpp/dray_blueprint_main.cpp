// dray's BlueprintReader and volume renders of dray's shipped data sets through the C++ classes
// (include/dray/drx_dray.hpp), written the way dray's tests are (t_dray_volume_render.cpp:60-108,
// t_dray_mpi_volume_render.cpp:20-62).
//
//   dray_blueprint_test load <root_file> <out_dir>        (host only, no GPU)
//       summary.txt: one line per domain "domain_id topology order n_el field:min:max ..."; per domain i
//       d<i>_values.bin (float32 xyz control points) and d<i>_<field>.bin (float32 coefficients)
//   dray_blueprint_test render <root_file> <field> <lit 0|1> <azimuth> <elevate> <out_base> [pos alpha]...
//       Spectral table with the given alpha points, 512 x 512 camera orbited (azimuth, then elevate if not 0)
//       and reset to the collection's bounds, Renderer::render, composite_background, save <out_base>.png
#include "dray/drx_dray.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <string>

using namespace dray;

template <typename T> static void write_raw (const std::string &path, const T *p, size_t n)
{
  std::ofstream f (path, std::ios::binary);
  f.write ((const char *)p, (std::streamsize)(n * sizeof (T)));
  if (!f) throw DRayError ("cannot write " + path);
}

static int load (const std::string &root, const std::string &out)
{
  Collection col = BlueprintReader::load (root);
  std::ofstream sum (out + "/summary.txt");
  for (int32 i = 0; i < col.size (); ++i)
  {
    DataSet ds = col.domain (i);
    Mesh *mesh = ds.mesh ();
    sum << ds.domain_id () << " " << mesh->name () << " " << mesh->order () << " " << mesh->cells ();
    std::string fields;
    auto dump = [&] (auto &m) {
      GridFunction<3> g = m.get_dof_data ();
      write_raw (out + "/d" + std::to_string (i) + "_values.bin", (const float *)g.m_values.get_host_ptr_const (),
                 3 * g.m_values.size ());
    };
    dispatch_3d (mesh, dump);
    for (const std::string &name : ds.fields ())
    {
      Field *f = ds.field (name);
      const Range r = f->range ()[0];
      sum << " " << name << ":" << r.min () << ":" << r.max ();
      auto *uf = dynamic_cast<UnstructuredField<HexScalar> *> (f);
      if (!uf) throw DRayError ("field '" + name + "' is not an UnstructuredField<HexScalar>");
      GridFunction<1> g = uf->get_dof_data ();
      write_raw (out + "/d" + std::to_string (i) + "_" + name + ".bin", (const float *)g.m_values.get_host_ptr_const (),
                 g.m_values.size ());
    }
    sum << "\n";
  }
  return 0;
}

static int render (int argc, char **argv)
{
  const std::string root = argv[2], field = argv[3];
  const bool lit = std::atoi (argv[4]) != 0;
  const float az = (float)std::atof (argv[5]), el = (float)std::atof (argv[6]);
  const std::string out = argv[7];
  Collection col = BlueprintReader::load (root);
  ColorTable table ("Spectral");
  for (int a = 8; a + 1 < argc; a += 2) table.add_alpha ((float)std::atof (argv[a]), (float)std::atof (argv[a + 1]));
  Camera camera;
  camera.set_width (512);
  camera.set_height (512);
  camera.azimuth (az);
  if (el != 0.f) camera.elevate (el);
  camera.reset_to_bounds (col.bounds ());
  auto volume = std::make_shared<Volume> (col);
  volume->field (field);
  volume->use_lighting (lit);
  volume->color_map ().color_table (table);
  Renderer renderer;
  renderer.volume (volume);
  Framebuffer fb = renderer.render (camera);
  fb.composite_background ();
  fb.save (out);
  return 0;
}

int main (int argc, char **argv)
{
  try
  {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "load" && argc == 4) return load (argv[2], argv[3]);
    if (mode == "render" && argc >= 8) return render (argc, argv);
    std::fprintf (stderr, "usage: dray_blueprint_test load <root> <out_dir> | render <root> <field> <lit> <az> <el> "
                          "<out_base> [pos alpha]...\n");
    return 2;
  }
  catch (const std::exception &e)
  {
    std::fprintf (stderr, "error: %s\n", e.what ());
    return 1;
  }
}

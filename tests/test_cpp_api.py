"""The C++ host layer (include/dray/drx_dray.hpp): dray's class names over the C ABI.

tests/cpp/dray_api_test drives mesh -> BVH -> locate -> field eval -> camera -> Volume::integrate ->
composite through those classes in a process of its own; every output must be bit-identical to the same
calls made through the C ABI from Python, whose parity with the oracle the other GPU tests establish, and
the BVH bit-identical to the oracle's.
"""
import os
import subprocess

import numpy as np
import pytest

from devil_ray_amd import build as B
from devil_ray_amd import synthetic as syn
from devil_ray_amd.synthetic import LOCATION_DTYPE, PARTIAL_DTYPE, RAY_DTYPE
from pngread import read_png

BIN = B.CPP_TEST_BIN


def _need_bin():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built (run devil_ray_amd/build.py)")


def test_cpp_program_links_and_reports_usage():
    _need_bin()
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_library_exports_dray_classes():
    out = subprocess.run(["nm", "-DC", "--defined-only", B.LIB], capture_output=True, text=True, check=True).stdout
    hex3 = "dray::UnstructuredMesh<dray::Element<3, 3, (dray::ElemType)0, -1> >::"
    for sym in (hex3 + "get_bvh()", hex3 + "locate(", "dray::UnstructuredMesh<dray::Element<3, 3, (dray::ElemType)0, 1> >::",
                "dray::UnstructuredMesh<dray::Element<3, 3, (dray::ElemType)0, 2> >::", "dray::Volume::integrate(",
                "dray::Volume::composite(", "dray::LinearBVHBuilder::construct(", "dray::Camera::create_rays(",
                "dray::UnstructuredField<dray::Element<3, 1, (dray::ElemType)0, -1> >::eval(", "dray::ColorTable::sample(",
                "dray::DataSet::mesh(int)", "dray::DataSet::field(", "dray::detail::cast_mesh_failed("):
        assert sym in out, f"libdrx.so does not export {sym}"


def test_camera_orbits_match_oracle(orc, tmp_path):
    """Camera::reset_to_bounds(pads), azimuth(-60), elevate(25) in the C++ layer (host code, no GPU)
    give the oracle's positions bit for bit (its closed forms are checked in test_oracle.py)."""
    _need_bin()
    r = subprocess.run([BIN, "--camera", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(tmp_path / "camera.bin", np.float32)
    up = (0.0, 1.0, 0.0)
    p, look, fov = orc.camera_orbit((0.3, 0.7, 4.0), (0.5, 0.5, 0.5), up, 30.0, "reset",
                                    bounds6=(0, 1, 0, 2, -1, 1), pads=(0.1, 0.2, 0.3))
    p2, _, _ = orc.camera_orbit(p, look, up, fov, "azimuth", -60.0)
    p3, _, _ = orc.camera_orbit(p2, look, up, fov, "elevate", 25.0)
    exp = np.array([*p, *look, fov, *p2, *p3], np.float32)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


CPP_MESHES = [("p3_n5", lambda: syn.hex_mesh(5, 3, 0.1, 0.01, 2)),
              ("p2_n4", lambda: syn.hex_mesh(4, 2, 0.05, 0.01, 7))]


@pytest.mark.gpu
@pytest.mark.parametrize("name,make", CPP_MESHES, ids=[c[0] for c in CPP_MESHES])
def test_cpp_api_matches_c_abi(gpu, orc, tmp_path, name, make):
    """The program holds its mesh and field as Mesh* / Field* in a DataSet (HexMesh and
    UnstructuredField<HexScalar>), reaches the BVH through dispatch_3d, locates and evaluates through the
    virtual calls, and for orders 1 and 2 checks HexMesh_P1 / HexMesh_P2 against HexMesh itself."""
    import torch
    _need_bin()
    mesh = make()
    w, h, samples, opacity = 40, 36, 96, 0.02
    alpha0 = syn.table_alpha_for_corrected(opacity, samples)
    pos, look = (0.5, 0.5, 2.5), (0.5, 0.5, 0.5)
    ind, outd = tmp_path / "in", tmp_path / "out"
    ind.mkdir()
    outd.mkdir()
    ob = orc.build_mesh_bvh(mesh)
    pts = syn.query_points(20000, ob.bounds, seed=5)
    mesh.ctrl_idx.astype(np.int32).tofile(ind / "ctrl.bin")
    mesh.values.astype(np.float32).tofile(ind / "coords.bin")
    mesh.field.astype(np.float32).tofile(ind / "field.bin")
    pts.astype(np.float32).tofile(ind / "points.bin")
    rng = np.random.default_rng(3)
    plp = np.concatenate([rng.random((2000, 3)) * [2.0, 1.0, 1.0], rng.random((50, 3)) + [0.0, 1.2, 0.0]])
    plp = plp.astype(np.float32)
    plp.tofile(ind / "pl_points.bin")
    (ind / "params.txt").write_text(" ".join(str(x) for x in (
        mesh.order, mesh.n_el, len(pts), w, h, samples, repr(float(alpha0)), 45.0, *pos, *look)))
    print(f"[cpp api] {name}, {len(pts)} points, {w}x{h} @ {samples} spp", flush=True)
    r = subprocess.run([BIN, str(ind), str(outd)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "errors ok" in r.stdout
    # element_name (elem_attr.hpp:408-427) of the mesh (Hex3) and field (HexScalar) types
    assert (outd / "types.txt").read_text().split() == ["3D_Tensor_C3_P-1", "3D_Tensor_C1_P-1"]

    # BVH: bit-exact with the oracle
    inner = np.fromfile(outd / "inner.bin", np.float32).reshape(-1, 16)
    assert np.array_equal(inner[:, :14].view(np.uint32), ob.inner[:, :14].view(np.uint32))
    assert np.array_equal(np.fromfile(outd / "leaf.bin", np.int32), ob.leaf_nodes)
    assert np.array_equal(np.fromfile(outd / "ids.bin", np.int32), ob.aabb_ids)
    assert np.array_equal(np.fromfile(outd / "bounds.bin", np.float32).view(np.uint32), ob.bounds.view(np.uint32))

    # the same calls from Python
    m = gpu.UnstructuredMesh(mesh)
    f = gpu.UnstructuredField(mesh)
    loc = m.locate(gpu.to_device(pts.reshape(-1)))
    vals = f.eval(loc)
    torch.cuda.synchronize()
    assert np.array_equal(np.fromfile(outd / "loc.bin", np.uint8), loc.cpu().numpy())
    assert np.array_equal(np.fromfile(outd / "vals.bin", np.float32).view(np.uint32),
                          vals.cpu().numpy().view(np.uint32))
    rays = gpu.Camera(w, h, 45.0, pos=pos, look_at=look).create_rays()
    assert np.array_equal(np.fromfile(outd / "rays.bin", np.uint8), rays.cpu().numpy())
    vol = gpu.Volume(m, f)
    vol.use_lighting = False
    vol.samples(samples)
    vol.table = gpu.ColorTable()
    vol.table.add_point(0.0, (0.5, 0.5, 0.5))
    vol.table.add_point(1.0, (1.0, 0.0, 0.0))
    vol.table.add_alpha(0.0, alpha0)
    vol.table.set_smooth(True)
    parts, n = vol.integrate(rays, tile_width=w, tile_height=h)
    img = gpu.composite(parts, n, w, h)
    torch.cuda.synchronize()
    cp = np.fromfile(outd / "partials.bin", np.uint8)
    assert len(cp) == n * PARTIAL_DTYPE.itemsize
    assert np.array_equal(cp, parts.cpu().numpy()[: n * PARTIAL_DTYPE.itemsize])
    assert np.array_equal(np.fromfile(outd / "image.bin", np.float32).view(np.uint32),
                          img.cpu().numpy().view(np.uint32))
    # Volume::save: the composited image as an 8-bit PNG, row 0 of the framebuffer at the bottom
    vs = read_png(outd / "vol_save.png")
    assert np.array_equal(vs, (img.cpu().numpy().reshape(h, w, 4) * np.float32(255)).astype(np.uint8)[::-1])
    # lit render with the default light: the same as the Python call
    light = np.fromfile(outd / "light.bin", np.float32)
    assert np.allclose(light, gpu.default_light(gpu.Camera(w, h, 45.0, pos=pos, look_at=look)))
    vol2 = gpu.Volume(m, f)
    vol2.samples(samples)
    vol2.table = vol.table
    parts2, n2 = vol2.integrate(rays, tile_width=w, tile_height=h, lights=light[None, :])
    img2 = gpu.composite(parts2, n2, w, h)
    torch.cuda.synchronize()
    assert np.array_equal(np.fromfile(outd / "lit_image.bin", np.float32).view(np.uint32),
                          img2.cpu().numpy().view(np.uint32))
    # PointLocation over two domains (the mesh, and a copy shifted by +1 in x with field + 10):
    # a point takes the last domain that locates it, else empty_val (-7)
    m2h = syn.HexMesh(n=mesh.n, order=mesh.order, ctrl_idx=mesh.ctrl_idx,
                      values=(mesh.values + np.float32([1.0, 0.0, 0.0])).astype(np.float32),
                      field=(mesh.field + np.float32(10.0)).astype(np.float32))
    m2, f2 = gpu.UnstructuredMesh(m2h), gpu.UnstructuredField(m2h)
    dpl = gpu.to_device(plp.reshape(-1))
    l0, l1 = m.locate(dpl), m2.locate(dpl)
    v0, v1 = f.eval(l0).cpu().numpy(), f2.eval(l1).cpu().numpy()
    c0 = l0.cpu().numpy().view(LOCATION_DTYPE)["cell"]
    c1 = l1.cpu().numpy().view(LOCATION_DTYPE)["cell"]
    exp = np.where(c1 >= 0, v1, np.where(c0 >= 0, v0, np.float32(-7.0)))
    got = np.fromfile(outd / "pl_values.bin", np.float32)
    assert np.array_equal(got.view(np.uint32), exp.astype(np.float32).view(np.uint32))
    assert (c0 >= 0).sum() > 500 and (c1 >= 0).sum() > 500 and (exp == -7).sum() >= 50
    b2 = np.fromfile(outd / "bounds2.bin", np.float32)
    assert b2[0] == ob.bounds[0] and np.isclose(b2[1], ob.bounds[1] + 1.0) and np.array_equal(b2[2:], ob.bounds[2:])
    r2 = np.fromfile(outd / "range2.bin", np.float32)
    assert r2[0] == mesh.field.min() and r2[1] == np.float32(mesh.field.max() + np.float32(10.0))
    # the two-domain render: each domain with the collection's bounds and field range, then composite_domains
    dparts = []
    for mm, ff in ((m, f), (m2, f2)):
        vv = gpu.Volume(mm, ff)
        vv.use_lighting = False
        vv.samples(samples)
        vv.sample_bounds = b2
        vv.color_map_range(float(r2[0]), float(r2[1]))
        vv.table = vol.table
        dparts.append(vv.integrate(rays))
    timg = gpu.composite_domains(dparts, w, h)
    torch.cuda.synchronize()
    assert np.array_equal(np.fromfile(outd / "two_image.bin", np.float32).view(np.uint32),
                          timg.cpu().numpy().view(np.uint32))
    # Renderer::render over the same collection -> Framebuffer colours and depths (the oracle's
    # partials_to_framebuffer of the two domains' partials), composite_background -> the image above,
    # Framebuffer::save -> the PNG of that image
    tp = [np.fromfile(outd / f"two_parts{d}.bin", PARTIAL_DTYPE) for d in (0, 1)]
    for d in (0, 1):
        assert np.array_equal(tp[d].view(np.uint8),
                              dparts[d][0].cpu().numpy()[: dparts[d][1] * PARTIAL_DTYPE.itemsize])
    col = np.zeros((w * h, 4), np.float32)
    dep = np.full(w * h, np.inf, np.float32)
    orc.partials_to_framebuffer(np.concatenate(tp), w, h, col, dep)
    assert np.array_equal(np.fromfile(outd / "fb_colors.bin", np.float32).view(np.uint32), col.view(np.uint32).ravel())
    assert np.array_equal(np.fromfile(outd / "fb_depths.bin", np.float32).view(np.uint32), dep.view(np.uint32))
    assert np.isfinite(dep).sum() > 100
    fin = np.fromfile(outd / "fb_final.bin", np.float32)
    assert np.array_equal(fin.view(np.uint32), timg.cpu().numpy().view(np.uint32))
    png = read_png(outd / "render.png")
    assert np.array_equal(png, np.clip(fin.reshape(h, w, 4) * np.float32(255), 0, 255).astype(np.uint8)[::-1])
    st = (outd / "stats.txt").read_text().split()
    assert int(st[0]) == vol.last_stats.located_samples and int(st[3]) == n
    assert LOCATION_DTYPE.itemsize == 16 and RAY_DTYPE.itemsize == 36

"""Host side of the data-set path: the HDF5 subset reader, the Blueprint/MFEM -> Bernstein conversion, named
colour tables and camera orbits; and the CPU oracle's renders of dray's two shipped-data volume tests against
dray's own baseline images (the reader, conversion, colour table, camera and render chain pinned together)."""
import os
import subprocess
from math import comb

import numpy as np
import pytest

import blueprint_cases as bc
from devil_ray_amd import blueprint as bp
from devil_ray_amd import build as B
from devil_ray_amd.hdf5_lite import H5File

BP_BIN = B.CPP_BP_BIN


def _bern_eval_hex(coef, p, u, v, w):
    """Tensor Bernstein sum at (u, v, w), coefficients lexicographic (i fastest), written out directly."""
    c = np.asarray(coef, np.float64).reshape(p + 1, p + 1, p + 1)   # [k, j, i]
    b = lambda t: np.array([comb(p, i) * t ** i * (1 - t) ** (p - i) for i in range(p + 1)])
    return float(np.einsum("kji,i,j,k->", c, b(u), b(v), b(w)))


# ---- HDF5 subset reader ------------------------------------------------------------------------------
def test_hdf5_reads_root_and_domain():
    r = H5File(bc.root("triple_vr"))
    assert r.read("/file_pattern") == "field_dump.cycle_006700/domain_%06d.hdf5"
    assert int(np.asarray(r.read("/number_of_files")).reshape(-1)[0]) == 1
    d = H5File(bc.GOLDEN + "/tripple_point/field_dump.cycle_006700/domain_000000.hdf5")
    assert d.read("/topologies/mesh/elements/shape") == "hex"
    conn = d.read("/topologies/mesh/elements/connectivity")
    nv = len(d.read("/coordsets/coords/values/x"))
    assert conn.size == 8 * 3276 and conn.min() == 0 and conn.max() == nv - 1
    with pytest.raises(KeyError):
        d.read("/no/such/path")


def test_hdf5_rejects_non_hdf5(tmp_path):
    f = tmp_path / "x.hdf5"
    f.write_bytes(b"not an hdf5 file at all")
    with pytest.raises(ValueError):
        H5File(str(f))


# ---- MFEM numbering and basis conversion ------------------------------------------------------------------
def test_h1_q2_numbering_counts_and_sharing():
    """Two hexes sharing the face x = 1: 12 vertices, 20 edges, 11 faces, 2 interiors = 45 dofs (the 5x3x3
    Q2 lattice), and the shared face's 9 dofs are the same global dofs from both elements."""
    conn = np.array([[0, 1, 2, 3, 4, 5, 6, 7], [1, 8, 9, 2, 5, 10, 11, 6]])
    dofs, n = bp.mfem_h1_element_dofs(conn, 12, 2)
    assert n == 45 and dofs.shape == (2, 27)
    lat = bp._local_lattice(2)
    g = np.empty((2, 27), np.int64)
    g[:, lat] = dofs
    g = g.reshape(2, 3, 3, 3)                 # [e, k, j, i]
    assert np.array_equal(g[0, :, :, 2], g[1, :, :, 0])
    assert len(np.unique(dofs)) == 45
    # edge 0-1 of element 0 is numbered first, the interiors last
    assert dofs[0, 8] == 12 and dofs[1, 26] == 44


def test_q2_lattice_positions():
    """The 27 local dofs sit at the lattice point their entity's midpoint gives: vertices at corners, edge
    dofs at edge midpoints, face dofs at face centres, the interior at (1, 1, 1)."""
    lat = bp._local_lattice(2)
    assert sorted(lat.tolist()) == list(range(27))
    assert lat[0] == 0 and lat[6] == 26 and lat[26] == 13     # vertex 0, vertex 6 (1,1,1), the centre
    assert lat[8] == 1                                        # edge 0-1 midpoint (1, 0, 0) in half units


def test_nodal_to_bernstein_interpolates():
    """Bernstein coefficients from nodal values reproduce those values at the nodes (GaussLobatto orders 1-3,
    GaussLegendre orders 0-2), and the order-2 GaussLobatto closed form agrees with the general solve."""
    rng = np.random.default_rng(3)
    for btype, orders in ((bp.GAUSS_LOBATTO, (1, 2, 3)), (bp.GAUSS_LEGENDRE, (0, 1, 2))):
        for p in orders:
            nodes = bp.nodes_1d(btype, p)
            f = rng.standard_normal((p + 1) ** 3)
            c = bp.nodal_to_bernstein(f[None], p, nodes)[0]
            for k in range(p + 1):
                for j in range(p + 1):
                    for i in range(p + 1):
                        val = _bern_eval_hex(c, p, nodes[i], nodes[j], nodes[k]) if p else float(c[0])
                        assert abs(val - f[i + (p + 1) * (j + (p + 1) * k)]) < 1e-12
    f = rng.standard_normal((5, 27))
    assert np.allclose(bp.lagrange_q2_to_bernstein(f), bp.nodal_to_bernstein(f, 2, np.array([0.0, 0.5, 1.0])),
                       atol=1e-13)
    assert np.allclose(bp.nodes_1d(bp.GAUSS_LOBATTO, 2), [0.0, 0.5, 1.0])
    assert np.allclose(bp.nodes_1d(bp.GAUSS_LEGENDRE, 1), [0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)])


def test_degree_elevation_is_exact():
    rng = np.random.default_rng(4)
    for p, q in ((1, 2), (1, 3), (2, 3)):
        c = rng.standard_normal((p + 1) ** 3)
        e = bp.elevate_bernstein_hex(c[None], p, q)[0]
        for u, v, w in rng.random((6, 3)):
            assert abs(_bern_eval_hex(c, p, u, v, w) - _bern_eval_hex(e, q, u, v, w)) < 1e-12
    # corners are kept; a bump's interior maximum is pulled in (why the imported range is kept separately)
    e = bp.elevate_bernstein_hex(np.array([0.0, 1.0, 0.0] * 9)[None], 2, 3)[0]
    assert e.max() < 1.0


def test_parse_basis():
    assert bp.parse_basis("H1_3D_P2") == ("H1", bp.GAUSS_LOBATTO, 2)
    assert bp.parse_basis("L2_T2_3D_P1") == ("L2", bp.POSITIVE, 1)
    assert bp.parse_basis("L2_3D_P0") == ("L2", bp.GAUSS_LEGENDRE, 0)
    assert bp.parse_basis("H1Pos_3D_P3") == ("H1", bp.POSITIVE, 3)
    with pytest.raises(ValueError):
        bp.parse_basis("ND_3D_P1")


def test_triple_point_domain():
    """The triple point: 3276 order-2 hexes in [0, 3] x [0, 3] x [0, 7]; the order-1 L2 density elevated to
    order 2; every element's Jacobian determinant positive at its corners and centre (a wrong dof numbering
    scrambles control points and folds elements)."""
    (d,) = bp.load(bc.root("triple_vr"))
    m = d.mesh
    assert m.order == 2 and m.n_el == 3276 and d.topology == "mesh"
    assert np.allclose(m.values.min(0), 0.0, atol=1e-5)
    assert np.allclose(m.values.max(0), [3.0, 3.0, 7.0], atol=1e-4)
    assert set(d.fields) == {"density"}
    lo, hi = d.ranges["density"]
    assert 0.12 < lo < 0.13 and 6.0 < hi < 6.1
    assert d.fields["density"].min() >= lo and d.fields["density"].max() <= hi
    nets = m.values.reshape(-1, 27, 3).astype(np.float64)
    pts = [(0, 0, 0), (1, 1, 1), (0.5, 0.5, 0.5), (1, 0, 1)]
    for e in range(0, 3276, 7):
        c = nets[e].reshape(3, 3, 3, 3)         # [k, j, i, xyz]
        for u, v, w in pts:
            J = np.empty((3, 3))
            h = 1e-6
            for a in range(3):
                dp = np.array([u, v, w], float)
                dm = dp.copy()
                dp[a] = min(1.0, dp[a] + h)
                dm[a] = max(0.0, dm[a] - h)
                fp = [_bern_eval_hex(c[..., x].reshape(-1), 2, *dp) for x in range(3)]
                fm = [_bern_eval_hex(c[..., x].reshape(-1), 2, *dm) for x in range(3)]
                J[:, a] = (np.array(fp) - np.array(fm)) / (dp[a] - dm[a])
            assert np.linalg.det(J) > 0, (e, u, v, w)


def test_laghos_domains():
    doms = bp.load(bc.root("tg_mpi_volume"))
    assert [d.domain_id for d in doms] == list(range(8))
    assert all(d.topology == "main" and d.mesh.order == 2 and d.mesh.n_el == 8 for d in doms)
    assert all(set(d.fields) == {"density", "specific_internal_energy"} for d in doms)   # velocity is a vector
    lo = min(d.mesh.values.min() for d in doms)
    hi = max(d.mesh.values.max() for d in doms)
    assert abs(lo) < 1e-5 and abs(hi - 1.0) < 1e-5


# ---- colour tables and camera (host functions of libdrx) --------------------------------------------------
def test_named_colour_tables():
    """Spectral is ColorBrewer's 11-class Spectral (9e0142 ... 5e4fa2, each channel byte / 255 to 4 places);
    named tables are smooth except "sharp"; "" and "default" give cool2warm; unknown names give YlOrRd."""
    from devil_ray_amd import api
    hexes = "9e0142 d53e4f f46d43 fdae61 fee08b ffffbf e6f598 abdda4 66c2a5 3288bd 5e4fa2".split()
    t = api.ColorTable("Spectral")
    pts = np.array(t.rgb)
    assert t.smooth and pts.shape == (11, 5)
    assert np.allclose(pts[:, 0], np.arange(11) / 10.0, atol=1e-7)
    want = np.array([[int(h[i:i + 2], 16) / 255.0 for i in (0, 2, 4)] for h in hexes])
    assert np.abs(pts[:, 1:4] - want).max() < 6e-5
    assert np.all(pts[:, 4] == 1.0)
    assert not api.ColorTable("sharp").smooth and api.ColorTable("dense").smooth
    assert api.ColorTable("").rgb == api.ColorTable("default").rgb == api.ColorTable("cool2warm").rgb
    assert len(api.ColorTable("cool2warm").rgb) == 256
    assert api.ColorTable("no such table").rgb == api.ColorTable("YlOrRd").rgb
    iso = np.array(api.ColorTable("IsoL").rgb)
    assert np.array_equal(iso[:, 0], (np.arange(6, dtype=np.float32) / np.float32(5)))
    assert api.ColorTable().rgb == [] and not api.ColorTable().smooth


def test_camera_orbits_match_oracle(orc):
    """Camera.azimuth / elevate / reset_to_bounds through libdrx equal the oracle's float32 restatement bit
    for bit (the same calls the volume tests make)."""
    from devil_ray_amd import api
    cam = api.Camera(512, 512)
    cam.azimuth(20.0)
    cam.elevate(10.0)
    b = np.array([0.0, 1.0, -0.5, 1.0, 0.0, 7.0], np.float32)
    cam.reset_to_bounds(b)
    pos, look, up, fov = (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 30.0
    pos, look, fov = orc.camera_orbit(pos, look, up, fov, "azimuth", 20.0)
    pos, look, fov = orc.camera_orbit(pos, look, up, fov, "elevate", 10.0)
    pos, look, fov = orc.camera_orbit(pos, look, up, fov, "reset", bounds6=b)
    assert np.array_equal(np.array(cam.pos, np.float32), pos)
    assert np.array_equal(np.array(cam.look_at, np.float32), look)
    assert cam.fov == fov == 60.0


# ---- the oracle's renders against dray's baseline images ------------------------------------------------------
@pytest.mark.parametrize("case", list(bc.CASES))
def test_oracle_render_matches_dray_baseline(orc, case):
    """dray's own bar is at most 20% differing pixels; the oracle stays within 5% overall, almost all of it the
    colour bar dray draws, and within 1% outside it."""
    doms = bp.load(bc.root(case))
    img = bc.oracle_render(orc, doms, case)
    base = bc.baseline(case)
    whole = bc.diff_fraction(img, base)
    outside = bc.diff_fraction_outside_bar(img, base)
    print(f"{case}: {whole:.4f} of pixels differ, {outside:.4f} outside the colour bar")
    assert whole <= 0.05
    assert outside <= 0.01
    # the picture is not blank: a planted mistake in the table (no colour) or the camera (turned away) fails
    assert (img[..., :3] < 250).any(-1).mean() > 0.2


# ---- the C++ BlueprintReader (include/dray/drx_dray.hpp, host code) ---------------------------------------------------
@pytest.mark.parametrize("case", list(bc.CASES))
def test_cpp_reader_matches_python_reader(case, tmp_path):
    """BlueprintReader::load through the C++ classes gives the Python reader's domains bit for bit: domain ids,
    topology names, orders, element counts, float32 control points, field coefficients and ranges."""
    if not os.path.exists(BP_BIN):
        pytest.fail(f"{BP_BIN} is not built (run devil_ray_amd/build.py)")
    r = subprocess.run([BP_BIN, "load", bc.root(case), str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    doms = bp.load(bc.root(case))
    lines = (tmp_path / "summary.txt").read_text().splitlines()
    assert len(lines) == len(doms)
    for i, (d, line) in enumerate(zip(doms, lines)):
        parts = line.split()
        assert (int(parts[0]), parts[1], int(parts[2]), int(parts[3])) == (d.domain_id, d.topology, d.mesh.order,
                                                                          d.mesh.n_el)
        fields = {f.split(":")[0]: (float(f.split(":")[1]), float(f.split(":")[2])) for f in parts[4:]}
        assert set(fields) == set(d.fields)
        got = np.fromfile(tmp_path / f"d{i}_values.bin", np.float32).reshape(-1, 3)
        assert np.array_equal(got.view(np.uint32), d.mesh.values.view(np.uint32))
        for name in d.fields:
            gf = np.fromfile(tmp_path / f"d{i}_{name}.bin", np.float32)
            assert np.array_equal(gf.view(np.uint32), d.fields[name].view(np.uint32))
            assert np.allclose(fields[name], d.ranges[name], rtol=1e-5)


def test_cpp_reader_errors(tmp_path):
    if not os.path.exists(BP_BIN):
        pytest.fail(f"{BP_BIN} is not built (run devil_ray_amd/build.py)")
    bad = tmp_path / "x.root"
    bad.write_bytes(b"not hdf5")
    r = subprocess.run([BP_BIN, "load", str(bad), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "not an HDF5 file" in r.stderr
    r = subprocess.run([BP_BIN, "load", str(tmp_path / "missing.root"), str(tmp_path)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_readers_reject_corrupted_files(tmp_path):
    """Seeded byte corruption of dray's Laghos root and domain files (lengths, offsets and counts set to random
    or all-ones bytes): the C++ BlueprintReader either loads the data set or exits with an error message
    (status 1), never with a signal, and the Python reader either loads it or raises an exception."""
    import shutil
    if not os.path.exists(BP_BIN):
        pytest.fail(f"{BP_BIN} is not built (run devil_ray_amd/build.py)")
    src_root = bc.root("tg_mpi_volume")
    src_dir = src_root[: -len(".root")]
    root = tmp_path / os.path.basename(src_root)
    dom_dir = tmp_path / os.path.basename(src_dir)
    shutil.copytree(src_dir, dom_dir)
    originals = {root: open(src_root, "rb").read(), dom_dir / "domain_000000.hdf5":
                 open(os.path.join(src_dir, "domain_000000.hdf5"), "rb").read()}
    out = tmp_path / "out"
    out.mkdir()
    rng = np.random.default_rng(2024)
    outcomes = {0: 0, 1: 0}
    for trial in range(120):
        for path, data in originals.items():
            path.write_bytes(data)
        target = root if trial % 3 == 0 else dom_dir / "domain_000000.hdf5"
        buf = bytearray(originals[target])
        for _ in range(int(rng.integers(1, 7))):
            pos = int(rng.integers(0, min(len(buf), 8192) if rng.random() < 0.8 else len(buf)))
            buf[pos] = 0xFF if rng.random() < 0.3 else int(rng.integers(0, 256))
        target.write_bytes(bytes(buf))
        r = subprocess.run([BP_BIN, "load", str(root), str(out)], capture_output=True, text=True, errors="replace",
                           timeout=60)
        assert r.returncode in (0, 1), (trial, r.returncode, r.stderr[-300:])
        if r.returncode == 1:
            assert r.stderr.startswith("error:"), (trial, r.stderr[-300:])
        outcomes[r.returncode] += 1
        try:
            with np.errstate(over="ignore"):   # corrupted coordinates may overflow float32: data, not a fault
                bp.load(str(root))
        except Exception:
            pass
    print(f"[corrupted files] loaded {outcomes[0]}, rejected {outcomes[1]} of 120")
    assert outcomes[1] > 0

"""GPU parity at the benchmarked sizes, against the CPU oracle.

Each workload is built exactly as the benchmark builds it; the checks the oracle can afford run in full
(split and LBVH at 8.65 M leaves and at more than 2^24 leaves), the rest on strided or sampled subsets:

* order-3 64^3 mesh (8,650,752 leaves), 1024^2 perspective frame, 512 samples/ray: the split and the
  BVH bit-exact; >= 2,000 evenly strided active rays integrated on the GPU (as part of the whole frame and
  on their own) against the FP32 oracle: equal partial counts, pixel ids and located-sample totals, RGBA
  within RGBA_TOL; 20,000 uniform points located: ids equal to the oracle's except points whose FP64
  reference point lies within 1e-6 plus the FP32 solution's own error of the acceptance band edge
  (parity_util.explain_mismatches), reference coordinates within 1e-5 of the FP64 inverse map;
* 48^3 mixed-order mesh, the furnace driver's query points (first 2^20 of the 2^26): ids and reference
  coordinates as above with the FP64 polish on;
* order-4 128^3 mesh (69,206,016 leaves) with warp 0.25, FP64 polish on: the split of 4,096 sampled
  elements bit-exact, 20,000 located points and 2,000 strided rays of the 2048^2 frame as above.
"""
import numpy as np
import pytest
import torch

from devil_ray_amd import synthetic as syn
from devil_ray_amd.synthetic import LOCATION_DTYPE, PARTIAL_DTYPE
from parity_util import explain_mismatches

pytestmark = pytest.mark.gpu

REF_TOL = 1e-5
RGBA_TOL = 1e-4


def _np(t, dtype=None):
    a = t.detach().cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def _oracle_bvh(orc, bvh):
    """The oracle's view of the GPU-built BVH arrays (their bit-exactness is checked separately)."""
    ob = orc.OracleBVH()
    ob.inner = _np(bvh.inner_nodes).reshape(-1, 16)
    ob.leaf_nodes = _np(bvh.leaf_nodes)
    ob.aabb_ids = _np(bvh.aabb_ids)
    ob.subrefs = _np(bvh.subrefs).reshape(-1, 6)
    ob.bounds = np.asarray(bvh.bounds, np.float32)
    return ob


def _volume(gpu, m, f, samples):
    """The benchmark's volume: unlit, grey -> red transfer function, corrected opacity 0.02 per sample."""
    vol = gpu.Volume(m, f)
    vol.use_lighting = False
    vol.samples(samples)
    vol.table = gpu.ColorTable()
    vol.table.add_point(0.0, (0.5, 0.5, 0.5))
    vol.table.add_point(1.0, (1.0, 0.0, 0.0))
    vol.table.add_alpha(0.0, syn.table_alpha_for_corrected(0.02, samples))
    vol.table.set_smooth(True)
    return vol


def _camera(gpu, bvh, res):
    c = bvh.bounds
    centre = (0.5 * (c[0] + c[1]), 0.5 * (c[2] + c[3]), 0.5 * (c[4] + c[5]))
    return gpu.Camera(res, res, 45.0, pos=(centre[0], centre[1], centre[2] + 2.5), look_at=centre)


def _check_locate(gpu, orc, mesh, m, ob, pts, opts, label):
    """GPU locate against the FP32 oracle over the same BVH: ids equal except at the band edge (1e-6 plus
    the FP32 solution's error); every GPU-located reference point within REF_TOL of the FP64 inverse map."""
    loc = _np(m.locate(gpu.to_device(pts.reshape(-1)), opts), LOCATION_DTYPE)
    gc, gr = loc["cell"], loc["ref"]
    oc, orf = orc.locate(mesh, ob, pts)
    mism, explained, edge, by64 = explain_mismatches(orc, mesh, ob, pts, gc, gr, oc, orf)
    hit = np.nonzero(gc >= 0)[0]
    ex = orc.refine_f64(mesh, pts[hit], gc[hit], gr[hit].astype(np.float64))
    err = np.abs(gr[hit].astype(np.float64) - ex).max(axis=1)
    print(f"  [{label}] {len(pts)} points, located {len(hit)}, id mismatches {len(mism)} (explained "
          f"{int(explained.sum())}, {int(by64.sum())} of them where the GPU agrees with the FP64 oracle), "
          f"max ref err {err.max():.3g} mean {err.mean():.3g}", flush=True)
    assert explained.all(), f"unexplained located-id mismatches at {mism[~explained][:10]} (excess over the allowance {edge[~explained][:10]})"
    assert err.max() <= REF_TOL, f"max reference-coordinate error {err.max()}"
    assert len(hit) > len(pts) // 3
    return gc


def _per_ray(parts, pix):
    """Partials grouped per ray (pixel id): (count per ray in the order of pix, index of each ray's first)."""
    order = np.argsort(pix)
    pos = np.searchsorted(pix[order], parts["pixel_id"])
    ray = order[np.minimum(pos, len(pix) - 1)]
    assert np.all(pix[ray] == parts["pixel_id"]), "partial of an unknown ray"
    return np.bincount(ray, minlength=len(pix)), ray


def _check_rays(gpu, orc, mesh, vol, f, ob, sub, res, samples, full=None):
    """GPU integrate of the rays `sub` against the oracle, ray by ray, and, if given, against the same rays'
    partials inside the whole frame `full` (bit for bit).  Every ray's partials (count, pixel, depth,
    RGBA) match the FP32 oracle's, except rays where a sample's location differs between the two FP32
    computations: those (at most 0.5% of the rays) must match the oracle that locates the same FP32
    sample points with the FP64 walk and Newton (exact arithmetic decides).  Located-sample totals are
    accounted the same way."""
    parts, cnt = vol.integrate(gpu.as_bytes(sub), active=(gpu.as_bytes(sub), len(sub)))
    st = vol.last_stats
    gp = _np(parts, np.uint8)[: cnt * 24].view(PARTIAL_DTYPE).copy()
    if full is not None:
        sel = np.isin(full["pixel_id"], sub["pixel_id"])
        assert np.array_equal(full[sel].view(np.uint8), gp.view(np.uint8)), "subset differs from the whole frame"
    lut = vol.table.sample(1024, opacity_factor=10.0 / samples)
    lo, hi = f.range()
    sd = orc.sample_dist(ob.bounds, samples)
    op_raw, oloc, ost = orc.integrate(mesh, ob, sub, sd, lut, lo, hi)
    op = orc.compact_partials(op_raw)
    pix = sub["pixel_id"]
    gn, _ = _per_ray(gp, pix)
    on, _ = _per_ray(op, pix)
    differ = np.nonzero(gn != on)[0]
    located_expected = int(oloc.sum())
    if len(differ):
        rays64 = sub[differ].copy()
        p64, loc64, _ = orc.integrate(mesh, ob, rays64, sd, lut, lo, hi, locate64=True)
        p64 = orc.compact_partials(p64)
        n64, _ = _per_ray(p64, rays64["pixel_id"])
        assert np.array_equal(gn[differ], n64), f"rays {differ[:8]}: partial counts differ from the FP64-located oracle"
        located_expected += int(loc64.sum()) - int(oloc[differ].sum())
        # replace those rays' partials by the FP64-located oracle's
        keep = ~np.isin(op["pixel_id"], rays64["pixel_id"])
        op = np.concatenate([op[keep], p64])
        _, ray_of = _per_ray(op, pix)
        op = op[np.argsort(ray_of, kind="stable")]   # ray order, each ray's segments in order
    print(f"  [rays] {len(sub)} rays: partials gpu {len(gp)} oracle {len(op)}; located samples gpu "
          f"{st.located_samples} oracle {located_expected}; {len(differ)} rays decided by the FP64-located oracle",
          flush=True)
    assert len(differ) <= max(2, len(sub) // 200)
    assert len(gp) == len(op), "partial counts differ"
    assert np.array_equal(gp["pixel_id"], op["pixel_id"])
    assert st.located_samples == located_expected, "located-sample totals differ"
    dd = np.abs(gp["depth"] - op["depth"]).max()
    d = np.abs(gp["rgba"] - op["rgba"]).max()
    print(f"  [rays] partial depth max diff {dd:.3g}, rgba max diff {d:.3g}", flush=True)
    assert dd <= 1e-5 * max(1.0, float(np.abs(op["depth"]).max()))
    assert d <= RGBA_TOL, f"partial RGBA differs by {d}"
    gimg = orc.composite(gp, res, res)
    oimg = orc.composite(op, res, res)
    di = np.abs(gimg[pix] - oimg[pix]).max()
    assert di <= RGBA_TOL, f"composited RGBA differs by {di}"
    return gp


# ---------------------------------------------------------------------------------------------------
# config 1: order-3 64^3, 1024^2, 512 samples/ray
# ---------------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cfg1(gpu):
    print("[config 1] generating the 64^3 order-3 mesh", flush=True)
    mesh = syn.hex_mesh(64, 3, 0.1, 0.01, 1234)
    m = gpu.UnstructuredMesh(mesh)
    bvh = m.get_bvh()
    torch.cuda.synchronize()
    return mesh, m, bvh


@pytest.mark.timeout(300)
def test_config1_split_and_bvh_bit_exact(gpu, orc, cfg1):
    mesh, m, bvh = cfg1
    print("[config 1] split + LBVH, 8,650,752 leaves", flush=True)
    aabbs, prims, subrefs = m.split_elements()
    oa, op, osr = orc.split_elements(mesh)
    assert np.array_equal(_np(aabbs).reshape(-1, 6).view(np.uint32), oa.view(np.uint32)), "sub-box AABBs differ"
    assert np.array_equal(_np(prims), op)
    assert np.array_equal(_np(subrefs).reshape(-1, 6).view(np.uint32), osr.view(np.uint32)), "subrefs differ"
    ob = orc.bvh_build(oa, op)
    gb = gpu.LinearBVHBuilder().construct(aabbs, prims, want_mcodes=True)
    torch.cuda.synchronize()
    assert np.array_equal(np.asarray(gb.bounds, np.float32).view(np.uint32), ob.bounds.view(np.uint32))
    assert np.array_equal(_np(gb.mcodes).view(np.uint32), ob.mcodes_sorted)
    assert np.array_equal(_np(gb.aabb_ids), ob.aabb_ids)
    assert np.array_equal(_np(gb.leaf_nodes), ob.leaf_nodes)
    gi = _np(gb.inner_nodes).reshape(-1, 16)
    assert np.array_equal(gi[:, :14].view(np.uint32), ob.inner[:, :14].view(np.uint32)), "inner nodes differ"
    # the mesh's cached BVH (the one every query walks) is the same tree
    assert np.array_equal(_np(bvh.inner_nodes).reshape(-1, 16)[:, :14].view(np.uint32), ob.inner[:, :14].view(np.uint32))
    assert np.array_equal(_np(bvh.leaf_nodes), ob.leaf_nodes) and np.array_equal(_np(bvh.aabb_ids), ob.aabb_ids)


@pytest.mark.timeout(300)
def test_config1_locate_sampled(gpu, orc, cfg1):
    mesh, m, bvh = cfg1
    ob = _oracle_bvh(orc, bvh)
    pts = syn.query_points(20000, bvh.bounds, seed=99)
    _check_locate(gpu, orc, mesh, m, ob, pts, gpu.LocateOptions(), "config 1 locate")


@pytest.mark.timeout(300)
def test_config1_frame_strided_rays(gpu, orc, cfg1):
    mesh, m, bvh = cfg1
    res, samples = 1024, 512
    f = gpu.UnstructuredField(mesh)
    vol = _volume(gpu, m, f, samples)
    cam = _camera(gpu, bvh, res)
    rays = cam.create_rays()
    parts, cnt = vol.integrate(rays, tile_width=res, tile_height=res)   # the benchmark's frame
    full = _np(parts, np.uint8)[: cnt * 24].view(PARTIAL_DTYPE)
    act, n = gpu.remove_missed_rays(rays, bvh.bounds)
    from devil_ray_amd.synthetic import RAY_DTYPE
    act = _np(act, np.uint8)[: n * 36].view(RAY_DTYPE)
    sub = act[:: max(1, n // 2400)].copy()
    print(f"[config 1] frame: {n} active rays, {cnt} partials, {vol.last_stats.located_samples} located; "
          f"checking {len(sub)} strided rays", flush=True)
    assert len(sub) >= 2000
    _check_rays(gpu, orc, mesh, vol, f, _oracle_bvh(orc, bvh), sub, res, samples, full=full)


# ---------------------------------------------------------------------------------------------------
# LBVH beyond 2^24 leaves: dray's float ceil(l / 2^k) in the split search is inexact there
# ---------------------------------------------------------------------------------------------------
@pytest.mark.timeout(300)
def test_bvh_bit_exact_beyond_2p24_leaves(gpu, orc):
    n = (1 << 24) + (1 << 20)
    print(f"[bvh] {n} random boxes", flush=True)
    rng = np.random.default_rng(2024)
    lo = rng.random((n, 3), dtype=np.float32)
    hi = lo + rng.random((n, 3), dtype=np.float32) * np.float32(1e-3)
    a = np.empty((n, 6), np.float32)
    a[:, 0::2], a[:, 1::2] = lo, hi
    a[n - 4096:] = a[n - 4096]          # a run of identical boxes high up: equal Morton codes, index tie-break
    a[(1 << 24) + 7] = a[3]
    prims = rng.permutation(n).astype(np.int32)
    ob = orc.bvh_build(a, prims)
    gb = gpu.LinearBVHBuilder().construct(gpu.to_device(a.reshape(-1)), gpu.to_device(prims), want_mcodes=True)
    torch.cuda.synchronize()
    assert np.array_equal(_np(gb.mcodes).view(np.uint32), ob.mcodes_sorted)
    assert np.array_equal(_np(gb.aabb_ids), ob.aabb_ids)
    assert np.array_equal(_np(gb.leaf_nodes), ob.leaf_nodes)
    gi = _np(gb.inner_nodes).reshape(-1, 16)
    assert np.array_equal(gi[:, :14].view(np.uint32), ob.inner[:, :14].view(np.uint32)), "inner nodes differ"


# ---------------------------------------------------------------------------------------------------
# config 2: 48^3 mixed order, furnace query points, FP64 polish
# ---------------------------------------------------------------------------------------------------
@pytest.mark.timeout(300)
def test_config2_subset_polish(gpu, orc):
    print("[config 2] 48^3 mixed-order mesh, first 2^20 furnace points", flush=True)
    mesh = syn.mixed_hex_mesh(48, 0.15, 0.01, 1234)
    m = gpu.UnstructuredMesh(mesh)
    bvh = m.get_bvh()
    pts = syn.furnace_points(1 << 20, bvh.bounds)
    ob = _oracle_bvh(orc, bvh)
    opts = gpu.LocateOptions(fp64_polish=True, cond_threshold=1e4)
    dp = gpu.to_device(pts.reshape(-1))
    loc = _np(m.locate(dp, opts), LOCATION_DTYPE)
    gc, gr = loc["cell"], loc["ref"]
    oc, orf = orc.locate(mesh, ob, pts)
    mism, explained, edge, by64 = explain_mismatches(orc, mesh, ob, pts, gc, gr, oc, orf)
    print(f"  {len(pts)} points, located {(gc >= 0).sum()}, id mismatches {len(mism)} (explained "
          f"{int(explained.sum())}, {int(by64.sum())} of them where the GPU agrees with the FP64 oracle)", flush=True)
    assert explained.all(), f"unexplained located-id mismatches at {mism[~explained][:10]}"
    # the benchmark's fastest variants return the same records
    for o in (gpu.LocateOptions(fp64_polish=True, traversal=3, coherent_order=True),
              gpu.LocateOptions(fp64_polish=True, traversal=1, coherent_order=True, launch=2)):
        assert np.array_equal(_np(m.locate(dp, o)).view(np.uint8), loc.view(np.uint8))
    idx = np.nonzero(gc >= 0)[0][:: 50]
    ex = orc.refine_f64(mesh, pts[idx], gc[idx], gr[idx].astype(np.float64))
    err = np.abs(gr[idx].astype(np.float64) - ex).max()
    print(f"  sampled {len(idx)} located points: max ref err {err:.3g}", flush=True)
    assert err <= REF_TOL


# ---------------------------------------------------------------------------------------------------
# config 4: order-4 128^3, warp 0.25, 2048^2, 1000 samples/ray, FP64 polish
# ---------------------------------------------------------------------------------------------------
@pytest.mark.timeout(600)
def test_config4_sampled_polish(gpu, orc):
    print("[config 4] generating the 128^3 order-4 mesh", flush=True)
    mesh = syn.hex_mesh(128, 4, 0.25, 0.01, 1234)
    m = gpu.UnstructuredMesh(mesh)
    print("[config 4] split of 4,096 sampled elements", flush=True)
    aabbs, prims, subrefs = m.split_elements()
    rng = np.random.default_rng(4)
    els = np.sort(rng.choice(mesh.n_el, 4096, replace=False))
    npe = 125
    sub = syn.HexMesh(n=0, order=4, ctrl_idx=mesh.ctrl_idx.reshape(-1, npe)[els].reshape(-1).copy(),
                      values=mesh.values, field=mesh.field)
    oa, _, osr = orc.split_elements(sub)
    rows = (els[:, None] * 33 + np.arange(33)[None, :]).reshape(-1)
    assert np.array_equal(_np(aabbs).reshape(-1, 6)[rows].view(np.uint32), oa.view(np.uint32))
    assert np.array_equal(_np(subrefs).reshape(-1, 6)[rows].view(np.uint32), osr.view(np.uint32))
    del aabbs, prims, subrefs
    print("[config 4] LBVH", flush=True)
    bvh = m.get_bvh()
    torch.cuda.synchronize()
    ob = _oracle_bvh(orc, bvh)
    opts = gpu.LocateOptions(fp64_polish=True, cond_threshold=1e4)
    pts = syn.query_points(20000, bvh.bounds, seed=99)
    _check_locate(gpu, orc, mesh, m, ob, pts, opts, "config 4 locate, polish")
    res, samples = 2048, 1000
    f = gpu.UnstructuredField(mesh)
    vol = _volume(gpu, m, f, samples)
    vol.opts = opts
    rays = _camera(gpu, bvh, res).create_rays()
    act, n = gpu.remove_missed_rays(rays, bvh.bounds)
    from devil_ray_amd.synthetic import RAY_DTYPE
    act = _np(act, np.uint8)[: n * 36].view(RAY_DTYPE)
    sub_rays = act[:: max(1, n // 2000)].copy()
    print(f"[config 4] {n} active rays; checking {len(sub_rays)} strided rays", flush=True)
    _check_rays(gpu, orc, mesh, vol, f, ob, sub_rays, res, samples)


# ---------------------------------------------------------------------------------------------------
# config 0: order-2 8^3, FP64 field, orthographic 256^2, 200 samples/ray (the whole frame)
# ---------------------------------------------------------------------------------------------------
@pytest.mark.timeout(300)
def test_config0_fp64_field_full_frame(gpu, orc):
    """The FP64 field (dray's Float = double field: double coefficients, evaluated and colour-mapped in
    double) over FP32 geometry, the whole 256^2 orthographic frame against the oracle's integrate with the
    same precision split: equal partial counts, pixel ids and located-sample totals, RGBA within RGBA_TOL;
    Field::eval of the FP64 field against the oracle's FP64 eval_d."""
    from devil_ray_amd.synthetic import RAY_DTYPE
    mesh = syn.hex_mesh(8, 2, 0.05, 0.005, 1234, jitter_scaled=False, keep_field64=True)
    m = gpu.UnstructuredMesh(mesh)
    bvh = m.get_bvh()
    f64 = gpu.UnstructuredField(mesh, fp64=True)
    lo, hi = f64.range()
    assert lo == float(mesh.field64.min()) and hi == float(mesh.field64.max())
    res, samples = 256, 200
    vol = _volume(gpu, m, f64, samples)
    rays = syn.orthographic_rays(res, res)
    parts, cnt = vol.integrate(gpu.as_bytes(rays), tile_width=res, tile_height=res)
    st = vol.last_stats
    gp = _np(parts, np.uint8)[: cnt * 24].view(PARTIAL_DTYPE).copy()
    ob = orc.build_mesh_bvh(mesh)
    assert np.array_equal(_np(bvh.inner_nodes).reshape(-1, 16)[:, :14].view(np.uint32), ob.inner[:, :14].view(np.uint32))
    act = orc.remove_missed_rays(rays, ob.bounds)
    lut = vol.table.sample(1024, opacity_factor=10.0 / samples)
    op, _, ost = orc.integrate(mesh, ob, act, orc.sample_dist(ob.bounds, samples), lut, lo, hi, field64=True)
    op = orc.compact_partials(op)
    d = np.abs(gp["rgba"] - op["rgba"]).max() if len(gp) == len(op) else np.inf
    print(f"[config 0] {len(act)} active rays: partials gpu {len(gp)} oracle {len(op)}, located gpu "
          f"{st.located_samples} oracle {int(ost[0])}, rgba max diff {d:.3g}", flush=True)
    assert len(act) == res * res
    assert len(gp) == len(op) and np.array_equal(gp["pixel_id"], op["pixel_id"])
    assert st.located_samples == int(ost[0])
    assert d <= RGBA_TOL
    di = np.abs(orc.composite(gp, res, res) - orc.composite(op, res, res)).max()
    assert di <= RGBA_TOL
    # the FP64 field is what ran: the FP32 field's frame differs from it, within the same bar
    f32 = gpu.UnstructuredField(mesh)
    vol32 = _volume(gpu, m, f32, samples)
    parts32, cnt32 = vol32.integrate(gpu.as_bytes(rays), tile_width=res, tile_height=res)
    g32 = _np(parts32, np.uint8)[: cnt32 * 24].view(PARTIAL_DTYPE)
    assert cnt32 == cnt and not np.array_equal(g32.view(np.uint8), gp.view(np.uint8))
    # Field::eval of the FP64 field
    pts = syn.query_points(5000, bvh.bounds, seed=3)
    loc = m.locate(gpu.to_device(pts.reshape(-1)))
    vals = _np(f64.eval(loc))
    l = _np(loc, LOCATION_DTYPE)
    hit = np.nonzero(l["cell"] >= 0)[0]
    ref = np.array([orc.eval_d(mesh.field64[mesh.elem_dofs(int(l["cell"][i]))].reshape(-1, 1), 2,
                               l["ref"][i].astype(np.float64))[0][0] for i in hit])
    assert np.abs(vals[hit] - ref.astype(np.float32)).max() <= 2 * np.spacing(np.float32(np.abs(ref).max()))


# ---------------------------------------------------------------------------------------------------
# config 2 as dray holds a mixed-order mesh: one domain per order in a Collection, PointLocation
# ---------------------------------------------------------------------------------------------------
@pytest.mark.timeout(300)
def test_config2_point_location_per_order_domains(gpu, orc):
    """The 48^3 mixed-order mesh as four uniform-order domains (dray's UnstructuredMesh has one order) in a
    Collection; PointLocation::execute over the first 2^20 furnace points against the oracle's restatement
    of point_location.cpp:178-287 over the same domains: per domain, located ids as the parity bar allows;
    every value equal to the oracle's to float rounding (1e-5 of the field's range) except points whose
    winning domain's id differs for an excused reason; empty_val where no domain locates."""
    mesh = syn.mixed_hex_mesh(48, 0.15, 0.01, 1234)
    doms = syn.split_by_order(mesh)
    col = gpu.Collection()
    gdoms, odoms = [], []
    for p, d, els in doms:
        m, f = gpu.UnstructuredMesh(d), gpu.UnstructuredField(d)
        ds = gpu.DataSet(m)
        ds.add_field("f", f)
        col.add_domain(ds)
        gdoms.append((d, m))
        odoms.append((d, _oracle_bvh(orc, m.get_bvh()), [d.field]))
    bb = col.bounds()
    pts = syn.furnace_points(1 << 20, bb)
    dp = gpu.to_device(pts.reshape(-1))
    pl = gpu.PointLocation()
    pl.add_var("f")
    pl.empty_val(-7.0)
    got = _np(pl.execute(col, dp)["values"][0])
    want, ocells = orc.point_location(odoms, pts, 1, -7.0)
    excused = np.zeros(len(pts), bool)
    for (d, m), (_, ob, _), oc in zip(gdoms, odoms, ocells):
        loc = _np(m.locate(dp), LOCATION_DTYPE)
        orf = orc.locate(d, ob, pts)[1]
        mism, explained, _, _ = explain_mismatches(orc, d, ob, pts, loc["cell"], loc["ref"], oc, orf)
        assert explained.all(), f"order-{d.order} domain: unexplained located-id mismatches at {mism[~explained][:8]}"
        excused[mism] = True
    span = float(mesh.field.max() - mesh.field.min())
    diff = np.abs(got - want[0])
    bad = (diff > 1e-5 * span) & ~excused
    print(f"[config 2 domains] {len(doms)} domains, {len(pts)} points: located {int((got != -7).sum())}, "
          f"excused {int(excused.sum())}, max |value - oracle| {diff[~excused].max():.3g}", flush=True)
    assert not bad.any(), f"values differ from the oracle at {np.nonzero(bad)[0][:8]}"
    assert np.array_equal(got == np.float32(-7.0), want[0] == np.float32(-7.0)) or excused.any()

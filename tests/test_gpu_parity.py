"""GPU parity: libdrx (through its C ABI) against the CPU oracle on identical seeded inputs.

Bars: BVH (boxes, Morton codes, permutation, emitted nodes) bit-exact; camera rays and ray set-up
bit-exact; located element ids identical except points the FP64 inverse map places within 1e-6 (plus the
FP32 solution's own error) of the [-1e-4, 1+1e-4] acceptance band edge of an element one side accepted
(parity_util.explain_mismatches); reference coordinates within 1e-5 of the
FP64-refined oracle; renders: equal partial counts and located-sample totals, RGBA within RGBA_TOL.
"""
import numpy as np
import pytest
import torch

from devil_ray_amd import synthetic as syn
from devil_ray_amd.synthetic import LOCATION_DTYPE, PARTIAL_DTYPE, RAY_DTYPE
from parity_util import explain_mismatches

pytestmark = pytest.mark.gpu

REF_TOL = 1e-5
RGBA_TOL = 1e-4      # per channel; FP32 GPU vs FP32 oracle differ by rounding only (observed <= 3e-5)
LIT_TOL = 2e-4       # lit renders: the shading gradient goes through J^-1, which amplifies rounding


def _np(t, dtype=None):
    a = t.detach().cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def _gpu_bvh(gpu, mesh):
    m = gpu.UnstructuredMesh(mesh)
    bvh = m.get_bvh()
    torch.cuda.synchronize()
    return m, bvh


MESHES = [
    ("p1_n3", lambda: syn.hex_mesh(3, 1, 0.05, 0.1, 11)),
    ("p2_n4", lambda: syn.hex_mesh(4, 2, 0.05, 0.005, 1, jitter_scaled=False)),
    ("p3_n5", lambda: syn.hex_mesh(5, 3, 0.1, 0.01, 2)),
    ("p4_n3", lambda: syn.hex_mesh(3, 4, 0.25, 0.01, 3)),
    ("mixed_n4", lambda: syn.mixed_hex_mesh(4, 0.15, 0.05, 5)),
]


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
def test_split_and_bvh_bit_exact(gpu, orc, name, make):
    mesh = make()
    print(f"[split/bvh] {name}", flush=True)
    m = gpu.UnstructuredMesh(mesh)
    aabbs, prims, subrefs = m.split_elements()
    oa, op, osr = orc.split_elements(mesh)
    ga, gp, gs = _np(aabbs).reshape(-1, 6), _np(prims), _np(subrefs).reshape(-1, 6)
    assert np.array_equal(ga.view(np.uint32), oa.view(np.uint32)), "sub-box AABBs differ"
    assert np.array_equal(gp, op)
    assert np.array_equal(gs.view(np.uint32), osr.view(np.uint32)), "subrefs differ"
    ob = orc.bvh_build(oa, op)
    gb = gpu.LinearBVHBuilder().construct(aabbs, prims, want_mcodes=True)
    torch.cuda.synchronize()
    assert np.array_equal(np.asarray(gb.bounds, np.float32).view(np.uint32), ob.bounds.view(np.uint32))
    assert np.array_equal(_np(gb.mcodes).view(np.uint32), ob.mcodes_sorted)
    assert np.array_equal(_np(gb.aabb_ids), ob.aabb_ids)
    assert np.array_equal(_np(gb.leaf_nodes), ob.leaf_nodes)
    gi = _np(gb.inner_nodes).reshape(-1, 16)
    assert np.array_equal(gi[:, :14].view(np.uint32), ob.inner[:, :14].view(np.uint32)), "inner nodes differ"


@pytest.mark.parametrize("n", [0, 1, 2, 3, 7, 64, 1000, 5000])
def test_bvh_small_and_degenerate(gpu, orc, n):
    rng = np.random.default_rng(n + 100)
    lo = rng.random((n, 3), dtype=np.float32)
    hi = lo + rng.random((n, 3), dtype=np.float32) * 0.1
    a = np.zeros((n, 6), np.float32)
    a[:, 0::2], a[:, 1::2] = lo, hi
    if n >= 7:
        a[3] = a[1]            # duplicate boxes -> equal Morton codes (tie order must be stable)
        a[5] = a[1]
    prims = np.arange(n, dtype=np.int32)[::-1].copy()
    ob = orc.bvh_build(a, prims)
    gb = gpu.LinearBVHBuilder().construct(gpu.to_device(a.reshape(-1)) if n else torch.empty(0, device="cuda"),
                                          gpu.to_device(prims) if n else None, want_mcodes=True)
    torch.cuda.synchronize()
    assert np.array_equal(_np(gb.aabb_ids), ob.aabb_ids)
    assert np.array_equal(_np(gb.leaf_nodes), ob.leaf_nodes)
    assert np.array_equal(_np(gb.mcodes).view(np.uint32), ob.mcodes_sorted)
    gi = _np(gb.inner_nodes).reshape(-1, 16)
    assert np.array_equal(gi[:, :14].view(np.uint32), ob.inner[:, :14].view(np.uint32))


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
def test_locate_parity(gpu, orc, name, make):
    mesh = make()
    print(f"[locate] {name}", flush=True)
    m, bvh = _gpu_bvh(gpu, mesh)
    ob = orc.build_mesh_bvh(mesh)
    pts = syn.query_points(40000, bvh.bounds, seed=7)
    oc, orf, ost, _ = orc.locate(mesh, ob, pts, want_steps=True)
    for polish in (False, True):
        loc = _np(m.locate(gpu.to_device(pts.reshape(-1)), gpu.LocateOptions(fp64_polish=polish)), LOCATION_DTYPE)
        gc, gr = loc["cell"], loc["ref"]
        mism, explained, edge, by64 = explain_mismatches(orc, mesh, ob, pts, gc, gr, oc, orf)
        assert explained.all(), f"unexplained located-id mismatches: {mism[~explained][:10]} excess {edge[~explained][:10]}"
        assert len(mism) <= max(3, 1e-3 * len(pts)), f"{len(mism)} located-id mismatches"
        both = np.nonzero((gc == oc) & (gc >= 0))[0]
        ex = orc.refine_f64(mesh, pts[both], gc[both], orf[both].astype(np.float64))
        err = np.abs(gr[both].astype(np.float64) - ex).max()
        print(f"  polish={polish} located={len(both)} mismatches={len(mism)} max ref err={err:.3g}", flush=True)
        assert err <= REF_TOL, f"max reference-coordinate error {err}"


def test_field_eval(gpu, orc):
    mesh = syn.hex_mesh(4, 3, 0.1, 0.01, 2)
    m, bvh = _gpu_bvh(gpu, mesh)
    f = gpu.UnstructuredField(mesh)
    pts = syn.query_points(5000, bvh.bounds, seed=3)
    loc = m.locate(gpu.to_device(pts.reshape(-1)))
    vals = _np(f.eval(loc))
    l = _np(loc, LOCATION_DTYPE)
    ok = l["cell"] >= 0
    ctrl, fv = mesh.ctrl_idx, mesh.field.astype(np.float64)
    ref = np.zeros(len(pts))
    for i in np.nonzero(ok)[0]:
        d = fv[mesh.elem_dofs(int(l["cell"][i]))].reshape(-1, 1)
        ref[i] = orc.eval_d(d, 3, l["ref"][i].astype(np.float64))[0][0]
    assert np.abs(vals[ok] - ref[ok]).max() < 1e-5


def test_camera_and_ray_setup_bit_exact(gpu, orc):
    cam = gpu.Camera(96, 80, 45.0, pos=(0.5, 0.5, 3.0), look_at=(0.5, 0.5, 0.5))
    g = _np(cam.create_rays(), RAY_DTYPE)
    o = orc.camera_perspective(96, 80, 45.0, (0.5, 0.5, 3.0), (0.5, 0.5, 0.5))
    assert np.array_equal(g.view(np.uint8), o.view(np.uint8))
    bounds = np.array([0, 1, 0, 1, 0, 1.0000001], np.float32)
    act, n = gpu.remove_missed_rays(gpu.as_bytes(o), bounds)
    oa = orc.remove_missed_rays(o, bounds)
    assert n == len(oa) and n > 0
    assert np.array_equal(_np(act)[: n * 36], oa.view(np.uint8))


def _render_both(gpu, orc, mesh, rays_np, w, h, samples, opacity=0.02, tile=True, lights=None):
    m, bvh = _gpu_bvh(gpu, mesh)
    f = gpu.UnstructuredField(mesh)
    vol = gpu.Volume(m, f)
    vol.use_lighting = lights is not None
    vol.samples(samples)
    vol.table = gpu.ColorTable()
    vol.table.add_point(0.0, (0.5, 0.5, 0.5))
    vol.table.add_point(1.0, (1.0, 0.0, 0.0))
    vol.table.add_alpha(0.0, syn.table_alpha_for_corrected(opacity, samples))
    vol.table.set_smooth(True)
    parts, n = vol.integrate(gpu.as_bytes(rays_np), tile_width=w if tile else 0, tile_height=h if tile else 0,
                             lights=lights)
    img = _np(gpu.composite(parts, n, w, h)).reshape(-1, 4)
    gp = _np(parts, np.uint8)[: n * 24].view(PARTIAL_DTYPE)
    # oracle
    ob = orc.build_mesh_bvh(mesh)
    lut = vol.table.sample(1024, opacity_factor=10.0 / samples)
    lo, hi = f.range()
    act = orc.remove_missed_rays(rays_np, ob.bounds)
    op, _, stats = orc.integrate(mesh, ob, act, orc.sample_dist(ob.bounds, samples), lut, lo, hi, lights=lights)
    op = orc.compact_partials(op)
    oimg = orc.composite(op, w, h)
    return gp, img, op, oimg, vol.last_stats, stats


@pytest.mark.parametrize("name,make", [MESHES[2], MESHES[4]], ids=["p3_n5", "mixed_n4"])
def test_locate_centre_start_parity(gpu, orc, name, make):
    """DRX_INIT_SUBREF_CENTRE (dray's subref_center start) against the oracle in the same mode: the same
    located-id and reference-coordinate bars as the default start, Newton step totals within 1%, and more
    steps than the corner-affine start."""
    mesh = make()
    print(f"[locate centre start] {name}", flush=True)
    m, bvh = _gpu_bvh(gpu, mesh)
    ob = orc.build_mesh_bvh(mesh)
    pts = syn.query_points(40000, bvh.bounds, seed=9)
    _, _, ost_affine, _ = orc.locate(mesh, ob, pts, want_steps=True)
    try:
        orc.set_newton_start("centre")
        oc, orf, ost, _ = orc.locate(mesh, ob, pts, want_steps=True)
    finally:
        orc.set_newton_start("affine")
    loc, steps, _ = m.locate(gpu.to_device(pts.reshape(-1)), gpu.LocateOptions(start="centre"), stats=True)
    loc = _np(loc, LOCATION_DTYPE)
    gc, gr, gst = loc["cell"], loc["ref"], _np(steps)
    mism, explained, edge, by64 = explain_mismatches(orc, mesh, ob, pts, gc, gr, oc, orf)
    assert explained.all(), f"unexplained located-id mismatches: {mism[~explained][:10]} excess {edge[~explained][:10]}"
    assert len(mism) <= max(3, 1e-3 * len(pts)), f"{len(mism)} located-id mismatches"
    both = np.nonzero((gc == oc) & (gc >= 0))[0]
    ex = orc.refine_f64(mesh, pts[both], gc[both], orf[both].astype(np.float64))
    err = np.abs(gr[both].astype(np.float64) - ex).max()
    print(f"  located={len(both)} mismatches={len(mism)} max ref err={err:.3g} steps gpu={gst.sum()} "
          f"oracle={ost.sum()} oracle affine={ost_affine.sum()}", flush=True)
    assert err <= REF_TOL
    assert abs(int(gst.sum()) - int(ost.sum())) <= 0.01 * ost.sum()
    assert ost.sum() > 1.1 * ost_affine.sum()


def test_fp64_polish_fires(gpu, orc):
    """With cond_threshold below 1 every accepted point takes the two FP64 Newton corrections: located
    cells are unchanged, and reference points match the FP64 oracle (iterated to 1e-14 on the same float32
    control points) to float32 output rounding, closer than the FP32 solution.  In a render every located
    sample is polished."""
    mesh = syn.hex_mesh(3, 4, 0.25, 0.01, 3)
    print("[fp64 polish] p4_n3 warp 0.25", flush=True)
    m, bvh = _gpu_bvh(gpu, mesh)
    pts = syn.query_points(40000, bvh.bounds, seed=13)
    d = gpu.to_device(pts.reshape(-1))
    l32 = _np(m.locate(d, gpu.LocateOptions()), LOCATION_DTYPE)
    l64 = _np(m.locate(d, gpu.LocateOptions(fp64_polish=True, cond_threshold=0.5)), LOCATION_DTYPE)
    assert np.array_equal(l32["cell"], l64["cell"])
    idx = np.nonzero(l32["cell"] >= 0)[0]
    ex = orc.refine_f64(mesh, pts[idx], l32["cell"][idx], l32["ref"][idx].astype(np.float64))
    e32 = np.abs(l32["ref"][idx].astype(np.float64) - ex).max(axis=1)
    e64 = np.abs(l64["ref"][idx].astype(np.float64) - ex).max(axis=1)
    print(f"  located={len(idx)} fp32 max {e32.max():.3g} mean {e32.mean():.3g}; "
          f"polished max {e64.max():.3g} mean {e64.mean():.3g}", flush=True)
    assert e64.max() < 1e-7
    assert e64.mean() < e32.mean() and e32.max() > 2 * e64.max()
    f = gpu.UnstructuredField(mesh)
    vol = gpu.Volume(m, f)
    vol.use_lighting = False
    vol.samples(64)
    vol.table = gpu.ColorTable()
    vol.table.add_point(0.0, (0.5, 0.5, 0.5))
    vol.table.add_point(1.0, (1.0, 0.0, 0.0))
    vol.table.add_alpha(0.0, 0.05)
    vol.opts = gpu.LocateOptions(fp64_polish=True, cond_threshold=0.5)
    rays = orc.camera_perspective(32, 32, 45.0, (0.5, 0.5, 2.5), (0.5, 0.5, 0.5))
    vol.integrate(gpu.as_bytes(rays), tile_width=32, tile_height=32)
    st = vol.last_stats
    print(f"  render: located {st.located_samples} polished {st.polished}", flush=True)
    assert st.located_samples > 0 and st.polished == st.located_samples


def _gate_mesh():
    """Three order-2 elements far apart: a slab 1e-5 thick (cond(J) ~ 1e5 everywhere), a unit cube
    (cond 1) and a wedge whose thickness grows from 1e-5 to 1e-3 along x (cond from ~1e5 down to ~1e3,
    crossing 1e4 near x = 0.09), with heights h(u) = 1e-5 + u (1e-3 - 1e-5)."""
    p = 2
    slab = syn.box_net((0, 0, 0), (1, 1, 1e-5), p)
    cube = syn.box_net((3, 0, 0), (4, 1, 1), p)
    wedge = syn.box_net((6, 0, 0), (7, 1, 1), p)
    u = (wedge[:, 0] - 6.0)
    wedge[:, 2] = wedge[:, 2] * (1e-5 + u * (1e-3 - 1e-5))
    return syn.elements_from_nets(np.stack([slab, cube, wedge]), p)


def test_fp64_polish_gate(gpu, orc):
    """The FP64 correction runs exactly for accepted points whose Jacobian condition number exceeds the
    threshold (1e4): the per-point polished flags equal (cond_64 > 1e4) for the FP64 Jacobian's inf-norm
    condition number at the located point (numpy.linalg.cond), except points whose cond_64 lies within 2%
    of the threshold (the FP32 estimate's rounding); the unit cube polishes nothing, the slab everything."""
    mesh = _gate_mesh()
    m = gpu.UnstructuredMesh(mesh)
    m.get_bvh()
    rng = np.random.default_rng(17)
    k = 3000
    u = rng.random((3, k, 3))
    slab = np.stack([u[0, :, 0], u[0, :, 1], u[0, :, 2] * 1e-5], 1)
    cube = np.stack([3 + u[1, :, 0], u[1, :, 1], u[1, :, 2]], 1)
    h = 1e-5 + u[2, :, 0] * (1e-3 - 1e-5)
    wedge = np.stack([6 + u[2, :, 0], u[2, :, 1], u[2, :, 2] * h], 1)
    pts = np.concatenate([slab, cube, wedge]).astype(np.float32)
    loc, pol = m.locate(gpu.to_device(pts.reshape(-1)), gpu.LocateOptions(fp64_polish=True, cond_threshold=1e4),
                        polished=True)
    loc, pol = _np(loc, LOCATION_DTYPE), _np(pol)
    cells = loc["cell"]
    assert np.array_equal(cells, np.repeat(np.arange(3, dtype=np.int32), k)), "points left their elements"
    cond = orc.jacobian_cond_f64(mesh, cells, loc["ref"].astype(np.float64))
    want = cond > 1e4
    near = np.abs(cond / 1e4 - 1) < 0.02
    bad = (pol.astype(bool) != want) & ~near
    print(f"[polish gate] polished {pol.sum()} of {len(pts)}; cond>1e4 (FP64) {want.sum()}; near threshold "
          f"{near.sum()}; slab cond {cond[:k].min():.3g}..{cond[:k].max():.3g}, wedge {cond[2 * k:].min():.3g}.."
          f"{cond[2 * k:].max():.3g}", flush=True)
    assert not bad.any(), f"{bad.sum()} points polished against the FP64 condition number, e.g. {np.nonzero(bad)[0][:5]}"
    assert pol[k:2 * k].sum() == 0 and pol[:k].all()
    assert 0.03 * k < pol[2 * k:].sum() < 0.2 * k   # the wedge crosses the threshold
    # without the polish option nothing is flagged
    _, pol0 = m.locate(gpu.to_device(pts.reshape(-1)), gpu.LocateOptions(), polished=True)
    assert _np(pol0).sum() == 0


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
def test_compact_nodes_same_candidates(gpu, name, make):
    """drx_bvh_quantize: the binary walk over the 32-byte compact nodes (cell ranges, then each leaf's exact
    box) gives dray's candidates in dray's order: identical cells, reference points (bitwise), Newton
    steps and candidate counts per point to the walk over dray's 64-byte nodes."""
    mesh = make()
    print(f"[compact nodes] {name}", flush=True)
    m, bvh = _gpu_bvh(gpu, mesh)
    assert bvh.qnodes is not None
    pts = syn.query_points(60000, bvh.bounds, seed=21)
    d = gpu.to_device(pts.reshape(-1))
    out = {}
    for use in (False, True):
        bvh.use_qnodes = use
        loc, steps, cands = m.locate(d, gpu.LocateOptions(), stats=True)
        out[use] = (_np(loc, LOCATION_DTYPE), _np(steps), _np(cands))
    bvh.use_qnodes = True
    (a, sa, ca), (b, sb, cb) = out[False], out[True]
    assert np.array_equal(a["cell"], b["cell"])
    assert np.array_equal(a["ref"].view(np.uint32), b["ref"].view(np.uint32))
    assert np.array_equal(sa, sb) and np.array_equal(ca, cb)
    print(f"  located {(a['cell'] >= 0).sum()} candidates {ca.sum()}", flush=True)


def test_compact_nodes_layout(gpu, orc):
    """Compact node k: the cells of both child boxes of dray's node k under cell(v) = clamp(floor((v - origin)
    * cells_per_unit)), children as uint4 offsets (2 * inner index) or dray's leaf code; every leaf's exact
    box; and points drawn inside each child box fall inside its cell range."""
    mesh = syn.hex_mesh(4, 2, 0.05, 0.005, 1, jitter_scaled=False)
    m, bvh = _gpu_bvh(gpu, mesh)
    inner = _np(bvh.inner_nodes).reshape(-1, 16)
    q = _np(bvh.qnodes).view(np.uint32).reshape(-1, 8)
    boxes = _np(bvh.leaf_boxes).reshape(-1, 8)
    org, inv = bvh.qgrid[:3].astype(np.float32), bvh.qgrid[3:].astype(np.float32)

    def cell(v, dd):
        c = np.floor((np.float32(v) - org[dd]).astype(np.float32) * inv[dd]).astype(np.float32)
        return np.clip(c, 0, 65535).astype(np.uint32)

    lo = [inner[:, 0:3], inner[:, 6:9]]
    hi = [inner[:, 3:6], inner[:, 9:12]]
    ch = inner[:, 12:14].view(np.int32)
    for k in range(2):
        ql = [cell(lo[k][:, dd], dd) for dd in range(3)]
        qh = [cell(hi[k][:, dd], dd) for dd in range(3)]
        w = q[:, 3 * k:3 * k + 3]
        assert np.array_equal(w[:, 0], ql[0] | (ql[1] << 16))
        assert np.array_equal(w[:, 1], ql[2] | (qh[0] << 16))
        assert np.array_equal(w[:, 2], qh[1] | (qh[2] << 16))
        c = ch[:, k]
        assert np.array_equal(q[:, 6 + k].view(np.int32), np.where(c >= 0, c // 2, c))
        leaf = -c[c < 0] - 1
        ex = np.concatenate([lo[k][c < 0], hi[k][c < 0]], axis=1)
        assert np.array_equal(boxes[leaf][:, :6].view(np.uint32), ex.view(np.uint32))
        # containment: random points inside each child box (and its corners) land in its cell range
        rng = np.random.default_rng(k)
        t = rng.random((len(inner), 3)).astype(np.float32)
        for frac in (np.zeros_like(t), np.ones_like(t), t):
            p = (lo[k] + frac * (hi[k] - lo[k])).astype(np.float32)
            p = np.minimum(np.maximum(p, lo[k]), hi[k])
            for dd in range(3):
                cp = cell(p[:, dd], dd)
                assert np.all(ql[dd] <= cp) and np.all(cp <= qh[dd])
    assert len(np.unique(-ch[ch < 0] - 1)) == bvh.n_leaves


def test_compact_nodes_render_identical(gpu, orc):
    """The order-3 render (the stage-in-LDS kernel) over compact nodes is bit-identical to dray's nodes."""
    mesh = syn.hex_mesh(5, 3, 0.1, 0.01, 2)
    m, bvh = _gpu_bvh(gpu, mesh)
    f = gpu.UnstructuredField(mesh)
    vol = gpu.Volume(m, f)
    vol.use_lighting = False
    vol.samples(128)
    vol.table = gpu.ColorTable()
    vol.table.add_point(0.0, (0.5, 0.5, 0.5))
    vol.table.add_point(1.0, (1.0, 0.0, 0.0))
    vol.table.add_alpha(0.0, 0.05)
    vol.opts = gpu.LocateOptions(traversal=3)   # the binary walk, which reads the compact nodes
    rays = gpu.as_bytes(orc.camera_perspective(64, 64, 45.0, (0.5, 0.5, 2.5), (0.5, 0.5, 0.5)))
    res = {}
    for use in (False, True):
        bvh.use_qnodes = use
        parts, n = vol.integrate(rays, tile_width=64, tile_height=64)
        st = vol.last_stats
        res[use] = (_np(parts, np.uint8)[: n * 24].copy(), st.located_samples, st.newton_steps, st.candidates)
    bvh.use_qnodes = True
    assert np.array_equal(res[False][0], res[True][0])
    assert res[False][1:] == res[True][1:]


@pytest.mark.parametrize("name,make", [MESHES[2], MESHES[4]], ids=["p3_n5", "mixed_n4"])
def test_locate_launch_forms_identical(gpu, name, make):
    """drx_locate's streamed walks give identical locations, Newton steps and candidate counts for every
    chunk size and launch form (chunked; persistent with a chunk queue), for both walks and with Morton
    order, on a point count that is no multiple of the chunk."""
    mesh = make()
    print(f"[launch forms] {name}", flush=True)
    m, bvh = _gpu_bvh(gpu, mesh)
    pts = syn.query_points(40037, bvh.bounds, seed=31)
    d = gpu.to_device(pts.reshape(-1))
    for trav, coherent in ((0, False), (1, False), (0, True)):
        ref = None
        for launch, chunk in ((1, 1024), (1, 64), (1, 256), (1, 4096), (2, 64), (2, 1024), (2, 4096)):
            opts = gpu.LocateOptions(traversal=trav, coherent_order=coherent, launch=launch, chunk=chunk)
            loc, steps, cands = m.locate(d, opts, stats=True)
            got = (_np(loc), _np(steps), _np(cands))
            if ref is None:
                ref = got
            else:
                assert all(np.array_equal(a, b) for a, b in zip(ref, got)), (trav, coherent, launch, chunk)
    for n in (1, 63, 64, 65, 1000):   # grids of one wavefront, partial chunks, a partial last refill
        dn = gpu.to_device(pts[:n].reshape(-1))
        ref = None
        for launch, chunk in ((1, 1024), (1, 64), (2, 64), (2, 128), (2, 4096)):
            got = _np(m.locate(dn, gpu.LocateOptions(launch=launch, chunk=chunk)))
            if ref is None:
                ref = got
            else:
                assert np.array_equal(ref, got), (n, launch, chunk)
    for bad in (gpu.LocateOptions(launch=3), gpu.LocateOptions(chunk=32), gpu.LocateOptions(chunk=-1)):
        with pytest.raises(gpu.L.DrxError):
            m.locate(d, bad)


def test_locate_autotune(gpu):
    """drx_locate_autotune times the 10 candidates, records the fastest as the default for this order and
    walk, and locate with the default options then returns the same results as before tuning."""
    mesh = syn.mixed_hex_mesh(4, 0.15, 0.05, 5)
    m, bvh = _gpu_bvh(gpu, mesh)
    d = gpu.to_device(syn.query_points(200000, bvh.bounds, seed=41).reshape(-1))
    before = _np(m.locate(d))
    t = m.locate_autotune(d)
    print(f"[autotune] {t}", flush=True)
    assert t["launch"] in ("chunked", "persistent") and t["chunk"] in (256, 512, 1024, 2048, 4096)
    assert len(t["ms"]) == 10 and all(v > 0 for v in t["ms"].values())
    assert t["ms"][f"{t['launch']}/{t['chunk']}"] == min(t["ms"].values())
    assert np.array_equal(before, _np(m.locate(d)))
    with pytest.raises(gpu.L.DrxError):
        m.locate_autotune(d, gpu.LocateOptions(traversal=2))


@pytest.mark.parametrize("name,make", [MESHES[1], MESHES[2], MESHES[4]], ids=["p2_n4", "p3_n5", "mixed_n4"])
def test_render_walks_identical(gpu, orc, name, make):
    """Volume::integrate gives bit-identical partials and the same located-sample, Newton-step and
    candidate counts with every walk: automatic, binary over the compact nodes, 8-wide per lane,
    wave-coherent."""
    mesh = make()
    m, bvh = _gpu_bvh(gpu, mesh)
    f = gpu.UnstructuredField(mesh)
    vol = gpu.Volume(m, f)
    vol.use_lighting = False
    vol.samples(96)
    vol.table = gpu.ColorTable()
    vol.table.add_point(0.0, (0.5, 0.5, 0.5))
    vol.table.add_point(1.0, (1.0, 0.0, 0.0))
    vol.table.add_alpha(0.0, 0.05)
    rays = gpu.as_bytes(orc.camera_perspective(48, 48, 45.0, (0.5, 0.5, 2.5), (0.5, 0.5, 0.5)))
    res = {}
    for trav in (0, 3, 1, 2):
        vol.opts = gpu.LocateOptions(traversal=trav)
        parts, n = vol.integrate(rays, tile_width=48, tile_height=48)
        st = vol.last_stats
        res[trav] = (_np(parts, np.uint8)[: n * 24].copy(), st.located_samples, st.newton_steps, st.candidates)
    print(f"[render walks] {name}: located {res[3][1]}", flush=True)
    for trav in (0, 1, 2):
        assert np.array_equal(res[trav][0], res[3][0]), trav
        assert res[trav][1:] == res[3][1:], trav


@pytest.mark.parametrize("name,make", MESHES[:4], ids=[m[0] for m in MESHES[:4]])
def test_volume_integrate_parity(gpu, orc, name, make):
    mesh = make()
    print(f"[integrate] {name}", flush=True)
    w = h = 48
    rays = orc.camera_perspective(w, h, 45.0, (0.5, 0.5, 2.5), (0.5, 0.5, 0.5))
    gp, img, op, oimg, gst, ost = _render_both(gpu, orc, mesh, rays, w, h, samples=128)
    print(f"  partials gpu={len(gp)} oracle={len(op)} located gpu={gst.located_samples} oracle={ost[0]}", flush=True)
    assert len(gp) == len(op) and np.array_equal(gp["pixel_id"], op["pixel_id"])
    assert gst.located_samples == ost[0]
    d = np.abs(img - oimg)
    print(f"  rgba max diff {d.max():.3g} mean {d.mean():.3g}", flush=True)
    assert d.max() <= RGBA_TOL


@pytest.mark.parametrize("name,make", [MESHES[1], MESHES[2], MESHES[4]], ids=["p2_n4", "p3_n5", "mixed_n4"])
def test_volume_lit_parity(gpu, orc, name, make):
    """use_lighting (dray's default): VolumeShader::shaded_color with the Renderer's default light, and a
    second light, against the oracle's lit render."""
    mesh = make()
    w = h = 48
    pos, look = (0.5, 0.5, 2.5), (0.5, 0.5, 0.5)
    rays = orc.camera_perspective(w, h, 45.0, pos, look)
    L0 = orc.default_light(pos, look)
    assert np.allclose(gpu.default_light(gpu.Camera(w, h, 45.0, pos=pos, look_at=look)), L0)
    L1 = L0.copy()
    L1[:3] = (-1.0, 2.0, 0.5)
    for lights in (L0[None, :], np.stack([L0, L1])):
        gp, img, op, oimg, gst, ost = _render_both(gpu, orc, mesh, rays, w, h, samples=128, lights=lights)
        d = np.abs(img - oimg)
        print(f"  {name} lights={len(lights)} partials {len(gp)}/{len(op)} rgba max {d.max():.3g} mean {d.mean():.3g}",
              flush=True)
        assert len(gp) == len(op) and np.array_equal(gp["pixel_id"], op["pixel_id"])
        assert gst.located_samples == ost[0]
        assert d.max() <= LIT_TOL
        assert img[:, :3].max() > 0.05   # the lit image is not black


def test_multi_domain_render(gpu, orc):
    """Renderer::render over a two-domain collection: each domain integrated on its own (sample distance
    from the collection's bounds), partials composited across domains by depth."""
    a = syn.hex_mesh(4, 2, 0.05, 0.01, 7)
    b = syn.HexMesh(n=a.n, order=a.order, ctrl_idx=a.ctrl_idx,
                    values=(a.values + np.float32([0.6, 0.0, -0.7])).astype(np.float32), field=(a.field * 0.5).astype(np.float32))
    w = h = 40
    rays = orc.camera_perspective(w, h, 50.0, (0.8, 0.5, 2.8), (0.8, 0.5, 0.1))
    lo = float(min(a.field.min(), b.field.min()))
    hi = float(max(a.field.max(), b.field.max()))
    doms = []
    for mesh in (a, b):
        m = gpu.UnstructuredMesh(mesh)
        doms.append((mesh, m, gpu.UnstructuredField(mesh), orc.build_mesh_bvh(mesh)))
    cb = np.array([min(d[3].bounds[0] for d in doms), max(d[3].bounds[1] for d in doms),
                   min(d[3].bounds[2] for d in doms), max(d[3].bounds[3] for d in doms),
                   min(d[3].bounds[4] for d in doms), max(d[3].bounds[5] for d in doms)], np.float32)
    gparts, oparts = [], []
    for mesh, m, f, ob in doms:
        vol = gpu.Volume(m, f)
        vol.use_lighting = False
        vol.samples(96)
        vol.sample_bounds = cb
        vol.color_map_range(lo, hi)
        vol.table = gpu.ColorTable()
        vol.table.add_point(0.0, (0.1, 0.3, 0.9))
        vol.table.add_point(1.0, (0.9, 0.2, 0.1))
        vol.table.add_alpha(0.0, 0.3)
        vol.table.set_smooth(True)
        parts, n = vol.integrate(gpu.as_bytes(rays), tile_width=w, tile_height=h)
        gparts.append((parts, n))
        lut = vol.table.sample(1024, opacity_factor=10.0 / 96)
        act = orc.remove_missed_rays(rays, ob.bounds)
        op, _, _ = orc.integrate(mesh, ob, act, orc.sample_dist(cb, 96), lut, lo, hi)
        oparts.append(orc.compact_partials(op))
    img = _np(gpu.composite_domains(gparts, w, h)).reshape(-1, 4)
    torch.cuda.synchronize()
    allg = np.concatenate([_np(p, np.uint8)[: n * 24].view(PARTIAL_DTYPE) for p, n in gparts])
    assert np.array_equal(img.view(np.uint32), orc.composite_domains(allg, w, h).view(np.uint32))
    oimg = orc.composite_domains(np.concatenate(oparts), w, h)
    d = np.abs(img - oimg)
    both = np.intersect1d(_np(gparts[0][0], np.uint8)[: gparts[0][1] * 24].view(PARTIAL_DTYPE)["pixel_id"],
                          _np(gparts[1][0], np.uint8)[: gparts[1][1] * 24].view(PARTIAL_DTYPE)["pixel_id"])
    print(f"  pixels covered by both domains {len(both)}; rgba max {d.max():.3g} mean {d.mean():.3g}", flush=True)
    assert len(both) > 50
    assert d.max() <= RGBA_TOL


def test_framebuffer_entry_points(gpu, orc):
    """drx_partials_to_framebuffer (replace and blend-over-surface), drx_framebuffer_clear and
    drx_framebuffer_composite_background against the oracle, bit for bit: 30,000 partials from three
    'domains' in shuffled order, many pixels with several partials, equal depths inside a pixel."""
    rng = np.random.default_rng(11)
    w, h, n = 97, 61, 30000
    p = np.zeros(n, PARTIAL_DTYPE)
    p["pixel_id"] = rng.integers(0, w * h, n)
    p["depth"] = rng.integers(0, 40, n).astype(np.float32) * np.float32(0.25)
    p["rgba"] = rng.uniform(0, 0.3, (n, 4)).astype(np.float32)
    surf = rng.uniform(0, 1, (w * h, 4)).astype(np.float32)
    for blend in (False, True):
        fb = gpu.Framebuffer(w, h)
        fb.colors.copy_(torch.from_numpy(surf.reshape(-1)).to(fb.colors.device))
        gpu.partials_to_framebuffer([(gpu.as_bytes(p[: n // 3]), n // 3), (gpu.as_bytes(p[n // 3:]), n - n // 3)], fb,
                                    blend=blend)
        col, dep = surf.copy(), np.full(w * h, np.inf, np.float32)
        orc.partials_to_framebuffer(p, w, h, col, dep, blend=blend)
        torch.cuda.synchronize()
        assert np.array_equal(_np(fb.colors).view(np.uint32), col.view(np.uint32).ravel())
        assert np.array_equal(_np(fb.depths).view(np.uint32), dep.view(np.uint32))
    fb.background_color((0.2, 0.4, 0.6, 1.0))
    fb.composite_background()
    orc.framebuffer_composite_background(col, (0.2, 0.4, 0.6, 1.0))
    torch.cuda.synchronize()
    assert np.array_equal(_np(fb.colors).view(np.uint32), col.view(np.uint32).ravel())
    fb.clear((0.5, 0.0, 0.25, 1.0))
    torch.cuda.synchronize()
    assert np.all(_np(fb.colors).reshape(-1, 4) == np.float32([0.5, 0.0, 0.25, 1.0])) and np.all(_np(fb.depths) == np.inf)


def test_renderer_two_domains(gpu, orc, tmp_path):
    """Renderer::render over two domains -> Framebuffer (colours and nearest depths) ->
    composite_background -> save: the framebuffer equals the oracle's compositing of the GPU partials
    bit for bit and the oracle's own render within the RGBA tolerance; the PNG holds the quantised image."""
    from pngread import read_png
    a = syn.hex_mesh(4, 2, 0.05, 0.01, 7)
    b = syn.HexMesh(n=a.n, order=a.order, ctrl_idx=a.ctrl_idx,
                    values=(a.values + np.float32([0.6, 0.0, -0.7])).astype(np.float32), field=(a.field * 0.5).astype(np.float32))
    w, h = 48, 40
    cam = gpu.Camera(w, h, 50.0, pos=(0.8, 0.5, 2.8), look_at=(0.8, 0.5, 0.1))
    vols, obs = [], []
    for mesh in (a, b):
        v = gpu.Volume(gpu.UnstructuredMesh(mesh), gpu.UnstructuredField(mesh))
        v.use_lighting = False
        v.samples(96)
        v.table = gpu.ColorTable()
        v.table.add_point(0.0, (0.1, 0.3, 0.9))
        v.table.add_point(1.0, (0.9, 0.2, 0.1))
        v.table.add_alpha(0.0, 0.3)
        v.table.set_smooth(True)
        vols.append(v)
        obs.append(orc.build_mesh_bvh(mesh))
    r = gpu.Renderer()
    r.volume(vols)
    fb = r.render(cam)
    colors, depths = _np(fb.colors).reshape(-1, 4).copy(), _np(fb.depths).copy()
    # the partials the renderer composited, recomputed through the same volumes
    rays = cam.create_rays()
    gparts = [v.integrate(rays, tile_width=w, tile_height=h) for v in vols]
    torch.cuda.synchronize()
    allg = np.concatenate([_np(p, np.uint8)[: n * 24].view(PARTIAL_DTYPE) for p, n in gparts])
    col, dep = np.zeros((w * h, 4), np.float32), np.full(w * h, np.inf, np.float32)
    orc.partials_to_framebuffer(allg, w, h, col, dep)
    assert np.array_equal(colors.view(np.uint32), col.view(np.uint32)) and np.array_equal(depths, dep)
    assert np.isfinite(depths).sum() > 300
    # the oracle's own render of both domains (collection bounds and range, as the renderer set them)
    cb, (lo, hi) = vols[0].sample_bounds, vols[0].range
    oparts = []
    orays = orc.camera_perspective(w, h, 50.0, (0.8, 0.5, 2.8), (0.8, 0.5, 0.1))
    for mesh, ob, v in zip((a, b), obs, vols):
        lut = v.table.sample(1024, opacity_factor=10.0 / 96)
        op, _, _ = orc.integrate(mesh, ob, orc.remove_missed_rays(orays, ob.bounds), orc.sample_dist(cb, 96), lut, lo, hi)
        oparts.append(orc.compact_partials(op))
    ocol, odep = np.zeros((w * h, 4), np.float32), np.full(w * h, np.inf, np.float32)
    orc.partials_to_framebuffer(np.concatenate(oparts), w, h, ocol, odep)
    d = np.abs(colors - ocol)
    assert d.max() <= RGBA_TOL
    assert np.array_equal(np.isfinite(depths), np.isfinite(odep))
    fb.composite_background()
    fb.save(str(tmp_path / "two"))
    fin = _np(fb.colors).reshape(h, w, 4)
    assert np.array_equal(fin.reshape(-1, 4), orc.composite_domains(allg, w, h))
    png = read_png(tmp_path / "two.png")
    assert np.array_equal(png, np.clip(fin * np.float32(255), 0, 255).astype(np.uint8)[::-1])


def test_volume_orthographic_config0_small(gpu, orc):
    mesh = syn.hex_mesh(8, 2, 0.05, 0.005, 1, jitter_scaled=False)
    w = h = 32
    rays = syn.orthographic_rays(w, h)
    gp, img, op, oimg, gst, ost = _render_both(gpu, orc, mesh, rays, w, h, samples=200, tile=False)
    assert len(gp) == len(op)
    assert np.array_equal(gp["pixel_id"], op["pixel_id"])
    assert gst.located_samples == ost[0]
    d = np.abs(img - oimg)
    assert d.max() <= RGBA_TOL


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
def test_wide_bvh_same_candidates_as_binary(gpu, name, make):
    """The 8-wide traversal must try exactly the binary depth-first candidate sequence: identical
    locations (bit for bit), Newton step counts and candidate counts per point."""
    mesh = make()
    m, bvh = _gpu_bvh(gpu, mesh)
    assert bvh.wide_nodes is not None
    pts = gpu.to_device(syn.query_points(30000, bvh.bounds, seed=21).reshape(-1))
    out = {}
    for trav in (1, 3):   # 8-wide per-lane walk, binary per-lane walk
        loc, steps, cands = m.locate(pts, gpu.LocateOptions(traversal=trav), stats=True)
        out[trav] = (_np(loc), _np(steps), _np(cands))
    for a, b in zip(out[1], out[3]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
def test_wave_traversal_same_as_lane(gpu, name, make):
    """Wave-coherent traversal (union of the lanes' paths, per-lane candidate queues) must reproduce the
    per-lane depth-first result exactly, on coherent (ray-ordered) and incoherent (random) points."""
    mesh = make()
    m, bvh = _gpu_bvh(gpu, mesh)
    rnd = syn.query_points(20000, bvh.bounds, seed=23)
    b = bvh.bounds
    t = np.linspace(0, 1, 20000, dtype=np.float32)
    ray = np.stack([b[0] + (b[1] - b[0]) * t, b[2] + (b[3] - b[2]) * (0.3 + 0.4 * t), b[4] + (b[5] - b[4]) * 0.55 + 0 * t], 1)
    for pts in (rnd, ray.astype(np.float32)):
        d = gpu.to_device(np.ascontiguousarray(pts).reshape(-1))
        res = {}
        for mode in (1, 2):
            loc, steps, cands = m.locate(d, gpu.LocateOptions(traversal=mode), stats=True)
            res[mode] = (_np(loc, LOCATION_DTYPE), _np(steps), _np(cands))
        for mode in (1, 2):   # Morton-ordered processing returns results in input order, unchanged
            loc, steps, cands = m.locate(d, gpu.LocateOptions(traversal=mode, coherent_order=True), stats=True)
            assert np.array_equal(_np(loc), res[mode][0].view(np.uint8).reshape(-1))
            assert np.array_equal(_np(steps), res[mode][1]) and np.array_equal(_np(cands), res[mode][2])
        (l1, s1, c1), (l2, s2, c2) = res[1], res[2]
        dc = int((l1["cell"] != l2["cell"]).sum())
        dr = float(np.abs(l1["ref"] - l2["ref"]).max())
        print(f"  {name}: cells differ {dc}, max ref diff {dr:.3g}, steps differ {int((s1 != s2).sum())}, "
              f"cands differ {int((c1 != c2).sum())}", flush=True)
        assert dc == 0 and np.array_equal(c1, c2)
        # the two entry points inline separately compiled Newton code: references may differ in the last bits
        assert dr <= 1e-6 and (s1 != s2).sum() <= 1e-4 * len(s1)


def test_ray_tile_order_longest_first(gpu, orc):
    """drx_ray_tile_order: a permutation of the active rays in which each 8x8 tile's rays form one
    contiguous run; tiles come longest [near, far] segment first, ties in tile-index order."""
    mesh = syn.hex_mesh(4, 2, 0.05, 0.01, 3)
    ob = orc.build_mesh_bvh(mesh)
    w, h = 72, 56
    cam = gpu.Camera(w, h, 40.0, pos=(0.5, 0.6, 2.2), look_at=(0.5, 0.5, 0.5))
    act, n = gpu.remove_missed_rays(cam.create_rays(), ob.bounds)
    order = _np(gpu.ray_tile_order(act, n, w, h))[:n]
    rays = _np(act, np.uint8)[: n * RAY_DTYPE.itemsize].view(RAY_DTYPE)
    assert n > 500 and np.array_equal(np.sort(order), np.arange(n))
    pid = rays["pixel_id"][order]
    tile = (pid // w // 8) * ((w + 7) // 8) + (pid % w) // 8
    seg = (rays["far"] - rays["near"])[order]
    starts = np.flatnonzero(np.r_[True, tile[1:] != tile[:-1]])
    assert len(np.unique(tile)) == len(starts)                       # each tile is one contiguous run
    longest = np.maximum.reduceat(seg, starts)
    assert np.all(np.diff(longest) <= 0)
    t_first = tile[starts]
    same = np.diff(longest) == 0
    assert np.all(np.diff(t_first)[same] > 0)                         # ties keep tile-index order


def test_leaf_start_maps_same_results(gpu, orc):
    """drx_bvh_leaf_starts: locating with the stored per-leaf Newton start maps gives the same cells as
    the start computed from the element corners at each candidate (reference points within 1e-6, Newton
    step totals within 1%), on a uniform-order and a mixed-order mesh; the maps invert an affine element
    exactly (start = the answer)."""
    for mesh in (syn.hex_mesh(6, 3, 0.12, 0.01, 21), syn.mixed_hex_mesh(5, 0.12, 0.02, 4)):
        m = gpu.UnstructuredMesh(mesh)
        bvh = m.get_bvh()
        assert bvh.leaf_starts is not None
        pts = syn.query_points(40000, bvh.bounds, seed=5)
        dp = gpu.to_device(pts.reshape(-1))
        res = {}
        for use in (True, False):
            bvh.use_starts = use
            out, steps, _ = m.locate(dp, stats=True)
            loc = _np(out, np.uint8).view(LOCATION_DTYPE)
            res[use] = (loc["cell"].copy(), loc["ref"].copy(), int(_np(steps).astype(np.int64).sum()))
        bvh.use_starts = True
        assert np.array_equal(res[True][0], res[False][0])
        hit = res[True][0] >= 0
        assert hit.sum() > 1000 and np.abs(res[True][1][hit] - res[False][1][hit]).max() < 1e-6
        assert abs(res[True][2] - res[False][2]) <= 0.01 * res[False][2]
    # affine element: the map's start is the answer itself
    a = syn.hex_mesh(3, 1, 0.0, 0.0, 1)
    m = gpu.UnstructuredMesh(a)
    bvh = m.get_bvh()
    st = _np(bvh.leaf_starts).reshape(-1, 3, 4)
    p = np.array([0.41, 0.27, 0.83], np.float32)
    x0 = st[:, :, :3] @ p + st[:, :, 3]
    el = _np(bvh.leaf_nodes)
    exp = (p[None, :] * 3 - np.stack([el % 3, (el // 3) % 3, el // 9], 1))
    assert np.abs(x0 - exp).max() < 1e-5


def test_point_location_collection(gpu, orc):
    """PointLocation over a two-domain Collection (a mesh and a copy shifted by (+0.6, +0.5, 0) whose
    field is f + 10; the copies overlap and leave two corners of the union box uncovered) against the
    oracle's restatement of point_location.cpp:178-287 (values within 1e-5, apart from points whose
    located id differs for a reason the parity bar excuses): a point takes the last domain that locates
    it, else empty_val; Collection range / bounds are the unions."""
    a = syn.hex_mesh(4, 2, 0.05, 0.01, 7)
    b = syn.HexMesh(n=a.n, order=a.order, ctrl_idx=a.ctrl_idx,
                    values=(a.values + np.float32([0.6, 0.5, 0.0])).astype(np.float32),
                    field=(a.field + np.float32(10.0)).astype(np.float32))
    col = gpu.Collection()
    meshes = []
    for mesh in (a, b):
        m, f = gpu.UnstructuredMesh(mesh), gpu.UnstructuredField(mesh)
        ds = gpu.DataSet(m)
        ds.add_field("f", f)
        col.add_domain(ds)
        meshes.append((m, f))
    bb = col.bounds()
    pts = syn.query_points(20000, bb, seed=9)
    dp = gpu.to_device(pts.reshape(-1))
    pl = gpu.PointLocation()
    pl.add_var("f")
    pl.add_var("nope")
    pl.empty_val(-7.0)
    res = pl.execute(col, dp)
    got = _np(res["values"][0])
    # the oracle's restatement of PointLocation::execute over the same two domains
    obs = [orc.build_mesh_bvh(a), orc.build_mesh_bvh(b)]
    exp_o, ocells = orc.point_location([(a, obs[0], [a.field]), (b, obs[1], [b.field])], pts, 1, -7.0)
    l0, l1 = meshes[0][0].locate(dp), meshes[1][0].locate(dp)
    c0 = _np(l0, np.uint8).view(LOCATION_DTYPE)["cell"]
    c1 = _np(l1, np.uint8).view(LOCATION_DTYPE)["cell"]
    excused = np.zeros(len(pts), bool)
    for mesh, ob, l, oc in ((a, obs[0], l0, ocells[0]), (b, obs[1], l1, ocells[1])):
        loc = _np(l, np.uint8).view(LOCATION_DTYPE)
        orf = orc.locate(mesh, ob, pts)[1]
        mism, explained, _, _ = explain_mismatches(orc, mesh, ob, pts, loc["cell"], loc["ref"], oc, orf)
        assert explained.all()
        excused[mism] = True
    d = np.abs(got - exp_o[0])
    assert d[~excused].max() <= 1e-5 and excused.sum() <= 10
    # and the composition rule itself on the GPU's own per-domain results: the last domain that locates wins
    v0, v1 = _np(meshes[0][1].eval(l0)), _np(meshes[1][1].eval(l1))
    exp = np.where(c1 >= 0, v1, np.where(c0 >= 0, v0, np.float32(-7.0))).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    assert (c0 >= 0).sum() > 1000 and (c1 >= 0).sum() > 1000 and (exp == -7).sum() > 100
    assert len(res["values"]) == 1 and res["vars"] == ["f", "nope"]
    lo, hi = col.range("f")
    assert lo == float(a.field.min()) and hi == float(np.float32(a.field.max() + np.float32(10.0)))
    ba, bbm = meshes[0][0].bounds(), meshes[1][0].bounds()
    assert bb[0] == ba[0] and bb[1] == bbm[1] and bb[2] == ba[2] and bb[3] == bbm[3]
    both = (c0 >= 0) & (c1 >= 0)
    assert both.sum() > 100 and np.all(got[both] == v1[both])


def test_scratch_per_stream(gpu, orc):
    """Library scratch is kept per stream: drx_partials_to_framebuffer (which returns before its kernels
    have read its sorted partials) issued on one stream and a volume render (whose partial slots use the
    same scratch slot) issued right after on another stream give exactly what each gives alone."""
    rng = np.random.default_rng(12)
    w, h, n = 64, 48, 200000
    p = np.zeros(n, PARTIAL_DTYPE)
    p["pixel_id"] = rng.integers(0, w * h, n)
    p["depth"] = rng.integers(0, 50, n).astype(np.float32) * np.float32(0.5)
    p["rgba"] = rng.uniform(0, 0.2, (n, 4)).astype(np.float32)
    dp = gpu.as_bytes(p)
    mesh = syn.hex_mesh(4, 2, 0.05, 0.01, 3)
    m = gpu.UnstructuredMesh(mesh)
    f = gpu.UnstructuredField(mesh)
    vol = gpu.Volume(m, f)
    vol.use_lighting = False
    vol.samples(128)
    vol.table = gpu.ColorTable()
    vol.table.add_point(0.0, (0.5, 0.5, 0.5))
    vol.table.add_point(1.0, (1.0, 0.0, 0.0))
    vol.table.add_alpha(0.0, 0.05)
    rays = gpu.as_bytes(orc.camera_perspective(96, 96, 45.0, (0.5, 0.5, 2.5), (0.5, 0.5, 0.5)))
    # each alone, on the current stream
    fb0 = gpu.Framebuffer(w, h)
    gpu.partials_to_framebuffer([(dp, n)], fb0)
    parts0, n0 = vol.integrate(rays)
    torch.cuda.synchronize()
    want_fb, want_parts = _np(fb0.colors).copy(), _np(parts0, np.uint8)[: n0 * 24].copy()
    # back to back on two streams
    fb1 = gpu.Framebuffer(w, h)
    torch.cuda.synchronize()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    with torch.cuda.stream(sa):
        gpu.partials_to_framebuffer([(dp, n)], fb1)
    with torch.cuda.stream(sb):
        parts1, n1 = vol.integrate(rays)
    torch.cuda.synchronize()
    assert n1 == n0 and np.array_equal(_np(parts1, np.uint8)[: n1 * 24], want_parts)
    assert np.array_equal(_np(fb1.colors).view(np.uint32), want_fb.view(np.uint32))


def test_auto_walk_choice(gpu):
    """DRX_TRAVERSE_AUTO with DRX_LAUNCH_AUTO in a render: the sample-parallel kernel (walk 5, binary walk)
    for frames of fewer than 300 active rays per SIMD; above that, and with one lane per ray asked for
    (DRX_LAUNCH_CHUNKED), the wave walk for a BVH of at most 65,536 leaves (whose 8-wide copy is built
    with it) and at least 48 active rays per SIMD, else the binary walk; a larger BVH gets no 8-wide copy
    until a walk asks for one.  Every form renders the same partials."""
    simds = 4 * torch.cuda.get_device_properties(0).multi_processor_count
    small = syn.hex_mesh(8, 2, 0.05, 0.005, 1234, jitter_scaled=False)
    m = gpu.UnstructuredMesh(small)
    assert m.get_bvh().n_leaves <= 1 << 16 and m.get_bvh().wide_nodes is not None
    vol = gpu.Volume(m, gpu.UnstructuredField(small))
    vol.use_lighting = False
    vol.samples(100)
    vol.table = gpu.ColorTable()
    vol.table.add_point(0.0, (0.5, 0.5, 0.5))
    vol.table.add_point(1.0, (1.0, 0.0, 0.0))
    vol.table.add_alpha(0.0, 0.05)
    side_big = int(np.ceil(np.sqrt(48 * simds))) + 8
    side_huge = int(np.ceil(np.sqrt(300 * simds))) + 8
    for side, launch, want in ((side_big, 0, 5), (96, 0, 5), (side_huge, 0, 2), (side_big, 1, 2), (96, 1, 0)):
        vol.opts = gpu.LocateOptions(launch=launch)
        parts, n = vol.integrate(gpu.as_bytes(syn.orthographic_rays(side, side)))
        assert vol.last_stats.walk == want, (side, launch, vol.last_stats.walk)
        if side == side_big:
            got = _np(parts, np.uint8)[: n * 24].copy()
            if launch == 0:
                first = got
            else:
                assert np.array_equal(got, first)
    big = syn.hex_mesh(20, 2, 0.05, 0.005, 3, jitter_scaled=False)
    mb = gpu.UnstructuredMesh(big)
    bvh = mb.get_bvh()
    assert bvh.n_leaves > 1 << 16 and bvh.wide_nodes is None
    vb = gpu.Volume(mb, gpu.UnstructuredField(big))
    vb.use_lighting = False
    vb.samples(100)
    vb.table = vol.table
    parts, n = vb.integrate(gpu.as_bytes(syn.orthographic_rays(side_big, side_big)))
    assert vb.last_stats.walk == 5 and bvh.wide_nodes is None
    vb.opts = gpu.LocateOptions(launch=1)
    parts1, n1 = vb.integrate(gpu.as_bytes(syn.orthographic_rays(side_big, side_big)))
    assert vb.last_stats.walk == 0 and bvh.wide_nodes is None
    vb.opts = gpu.LocateOptions(traversal=2)
    parts2, n2 = vb.integrate(gpu.as_bytes(syn.orthographic_rays(side_big, side_big)))
    assert vb.last_stats.walk == 2 and bvh.wide_nodes is not None
    assert n1 == n and n2 == n
    assert np.array_equal(_np(parts2, np.uint8)[: n * 24], _np(parts, np.uint8)[: n * 24])
    assert np.array_equal(_np(parts1, np.uint8)[: n * 24], _np(parts, np.uint8)[: n * 24])


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
def test_render_sample_parallel_identical(gpu, orc, name, make):
    """The sample-parallel launch (DRX_LAUNCH_SAMPLES: four wavefronts per 64 rays locate and shade
    consecutive samples, one replays each ray's march over them) renders the partials of one lane per ray
    (DRX_LAUNCH_CHUNKED, binary walk) bit for bit, with the same located-sample, Newton-step, candidate
    and polished counts: unlit, lit, with FP64 polish on every accepted point, and with fewer samples than
    one round (the march ends inside the first round)."""
    mesh = make()
    m, bvh = _gpu_bvh(gpu, mesh)
    f = gpu.UnstructuredField(mesh)
    vol = gpu.Volume(m, f)
    vol.table = gpu.ColorTable()
    vol.table.add_point(0.0, (0.5, 0.5, 0.5))
    vol.table.add_point(1.0, (1.0, 0.0, 0.0))
    vol.table.add_alpha(0.0, 0.05)
    vol.table.add_alpha(1.0, 0.6)
    w = h = 40
    rays = gpu.as_bytes(orc.camera_perspective(w, h, 45.0, (0.5, 0.5, 2.5), (0.5, 0.5, 0.5)))
    for lit, polish, samples in ((False, False, 96), (True, False, 96), (False, True, 96), (False, False, 5)):
        vol.use_lighting = lit
        vol.samples(samples)
        res = {}
        for launch in (1, 3):
            vol.opts = gpu.LocateOptions(launch=launch, traversal=3, fp64_polish=polish,
                                         cond_threshold=0.5 if polish else 1e4)
            parts, n = vol.integrate(rays, tile_width=w, tile_height=h)
            st = vol.last_stats
            res[launch] = (_np(parts, np.uint8)[: n * 24].copy(),
                           (st.located_samples, st.newton_steps, st.candidates, st.polished), st.walk)
        print(f"[sample-parallel] {name} lit={lit} polish={polish} samples={samples}: partials "
              f"{len(res[1][0]) // 24}, counts {res[1][1]}", flush=True)
        assert res[1][2] == 0 and res[3][2] == 5
        assert res[1][1][0] > 0 and (not polish or res[1][1][3] > 0)
        assert np.array_equal(res[3][0], res[1][0])
        assert res[3][1] == res[1][1]


def test_render_sample_parallel_refuses_other_forms(gpu):
    """DRX_LAUNCH_SAMPLES runs the binary walk with the plain start: the 8-wide and wave walks and the
    learned start are refused with DRX_ERR_UNSUPPORTED."""
    mesh = syn.hex_mesh(3, 2, 0.05, 0.005, 1, jitter_scaled=False)
    m, bvh = _gpu_bvh(gpu, mesh)
    bvh.ensure_wide()
    vol = gpu.Volume(m, gpu.UnstructuredField(mesh))
    vol.use_lighting = False
    vol.samples(32)
    vol.table = gpu.ColorTable()
    vol.table.add_alpha(0.0, 0.05)
    rays = gpu.as_bytes(syn.orthographic_rays(16, 16))
    for trav in (1, 2):
        vol.opts = gpu.LocateOptions(launch=3, traversal=trav)
        with pytest.raises(gpu.L.DrxError, match="sample-parallel"):
            vol.integrate(rays)


def test_given_guess_locate(gpu, orc):
    """DRX_INIT_GUESS (dray's use_init_guess): located with the converged locations of a first pass as
    guesses, every point lands in the same cell, its reference point moves by float rounding only, and
    the candidate that accepts takes exactly one Newton step (the start is already converged)."""
    mesh = syn.hex_mesh(5, 3, 0.1, 0.01, 2)
    m = gpu.UnstructuredMesh(mesh)
    pts = syn.query_points(40000, m.bounds(), seed=17)
    dp = gpu.to_device(pts.reshape(-1))
    first, s1, c1 = m.locate(dp, gpu.LocateOptions(), stats=True)
    again, s2, c2 = m.locate(dp, gpu.LocateOptions(guess=first), stats=True)
    a, b = _np(first, LOCATION_DTYPE), _np(again, LOCATION_DTYPE)
    assert np.array_equal(a["cell"], b["cell"]) and np.array_equal(_np(c1), _np(c2))
    hit = a["cell"] >= 0
    assert np.abs(a["ref"][hit] - b["ref"][hit]).max() <= 2e-6
    one = hit & (_np(c1) == 1)
    assert one.sum() > 10000 and np.all(_np(s2)[one] == 1)
    print(f"[given guess] Newton steps per located point {_np(s1)[hit].mean():.3f} -> {_np(s2)[hit].mean():.3f}",
          flush=True)
    # DRX_INIT_GUESS without guesses is refused
    import ctypes as C
    o = gpu.LocateOptions().c()
    o.init_mode = gpu.L.DRX_INIT_GUESS
    mv, bv = m.gf.view(), m.get_bvh().view()
    out = torch.empty(16 * len(pts), dtype=torch.uint8, device="cuda")
    rc = gpu.L.lib().drx_locate(C.byref(mv), C.byref(bv), C.c_void_p(dp.data_ptr()), len(pts),
                                C.c_void_p(out.data_ptr()), C.byref(o), None)
    assert rc == gpu.L.DRX_ERR_ARG


def test_render_record_replay(gpu, orc):
    """drx_volume_integrate with opts.record writes the location of every (ray, sample) and renders the
    same partials as the plain binary walk; replayed as given starts (DRX_INIT_GUESS), the render is the
    same to float rounding with fewer Newton steps (one per located sample on the accepting candidate)."""
    mesh = syn.hex_mesh(5, 3, 0.1, 0.01, 2)
    m, f = gpu.UnstructuredMesh(mesh), gpu.UnstructuredField(mesh)
    vol = gpu.Volume(m, f)
    vol.use_lighting = False
    vol.samples(128)
    vol.table = gpu.ColorTable()
    vol.table.add_point(0.0, (0.5, 0.5, 0.5))
    vol.table.add_point(1.0, (1.0, 0.0, 0.0))
    vol.table.add_alpha(0.0, 0.05)
    w = h = 48
    rays = gpu.as_bytes(orc.camera_perspective(w, h, 45.0, (0.5, 0.5, 2.5), (0.5, 0.5, 0.5)))
    vol.opts = gpu.LocateOptions(traversal=3)
    p0, n0 = vol.integrate(rays, tile_width=w, tile_height=h)
    s0 = vol.last_stats
    spr = 300
    rec = torch.full((w * h * spr * 4,), -1, dtype=torch.int32, device="cuda")
    vol.opts = gpu.LocateOptions(record=rec, samples_per_ray=spr)
    p1, n1 = vol.integrate(rays, tile_width=w, tile_height=h)
    s1 = vol.last_stats
    assert s1.walk == 4 and n1 == n0 and np.array_equal(_np(p1, np.uint8)[: n1 * 24], _np(p0, np.uint8)[: n0 * 24])
    assert (s1.located_samples, s1.newton_steps) == (s0.located_samples, s0.newton_steps)
    r = _np(rec).view(LOCATION_DTYPE)
    assert (r["cell"] >= 0).sum() == s0.located_samples
    vol.opts = gpu.LocateOptions(guess=rec, record=None, samples_per_ray=spr)
    p2, n2 = vol.integrate(rays, tile_width=w, tile_height=h)
    s2 = vol.last_stats
    g0 = _np(p0, np.uint8)[: n0 * 24].view(PARTIAL_DTYPE)
    g2 = _np(p2, np.uint8)[: n2 * 24].view(PARTIAL_DTYPE)
    print(f"[record/replay] located {s0.located_samples}/{s2.located_samples} newton {s0.newton_steps} -> "
          f"{s2.newton_steps}", flush=True)
    assert n2 == n0 and s2.located_samples == s0.located_samples and s2.newton_steps < s0.newton_steps
    assert np.abs(g2["rgba"] - g0["rgba"]).max() <= 1e-4


def _slab_stack(n_slabs=6, gap=0.06):
    """Order-2 slabs of 3 x 3 x 3 elements stacked along z with empty gaps between them: a ray along z
    crosses n_slabs segments, more than the five partials a ray keeps."""
    thick = (1.0 - gap * (n_slabs - 1)) / n_slabs
    idx, vals, fld, off = [], [], [], 0
    for k in range(n_slabs):
        m = syn.hex_mesh(3, 2, 0.05, 0.005, 40 + k, jitter_scaled=False)
        v = m.values.astype(np.float64).copy()
        v[:, 2] = v[:, 2] * thick + k * (thick + gap)
        idx.append(m.ctrl_idx + off)
        vals.append(v.astype(np.float32))
        fld.append(m.field)
        off += len(m.values)
    return syn.HexMesh(n=0, order=2, ctrl_idx=np.concatenate(idx).astype(np.int32), values=np.concatenate(vals),
                       field=np.concatenate(fld).astype(np.float32))


def test_render_segments_and_partial_limit(gpu, orc):
    """Rays through six separated slabs: each segment closes at the first miss, the march re-locates at the
    same distance, and a ray keeps its first five partials.  Both launches (one lane per ray, sample-parallel)
    give the same partials and counts bit for bit, and both match the oracle's integrate (partial count,
    pixel ids, located samples, RGBA within RGBA_TOL)."""
    mesh = _slab_stack()
    w = h = 32
    rays = syn.orthographic_rays(w, h)
    gp, img, op, oimg, gst, ost = _render_both(gpu, orc, mesh, rays, w, h, samples=200)
    assert gst.walk == 5
    per_ray = np.bincount(gp["pixel_id"], minlength=w * h)
    print(f"[segments] partials {len(gp)} (oracle {len(op)}), rays with 5 partials {(per_ray == 5).sum()}, "
          f"located {gst.located_samples} (oracle {ost[0]})", flush=True)
    assert (per_ray == 5).sum() > w * h // 2 and per_ray.max() == 5
    assert len(gp) == len(op) and np.array_equal(gp["pixel_id"], op["pixel_id"])
    assert gst.located_samples == ost[0]
    assert np.abs(img - oimg).max() <= RGBA_TOL
    m, bvh = _gpu_bvh(gpu, mesh)
    vol = gpu.Volume(m, gpu.UnstructuredField(mesh))
    vol.use_lighting = False
    vol.samples(200)
    vol.table = gpu.ColorTable()
    vol.table.add_point(0.0, (0.5, 0.5, 0.5))
    vol.table.add_point(1.0, (1.0, 0.0, 0.0))
    vol.table.add_alpha(0.0, syn.table_alpha_for_corrected(0.02, 200))
    vol.table.set_smooth(True)
    res = {}
    for launch in (1, 3):
        vol.opts = gpu.LocateOptions(launch=launch)
        parts, n = vol.integrate(gpu.as_bytes(rays), tile_width=w, tile_height=h)
        st = vol.last_stats
        res[launch] = (_np(parts, np.uint8)[: n * 24].copy(), (st.located_samples, st.newton_steps, st.candidates))
    assert np.array_equal(res[1][0], res[3][0]) and res[1][1] == res[3][1]
    assert np.array_equal(res[3][0].view(PARTIAL_DTYPE)["pixel_id"], gp["pixel_id"])

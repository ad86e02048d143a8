"""Element orders 5-8 (DRX_MAX_ORDER; dray allows up to 20, constants.hpp:14) against the oracle: the split
and the LBVH bit-exact (orders 7 and 8 keep the split's nets in global memory), located ids and reference
coordinates, field evaluation, unlit and lit renders, with the same bars as the orders 1-4 tests.  Orders
above 4 run the binary walk with the plain Newton start; the other forms are refused."""
import numpy as np
import pytest
import torch

from devil_ray_amd import synthetic as syn
from devil_ray_amd.synthetic import LOCATION_DTYPE
from parity_util import explain_mismatches
from test_gpu_parity import LIT_TOL, REF_TOL, RGBA_TOL, _gpu_bvh, _np, _render_both

pytestmark = pytest.mark.gpu

HIGH = [(f"p{p}_n2", p) for p in (5, 6, 7, 8)]


def _mesh(p):
    return syn.hex_mesh(2, p, 0.15, 0.01, 30 + p)


@pytest.mark.parametrize("name,p", HIGH, ids=[h[0] for h in HIGH])
def test_high_order_split_and_bvh_bit_exact(gpu, orc, name, p):
    print(f"[high order split/bvh] {name}", flush=True)
    mesh = _mesh(p)
    m = gpu.UnstructuredMesh(mesh)
    aabbs, prims, subrefs = m.split_elements()
    oa, op, osr = orc.split_elements(mesh)
    assert np.array_equal(_np(aabbs).reshape(-1, 6).view(np.uint32), oa.view(np.uint32))
    assert np.array_equal(_np(prims), op)
    assert np.array_equal(_np(subrefs).reshape(-1, 6).view(np.uint32), osr.view(np.uint32))
    ob = orc.bvh_build(oa, op)
    gb = gpu.LinearBVHBuilder().construct(aabbs, prims, want_mcodes=True)
    torch.cuda.synchronize()
    assert np.array_equal(_np(gb.leaf_nodes), ob.leaf_nodes)
    gi = _np(gb.inner_nodes).reshape(-1, 16)
    assert np.array_equal(gi[:, :14].view(np.uint32), ob.inner[:, :14].view(np.uint32))


@pytest.mark.parametrize("name,p", HIGH, ids=[h[0] for h in HIGH])
def test_high_order_locate_and_field(gpu, orc, name, p):
    print(f"[high order locate] {name}", flush=True)
    mesh = _mesh(p)
    m, bvh = _gpu_bvh(gpu, mesh)
    ob = orc.build_mesh_bvh(mesh)
    pts = syn.query_points(20000, bvh.bounds, seed=5)
    oc, orf, _, _ = orc.locate(mesh, ob, pts, want_steps=True)
    for polish in (False, True):
        loc = m.locate(gpu.to_device(pts.reshape(-1)), gpu.LocateOptions(fp64_polish=polish))
        l = _np(loc, LOCATION_DTYPE)
        gc, gr = l["cell"], l["ref"]
        mism, explained, edge, _ = explain_mismatches(orc, mesh, ob, pts, gc, gr, oc, orf)
        assert explained.all(), f"unexplained located-id mismatches: {mism[~explained][:10]}"
        both = np.nonzero((gc == oc) & (gc >= 0))[0]
        assert len(both) > 0.9 * len(pts)
        ex = orc.refine_f64(mesh, pts[both], gc[both], orf[both].astype(np.float64))
        err = np.abs(gr[both].astype(np.float64) - ex).max()
        print(f"  polish={polish} located={len(both)} mismatches={len(mism)} max ref err={err:.3g}", flush=True)
        assert err <= REF_TOL
    # Field::eval at the located points against the oracle's eval_d
    f = gpu.UnstructuredField(mesh)
    vals = _np(f.eval(loc))
    fv = mesh.field.astype(np.float64)
    ok = np.nonzero(l["cell"] >= 0)[0][::50]
    ref = np.array([orc.eval_d(fv[mesh.elem_dofs(int(l["cell"][i]))].reshape(-1, 1), p,
                               l["ref"][i].astype(np.float64))[0][0] for i in ok])
    assert np.abs(vals[ok] - ref).max() < 1e-5


@pytest.mark.parametrize("name,p", HIGH[::3], ids=[h[0] for h in HIGH[::3]])
def test_high_order_render_parity(gpu, orc, name, p):
    mesh = _mesh(p)
    w = h = 40
    pos, look = (0.5, 0.5, 2.5), (0.5, 0.5, 0.5)
    rays = orc.camera_perspective(w, h, 45.0, pos, look)
    for lights, tol in ((None, RGBA_TOL), (orc.default_light(pos, look)[None, :], LIT_TOL)):
        print(f"[high order render] {name} lit={lights is not None}", flush=True)
        gp, img, op, oimg, gst, ost = _render_both(gpu, orc, mesh, rays, w, h, samples=96, lights=lights)
        d = np.abs(img - oimg)
        print(f"  partials {len(gp)}/{len(op)} located {gst.located_samples}/{ost[0]} rgba max {d.max():.3g}",
              flush=True)
        assert len(gp) == len(op) and np.array_equal(gp["pixel_id"], op["pixel_id"])
        assert gst.located_samples == ost[0]
        assert d.max() <= tol


def test_high_order_refuses_other_forms(gpu):
    """Orders above 4: the wave and 8-wide walks and the given start are refused (DRX_ERR_UNSUPPORTED);
    a mixed-order mesh with an element above order 4 is refused when it is packed."""
    mesh = _mesh(5)
    m, bvh = _gpu_bvh(gpu, mesh)
    bvh.ensure_wide()
    d = gpu.to_device(syn.query_points(1000, bvh.bounds, seed=2).reshape(-1))
    for opts in (gpu.LocateOptions(traversal=2), gpu.LocateOptions(traversal=1)):
        with pytest.raises(gpu.L.DrxError, match="binary walk"):
            m.locate(d, opts)
    mixed = syn.mixed_hex_mesh(2, 0.1, 0.01, 3)
    bad = mixed.orders.copy()
    bad[0] = 5
    with pytest.raises(gpu.L.DrxError, match="mixed-order meshes take"):
        gpu.DeviceGridFunction(mixed.ctrl_idx, mixed.values, 3, 0, bad, mixed.offsets)

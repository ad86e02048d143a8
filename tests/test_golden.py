"""Committed fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py) pin the synthetic
generator and the oracle bit-for-bit; the GPU half checks libdrx against the same fixtures."""
import os

import numpy as np
import pytest

from devil_ray_amd import synthetic as syn

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {
    "p2_n2": lambda: syn.hex_mesh(2, 2, 0.05, 0.005, 1, jitter_scaled=False),
    "p3_n2": lambda: syn.hex_mesh(2, 3, 0.1, 0.01, 2),
    "mixed_n2": lambda: syn.mixed_hex_mesh(2, 0.15, 0.05, 3),
}


def _load(name):
    return np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


@pytest.mark.parametrize("name", sorted(CASES))
def test_generator_and_oracle_reproduce_fixture(orc, name):
    g = _load(name)
    mesh = CASES[name]()
    assert _same(mesh.ctrl_idx, g["ctrl_idx"]) and _same(mesh.values, g["values"]) and _same(mesh.field, g["field"])
    bvh = orc.build_mesh_bvh(mesh)
    for k in ("aabbs", "subrefs", "inner", "leaf_nodes", "aabb_ids", "bounds"):
        assert _same(getattr(bvh, k), g[k]), k
    assert _same(bvh.mcodes_sorted, g["mcodes"])
    cells, refs = orc.locate(mesh, bvh, g["points"])
    assert _same(cells, g["cells"]) and _same(refs, g["refs"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_fixture(gpu, name):
    import torch
    from devil_ray_amd.synthetic import LOCATION_DTYPE, PARTIAL_DTYPE
    g = _load(name)
    mesh = CASES[name]()
    m = gpu.UnstructuredMesh(mesh)
    bvh = m.get_bvh()
    torch.cuda.synchronize()
    inner = bvh.inner_nodes.cpu().numpy().reshape(-1, 16)
    assert np.array_equal(inner[:, :14].view(np.uint32), g["inner"][:, :14].view(np.uint32))
    assert np.array_equal(bvh.leaf_nodes.cpu().numpy(), g["leaf_nodes"])
    assert np.array_equal(bvh.aabb_ids.cpu().numpy(), g["aabb_ids"])
    loc = m.locate(gpu.to_device(g["points"].reshape(-1))).cpu().numpy().view(LOCATION_DTYPE)
    assert np.array_equal(loc["cell"], g["cells"])
    ok = loc["cell"] >= 0
    assert np.abs(loc["ref"][ok].astype(np.float64) - g["refs64"][ok]).max() < 1e-5
    f = gpu.UnstructuredField(mesh)
    vol = gpu.Volume(m, f)
    vol.use_lighting = False
    vol.samples(64)
    vol.table = gpu.ColorTable()
    vol.table.add_point(0.0, (0.5, 0.5, 0.5))
    vol.table.add_point(1.0, (1.0, 0.0, 0.0))
    vol.table.add_alpha(0.0, syn.table_alpha_for_corrected(0.02, 64))
    vol.table.set_smooth(True)
    assert np.array_equal(vol.table.sample(1024, opacity_factor=10.0 / 64), g["lut"])
    parts, n = vol.integrate(gpu.as_bytes(syn.orthographic_rays(8, 8)))
    gp = parts.cpu().numpy()[: n * 24].view(PARTIAL_DTYPE)
    op = g["partials"].view(PARTIAL_DTYPE)
    assert n == len(op) and np.array_equal(gp["pixel_id"], op["pixel_id"]) and np.array_equal(gp["depth"], op["depth"])
    assert np.abs(gp["rgba"] - op["rgba"]).max() < 2e-3
    img = gpu.composite(parts, n, 8, 8).cpu().numpy().reshape(-1, 4)
    assert np.abs(img - g["image"]).max() < 2e-3

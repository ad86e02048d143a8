"""Writes the committed golden fixtures (tests/golden/*.npz) from the synthetic generator and the CPU
oracle.  Inputs and expected outputs only (no code).  Re-run after a deliberate change to either:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from devil_ray_amd import synthetic as syn  # noqa: E402
from oracle import orc  # noqa: E402

CASES = {
    "p2_n2": lambda: syn.hex_mesh(2, 2, 0.05, 0.005, 1, jitter_scaled=False),
    "p3_n2": lambda: syn.hex_mesh(2, 3, 0.1, 0.01, 2),
    "mixed_n2": lambda: syn.mixed_hex_mesh(2, 0.15, 0.05, 3),
}


def make(name, mesh):
    bvh = orc.build_mesh_bvh(mesh)
    pts = syn.query_points(256, bvh.bounds, seed=17)
    cells, refs = orc.locate(mesh, bvh, pts)
    cells64, refs64 = orc.locate(mesh, bvh, pts, dtype=np.float64)
    samples, w = 64, 8
    rays = syn.orthographic_rays(w, w)
    a = syn.table_alpha_for_corrected(0.02, samples)
    lut = orc.color_table_sample(syn.grey_to_red_table(), [[0.0, a]], True, 10.0 / samples)
    act = orc.remove_missed_rays(rays, bvh.bounds)
    lo, hi = float(mesh.field.min()), float(mesh.field.max())
    parts, _, _ = orc.integrate(mesh, bvh, act, orc.sample_dist(bvh.bounds, samples), lut, lo, hi)
    parts = orc.compact_partials(parts)
    img = orc.composite(parts, w, w)
    out = dict(ctrl_idx=mesh.ctrl_idx, values=mesh.values, field=mesh.field,
               aabbs=bvh.aabbs, subrefs=bvh.subrefs, inner=bvh.inner, leaf_nodes=bvh.leaf_nodes,
               aabb_ids=bvh.aabb_ids, bounds=bvh.bounds, mcodes=bvh.mcodes_sorted,
               points=pts, cells=cells, refs=refs, cells64=cells64, refs64=refs64,
               partials=parts.view(np.uint8), image=img, lut=lut)
    if mesh.orders is not None:
        out.update(orders=mesh.orders, offsets=mesh.offsets)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


if __name__ == "__main__":
    for name, fn in CASES.items():
        make(name, fn())
        print("wrote", name)

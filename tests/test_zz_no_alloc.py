"""drx_locate_noalloc (C ABI) against drx_locate: identical Location records, Newton-step and candidate
counts on two calls in a row, the second with a smaller input, across the launch forms and walks that
use its static scratch (persistent chunk queue, Morton order); sizes beyond the scratch limit are refused
with a clear error and leave the library usable."""
import ctypes as C

import numpy as np
import pytest
import torch

from devil_ray_amd import _lib as L
from devil_ray_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _call(fn, m, pts, opts):
    bvh = m.get_bvh()
    n = int(pts.numel() // 3)
    dev = pts.device
    out = torch.empty(max(n, 1) * 16, dtype=torch.uint8, device=dev)
    steps = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
    cands = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
    o = opts.c(steps, cands)
    mv, bv = m.gf.view(), bvh.view()
    rc = fn(C.byref(mv), C.byref(bv), C.c_void_p(pts.data_ptr()), n, C.c_void_p(out.data_ptr()), C.byref(o), None)
    torch.cuda.synchronize()
    return rc, out.cpu().numpy(), steps.cpu().numpy(), cands.cpu().numpy()


def test_locate_noalloc_matches_locate(gpu):
    lib = L.lib()
    mesh = syn.mixed_hex_mesh(5, 0.15, 0.05, 5)
    m = gpu.UnstructuredMesh(mesh)
    bvh = m.get_bvh()
    big = syn.query_points(150001, bvh.bounds, seed=61)
    small = big[:7777].copy()
    cases = [gpu.LocateOptions(),
             gpu.LocateOptions(launch=2, chunk=256),
             gpu.LocateOptions(coherent_order=True, launch=2, chunk=512),
             gpu.LocateOptions(traversal=1, launch=1, chunk=1024),
             gpu.LocateOptions(traversal=2, coherent_order=True),
             gpu.LocateOptions(fp64_polish=True, cond_threshold=0.5, launch=2, chunk=1024)]
    for i, opts in enumerate(cases):
        for pts_h in (big, small):   # the second call reuses the scratch with fewer points
            print(f"[noalloc] case {i} {opts} n={len(pts_h)}", flush=True)
            pts = gpu.to_device(pts_h.reshape(-1))
            rc0, l0, s0, c0 = _call(lib.drx_locate, m, pts, opts)
            rc1, l1, s1, c1 = _call(lib.drx_locate_noalloc, m, pts, opts)
            assert rc0 == 0 and rc1 == 0, lib.drx_last_error()
            assert np.array_equal(l0, l1), f"case {i}: locations differ"
            assert np.array_equal(s0, s1) and np.array_equal(c0, c1), f"case {i}: step / candidate counts differ"
            assert (l1.view(np.int32).reshape(-1, 4)[:, 0] >= 0).sum() > len(pts_h) // 4


def test_locate_noalloc_size_limit(gpu):
    lib = L.lib()
    mesh = syn.hex_mesh(4, 2, 0.05, 0.005, 1, jitter_scaled=False)
    m = gpu.UnstructuredMesh(mesh)
    bvh = m.get_bvh()
    limit = L.LOCATE_NOALLOC_MAX_ORDERED
    pts = gpu.to_device(syn.query_points(limit + 1, bvh.bounds, seed=3).reshape(-1))
    rc, _, _, _ = _call(lib.drx_locate_noalloc, m, pts, gpu.LocateOptions(coherent_order=True))
    assert rc == L.DRX_ERR_ARG and b"DRX_LOCATE_NOALLOC_MAX_ORDERED" in lib.drx_last_error()
    # at the limit, and beyond it without the Morton order, the call runs and agrees with drx_locate
    for n, coherent in ((limit, True), (limit + 1, False)):
        p = pts[: 3 * n]
        rc0, l0, _, _ = _call(lib.drx_locate, m, p, gpu.LocateOptions(coherent_order=coherent))
        rc1, l1, _, _ = _call(lib.drx_locate_noalloc, m, p, gpu.LocateOptions(coherent_order=coherent))
        assert rc0 == 0 and rc1 == 0, lib.drx_last_error()
        assert np.array_equal(l0, l1)

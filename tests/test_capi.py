"""CPU checks of the C ABI: libdrx.so loads, exports exactly what include/drx.h declares, its host-side
functions agree with the oracle, and argument errors come back as status codes with messages (no GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from devil_ray_amd import _lib as L
from devil_ray_amd import synthetic as syn
from pngread import read_png

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "drx.h")


def _declared():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(drx_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = L.lib()
    decl = _declared()
    assert len(decl) >= 15
    for name in decl:
        assert hasattr(lib, name), f"{name} declared in include/drx.h but not exported"
    assert decl == L.EXPORTED, "ctypes signatures out of sync with include/drx.h"


def test_version_and_strides():
    lib = L.lib()
    assert b"gfx950" in lib.drx_version()
    for p in range(1, 5):
        npe = (p + 1) ** 3
        pad = (npe + 3) // 4 * 4
        assert lib.drx_mesh_stride(p, 3) == 4 + 3 * pad
        assert lib.drx_mesh_stride(p, 1) == pad
    orders = np.array([1, 4, 2, 3], np.int32)
    offs = np.zeros(4, np.int64)
    tot = lib.drx_mesh_pack_layout(orders.ctypes.data, 4, 3, offs.ctypes.data)
    strides = [lib.drx_mesh_stride(int(p), 3) for p in orders]
    assert tot == sum(strides) and list(offs) == [0] + list(np.cumsum(strides)[:-1])


@pytest.mark.parametrize("smooth", [True, False])
def test_color_table_matches_oracle(orc, smooth):
    rgb = np.array([[0.0, 0.2, 0.3, 0.4, 1.0], [0.3, 0.9, 0.1, 0.0, 1.0], [1.0, 0.0, 0.0, 1.0, 1.0]], np.float32)
    alpha = np.array([[0.1, 0.0], [1.0, 0.7]], np.float32)
    out = np.zeros((1024, 4), np.float32)
    rc = L.lib().drx_color_table_sample(rgb.ctypes.data, 3, alpha.ctypes.data, 2, int(smooth), 0.05, 1024, 1.0,
                                        out.ctypes.data)
    assert rc == 0
    ref = orc.color_table_sample(rgb, alpha, smooth, 0.05, 1024)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_argument_errors_are_status_codes():
    lib = L.lib()
    loc = C.create_string_buffer(16)
    rc = lib.drx_locate(None, None, None, 1, C.cast(loc, C.c_void_p), None, None)
    assert rc == 1 and b"null mesh" in lib.drx_last_error()
    cam = L.DrxCamera()
    cam.width, cam.height, cam.fov_y = 0, 10, 45.0
    assert lib.drx_camera_rays(C.byref(cam), C.cast(loc, C.c_void_p), None) == 1
    assert b"width" in lib.drx_last_error()
    m = L.DrxMesh()
    m.data, m.n_el, m.order, m.ncomp, m.stride = 1, 1, 7, 3, 0
    assert lib.drx_locate(C.byref(m), None, None, 1, None, None, None) == 1
    assert b"order 7" in lib.drx_last_error()
    assert lib.drx_mesh_pack(None, None, 2, None, None, None, 3, 1, None, None) == 1
    cnt = C.c_int32(5)
    b = np.array([1, 0, 1, 0, 1, 0], np.float32)         # empty bounds -> no active rays, no GPU work
    assert lib.drx_remove_missed_rays(C.c_void_p(8), 3, b.ctypes.data, C.c_void_p(8), C.byref(cnt), None) == 0
    assert cnt.value == 0


def test_init_without_gpu_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible; drx_init succeeds there")
    lib = L.lib()
    rc = lib.drx_init(0)
    assert rc != 0 and len(lib.drx_last_error()) > 0


def test_api_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from devil_ray_amd import api
    with pytest.raises(L.DrxError):
        api.UnstructuredMesh(syn.hex_mesh(1, 1, 0.0, 0.0, 0))


def test_save_png_pixels(tmp_path):
    """Framebuffer::save: row 0 of the framebuffer is the last PNG row; each channel is
    (uint8)(v * 255) (truncation), clamped to [0, 255]; 300 x 250 spans several 64 KiB deflate blocks."""
    lib = L.lib()
    w, h = 300, 250
    rng = np.random.default_rng(3)
    rgba = rng.uniform(-0.1, 1.1, (h, w, 4)).astype(np.float32)
    rgba[0, 0] = (0.0, 1.0, 0.5, 254.9 / 255)
    path = str(tmp_path / "img.png")
    assert lib.drx_save_png(rgba.ctypes.data, w, h, path.encode()) == 0
    got = read_png(path)
    exp = np.clip(rgba * np.float32(255), 0, 255).astype(np.uint8)[::-1]
    assert np.array_equal(got, exp)
    assert tuple(got[h - 1, 0]) == (0, 255, 127, 254)
    assert lib.drx_save_png(rgba.ctypes.data, 0, h, path.encode()) == L.DRX_ERR_ARG
    assert lib.drx_save_png(rgba.ctypes.data, w, h, str(tmp_path / "no" / "dir.png").encode()) == L.DRX_ERR_ARG
    assert b"cannot open" in lib.drx_last_error()

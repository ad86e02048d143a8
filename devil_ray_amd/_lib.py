"""ctypes binding of libdrx.so (the C ABI declared in include/drx.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# DRX_LIB overrides the library path (e.g. to compare builds made with different compile-time options)
LIB_PATH = os.environ.get("DRX_LIB", os.path.join(_HERE, "libdrx.so"))

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
F32 = C.c_float
F64 = C.c_double

DRX_OK = 0
DRX_ERR_ARG = 1
DRX_INIT_AFFINE, DRX_INIT_MLP, DRX_INIT_SUBREF_CENTRE, DRX_INIT_GUESS = 0, 1, 2, 3   # drx_init_mode
SPLITS_PER_ELEMENT = 33
MAX_ORDER = 8          # DRX_MAX_ORDER: one order per mesh
MAX_MIXED_ORDER = 4    # DRX_MAX_MIXED_ORDER: mixed-order meshes, learned start
LOCATE_NOALLOC_MAX_ORDERED = 262144   # DRX_LOCATE_NOALLOC_MAX_ORDERED


class DrxMesh(C.Structure):
    _fields_ = [("data", P), ("elem_offset", P), ("elem_order", P), ("n_el", I32), ("order", I32),
                ("ncomp", I32), ("stride", I32), ("data64", P)]


class DrxBVH(C.Structure):
    _fields_ = [("inner_nodes", P), ("leaf_nodes", P), ("aabb_ids", P), ("subrefs", P), ("leaf_pack", P),
                ("wide_nodes", P), ("n_leaves", I32), ("bounds", F32 * 6), ("leaf_starts", P), ("qnodes", P),
                ("leaf_boxes", P), ("qgrid", F32 * 6)]


class DrxLocateOpts(C.Structure):
    _fields_ = [("init_mode", I32), ("fp64_polish", I32), ("cond_threshold", F32), ("newton_steps", P),
                ("candidates", P), ("traversal", I32), ("coherent_order", I32), ("mlp", P),
                ("mlp_newton_steps", I32), ("launch", I32), ("chunk", I32), ("polished", P),
                ("mlp_fallback", I32), ("guess", P), ("record", P), ("samples_per_ray", I32)]


class DrxMlpTrainOpts(C.Structure):
    _fields_ = [("batch", I32), ("steps", I32), ("step0", I32), ("total_steps", I32), ("sample0", I64),
                ("lr", F32), ("lr_final", F32), ("beta1", F32), ("beta2", F32), ("eps", F32)]


class DrxIntegrateStats(C.Structure):
    _fields_ = [("located_samples", I64), ("newton_steps", I64), ("candidates", I64), ("polished", I64),
                ("kernel_ms", F32), ("total_ms", F32), ("walk", I32)]


class DrxCamera(C.Structure):
    _fields_ = [("width", I32), ("height", I32), ("fov_y", F32), ("pos", F32 * 3), ("look_at", F32 * 3),
                ("up", F32 * 3), ("zoom", F32)]


class DrxPointLight(C.Structure):
    _fields_ = [("pos", F32 * 3), ("amb", F32 * 3), ("diff", F32 * 3), ("spec", F32 * 3), ("spec_pow", F32)]


class DrxShading(C.Structure):
    _fields_ = [("use_lighting", I32), ("n_lights", I32), ("lights", C.POINTER(DrxPointLight))]


MAX_LIGHTS = 8


_SIGS = {
    "drx_init": (I32, [I32]),
    "drx_last_error": (C.c_char_p, []),
    "drx_version": (C.c_char_p, []),
    "drx_release": (I32, []),
    "drx_mesh_stride": (I32, [I32, I32]),
    "drx_mesh_pack_layout": (I64, [P, I32, I32, P]),
    "drx_mesh_pack": (I32, [P, P, I32, P, P, P, I32, I32, P, P]),
    "drx_mesh_pack_f64": (I32, [P, P, P, P, P, I32, I32, P, P]),
    "drx_bvh_split_elements": (I32, [P, P, P, P, I32, I32, P, P, P, P]),
    "drx_bvh_build": (I32, [P, P, I32, P, P, P, P, P, P]),
    "drx_bvh_pack_leaves": (I32, [C.POINTER(DrxBVH), P, P]),
    "drx_bvh_leaf_starts": (I32, [C.POINTER(DrxMesh), C.POINTER(DrxBVH), P, P]),
    "drx_bvh_quantize": (I32, [C.POINTER(DrxBVH), P, P, P, P]),
    "drx_locate_autotune": (I32, [C.POINTER(DrxMesh), C.POINTER(DrxBVH), P, I64, P, C.POINTER(DrxLocateOpts),
                                  C.POINTER(I32), C.POINTER(I32), P, P]),
    "drx_bvh_widen": (I32, [C.POINTER(DrxBVH), P, I64, C.POINTER(I64), P]),
    "drx_locate": (I32, [C.POINTER(DrxMesh), C.POINTER(DrxBVH), P, I64, P, C.POINTER(DrxLocateOpts), P]),
    "drx_locate_noalloc": (I32, [C.POINTER(DrxMesh), C.POINTER(DrxBVH), P, I64, P, C.POINTER(DrxLocateOpts), P]),
    "drx_field_eval": (I32, [C.POINTER(DrxMesh), P, I64, P, P]),
    "drx_camera_rays": (I32, [C.POINTER(DrxCamera), P, P]),
    "drx_remove_missed_rays": (I32, [P, I32, P, P, C.POINTER(I32), P]),
    "drx_ray_tile_order": (I32, [P, I32, I32, I32, P, P]),
    "drx_volume_integrate": (I32, [C.POINTER(DrxMesh), C.POINTER(DrxMesh), C.POINTER(DrxBVH), P, I32, P, I32, F64,
                                   F64, F32, P, C.POINTER(DrxLocateOpts), C.POINTER(DrxShading), P,
                                   C.POINTER(I32), C.POINTER(DrxIntegrateStats), P]),
    "drx_default_light": (I32, [C.POINTER(DrxCamera), C.POINTER(DrxPointLight)]),
    "drx_composite": (I32, [P, I32, I32, I32, P, P]),
    "drx_composite_domains": (I32, [P, I32, I32, I32, P, P]),
    "drx_framebuffer_clear": (I32, [P, P, I32, P, P]),
    "drx_partials_to_framebuffer": (I32, [P, I32, I32, I32, P, P, I32, P]),
    "drx_framebuffer_composite_background": (I32, [P, I32, P, P]),
    "drx_save_png": (I32, [P, I32, I32, C.c_char_p]),
    "drx_color_table_sample": (I32, [P, I32, P, I32, I32, F32, I32, F32, P]),
    "drx_color_table_preset": (I32, [C.c_char_p, P, I32, C.POINTER(I32), C.POINTER(I32)]),
    "drx_camera_reset_to_bounds": (I32, [C.POINTER(DrxCamera), P, P]),
    "drx_camera_elevate": (I32, [C.POINTER(DrxCamera), F32]),
    "drx_camera_azimuth": (I32, [C.POINTER(DrxCamera), F32]),
    "drx_mlp_param_count": (I64, []),
    "drx_mlp_state_bytes": (I64, []),
    "drx_mlp_packed_bytes": (I64, []),
    "drx_mlp_init": (I32, [P, C.c_uint64, P]),
    "drx_mlp_features": (I32, [C.POINTER(DrxMesh), P, P, I64, P, P, P, P]),
    "drx_mlp_train": (I32, [P, P, P, I64, C.POINTER(DrxMlpTrainOpts), P, P]),
    "drx_mlp_pack": (I32, [P, P, P]),
    "drx_mlp_infer": (I32, [P, P, P, I64, P, P]),
}

EXPORTED = sorted(_SIGS)

_lib = None


class DrxError(RuntimeError):
    pass


def lib():
    """Load libdrx.so (built by devil_ray_amd/build.py); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DrxError(f"{LIB_PATH} is missing: run devil_ray_amd/build.py (or __graft_entry__.build()) first")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(handle, name)
            f.restype = res
            f.argtypes = args
        _lib = handle
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != DRX_OK:
        msg = lib().drx_last_error().decode(errors="replace")
        raise DrxError(f"{what}: status {rc}: {msg}")

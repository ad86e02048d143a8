"""ctypes bindings for the CPU oracle (oracle/_build/liborc.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
always as the checker.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")
_lib = None

P = C.c_void_p
I32 = C.c_int
I64 = C.c_int64
F32 = C.c_float
F64 = C.c_double

_SIGS = {
    "orc_num_threads_max": (I32, []),
    "orc_split_elements": (I32, [P, P, P, P, I32, I32, P, P, P, I32]),
    "orc_bvh_build": (I32, [P, P, I32, P, P, P, P, P, P, P, P, P]),
    "orc_morton_3d": (C.c_uint32, [F32, F32, F32]),
    "orc_clz": (I32, [C.c_uint32]),
    "orc_eval_d_f32": (None, [P, I32, I32, P, P, P]),
    "orc_eval_d_f64": (None, [P, I32, I32, P, P, P]),
    "orc_lup_solve_f32": (I32, [P, P, P]),
    "orc_lup_solve_f64": (I32, [P, P, P]),
    "orc_newton_f32": (I32, [P, I32, P, P, P]),
    "orc_newton_f64": (I32, [P, I32, P, P, P]),
    "orc_locate_f32": (I32, [P, P, P, P, I32, I32, P, P, P, P, P, I64, P, P, P, P, I32]),
    "orc_set_newton_start": (I32, [I32]),
    "orc_locate_f64": (I32, [P, P, P, P, I32, I32, P, P, P, P, P, I64, P, P, P, P, I32]),
    "orc_locate_brute_f64": (I32, [P, P, P, P, I32, I32, P, I64, P, P]),
    "orc_refine_f64": (I32, [P, P, P, P, I32, I32, P, P, P, I64, P, I32]),
    "orc_field_eval_f32": (I32, [P, P, P, P, I32, I32, P, P, I64, P]),
    "orc_camera_perspective": (None, [I32, I32, F32, P, P, P, F32, P]),
    "orc_remove_missed_rays": (I32, [P, I32, P, P]),
    "orc_sample_dist": (F32, [P, I32]),
    "orc_integrate_f32": (I32, [P, P, P, P, P, P, P, I32, I32, P, P, P, P, P, I64, F32, P, I32, F64, F64, P, P, P, I32,
                                P, I32]),
    "orc_integrate_f64": (I32, [P, P, P, P, P, P, P, I32, I32, P, P, P, P, P, I64, F32, P, I32, F64, F64, P, P, P, I32,
                                P, I32]),
    "orc_integrate_f32_field64": (I32, [P, P, P, P, P, P, P, I32, I32, P, P, P, P, P, I64, F32, P, I32, F64, F64, P, P,
                                        P, I32, P, I32]),
    "orc_integrate_f32_locate64": (I32, [P, P, P, P, P, P, P, I32, I32, P, P, P, P, P, I64, F32, P, I32, F64, F64, P, P,
                                         P, I32, P, I32, P]),
    "orc_default_light": (None, [P, P, P, P]),
    "orc_shade_f32": (None, [P, P, P, P, P, P, P, I32, I32, I32, P, P, P, I32, F64, F64, P, I32, P]),
    "orc_shade_f64": (None, [P, P, P, P, P, P, P, I32, I32, I32, P, P, P, I32, F64, F64, P, I32, P]),
    "orc_compact_partials": (I32, [P, I64, P]),
    "orc_composite": (None, [P, I32, I32, I32, P]),
    "orc_composite_domains": (None, [P, I32, I32, I32, P]),
    "orc_color_table_sample": (None, [P, I32, P, I32, I32, F32, I32, F32, P]),
    "orc_partials_to_framebuffer": (None, [P, I32, I32, I32, P, P, I32]),
    "orc_framebuffer_composite_background": (None, [P, I32, P]),
    "orc_camera_orbit": (None, [P, P, I32, F32, P, P]),
}


def build() -> None:
    subprocess.check_call(["bash", os.path.join(_HERE, "build_oracle.sh")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the oracle must be contiguous"
    return a.ctypes.data


def max_threads() -> int:
    return lib().orc_num_threads_max()


# ---- convenience wrappers ---------------------------------------------------------------------
def _mesh_args(mesh, dtype=np.float32, field=False, field64=False):
    src = (mesh.field64 if field64 else mesh.field) if field else mesh.values
    vals = np.ascontiguousarray(src.astype(dtype))
    ctrl = np.ascontiguousarray(mesh.ctrl_idx)
    orders = None if mesh.orders is None else np.ascontiguousarray(mesh.orders.astype(np.int32))
    offs = None if mesh.offsets is None else np.ascontiguousarray(mesh.offsets.astype(np.int64))
    return ctrl, vals, orders, offs


def split_elements(mesh, nthreads=0):
    n_el = mesh.n_el
    aabbs = np.zeros((n_el * 33, 6), np.float32)
    prims = np.zeros(n_el * 33, np.int32)
    subrefs = np.zeros((n_el * 33, 6), np.float32)
    ctrl, vals, orders, offs = _mesh_args(mesh)
    lib().orc_split_elements(ptr(ctrl), ptr(vals), ptr(orders), ptr(offs), mesh.order, n_el,
                             ptr(aabbs), ptr(prims), ptr(subrefs), nthreads)
    return aabbs, prims, subrefs


class OracleBVH:
    pass


def bvh_build(aabbs: np.ndarray, prims: np.ndarray) -> OracleBVH:
    n = int(aabbs.shape[0])
    m = max(n, 2)
    b = OracleBVH()
    b.inner = np.zeros(((m - 1), 16), np.float32)
    b.leaf_nodes = np.zeros(m, np.int32)
    b.aabb_ids = np.zeros(m, np.int32)
    b.bounds = np.zeros(6, np.float32)
    b.mcodes_sorted = np.zeros(m, np.uint32)
    b.mcodes_unsorted = np.zeros(m, np.uint32)
    b.left = np.zeros(m - 1, np.int32)
    b.right = np.zeros(m - 1, np.int32)
    b.parent = np.zeros(2 * m - 1, np.int32)
    a = np.ascontiguousarray(aabbs, dtype=np.float32)
    p = np.ascontiguousarray(prims, dtype=np.int32)
    lib().orc_bvh_build(ptr(a), ptr(p), n, ptr(b.inner), ptr(b.leaf_nodes), ptr(b.aabb_ids), ptr(b.bounds),
                        ptr(b.mcodes_sorted), ptr(b.mcodes_unsorted), ptr(b.left), ptr(b.right), ptr(b.parent))
    return b


def build_mesh_bvh(mesh, nthreads=0):
    aabbs, prims, subrefs = split_elements(mesh, nthreads)
    b = bvh_build(aabbs, prims)
    b.aabbs, b.prims, b.subrefs = aabbs, prims, subrefs
    return b


def set_newton_start(mode: str):
    """Newton start rule of locate/integrate: "affine" (corner_start, the default) or "centre" (the
    candidate's SubRef centre, data_model/subref.hpp:52-56)."""
    code = {"affine": 0, "centre": 2}.get(mode)
    if code is None or lib().orc_set_newton_start(code) != 0:
        raise ValueError(f"unknown Newton start {mode!r}")


def locate(mesh, bvh, pts, dtype=np.float32, nthreads=0, want_steps=False):
    npts = int(pts.shape[0])
    ctrl, vals, orders, offs = _mesh_args(mesh, dtype)
    P_ = np.ascontiguousarray(pts.astype(dtype))
    cells = np.zeros(npts, np.int32)
    refs = np.zeros((npts, 3), dtype)
    steps = np.zeros(npts, np.int32) if want_steps else None
    stats = np.zeros(2, np.int64)
    fn = lib().orc_locate_f32 if dtype == np.float32 else lib().orc_locate_f64
    fn(ptr(ctrl), ptr(vals), ptr(orders), ptr(offs), mesh.order, mesh.n_el, ptr(bvh.inner), ptr(bvh.leaf_nodes),
       ptr(bvh.aabb_ids), ptr(bvh.subrefs), ptr(P_), npts, ptr(cells), ptr(refs), ptr(steps), ptr(stats), nthreads)
    return (cells, refs, steps, stats) if want_steps else (cells, refs)


def field_eval(mesh, cells, refs, values=None, out=None):
    """Field::eval (unstructured_field.cpp:213-246) in float32: out[i] = the field at (cells[i], refs[i]);
    points with cells[i] == -1 keep out[i].  values: per-control-point field (default mesh.field)."""
    ctrl, _, orders, offs = _mesh_args(mesh)
    vals = np.ascontiguousarray((mesh.field if values is None else values).astype(np.float32))
    c = np.ascontiguousarray(np.asarray(cells, np.int32))
    r = np.ascontiguousarray(np.asarray(refs, np.float32))
    o = np.zeros(len(c), np.float32) if out is None else out
    lib().orc_field_eval_f32(ptr(ctrl), ptr(vals), ptr(orders), ptr(offs), mesh.order, mesh.n_el, ptr(c), ptr(r),
                             len(c), ptr(o))
    return o


def point_location(domains, pts, n_vars, empty_val):
    """PointLocation::execute on one rank (queries/point_location.cpp:178-287): every value starts at
    empty_val; the domains locate every point in collection order and, for a domain that located any
    point, each variable's Field::eval overwrites the values of the points it located (misses are left
    untouched), so a point keeps the last domain that located it.  domains: [(mesh, OracleBVH, [field
    values per variable])].  Returns (values [n_vars, n], cells per domain)."""
    n = int(pts.shape[0])
    values = np.full((n_vars, n), np.float32(empty_val), np.float32)
    cells_all = []
    for mesh, ob, fields in domains:
        cells, refs = locate(mesh, ob, pts)
        cells_all.append(cells)
        if not (cells != -1).any():
            continue
        for v in range(n_vars):
            field_eval(mesh, cells, refs, values=fields[v], out=values[v])
    return values, cells_all


def locate_brute_f64(mesh, pts):
    npts = int(pts.shape[0])
    ctrl, vals, orders, offs = _mesh_args(mesh, np.float64)
    P_ = np.ascontiguousarray(pts.astype(np.float64))
    cells = np.zeros(npts, np.int32)
    refs = np.zeros((npts, 3), np.float64)
    lib().orc_locate_brute_f64(ptr(ctrl), ptr(vals), ptr(orders), ptr(offs), mesh.order, mesh.n_el, ptr(P_), npts,
                               ptr(cells), ptr(refs))
    return cells, refs


def refine_f64(mesh, pts, cells, refs, nthreads=0):
    npts = int(pts.shape[0])
    ctrl, vals, orders, offs = _mesh_args(mesh, np.float64)
    P_ = np.ascontiguousarray(pts.astype(np.float64))
    c = np.ascontiguousarray(cells.astype(np.int32))
    r = np.ascontiguousarray(refs.astype(np.float64))
    out = np.zeros((npts, 3), np.float64)
    lib().orc_refine_f64(ptr(ctrl), ptr(vals), ptr(orders), ptr(offs), mesh.order, mesh.n_el, ptr(P_), ptr(c), ptr(r),
                         npts, ptr(out), nthreads)
    return out


def eval_d(dofs: np.ndarray, p: int, ref, dtype=np.float64):
    """dofs: [npe, ncomp] in dray dof order.  Returns (val[ncomp], deriv[3, ncomp])."""
    ncomp = dofs.shape[1]
    d = np.ascontiguousarray(dofs.astype(dtype))
    r = np.ascontiguousarray(np.asarray(ref, dtype=dtype))
    val = np.zeros(ncomp, dtype)
    der = np.zeros((3, ncomp), dtype)
    fn = lib().orc_eval_d_f64 if dtype == np.float64 else lib().orc_eval_d_f32
    fn(ptr(d), p, ncomp, ptr(r), ptr(val), ptr(der))
    return val, der


def jacobian_cond_f64(mesh, cells, refs) -> np.ndarray:
    """inf-norm condition number ||J|| ||J^-1|| (numpy.linalg.cond) of each element's FP64 Jacobian at the
    given reference point (columns dphi/du, dphi/dv, dphi/dw, element.hpp:458-459): the quantity the
    FP64-polish gate estimates.  Unlocated points (cell < 0) get nan."""
    cells = np.asarray(cells)
    out = np.full(len(cells), np.nan)
    vals = mesh.values.astype(np.float64)
    for i in np.nonzero(cells >= 0)[0]:
        e = int(cells[i])
        p = int(mesh.orders[e]) if mesh.orders is not None else mesh.order
        _, der = eval_d(vals[mesh.elem_dofs(e)], p, np.asarray(refs[i], np.float64))
        out[i] = np.linalg.cond(der.T, p=np.inf)
    return out


def lup_solve(A, b, dtype=np.float64):
    A_ = np.ascontiguousarray(np.asarray(A, dtype=dtype))
    b_ = np.ascontiguousarray(np.asarray(b, dtype=dtype))
    x = np.zeros(3, dtype)
    fn = lib().orc_lup_solve_f64 if dtype == np.float64 else lib().orc_lup_solve_f32
    valid = fn(ptr(A_), ptr(b_), ptr(x))
    return x, bool(valid)


def newton(dofs, p, target, guess, dtype=np.float64):
    d = np.ascontiguousarray(np.asarray(dofs, dtype=dtype))
    t = np.ascontiguousarray(np.asarray(target, dtype=dtype))
    x = np.ascontiguousarray(np.asarray(guess, dtype=dtype)).copy()
    steps = np.zeros(1, np.int32)
    fn = lib().orc_newton_f64 if dtype == np.float64 else lib().orc_newton_f32
    ok = fn(ptr(d), p, ptr(t), ptr(x), ptr(steps))
    return bool(ok), x, int(steps[0])


def camera_perspective(w, h, fov_y, pos, look_at, up=(0.0, 1.0, 0.0), zoom=1.0):
    from devil_ray_amd.synthetic import RAY_DTYPE
    out = np.zeros(w * h, dtype=RAY_DTYPE)
    pos_ = np.asarray(pos, np.float32)
    la = np.asarray(look_at, np.float32)
    up_ = np.asarray(up, np.float32)
    lib().orc_camera_perspective(w, h, float(fov_y), ptr(pos_), ptr(la), ptr(up_), float(zoom), ptr(out))
    return out


def remove_missed_rays(rays, bounds6):
    out = np.zeros_like(rays)
    b = np.ascontiguousarray(np.asarray(bounds6, np.float32))
    k = lib().orc_remove_missed_rays(ptr(rays), int(rays.shape[0]), ptr(b), ptr(out))
    return out[:k].copy()


def sample_dist(bounds6, samples):
    b = np.ascontiguousarray(np.asarray(bounds6, np.float32))
    return float(lib().orc_sample_dist(ptr(b), int(samples)))


def color_table_sample(rgb_pts, alpha_pts, smooth=True, factor=-1.0, num_samples=1024, alpha_scale=1.0):
    rgb = np.ascontiguousarray(np.asarray(rgb_pts, np.float32).reshape(-1, 5))
    al = np.ascontiguousarray(np.asarray(alpha_pts, np.float32).reshape(-1, 2))
    out = np.zeros((num_samples, 4), np.float32)
    lib().orc_color_table_sample(ptr(rgb), rgb.shape[0], ptr(al), al.shape[0], int(bool(smooth)), float(factor),
                                 num_samples, float(alpha_scale), ptr(out))
    return out


def default_light(pos, look_at, up=(0, 1, 0)) -> np.ndarray:
    """Renderer's default PointLight for a camera: 13 floats (pos, amb, diff, spec, spec_pow)."""
    out = np.zeros(13, np.float32)
    a = [np.ascontiguousarray(np.asarray(v, np.float32)) for v in (pos, look_at, up)]
    lib().orc_default_light(ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(out))
    return out


def shade(mesh, cell, ref, ray_dir, lut, fmin, fmax, lights, dtype=np.float64, field=None):
    """VolumeShader::shaded_color for one sample -> RGBA float32[4].  field: per-control-point scalars
    (default mesh.field)."""
    ctrl, vals, orders, offs = _mesh_args(mesh, dtype)
    fctrl, fvals, _, foffs = _mesh_args(mesh, dtype, field=True)
    if field is not None:
        fvals = np.ascontiguousarray(np.asarray(field, dtype))
    out = np.zeros(4, np.float32)
    r = np.ascontiguousarray(np.asarray(ref, np.float64))
    d = np.ascontiguousarray(np.asarray(ray_dir, np.float32))
    lut_ = np.ascontiguousarray(lut.astype(np.float32))
    lt = np.ascontiguousarray(np.asarray(lights, np.float32).reshape(-1, 13))
    fn = lib().orc_shade_f32 if dtype == np.float32 else lib().orc_shade_f64
    fn(ptr(ctrl), ptr(vals), ptr(fctrl), ptr(fvals), ptr(orders), ptr(offs), ptr(foffs), mesh.order, mesh.n_el,
       int(cell), ptr(r), ptr(d), ptr(lut_), lut_.shape[0], float(fmin), float(fmax), ptr(lt), lt.shape[0], ptr(out))
    return out


def integrate(mesh, bvh, active_rays, sample_dist_, lut, fmin, fmax, dtype=np.float32, nthreads=0, lights=None,
              field64=False, locate64=False):
    """Returns (partials[5*n] uncompacted, per-ray located counts, stats[located, newton, candidates]).
    lights: None = unlit (VolumeShader::color); else float32 [k, 13] PointLights (shaded_color).
    field64 (unlit, dtype float32): geometry in float32, the field (mesh.field64) and colour map in double.
    locate64 (dtype float32): the FP32 march, each sample point located by the FP64 walk and Newton."""
    from devil_ray_amd.synthetic import PARTIAL_DTYPE
    n = int(active_rays.shape[0])
    ctrl, vals, orders, offs = _mesh_args(mesh, dtype)
    fctrl, fvals, _, foffs = _mesh_args(mesh, np.float64 if field64 else dtype, field=True, field64=field64)
    parts = np.zeros(5 * n, dtype=PARTIAL_DTYPE)
    located = np.zeros(n, np.int32)
    stats = np.zeros(3, np.int64)
    lut_ = np.ascontiguousarray(lut.astype(np.float32))
    rays = np.ascontiguousarray(active_rays)
    lt = None if lights is None else np.ascontiguousarray(np.asarray(lights, np.float32).reshape(-1, 13))
    extra = ()
    if field64:
        if dtype != np.float32 or lights is not None or locate64:
            raise ValueError("field64 integrates unlit over float32 geometry")
        fn = lib().orc_integrate_f32_field64
    elif locate64:
        if dtype != np.float32:
            raise ValueError("locate64 is the float32 march with FP64 location")
        fn = lib().orc_integrate_f32_locate64
        v64 = np.ascontiguousarray(mesh.values.astype(np.float64))
        extra = (ptr(v64),)
    else:
        fn = lib().orc_integrate_f32 if dtype == np.float32 else lib().orc_integrate_f64
    fn(ptr(ctrl), ptr(vals), ptr(fctrl), ptr(fvals), ptr(orders), ptr(offs), ptr(foffs), mesh.order, mesh.n_el,
       ptr(bvh.inner), ptr(bvh.leaf_nodes), ptr(bvh.aabb_ids), ptr(bvh.subrefs), ptr(rays), n, float(sample_dist_),
       ptr(lut_), lut_.shape[0], float(fmin), float(fmax), ptr(parts), ptr(located), ptr(stats), nthreads,
       ptr(lt), 0 if lt is None else lt.shape[0], *extra)
    return parts, located, stats


def compact_partials(parts):
    out = np.zeros_like(parts)
    k = lib().orc_compact_partials(ptr(parts), int(parts.shape[0]), ptr(out))
    return out[:k].copy()


def composite_domains(parts, w, h):
    """Cross-domain compositing: sort by (pixel, depth), then composite like composite()."""
    out = np.zeros((w * h, 4), np.float32)
    p = np.ascontiguousarray(parts)
    lib().orc_composite_domains(ptr(p), int(p.shape[0]), w, h, ptr(out))
    return out


def partials_to_framebuffer(parts, w, h, colors, depths, blend=False):
    """Renderer's volume compositing into a framebuffer (colors float32 [w*h, 4], depths float32 [w*h],
    both updated in place)."""
    p = np.ascontiguousarray(parts)
    assert colors.dtype == np.float32 and colors.flags.c_contiguous and colors.size == 4 * w * h
    assert depths.dtype == np.float32 and depths.flags.c_contiguous and depths.size == w * h
    lib().orc_partials_to_framebuffer(ptr(p), int(p.shape[0]), w, h, ptr(colors), ptr(depths), 1 if blend else 0)


def framebuffer_composite_background(colors, bg=(1.0, 1.0, 1.0, 1.0)):
    """Framebuffer::composite_background in place on colors float32 [n, 4]."""
    assert colors.dtype == np.float32 and colors.flags.c_contiguous
    b = np.ascontiguousarray(np.asarray(bg, np.float32))
    lib().orc_framebuffer_composite_background(ptr(colors), int(colors.size // 4), ptr(b))


def camera_orbit(pos, look_at, up, fov, mode, arg=0.0, bounds6=None, pads=(0.0, 0.0, 0.0)):
    """Camera::reset_to_bounds (mode "reset"), elevate ("elevate", arg degrees) or azimuth ("azimuth")
    -> (pos, look_at, fov) as float32."""
    cam = np.ascontiguousarray(np.array(list(pos) + list(look_at) + [fov], np.float32))
    u = np.ascontiguousarray(np.asarray(up, np.float32))
    b = np.ascontiguousarray(np.zeros(6, np.float32) if bounds6 is None else np.asarray(bounds6, np.float32))
    pd = np.ascontiguousarray(np.asarray(pads, np.float64))
    m = {"reset": 0, "elevate": 1, "azimuth": 2}[mode]
    lib().orc_camera_orbit(ptr(cam), ptr(u), m, float(arg), ptr(b), ptr(pd))
    return cam[0:3].copy(), cam[3:6].copy(), float(cam[6])


def composite(parts, w, h):
    out = np.zeros((w * h, 4), np.float32)
    p = np.ascontiguousarray(parts)
    lib().orc_composite(ptr(p), int(p.shape[0]), w, h, ptr(out))
    return out

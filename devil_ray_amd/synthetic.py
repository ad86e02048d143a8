"""Seeded synthetic Bernstein hex meshes, fields, query points and cameras.

dray's own datasets are read from Conduit/HDF5; the benchmark workloads are synthetic instead, so
this module pins every generator detail (lattice, warp, jitter, field, RNG) in code.

Data layout produced here is dray's ``GridFunction`` (src/dray/data_model/grid_function.hpp:22-23):
``ctrl_idx`` (int32, ``(p+1)^3`` entries per element, dof index ``i + (p+1)(j + (p+1)k)`` with u
fastest, src/dray/data_model/pos_tensor_element.tcc:159-180) into ``values`` (float32 Vec3 per
control point).  Elements are numbered ``e = ex + n*(ey + n*ez)``.

Uniform-order meshes share one global control-point lattice of ``(p*n+1)^3`` points (conforming).
The warp is ``A*sin(2*pi*x)*sin(2*pi*y)*sin(2*pi*z)`` along ``(1,1,1)/sqrt(3)``; it vanishes on the cube
faces.  Jitter is uniform in ``[-j, j]`` per component from a counter-based hash (splitmix64 of
``seed``, lattice index and component), with the component normal to a cube face zeroed for points on
that face, so the mesh keeps tiling the unit cube.

Mixed-order meshes (one order per element, drawn from ``orders``) keep faces conforming by making every
element face the bilinear patch of its (warped, jittered, shared) corner vertices; strictly interior control
points add the warp's curvature (its deviation from the trilinear interpolant) and their own jitter.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser applied elementwise to uint64 counters."""
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64) + _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_pm1(seed: int, stream: int, counters: np.ndarray) -> np.ndarray:
    """Uniform doubles in [-1, 1) from (seed, stream, counter)."""
    with np.errstate(over="ignore"):
        base = splitmix64(np.array([seed * 0x100000001B3 + stream], dtype=np.uint64))[0]
        z = splitmix64(counters.astype(np.uint64) ^ base)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def warp_displacement(pts: np.ndarray, amp: float) -> np.ndarray:
    s = amp * np.sin(2 * np.pi * pts[:, 0]) * np.sin(2 * np.pi * pts[:, 1]) * np.sin(2 * np.pi * pts[:, 2])
    return (s / math.sqrt(3.0))[:, None] * np.ones((1, 3))


def field_function(pts: np.ndarray) -> np.ndarray:
    """f = sin(2 pi x) cos(2 pi y) z, sampled at control points."""
    return np.sin(2 * np.pi * pts[:, 0]) * np.cos(2 * np.pi * pts[:, 1]) * pts[:, 2]


@dataclass
class HexMesh:
    n: int                      # elements per axis
    order: int                  # uniform order (0 for mixed)
    ctrl_idx: np.ndarray        # int32 [sum npe]
    values: np.ndarray          # float32 [n_pts, 3]
    field: np.ndarray           # float32 [n_pts] (same ctrl_idx)
    orders: Optional[np.ndarray] = None    # int32 [n_el] for mixed-order meshes
    offsets: Optional[np.ndarray] = None   # int64 [n_el] start of element e in ctrl_idx
    field64: Optional[np.ndarray] = None   # float64 [n_pts]: the field unrounded (an FP64 field), if kept

    @property
    def n_el(self) -> int:
        if self.orders is not None:
            return int(len(self.orders))
        return int(len(self.ctrl_idx) // (self.order + 1) ** 3)

    def elem_dofs(self, e: int) -> np.ndarray:
        if self.orders is None:
            npe = (self.order + 1) ** 3
            return self.ctrl_idx[e * npe:(e + 1) * npe]
        p = int(self.orders[e])
        o = int(self.offsets[e])
        return self.ctrl_idx[o:o + (p + 1) ** 3]


def elements_from_nets(nets, p: int, field=None) -> HexMesh:
    """A mesh of independent order-p elements, one control net [(p+1)^3, 3] each (dof i + (p+1)(j + (p+1)k),
    u fastest); no control points are shared.  field: per-element nets [(p+1)^3] (default zero)."""
    nets = np.asarray(nets, dtype=np.float32).reshape(-1, (p + 1) ** 3, 3)
    values = np.ascontiguousarray(nets.reshape(-1, 3))
    f = np.zeros(len(values), np.float32) if field is None else np.asarray(field, np.float32).reshape(-1)
    return HexMesh(n=0, order=p, ctrl_idx=np.arange(len(values), dtype=np.int32), values=values, field=f)


def box_net(lo, hi, p: int) -> np.ndarray:
    """The control net of the affine map of [0,1]^3 onto the box [lo, hi] (a lattice net), float64."""
    t = np.arange(p + 1, dtype=np.float64) / p
    k, j, i = np.meshgrid(t, t, t, indexing="ij")
    u = np.stack([i.reshape(-1), j.reshape(-1), k.reshape(-1)], axis=1)
    return np.asarray(lo, np.float64) + u * (np.asarray(hi, np.float64) - np.asarray(lo, np.float64))


def _element_lattice_index(n: int, p: int) -> np.ndarray:
    """ctrl_idx for the shared (p*n+1)^3 lattice, element-major, dof i + (p+1)(j + (p+1)k)."""
    m = p * n + 1
    e = np.arange(n ** 3, dtype=np.int64)
    ex, ey, ez = e % n, (e // n) % n, e // (n * n)
    loc = np.arange((p + 1) ** 3, dtype=np.int64)
    i, j, k = loc % (p + 1), (loc // (p + 1)) % (p + 1), loc // ((p + 1) ** 2)
    gi = ex[:, None] * p + i[None, :]
    gj = ey[:, None] * p + j[None, :]
    gk = ez[:, None] * p + k[None, :]
    return (gi + m * (gj + m * gk)).reshape(-1)


def hex_mesh(n: int, p: int, amp: float, jitter: float, seed: int, jitter_scaled: bool = True,
             keep_field64: bool = False) -> HexMesh:
    """Uniform order-p mesh of n^3 elements on the unit cube (shared global lattice).

    jitter_scaled: jitter is ``jitter * (h/p)`` with ``h = 1/n`` (control-point spacing units); otherwise
    it is absolute.  keep_field64: also keep the field's coefficients unrounded (float64) as ``field64``.
    """
    m = p * n + 1
    g = np.arange(m, dtype=np.float64) / (p * n)
    I, J, K = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    # lattice index = I + m*(J + m*K)  (x fastest)
    I = I.transpose(2, 1, 0).reshape(-1)
    J = J.transpose(2, 1, 0).reshape(-1)
    K = K.transpose(2, 1, 0).reshape(-1)
    pts = np.stack([g[I], g[J], g[K]], axis=1)
    pts = pts + warp_displacement(pts, amp)
    jit = jitter * ((1.0 / n) / p) if jitter_scaled else jitter
    if jit != 0.0:
        cnt = np.arange(pts.shape[0], dtype=np.uint64) * np.uint64(3)
        for c, idx in enumerate((I, J, K)):
            d = uniform_pm1(seed, 1, cnt + np.uint64(c)) * jit
            d[(idx == 0) | (idx == m - 1)] = 0.0
            pts[:, c] += d
    values = pts.astype(np.float32)
    f64 = field_function(values.astype(np.float64))
    field = f64.astype(np.float32)
    ctrl = _element_lattice_index(n, p).astype(np.int32)
    return HexMesh(n=n, order=p, ctrl_idx=ctrl, values=values, field=field, field64=f64 if keep_field64 else None)


def mixed_hex_mesh(n: int, amp: float, jitter: float, seed: int, orders=(1, 2, 3, 4)) -> HexMesh:
    """Mixed-order mesh: per-element order drawn uniformly from ``orders`` (seeded), conforming faces.

    Corner vertices are shared and carry the global warp plus jitter (``jitter`` x spacing 1/n, zero normal
    component on the cube faces).  Every edge and face control point is the trilinear interpolation of the
    element's corners, so each face is the bilinear patch of its four corners in every order and faces
    between elements of different orders coincide.  Strictly interior control points add the warp's
    deviation from its own trilinear interpolant (the curvature the corners cannot carry) and a jitter of
    ``jitter`` x spacing / order.
    """
    n_el = n ** 3
    ords = np.asarray(orders, dtype=np.int32)
    pick = (uniform_pm1(seed, 7, np.arange(n_el, dtype=np.uint64)) + 1.0) * 0.5
    el_order = ords[np.minimum((pick * len(ords)).astype(np.int64), len(ords) - 1)]
    m = n + 1
    g = np.arange(m, dtype=np.float64) / n
    vi = np.arange(m ** 3, dtype=np.int64)
    VI, VJ, VK = vi % m, (vi // m) % m, vi // (m * m)
    lattice = np.stack([g[VI], g[VJ], g[VK]], axis=1)
    corner_warp = warp_displacement(lattice, amp)
    corners = lattice + corner_warp
    jit = jitter * (1.0 / n)
    if jit != 0.0:
        cnt = vi.astype(np.uint64) * np.uint64(3)
        for c, idx in enumerate((VI, VJ, VK)):
            d = uniform_pm1(seed, 1, cnt + np.uint64(c)) * jit
            d[(idx == 0) | (idx == m - 1)] = 0.0
            corners[:, c] += d
    npes = (el_order.astype(np.int64) + 1) ** 3
    offsets = np.zeros(n_el, dtype=np.int64)
    offsets[1:] = np.cumsum(npes)[:-1]
    total = int(npes.sum())
    values = np.zeros((total, 3), dtype=np.float64)
    e_all = np.arange(n_el, dtype=np.int64)
    ex, ey, ez = e_all % n, (e_all // n) % n, e_all // (n * n)
    for p in np.unique(el_order):
        p = int(p)
        sel = np.nonzero(el_order == p)[0]
        loc = np.arange((p + 1) ** 3)
        li, lj, lk = loc % (p + 1), (loc // (p + 1)) % (p + 1), loc // ((p + 1) ** 2)
        u, v, w = li / p, lj / p, lk / p
        c = np.zeros((len(sel), len(loc), 3))
        cw = np.zeros((len(sel), len(loc), 3))
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    vid = (ex[sel] + dx) + m * ((ey[sel] + dy) + m * (ez[sel] + dz))
                    wt = (u if dx else 1 - u) * (v if dy else 1 - v) * (w if dz else 1 - w)
                    c += wt[None, :, None] * corners[vid][:, None, :]
                    cw += wt[None, :, None] * corner_warp[vid][:, None, :]
        interior = (li > 0) & (li < p) & (lj > 0) & (lj < p) & (lk > 0) & (lk < p)
        if interior.any():
            x0 = np.stack([(ex[sel][:, None] + u[None, interior]) / n, (ey[sel][:, None] + v[None, interior]) / n,
                           (ez[sel][:, None] + w[None, interior]) / n], axis=2).reshape(-1, 3)
            disp = warp_displacement(x0, amp) - cw[:, interior, :].reshape(-1, 3)
            if jit != 0.0:
                gid = (offsets[sel][:, None] + loc[interior][None, :]).reshape(-1).astype(np.uint64) * np.uint64(3)
                for cc in range(3):
                    disp[:, cc] += uniform_pm1(seed, 2, gid + np.uint64(cc)) * (jit / p)
            c[:, interior, :] += disp.reshape(len(sel), -1, 3)
        idx = (offsets[sel][:, None] + loc[None, :]).reshape(-1)
        values[idx] = c.reshape(-1, 3)
    values32 = values.astype(np.float32)
    field = field_function(values32.astype(np.float64)).astype(np.float32)
    return HexMesh(n=n, order=0, ctrl_idx=np.arange(total, dtype=np.int32), values=values32, field=field,
                   orders=el_order, offsets=offsets)


def split_by_order(mesh: HexMesh):
    """A mixed-order mesh as dray holds it, one uniform-order mesh per order (dray's UnstructuredMesh has
    one order, data_model/unstructured_mesh.hpp:48; a mixed mesh is a Collection of such domains):
    [(order, HexMesh of that order's elements in their original order, their original element ids)].  The
    domains share the control-point arrays."""
    if mesh.orders is None:
        return [(mesh.order, mesh, np.arange(mesh.n_el, dtype=np.int32))]
    out = []
    for p in np.unique(mesh.orders):
        els = np.nonzero(mesh.orders == p)[0].astype(np.int32)
        npe = (int(p) + 1) ** 3
        idx = (mesh.offsets[els][:, None] + np.arange(npe)[None, :]).reshape(-1)
        out.append((int(p), HexMesh(n=mesh.n, order=int(p), ctrl_idx=mesh.ctrl_idx[idx].astype(np.int32),
                                    values=mesh.values, field=mesh.field), els))
    return out


def query_points(n_pts: int, bounds6, seed: int = 0) -> np.ndarray:
    """Uniform points in the box [xmin,xmax]x[ymin,ymax]x[zmin,zmax] from the counter hash (float32 [n,3])."""
    b = np.asarray(bounds6, dtype=np.float64)
    cnt = np.arange(n_pts, dtype=np.uint64) * np.uint64(3)
    out = np.empty((n_pts, 3), dtype=np.float32)
    for c in range(3):
        u = (uniform_pm1(seed, 11, cnt + np.uint64(c)) + 1.0) * 0.5
        out[:, c] = (b[2 * c] + u * (b[2 * c + 1] - b[2 * c])).astype(np.float32)
    return out


# ---- cameras --------------------------------------------------------------------------------
RAY_DTYPE = np.dtype([("dir", np.float32, 3), ("near", np.float32), ("orig", np.float32, 3),
                      ("far", np.float32), ("pixel_id", np.int32)])           # src/dray/ray.hpp:20-28
PARTIAL_DTYPE = np.dtype([("pixel_id", np.int32), ("depth", np.float32), ("rgba", np.float32, 4)])  # volume_partial.hpp:16-21
LOCATION_DTYPE = np.dtype([("cell", np.int32), ("ref", np.float32, 3)])       # src/dray/location.hpp:16-21
assert RAY_DTYPE.itemsize == 36 and PARTIAL_DTYPE.itemsize == 24 and LOCATION_DTYPE.itemsize == 16


def orthographic_rays(w: int, h: int, z0: float = -1.0) -> np.ndarray:
    """Rays along +z with origins at pixel centres over [0,1]^2 (pixel_id = j*w + i)."""
    r = np.zeros(w * h, dtype=RAY_DTYPE)
    i = np.arange(w * h) % w
    j = np.arange(w * h) // w
    r["orig"][:, 0] = ((i + 0.5) / w).astype(np.float32)
    r["orig"][:, 1] = ((j + 0.5) / h).astype(np.float32)
    r["orig"][:, 2] = np.float32(z0)
    r["dir"][:, 2] = 1.0
    r["dir"][:, 0] = np.float32(1e-7)   # camera.cpp:472-475 nudges zero components
    r["dir"][:, 1] = np.float32(1e-7)
    r["near"] = 0.0
    r["far"] = np.inf
    r["pixel_id"] = np.arange(w * h, dtype=np.int32)
    return r


def grey_to_red_table():
    """Linear grey (0.5,0.5,0.5) -> red (1,0,0) transfer function, smooth, one constant alpha.

    Returns (rgb_points [n,5] = pos,r,g,b,a ; alpha_points [n,2]).
    """
    rgb = np.array([[0.0, 0.5, 0.5, 0.5, 1.0], [1.0, 1.0, 0.0, 0.0, 1.0]], dtype=np.float32)
    return rgb


def table_alpha_for_corrected(opacity: float, samples: int) -> float:
    """Table alpha a such that 1-(1-a)^(10/samples) == opacity (per-sample corrected opacity)."""
    return 1.0 - (1.0 - opacity) ** (samples / 10.0)


# ---- furnace query points ------------------------------------------------------------------------
_MINSTD_A = 48271
_MINSTD_M = 2147483647


def minstd_outputs(n: int, seed: int = 0) -> np.ndarray:
    """The first n outputs of std::minstd_rand (x_{k+1} = 48271 x_k mod 2^31 - 1; a seed of 0 mod m starts
    the state at 1, as the standard's linear_congruential_engine does for an increment of 0), uint64.
    Computed in blocks: x_k = x_0 a^k mod m with a^k = a^(jB) a^(i+1) for k = jB + i + 1."""
    m, a = _MINSTD_M, _MINSTD_A
    x0 = seed % m or 1
    B = 1 << 16
    pw = np.empty(B, np.uint64)
    v = 1
    for i in range(B):
        v = v * a % m
        pw[i] = v                      # a^(i+1)
    nb = -(-n // B)
    starts = np.empty(nb, np.uint64)
    s, aB = x0, pow(a, B, m)
    for j in range(nb):
        starts[j] = s                  # x_0 a^(jB)
        s = s * aB % m
    out = (starts[:, None] * pw[None, :]) % np.uint64(m)
    return out.reshape(-1)[:n]


def furnace_points(n_pts: int, bounds6) -> np.ndarray:
    """The furnace point-location driver's query points (utilities/furnace/point_location.cpp:49-68): one
    std::minstd_rand seeded 0 and three std::uniform_real_distribution<float> over the bounds' axes, drawn
    x, y, z per point.  libstdc++'s uniform_real_distribution<float> takes one engine output per draw
    (generate_canonical<float, 24>: u = float(x - 1) / float(2^31 - 2), where float(2^31 - 2) = 2^31, and
    1 - 2^-24 if that rounds to 1) and
    returns u * (b - a) + a in float32.  float32 [n, 3]."""
    b = np.asarray(bounds6, np.float32)
    x = minstd_outputs(3 * n_pts).reshape(n_pts, 3)
    u = (x - np.uint64(1)).astype(np.float32) / np.float32(float(np.float32(_MINSTD_M - 1)))
    u = np.where(u >= np.float32(1), np.nextafter(np.float32(1), np.float32(0)), u).astype(np.float32)
    out = np.empty((n_pts, 3), np.float32)
    for c in range(3):
        lo, hi = b[2 * c], b[2 * c + 1]
        out[:, c] = u[:, c] * (hi - lo) + lo
    return out

"""Python mirror of dray's host API for this path, over the libdrx C ABI.

Names and argument meaning follow the reference (src/dray/...): ``LinearBVHBuilder.construct``
(linear_bvh_builder.hpp:20-21), ``UnstructuredMesh.locate`` (data_model/unstructured_mesh.cpp:123-148),
``UnstructuredField.eval`` (data_model/unstructured_field.cpp:213-246), ``Camera.create_rays``
(rendering/camera.cpp:208-243), ``ColorTable`` (color_table.cpp), ``Volume.integrate`` / ``Volume.save``
(rendering/volume.cpp:274-416).  Device memory comes from torch (ROCm); every computation runs in the
HIP kernels of libdrx.so and raises ``DrxError`` when the library or the GPU is unavailable.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib as L
from .synthetic import LOCATION_DTYPE, PARTIAL_DTYPE, RAY_DTYPE, HexMesh

_initialised = set()


def _dev() -> torch.device:
    if not torch.cuda.is_available():
        raise L.DrxError("no GPU visible: libdrx runs on gfx950 only")
    d = torch.cuda.current_device()
    if d not in _initialised:
        L.check(L.lib().drx_init(d), "drx_init")
        _initialised.add(d)
    return torch.device("cuda", d)


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def to_device(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to(_dev())


def as_bytes(a: np.ndarray) -> torch.Tensor:
    """A structured numpy array as a uint8 device tensor."""
    return to_device(np.ascontiguousarray(a).view(np.uint8).reshape(-1))


def from_bytes(t: torch.Tensor, dtype, n: int) -> np.ndarray:
    return t[: n * np.dtype(dtype).itemsize].cpu().numpy().view(dtype)[:n].copy()


# ---------------------------------------------------------------------------------------------
class BVH:
    """dray::BVH (bvh.hpp:15-39) in device memory, plus the fused leaf records used for traversal."""

    def __init__(self, inner, leaf_nodes, aabb_ids, bounds, subrefs=None, mcodes=None):
        self.inner_nodes = inner
        self.leaf_nodes = leaf_nodes
        self.aabb_ids = aabb_ids
        self.bounds = np.asarray(bounds, np.float32)
        self.subrefs = subrefs
        self.mcodes = mcodes
        self.leaf_pack = None
        self.wide_nodes = None
        self.leaf_starts = None
        self.qnodes = None          # compact traversal nodes (drx_bvh_quantize) + exact leaf boxes + grid
        self.leaf_boxes = None
        self.qgrid = np.zeros(6, np.float32)
        self.use_wide = True
        self.use_starts = True
        self.use_qnodes = True

    @property
    def n_leaves(self) -> int:
        return int(self.leaf_nodes.numel())

    def view(self) -> L.DrxBVH:
        v = L.DrxBVH()
        v.inner_nodes = self.inner_nodes.data_ptr()
        v.leaf_nodes = self.leaf_nodes.data_ptr()
        v.aabb_ids = self.aabb_ids.data_ptr()
        v.subrefs = self.subrefs.data_ptr() if self.subrefs is not None else None
        v.leaf_pack = self.leaf_pack.data_ptr() if self.leaf_pack is not None else None
        v.wide_nodes = self.wide_nodes.data_ptr() if (self.wide_nodes is not None and self.use_wide) else None
        v.n_leaves = self.n_leaves
        for q in range(6):
            v.bounds[q] = float(self.bounds[q])
        v.leaf_starts = self.leaf_starts.data_ptr() if (self.leaf_starts is not None and self.use_starts) else None
        if self.qnodes is not None and self.use_qnodes:
            v.qnodes = self.qnodes.data_ptr()
            v.leaf_boxes = self.leaf_boxes.data_ptr()
            for q in range(6):
                v.qgrid[q] = float(self.qgrid[q])
        return v

    def quantize(self) -> None:
        """Compact traversal nodes (drx_bvh_quantize): 32 bytes per inner node instead of dray's 64, plus each
        leaf's exact box; binary walks give the same candidates in the same order with them."""
        dev = _dev()
        self.qnodes = torch.empty(max(self.n_leaves - 1, 1) * 8, dtype=torch.int32, device=dev)
        self.leaf_boxes = torch.empty(self.n_leaves * 8, dtype=torch.float32, device=dev)
        grid = (C.c_float * 6)()
        v = self.view()
        L.check(L.lib().drx_bvh_quantize(C.byref(v), _ptr(self.qnodes), _ptr(self.leaf_boxes),
                                         C.cast(grid, C.c_void_p), _stream()),
                "drx_bvh_quantize")
        self.qgrid = np.array(list(grid), np.float32)

    def compute_leaf_starts(self, mesh_view: L.DrxMesh) -> None:
        """Per-leaf Newton start maps (drx_bvh_leaf_starts) from the geometry the BVH was built on."""
        self.leaf_starts = torch.empty(self.n_leaves * 12, dtype=torch.float32, device=_dev())
        v = self.view()
        L.check(L.lib().drx_bvh_leaf_starts(C.byref(mesh_view), C.byref(v), _ptr(self.leaf_starts), _stream()),
                "drx_bvh_leaf_starts")

    def pack_leaves(self) -> None:
        dev = _dev()
        self.leaf_pack = torch.empty(self.n_leaves * 8, dtype=torch.float32, device=dev)
        v = self.view()
        L.check(L.lib().drx_bvh_pack_leaves(C.byref(v), _ptr(self.leaf_pack), _stream()), "drx_bvh_pack_leaves")

    def ensure_wide(self) -> None:
        """The 8-wide copy, built on first use (the default walk does not read it)."""
        if self.wide_nodes is None:
            self.widen()

    def widen(self) -> None:
        """Build the 8-wide traversal copy (drx_bvh_widen): count, allocate, fill."""
        v = self.view()
        cnt = C.c_int64(0)
        L.check(L.lib().drx_bvh_widen(C.byref(v), None, 0, C.byref(cnt), _stream()), "drx_bvh_widen(count)")
        self.wide_nodes = torch.empty(int(cnt.value) * 64, dtype=torch.float32, device=_dev())
        L.check(L.lib().drx_bvh_widen(C.byref(v), _ptr(self.wide_nodes), int(cnt.value), C.byref(cnt), _stream()),
                "drx_bvh_widen")


class LinearBVHBuilder:
    """LinearBVHBuilder::construct(aabbs, prim_ids) (linear_bvh_builder.cpp:399-518)."""

    def construct(self, aabbs: torch.Tensor, prim_ids: Optional[torch.Tensor] = None, want_mcodes=False) -> BVH:
        dev = _dev()
        n = int(aabbs.numel() // 6)
        if prim_ids is None:
            prim_ids = torch.arange(n, dtype=torch.int32, device=dev)
        m = max(n, 2)
        inner = torch.empty((m - 1) * 16, dtype=torch.float32, device=dev)
        leaf_nodes = torch.empty(m, dtype=torch.int32, device=dev)
        aabb_ids = torch.empty(m, dtype=torch.int32, device=dev)
        mcodes = torch.empty(m, dtype=torch.int32, device=dev) if want_mcodes else None
        bounds = (C.c_float * 6)()
        L.check(L.lib().drx_bvh_build(_ptr(aabbs) if n else None, _ptr(prim_ids) if n else None, n, _ptr(inner),
                                      _ptr(leaf_nodes), _ptr(aabb_ids), C.cast(bounds, C.c_void_p), _ptr(mcodes),
                                      _stream()), "drx_bvh_build")
        return BVH(inner, leaf_nodes, aabb_ids, list(bounds), mcodes=mcodes)


# ---------------------------------------------------------------------------------------------
class DeviceGridFunction:
    """dray GridFunction (ctrl_idx + values) uploaded, plus the element-local packed blocks."""

    def __init__(self, ctrl_idx: np.ndarray, values: np.ndarray, ncomp: int, order: int,
                 orders: Optional[np.ndarray] = None, offsets: Optional[np.ndarray] = None, fp64: bool = False):
        """fp64 (scalar fields only): keep the values in double (dray's Float = double build) and pack
        FP64 blocks (drx_mesh_pack_f64)."""
        dev = _dev()
        if fp64 and ncomp != 1:
            raise L.DrxError("FP64 blocks are for scalar fields only")
        self.fp64 = fp64
        self.ncomp = ncomp
        self.order = int(order) if orders is None else int(orders.max())
        self.n_el = int(len(orders)) if orders is not None else int(len(ctrl_idx) // (order + 1) ** 3)
        self.ctrl_idx = to_device(ctrl_idx.astype(np.int32))
        self.values = to_device(values.astype(np.float64 if fp64 else np.float32).reshape(-1))
        self.orders = to_device(orders.astype(np.int32)) if orders is not None else None
        self.offsets = to_device(offsets.astype(np.int64)) if offsets is not None else None
        lib = L.lib()
        if orders is None:
            if not (1 <= order <= L.MAX_ORDER):
                raise L.DrxError(f"order {order} outside 1..{L.MAX_ORDER}")
            self.stride = lib.drx_mesh_stride(order, ncomp)
            total = self.stride * self.n_el
            self.pack_offsets = None
        else:
            h_off = np.zeros(self.n_el, np.int64)
            o32 = np.ascontiguousarray(orders.astype(np.int32))
            total = lib.drx_mesh_pack_layout(o32.ctypes.data, self.n_el, ncomp, h_off.ctypes.data)
            if total < 0:
                raise L.DrxError(L.lib().drx_last_error().decode(errors="replace"))
            self.stride = 0
            self.pack_offsets = to_device(h_off)
        if fp64:
            self.packed = torch.empty(max(int(total), 1), dtype=torch.float64, device=dev)
            L.check(lib.drx_mesh_pack_f64(_ptr(self.ctrl_idx), _ptr(self.values), _ptr(self.orders),
                                          _ptr(self.offsets), _ptr(self.pack_offsets), int(order), self.n_el,
                                          _ptr(self.packed), _stream()), "drx_mesh_pack_f64")
            return
        self.packed = torch.empty(max(int(total), 1), dtype=torch.float32, device=dev)
        L.check(lib.drx_mesh_pack(_ptr(self.ctrl_idx), _ptr(self.values), ncomp, _ptr(self.orders), _ptr(self.offsets),
                                  _ptr(self.pack_offsets), int(order), self.n_el, _ptr(self.packed), _stream()),
                "drx_mesh_pack")

    def view(self) -> L.DrxMesh:
        v = L.DrxMesh()
        if self.fp64:
            v.data = None
            v.data64 = self.packed.data_ptr()
        else:
            v.data = self.packed.data_ptr()
        v.elem_offset = self.pack_offsets.data_ptr() if self.pack_offsets is not None else None
        v.elem_order = self.orders.data_ptr() if self.orders is not None else None
        v.n_el = self.n_el
        v.order = self.order
        v.ncomp = self.ncomp
        v.stride = self.stride
        return v


@dataclass
class LocateOptions:
    fp64_polish: bool = False
    cond_threshold: float = 1e4
    traversal: int = 0          # 0 auto (= 3), 1 per lane (8-wide), 2 wave-coherent (8-wide), 3 per lane (binary)
    coherent_order: bool = False  # locate: process queries in Morton order
    mlp: Optional[torch.Tensor] = None   # packed learned initializer (LearnedInitializer.pack())
    mlp_newton_steps: int = 4            # Newton budget from the learned guess before the plain fallback
    start: str = "affine"   # plain Newton start: "affine" (corner-affine estimate) or "centre" (SubRef centre)
    launch: int = 0         # streamed walks: 0 tuned (drx_locate_autotune) else chunked, 1 chunked, 2 persistent;
                            # renders: 0 sample-parallel below 300 active rays per SIMD, 3 always, 1/2 never
    chunk: int = 0          # points per wavefront chunk (0: tuned, else 1024)
    mlp_fallback: str = "affine"   # learned start: the plain start that takes over, "affine" or "centre"
    guess: Optional[torch.Tensor] = None    # given starts (Location records): DRX_INIT_GUESS
    record: Optional[torch.Tensor] = None   # integrate: Location of every (ray, sample)
    samples_per_ray: int = 0                # integrate: guess / record entries per ray

    def c(self, steps=None, cands=None, polished=None) -> L.DrxLocateOpts:
        o = L.DrxLocateOpts()
        modes = {"affine": L.DRX_INIT_AFFINE, "centre": L.DRX_INIT_SUBREF_CENTRE}
        if self.start not in modes:
            raise ValueError(f"LocateOptions.start must be 'affine' or 'centre', not {self.start!r}")
        o.init_mode = modes[self.start]
        o.fp64_polish = 1 if self.fp64_polish else 0
        o.cond_threshold = float(self.cond_threshold)
        o.newton_steps = steps.data_ptr() if steps is not None else None
        o.candidates = cands.data_ptr() if cands is not None else None
        o.polished = polished.data_ptr() if polished is not None else None
        o.traversal = int(self.traversal)
        o.coherent_order = 1 if self.coherent_order else 0
        o.launch = int(self.launch)
        o.chunk = int(self.chunk)
        if self.mlp is not None:
            o.init_mode = 1
            o.mlp = self.mlp.data_ptr()
            o.mlp_newton_steps = int(self.mlp_newton_steps)
            fb = {"affine": 0, "centre": 1}
            if self.mlp_fallback not in fb:
                raise ValueError(f"LocateOptions.mlp_fallback must be 'affine' or 'centre', not {self.mlp_fallback!r}")
            o.mlp_fallback = fb[self.mlp_fallback]
        if self.guess is not None:
            o.init_mode = L.DRX_INIT_GUESS
            o.guess = self.guess.data_ptr()
        o.record = self.record.data_ptr() if self.record is not None else None
        o.samples_per_ray = int(self.samples_per_ray)
        return o


# BVHs of at most this many leaves get their 8-wide copy when they are built (the automatic walk of a render
# may take the wave walk over it); larger ones build it when a walk asks for it.
WIDE_EAGER_MAX_LEAVES = 1 << 16


class UnstructuredMesh:
    """UnstructuredMesh<Hex> (data_model/unstructured_mesh.hpp:117-135): geometry + lazily built BVH."""

    def __init__(self, mesh: HexMesh):
        self.host = mesh
        self.gf = DeviceGridFunction(mesh.ctrl_idx, mesh.values, 3, mesh.order, mesh.orders, mesh.offsets)
        self._bvh: Optional[BVH] = None

    @property
    def n_el(self) -> int:
        return self.gf.n_el

    def split_elements(self):
        """construct_bvh's per-element split: (aabbs [33 n_el, 6], prim_ids, subrefs [33 n_el, 6])."""
        dev = _dev()
        k = self.n_el * L.SPLITS_PER_ELEMENT
        aabbs = torch.empty(k * 6, dtype=torch.float32, device=dev)
        prims = torch.empty(k, dtype=torch.int32, device=dev)
        subrefs = torch.empty(k * 6, dtype=torch.float32, device=dev)
        L.check(L.lib().drx_bvh_split_elements(_ptr(self.gf.ctrl_idx), _ptr(self.gf.values), _ptr(self.gf.orders),
                                               _ptr(self.gf.offsets), self.host.order, self.n_el, _ptr(aabbs),
                                               _ptr(prims), _ptr(subrefs), _stream()), "drx_bvh_split_elements")
        return aabbs, prims, subrefs

    def get_bvh(self) -> BVH:
        """UnstructuredMesh::get_bvh (unstructured_mesh.cpp:23-31): built on first use, then cached."""
        if self._bvh is None:
            aabbs, prims, subrefs = self.split_elements()
            bvh = LinearBVHBuilder().construct(aabbs, prims)
            bvh.subrefs = subrefs
            bvh.pack_leaves()
            bvh.compute_leaf_starts(self.gf.view())
            bvh.quantize()
            if bvh.n_leaves <= WIDE_EAGER_MAX_LEAVES:
                bvh.widen()   # small BVHs: the automatic walk of a render may take the wave walk
            self._bvh = bvh
        return self._bvh

    def bounds(self) -> np.ndarray:
        return self.get_bvh().bounds

    def locate(self, points: torch.Tensor, opts: Optional[LocateOptions] = None, stats: bool = False,
               polished: bool = False):
        """Mesh::locate: points float32 [n,3] (device) -> Location records (uint8 tensor, 16 B each).
        stats: also per-point Newton steps and candidate counts; polished: also the per-point flag of the
        FP64 correction (appended last)."""
        bvh = self.get_bvh()
        n = int(points.numel() // 3)
        dev = _dev()
        out = torch.empty(max(n, 1) * 16, dtype=torch.uint8, device=dev)
        steps = torch.zeros(max(n, 1), dtype=torch.int32, device=dev) if stats else None
        cands = torch.zeros(max(n, 1), dtype=torch.int32, device=dev) if stats else None
        pol = torch.zeros(max(n, 1), dtype=torch.int32, device=dev) if polished else None
        if opts is not None and opts.traversal in (1, 2):
            bvh.ensure_wide()
        o = (opts or LocateOptions()).c(steps, cands, pol)
        mv, bv = self.gf.view(), bvh.view()
        L.check(L.lib().drx_locate(C.byref(mv), C.byref(bv), _ptr(points), n, _ptr(out), C.byref(o), _stream()),
                "drx_locate")
        res = (out, steps, cands) if stats else (out,)
        if polished:
            res = res + (pol,)
        return res if len(res) > 1 else out

    def locate_autotune(self, points: torch.Tensor, opts: Optional[LocateOptions] = None) -> dict:
        """drx_locate_autotune: time every chunk size x launch form of the streamed walk on these points and
        options, and make the fastest the default (launch 0, chunk 0) for this device, order, polish and walk."""
        n = int(points.numel() // 3)
        out = torch.empty(max(n, 1) * 16, dtype=torch.uint8, device=_dev())
        if opts is not None and opts.traversal in (1, 2):
            self.get_bvh().ensure_wide()
        o = (opts or LocateOptions()).c()
        launch, chunk = C.c_int32(0), C.c_int32(0)
        ms = (C.c_float * 10)()
        mv, bv = self.gf.view(), self.get_bvh().view()
        L.check(L.lib().drx_locate_autotune(C.byref(mv), C.byref(bv), _ptr(points), n, _ptr(out), C.byref(o),
                                            C.byref(launch), C.byref(chunk), C.cast(ms, C.c_void_p), _stream()),
                "drx_locate_autotune")
        chunks = (256, 512, 1024, 2048, 4096)
        return {"launch": {1: "chunked", 2: "persistent"}[launch.value], "chunk": chunk.value,
                "ms": {f"{'persistent' if l else 'chunked'}/{c}": float(ms[2 * i + l])
                       for i, c in enumerate(chunks) for l in (0, 1)}}


class UnstructuredField:
    """UnstructuredField<Hex scalar> (data_model/unstructured_field.cpp)."""

    def __init__(self, mesh: HexMesh, values: Optional[np.ndarray] = None, fp64: bool = False, value_range=None):
        """fp64: an FP64 field (dray's Float = double): values default to mesh.field64, kept and evaluated
        in double; its colour range is the double min/max.  value_range: the range to report instead of the
        values' min/max (a field degree-elevated at import keeps the range of its original coefficients)."""
        if values is None:
            values = mesh.field64 if fp64 else mesh.field
            if values is None:
                raise L.DrxError("UnstructuredField(fp64=True) needs FP64 values (mesh.field64)")
        vals = values
        self.host_values = vals
        self.fp64 = fp64
        self.gf = DeviceGridFunction(mesh.ctrl_idx, vals, 1, mesh.order, mesh.orders, mesh.offsets, fp64=fp64)
        # the device copy is a snapshot of vals, so its range is fixed here (not recomputed per render)
        v = np.asarray(vals, np.float64 if fp64 else np.float32)
        self._range = (float(v.min()), float(v.max())) if v.size else (float("inf"), float("-inf"))
        if value_range is not None:
            self._range = (float(value_range[0]), float(value_range[1]))

    def range(self):
        """get_range (unstructured_field.cpp:35-90): min/max over the control values (in the field's
        precision: float32, or double for an FP64 field)."""
        return self._range

    def eval(self, locations: torch.Tensor, values: Optional[torch.Tensor] = None) -> torch.Tensor:
        n = int(locations.numel() // 16)
        if values is None or values.numel() != n:
            values = torch.zeros(max(n, 1), dtype=torch.float32, device=_dev())
        fv = self.gf.view()
        L.check(L.lib().drx_field_eval(C.byref(fv), _ptr(locations), n, _ptr(values), _stream()), "drx_field_eval")
        return values


# ---------------------------------------------------------------------------------------------
class DataSet:
    """dray::DataSet (data_model/data_set.hpp:20-59): one domain's mesh and its named fields."""

    def __init__(self, mesh: "UnstructuredMesh"):
        self.mesh = mesh
        self._fields = {}

    def add_field(self, name: str, field: "UnstructuredField"):
        self._fields[name] = field

    def has_field(self, name: str) -> bool:
        return name in self._fields

    def field(self, name: str) -> "UnstructuredField":
        if name not in self._fields:
            raise L.DrxError(f"DataSet: no field named '{name}'")
        return self._fields[name]

    def fields(self):
        return list(self._fields)


class Collection:
    """dray::Collection (data_model/collection.hpp:16-55) on one rank: every domain is local."""

    def __init__(self):
        self.domains = []

    def add_domain(self, ds: DataSet):
        self.domains.append(ds)

    def domain(self, i: int) -> DataSet:
        return self.domains[i]

    def size(self) -> int:
        return len(self.domains)

    def has_field(self, name: str) -> bool:
        """True when every domain has the field (Collection::has_field)."""
        return len(self.domains) > 0 and all(d.has_field(name) for d in self.domains)

    def range(self, name: str):
        """Union of the domains' field ranges (Collection::range)."""
        rs = [d.field(name).range() for d in self.domains if d.has_field(name)]
        if not rs:
            raise L.DrxError(f"Collection: no domain has field '{name}'")
        return float(np.float32(min(r[0] for r in rs))), float(np.float32(max(r[1] for r in rs)))

    def bounds(self) -> np.ndarray:
        """Union of the domains' mesh bounds (Collection::bounds), [xmin, xmax, ymin, ymax, zmin, zmax]."""
        b = np.stack([d.mesh.bounds() for d in self.domains]).astype(np.float32)
        return np.array([b[:, 0].min(), b[:, 1].max(), b[:, 2].min(), b[:, 3].max(), b[:, 4].min(), b[:, 5].max()],
                        np.float32)

    def field_list(self) -> str:
        names = sorted({n for d in self.domains for n in d.fields()})
        return "[" + " ".join(names) + "]"


class BlueprintReader:
    """dray::BlueprintReader (io/blueprint_reader.cpp:280-440) for Conduit-Blueprint HDF5 data sets of
    high-order hex meshes: every domain the root file names becomes a DataSet of an UnstructuredMesh and its
    scalar fields (converted to Bernstein form and to one order per domain on the host)."""

    @staticmethod
    def load(root_file: str) -> Collection:
        from . import blueprint
        col = Collection()
        for d in blueprint.load(root_file):
            ds = DataSet(UnstructuredMesh(d.mesh))
            ds.domain_id = d.domain_id
            for name in d.fields:
                ds.add_field(name, UnstructuredField(d.mesh, values=d.fields[name], value_range=d.ranges[name]))
            col.add_domain(ds)
        return col


def collection_volumes(collection: Collection, field: str) -> list:
    """dray::Volume (Collection&) + field(name) as one Volume per domain: every domain samples with the
    Collection's bounds (volume.cpp:89-91) and colours with the Collection's field range; set the colour
    table, samples and lighting on each (or share one ColorTable object)."""
    if not collection.has_field(field):
        raise L.DrxError(f"Volume: field '{field}' is not in every domain {collection.field_list()}")
    bounds = collection.bounds()
    rng = collection.range(field)
    vols = []
    for d in collection.domains:
        v = Volume(d.mesh, d.field(field))
        v.sample_bounds = bounds
        v.range = rng
        vols.append(v)
    return vols


class PointLocation:
    """dray::PointLocation (queries/point_location.cpp:176-287) on one rank: every domain locates every
    point; a point located in a domain takes that domain's field values (later domains overwrite
    earlier ones), the others keep empty_val."""

    def __init__(self):
        self.m_vars = []
        self.m_empty_val = 0.0

    def add_var(self, var: str):
        self.m_vars.append(var)

    def empty_val(self, v: float):
        self.m_empty_val = float(v)

    def execute(self, collection: Collection, points: torch.Tensor) -> dict:
        """points: float32 device tensor of n*3 values -> {"points", "values" (one device tensor per valid
        variable), "vars", "empty_val"}.  Raises DrxError as dray does for missing variables or points."""
        if not self.m_vars:
            raise L.DrxError("PointLocation: must specify at least 1 variables.")
        n = int(points.numel() // 3)
        if n == 0:
            raise L.DrxError("PointLocation: must specify at least 1 point.")
        valid = [v for v in self.m_vars if collection.has_field(v)]
        if not valid:
            raise L.DrxError("PointLocation: must specify at least one valid field. Known fields "
                             + collection.field_list())
        values = [torch.full((n,), self.m_empty_val, dtype=torch.float32, device=_dev()) for _ in valid]
        for ds in collection.domains:
            locs = ds.mesh.locate(points)
            cells = locs.view(torch.int32).view(n, 4)[:, 0]
            if not bool((cells != -1).any()):
                continue
            for f, name in enumerate(valid):
                ds.field(name).eval(locs, values[f])
        return {"points": points, "values": values, "vars": list(self.m_vars), "empty_val": self.m_empty_val}


# ---------------------------------------------------------------------------------------------
class ColorTable:
    """dray::ColorTable (color_table.cpp): rgb + alpha control points and a smooth flag.  ColorTable() is
    empty and not smooth (:67-71); ColorTable(name) holds the named table's rgb points, smooth except
    "sharp", "" or "default" naming cool2warm and an unknown name falling back to YlOrRd (:378-1330)."""

    def __init__(self, name: Optional[str] = None):
        self.rgb = []      # (pos, r, g, b, a)
        self.alpha = []    # (pos, alpha)
        self.smooth = False
        if name is not None:
            n, sm = C.c_int32(0), C.c_int32(0)
            L.check(L.lib().drx_color_table_preset(name.encode(), None, 0, C.byref(n), C.byref(sm)),
                    "drx_color_table_preset")
            pts = np.zeros((n.value, 5), np.float32)
            L.check(L.lib().drx_color_table_preset(name.encode(), pts.ctypes.data, n.value, C.byref(n),
                                                   C.byref(sm)), "drx_color_table_preset")
            self.rgb = pts.tolist()
            self.smooth = bool(sm.value)

    def add_point(self, pos, rgb):
        c = list(rgb) + ([1.0] if len(rgb) == 3 else [])
        self.rgb.append([float(pos)] + [float(x) for x in c])

    def add_alpha(self, pos, a):
        self.alpha.append([float(pos), float(a)])

    def set_smooth(self, s: bool):
        self.smooth = bool(s)

    def sample(self, num_samples=1024, opacity_factor=-1.0, alpha_scale=1.0) -> np.ndarray:
        """correct_opacity(opacity_factor) (<=0: none) then sample(num_samples) -> RGBA [n,4] float32."""
        rgb = np.ascontiguousarray(np.asarray(self.rgb, np.float32).reshape(-1, 5))
        al = np.ascontiguousarray(np.asarray(self.alpha, np.float32).reshape(-1, 2))
        out = np.zeros((num_samples, 4), np.float32)
        L.check(L.lib().drx_color_table_sample(rgb.ctypes.data if rgb.size else None, rgb.shape[0],
                                               al.ctypes.data if al.size else None, al.shape[0],
                                               1 if self.smooth else 0, float(opacity_factor), num_samples,
                                               float(alpha_scale), out.ctypes.data), "drx_color_table_sample")
        return out


class Camera:
    """dray::Camera (rendering/camera.cpp): perspective rays via gen_perspective."""

    def __init__(self, width=500, height=500, fov=30.0, pos=(0, 0, 0), look_at=(0, 0, -1), up=(0, 1, 0), zoom=1.0):
        self.width, self.height, self.fov = int(width), int(height), float(fov)
        self.pos, self.look_at, self.up, self.zoom = tuple(pos), tuple(look_at), tuple(up), float(zoom)

    def c(self) -> L.DrxCamera:
        c = L.DrxCamera()
        c.width, c.height, c.fov_y, c.zoom = self.width, self.height, self.fov, self.zoom
        for q in range(3):
            c.pos[q], c.look_at[q], c.up[q] = self.pos[q], self.look_at[q], self.up[q]
        return c

    def _take(self, cc: L.DrxCamera):
        self.pos = tuple(float(cc.pos[q]) for q in range(3))
        self.look_at = tuple(float(cc.look_at[q]) for q in range(3))
        self.fov = float(cc.fov_y)

    def reset_to_bounds(self, bounds6, xpad: float = 0.0, ypad: float = 0.0, zpad: float = 0.0):
        """Camera::reset_to_bounds (camera.cpp:485-514): bounds (xmin, xmax, ymin, ymax, zmin, zmax)."""
        cc = self.c()
        b = np.ascontiguousarray(np.asarray(bounds6, np.float32).reshape(6))
        pads = np.ascontiguousarray(np.array([xpad, ypad, zpad], np.float64))
        L.check(L.lib().drx_camera_reset_to_bounds(C.byref(cc), b.ctypes.data, pads.ctypes.data),
                "drx_camera_reset_to_bounds")
        self._take(cc)

    def elevate(self, degrees: float):
        """Camera::elevate (camera.cpp:516-534)."""
        cc = self.c()
        L.check(L.lib().drx_camera_elevate(C.byref(cc), float(degrees)), "drx_camera_elevate")
        self._take(cc)

    def azimuth(self, degrees: float):
        """Camera::azimuth (camera.cpp:536-552)."""
        cc = self.c()
        L.check(L.lib().drx_camera_azimuth(C.byref(cc), float(degrees)), "drx_camera_azimuth")
        self._take(cc)

    def create_rays(self) -> torch.Tensor:
        out = torch.empty(self.width * self.height * RAY_DTYPE.itemsize, dtype=torch.uint8, device=_dev())
        cc = self.c()
        L.check(L.lib().drx_camera_rays(C.byref(cc), _ptr(out), _stream()), "drx_camera_rays")
        return out


def remove_missed_rays(rays: torch.Tensor, bounds6) -> tuple:
    """remove_missed_rays (ray.cpp:102-115) -> (active rays uint8 tensor, count)."""
    n = int(rays.numel() // RAY_DTYPE.itemsize)
    out = torch.empty(max(n, 1) * RAY_DTYPE.itemsize, dtype=torch.uint8, device=_dev())
    b = np.ascontiguousarray(np.asarray(bounds6, np.float32))
    cnt = C.c_int32(0)
    L.check(L.lib().drx_remove_missed_rays(_ptr(rays), n, b.ctypes.data, _ptr(out), C.byref(cnt), _stream()),
            "drx_remove_missed_rays")
    return out, int(cnt.value)


def ray_tile_order(rays: torch.Tensor, n: int, width: int, height: int) -> torch.Tensor:
    order = torch.empty(max(n, 1), dtype=torch.int32, device=_dev())
    L.check(L.lib().drx_ray_tile_order(_ptr(rays), n, width, height, _ptr(order), _stream()), "drx_ray_tile_order")
    return order


def sample_distance(bounds6, samples: int) -> float:
    """|bounds diagonal| / samples in float32 (volume.cpp:89-91)."""
    b = np.asarray(bounds6, np.float32)
    d = np.array([b[1] - b[0], b[3] - b[2], b[5] - b[4]], np.float32)
    s = np.float32(0)
    for x in d:
        s = np.float32(s + np.float32(x * x))
    return float(np.float32(np.sqrt(s)) / np.float32(samples))


class Volume:
    """dray::Volume (rendering/volume.hpp:25-60) for one hex-mesh domain and one scalar field."""

    def __init__(self, mesh: UnstructuredMesh, field: UnstructuredField):
        self.mesh = mesh
        self.field_ = field
        self.m_samples = 100
        self.use_lighting = True   # dray default (VolumeShader::shaded_color with the given lights)
        # ColorMap's default table (color_map.cpp:11-18) with the alpha ramp Volume's constructor adds
        # (volume.cpp:274-286)
        self.table = ColorTable("cool2warm")
        self.table.add_alpha(0.1, 0.0)
        self.table.add_alpha(1.0, 0.7)
        self.range = None
        self.opts = LocateOptions()
        self.last_stats = None
        self.sample_bounds = None   # dray's m_bounds (the Collection's bounds) for the sample distance
        self._lut = (None, None)    # (colour-table state, device LUT): resampled only when the state changes

    def samples(self, n: int):
        self.m_samples = int(n)

    def color_map_range(self, lo: float, hi: float):
        self.range = (float(lo), float(hi))

    def integrate(self, rays: torch.Tensor, tile_width: int = 0, tile_height: int = 0, active=None, lights=None):
        """Volume::integrate(rays, lights) (volume.cpp:323-351) -> (compacted partials uint8 tensor, count).
        lights: PointLights (13 floats each: pos, amb, diff, spec, spec_pow), used when use_lighting; with
        no lights a lit sample contributes nothing, as in dray (the Renderer passes default_light)."""
        lt = np.zeros((0, 13), np.float32) if lights is None else np.asarray(lights, np.float32).reshape(-1, 13)
        if self.use_lighting and lt.shape[0] > L.MAX_LIGHTS:
            raise L.DrxError(f"Volume::integrate: at most {L.MAX_LIGHTS} lights")
        larr = (L.DrxPointLight * max(1, lt.shape[0]))()
        for i, row in enumerate(lt):
            larr[i].pos[:] = row[0:3].tolist()
            larr[i].amb[:] = row[3:6].tolist()
            larr[i].diff[:] = row[6:9].tolist()
            larr[i].spec[:] = row[9:12].tolist()
            larr[i].spec_pow = float(row[12])
        shading = L.DrxShading(1 if self.use_lighting else 0, int(lt.shape[0]), larr)
        bounds = self.mesh.bounds()
        if active is None:
            active, n_act = remove_missed_rays(rays, bounds)
        else:
            active, n_act = active
        lo, hi = self.range if self.range is not None else self.field_.range()
        key = (tuple(map(tuple, self.table.rgb)), tuple(map(tuple, self.table.alpha)), self.table.smooth,
               self.m_samples)
        if self._lut[0] != key:
            self._lut = (key, to_device(self.table.sample(1024, opacity_factor=10.0 / self.m_samples).reshape(-1)))
        lut = self._lut[1]
        sd = sample_distance(bounds if self.sample_bounds is None else self.sample_bounds, self.m_samples)
        order = ray_tile_order(active, n_act, tile_width, tile_height) if (tile_width and n_act) else None
        out = torch.empty(max(5 * n_act, 1) * PARTIAL_DTYPE.itemsize, dtype=torch.uint8, device=_dev())
        cnt = C.c_int32(0)
        st = L.DrxIntegrateStats()
        if self.opts.traversal in (1, 2):
            self.mesh.get_bvh().ensure_wide()
        o = self.opts.c()
        mv, fv, bv = self.mesh.gf.view(), self.field_.gf.view(), self.mesh.get_bvh().view()
        L.check(L.lib().drx_volume_integrate(C.byref(mv), C.byref(fv), C.byref(bv), _ptr(active), n_act, _ptr(lut),
                                             1024, lo, hi, sd, _ptr(order), C.byref(o), C.byref(shading), _ptr(out),
                                             C.byref(cnt), C.byref(st), _stream()), "drx_volume_integrate")
        self.last_stats = st
        return out, int(cnt.value)


def default_light(camera: "Camera") -> np.ndarray:
    """detail::default_light (rendering/renderer.cpp:141-159): 13 floats (pos, amb, diff, spec, spec_pow)."""
    cc = camera.c()
    out = L.DrxPointLight()
    L.check(L.lib().drx_default_light(C.byref(cc), C.byref(out)), "drx_default_light")
    return np.array(list(out.pos) + list(out.amb) + list(out.diff) + list(out.spec) + [out.spec_pow], np.float32)


def composite_domains(domain_partials, width: int, height: int) -> torch.Tensor:
    """Renderer::render's cross-domain compositing: [(partials uint8 tensor, count), ...] of several
    domains -> RGBA float32 [w*h*4] (partials sorted by pixel and depth, blended front to back)."""
    parts = [p[: n * PARTIAL_DTYPE.itemsize] for p, n in domain_partials if n > 0]
    n = sum(int(q.numel()) for q in parts) // PARTIAL_DTYPE.itemsize
    cat = torch.cat(parts) if parts else torch.empty(PARTIAL_DTYPE.itemsize, dtype=torch.uint8, device=_dev())
    img = torch.empty(width * height * 4, dtype=torch.float32, device=_dev())
    L.check(L.lib().drx_composite_domains(_ptr(cat), n, width, height, _ptr(img), _stream()), "drx_composite_domains")
    return img


def composite(partials: torch.Tensor, n: int, width: int, height: int) -> torch.Tensor:
    """Volume::save's compositing + Framebuffer::composite_background -> RGBA float32 [w*h*4]."""
    img = torch.empty(width * height * 4, dtype=torch.float32, device=_dev())
    L.check(L.lib().drx_composite(_ptr(partials), n, width, height, _ptr(img), _stream()), "drx_composite")
    return img


class Framebuffer:
    """dray::Framebuffer (rendering/framebuffer.cpp): RGBA colours (4 floats per pixel) and depths on the
    device, pixel id = x + y * width with row 0 at the bottom; cleared to transparent black and +inf."""

    def __init__(self, width: int = 1024, height: int = 1024):
        if width <= 0 or height <= 0:
            raise L.DrxError("Framebuffer: width and height must be positive")
        self.width, self.height = int(width), int(height)
        n = self.width * self.height
        self.colors = torch.empty(n * 4, dtype=torch.float32, device=_dev())
        self.depths = torch.empty(n, dtype=torch.float32, device=_dev())
        self.bg_color = (1.0, 1.0, 1.0, 1.0)
        self.fg_color = (0.0, 0.0, 0.0, 1.0)
        self.clear()

    def clear(self, color=(0.0, 0.0, 0.0, 0.0)):
        c = np.ascontiguousarray(np.asarray(color, np.float32).reshape(4))
        L.check(L.lib().drx_framebuffer_clear(_ptr(self.colors), _ptr(self.depths), self.width * self.height,
                                              c.ctypes.data, _stream()), "drx_framebuffer_clear")

    def background_color(self, color):
        self.bg_color = tuple(float(x) for x in color)

    def composite_background(self):
        c = np.ascontiguousarray(np.asarray(self.bg_color, np.float32).reshape(4))
        L.check(L.lib().drx_framebuffer_composite_background(_ptr(self.colors), self.width * self.height,
                                                             c.ctypes.data, _stream()),
                "drx_framebuffer_composite_background")

    def save(self, name: str):
        """Framebuffer::save: writes name + ".png" (8-bit RGBA, row 0 at the bottom of the picture)."""
        host = np.ascontiguousarray(self.colors.cpu().numpy())
        L.check(L.lib().drx_save_png(host.ctypes.data, self.width, self.height, (name + ".png").encode()),
                "drx_save_png")


def partials_to_framebuffer(domain_partials, fb: Framebuffer, blend: bool = False):
    """apcomp PartialCompositor::composite + detail::partials_to_framebuffer (renderer.cpp:97-138,
    307-330): [(partials uint8 tensor, count), ...] of any number of domains -> fb's colours and depths."""
    parts = [p[: n * PARTIAL_DTYPE.itemsize] for p, n in domain_partials if n > 0]
    n = sum(int(q.numel()) for q in parts) // PARTIAL_DTYPE.itemsize
    if n == 0:
        return
    cat = torch.cat(parts) if len(parts) > 1 else parts[0]
    L.check(L.lib().drx_partials_to_framebuffer(_ptr(cat), n, fb.width, fb.height, _ptr(fb.colors), _ptr(fb.depths),
                                                1 if blend else 0, _stream()), "drx_partials_to_framebuffer")


class Renderer:
    """dray::Renderer::render's volume branch (rendering/renderer.cpp:229-337): one Volume per domain;
    every domain integrates the camera's rays with the renderer's lights (default_light(camera) when
    none were added), and the partials of all domains are composited by (pixel, depth) into a fresh
    Framebuffer.  With several domains, volumes whose sample bounds or colour-map range are unset take
    the union over the domains (dray's Volume uses its Collection's bounds and field range)."""

    def __init__(self):
        self.volumes = []
        self.lights = []
        # dray draws a colour bar (and optionally world annotations) over the image (renderer.cpp:166-186);
        # the switches are kept, nothing is drawn
        self.m_color_bar = True
        self.m_world_annotations = False

    def color_bar(self, on: bool):
        self.m_color_bar = bool(on)

    def world_annotations(self, on: bool):
        self.m_world_annotations = bool(on)

    def volume(self, volumes):
        self.volumes = list(volumes) if isinstance(volumes, (list, tuple)) else [volumes]

    def add_light(self, light):
        self.lights.append(np.asarray(light, np.float32).reshape(13))

    def clear_lights(self):
        self.lights = []

    def render(self, camera: Camera) -> Framebuffer:
        fb = Framebuffer(camera.width, camera.height)
        if not self.volumes:
            return fb
        rays = camera.create_rays()
        lights = self.lights if self.lights else [default_light(camera)]
        if len(self.volumes) > 1:
            b = np.stack([v.mesh.bounds() for v in self.volumes]).astype(np.float32)
            union_b = np.array([b[:, 0].min(), b[:, 1].max(), b[:, 2].min(), b[:, 3].max(), b[:, 4].min(),
                                b[:, 5].max()], np.float32)
            rg = [v.field_.range() for v in self.volumes]
            union_r = (min(r[0] for r in rg), max(r[1] for r in rg))
            for v in self.volumes:
                if v.sample_bounds is None:
                    v.sample_bounds = union_b
                if v.range is None:
                    v.range = union_r
        parts = [v.integrate(rays, tile_width=camera.width, tile_height=camera.height, lights=lights)
                 for v in self.volumes]
        partials_to_framebuffer(parts, fb, blend=False)
        return fb


# ---------------------------------------------------------------------------------------------
class LearnedInitializer:
    """Learned Newton initial guess: MLP 27 -> 64 x4 (GELU) -> 3, BF16 on MFMA (drx_mlp_* in libdrx).

    Not part of dray: dray starts Newton from LocateHack::eval_inverse's guess
    (data_model/device_mesh.hpp:131-165).  Where Newton from the learned guess has not converged within its
    budget, the plain method's Newton from its own start (the corner-affine map) decides.
    """

    def __init__(self, seed: int = 0):
        dev = _dev()
        lib = L.lib()
        self.state = torch.empty(int(lib.drx_mlp_state_bytes()), dtype=torch.uint8, device=dev)
        self.n_params = int(lib.drx_mlp_param_count())
        self.steps_taken = 0
        L.check(lib.drx_mlp_init(_ptr(self.state), int(seed), _stream()), "drx_mlp_init")

    def params(self) -> torch.Tensor:
        """fp32 master parameters (W1 [64,32], b1, W2..W4 [64,64], b2..b4, W5 [16,64], b5 [16]) - a view."""
        return self.state[: 4 * self.n_params].view(torch.float32)

    @staticmethod
    def features(mesh: "UnstructuredMesh", points: torch.Tensor, locations: torch.Tensor):
        """(x bf16 [n,32], target float32 [n,3] = located ref - base, base float32 [n,3])."""
        n = int(points.numel() // 3)
        dev = _dev()
        x = torch.empty((max(n, 1), 32), dtype=torch.bfloat16, device=dev)
        t = torch.empty((max(n, 1), 3), dtype=torch.float32, device=dev)
        b = torch.empty((max(n, 1), 3), dtype=torch.float32, device=dev)
        mv = mesh.gf.view()
        L.check(L.lib().drx_mlp_features(C.byref(mv), _ptr(points), _ptr(locations), n, _ptr(x), _ptr(t), _ptr(b),
                                         _stream()), "drx_mlp_features")
        return x[:n], t[:n], b[:n]

    def train(self, x: torch.Tensor, target: torch.Tensor, steps: int, batch: int = 16384, lr: float = 2e-3,
              lr_final: float = 2e-5, total_steps: int = 0, beta1: float = 0.9, beta2: float = 0.999,
              eps: float = 1e-8, record_loss: bool = True):
        """Run `steps` Adam steps (continuing the schedule of earlier calls); returns per-step MSE or None."""
        n = int(x.shape[0])
        o = L.DrxMlpTrainOpts()
        o.batch, o.steps, o.step0 = int(batch), int(steps), int(self.steps_taken)
        o.total_steps = int(total_steps) if total_steps else int(self.steps_taken + steps)
        o.sample0 = (self.steps_taken * int(batch)) % max(n, 1)
        o.lr, o.lr_final, o.beta1, o.beta2, o.eps = float(lr), float(lr_final), float(beta1), float(beta2), float(eps)
        loss = torch.zeros(max(steps, 1), dtype=torch.float32, device=_dev()) if record_loss else None
        L.check(L.lib().drx_mlp_train(_ptr(self.state), _ptr(x), _ptr(target), n, C.byref(o), _ptr(loss), _stream()),
                "drx_mlp_train")
        self.steps_taken += int(steps)
        return loss[:steps] if record_loss else None

    def pack(self) -> torch.Tensor:
        out = torch.empty(int(L.lib().drx_mlp_packed_bytes()), dtype=torch.uint8, device=_dev())
        L.check(L.lib().drx_mlp_pack(_ptr(self.state), _ptr(out), _stream()), "drx_mlp_pack")
        return out

    @staticmethod
    def infer(packed: torch.Tensor, x: torch.Tensor, base: Optional[torch.Tensor] = None) -> torch.Tensor:
        n = int(x.shape[0])
        out = torch.empty((max(n, 1), 3), dtype=torch.float32, device=_dev())
        L.check(L.lib().drx_mlp_infer(_ptr(packed), _ptr(x), _ptr(base), n, _ptr(out), _stream()), "drx_mlp_infer")
        return out[:n]

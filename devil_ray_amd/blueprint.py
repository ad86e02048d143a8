"""Conduit-Blueprint MFEM data sets as dray's BlueprintReader loads them (io/blueprint_reader.cpp:280-415) and
converts them to Bernstein form (mfem2dray.cpp:392-1160), for hexahedral meshes.

A root file names the domain files (`file_pattern`, `number_of_files`).  Each domain holds an unstructured
hex topology (the first one listed, as dray takes `child_names()[0]`) whose vertex connectivity defines an
MFEM mesh, the high-order node field named by the topology's `grid_function`, and fields.  Fields whose
name contains "_attribute" and fields with more than one component are skipped (dray imports the latter as
vectors, which the volume path does not render).

H1 degrees of freedom follow MFEM's numbering for a mesh built from that connectivity: vertices, then edges,
then faces, then element interiors, where edges and faces are numbered in the order the elements first reach
them (`Mesh::GetVertexToVertexTable` / `GetElementToEdgeTable` push each element's 12 edges, 0-1 1-2 3-2 0-3
4-5 5-6 7-6 4-7 0-4 1-5 2-6 3-7, into a table keyed by the vertex pair; `GetElementToFaceTable` pushes its 6
faces, 3210 0154 1265 2376 3047 4567, keyed by the three lowest vertex ids).  An element's local dofs are
its 8 vertices, 12 edges, 6 faces and its interior, in that order; with one dof per entity (orders 1 and 2)
no entity orientation enters.  L2 dofs are lexicographic per element (i fastest).

dray projects every non-positive basis onto MFEM's positive (Bernstein) basis of the same order
(`project_to_pos_basis`, `GridFunction::ProjectGridFunction`); within one space that projection is
interpolation, so here the nodal values at the basis' 1D nodes (GaussLobatto: 0, 1/2, 1 for order 2;
GaussLegendre: the Gauss points) are solved for Bernstein coefficients axis by axis.  The GPU kernels
evaluate mesh and field at one order, so a domain's mesh and fields are degree-elevated to the highest order
among them (exact in exact arithmetic).
"""
from __future__ import annotations

import os
import re
from math import comb
from typing import Optional

import numpy as np

from .hdf5_lite import H5File
from .synthetic import HexMesh

# MFEM's reference hexahedron: vertex coordinates, edge and face vertex lists
_HEX_V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
_HEX_E = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7), (0, 4), (1, 5), (2, 6), (3, 7)]
_HEX_F = [(3, 2, 1, 0), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7)]

GAUSS_LEGENDRE, GAUSS_LOBATTO, POSITIVE = 0, 1, 2   # MFEM BasisType values


def _local_lattice(order: int) -> np.ndarray:
    """Lattice index i + (p+1) (j + (p+1) k) of each local dof of an order-1 or order-2 MFEM H1 hex."""
    if order == 1:
        return np.array([v[0] + 2 * (v[1] + 2 * v[2]) for v in _HEX_V])
    if order != 2:
        raise ValueError(f"H1 order {order}: orders 1 and 2 are supported")
    pos = [2 * _HEX_V[c] for c in range(8)]
    pos += [_HEX_V[a] + _HEX_V[b] for a, b in _HEX_E]
    pos += [_HEX_V[list(f)].sum(axis=0) // 2 for f in _HEX_F]
    pos += [np.array([1, 1, 1])]
    return np.array([p[0] + 3 * (p[1] + 3 * p[2]) for p in pos])


def mfem_h1_element_dofs(conn: np.ndarray, n_vertices: int, order: int = 2):
    """Global H1 dof of each element's local dofs (vertices[, edges, faces, interior]) -> ([n_el, (p+1)^3],
    number of dofs)."""
    conn = np.asarray(conn, np.int64).reshape(-1, 8)
    n_el = len(conn)
    if order == 1:
        return conn.copy(), n_vertices
    if order != 2:
        raise ValueError(f"H1 order {order}: orders 1 and 2 are supported")
    edge_id, face_id = {}, {}
    el_edges = np.empty((n_el, 12), np.int64)
    el_faces = np.empty((n_el, 6), np.int64)
    for e in range(n_el):
        v = conn[e]
        for j, (a, b) in enumerate(_HEX_E):
            key = (v[a], v[b]) if v[a] < v[b] else (v[b], v[a])
            el_edges[e, j] = edge_id.setdefault(key, len(edge_id))
    for e in range(n_el):
        v = conn[e]
        for j, f in enumerate(_HEX_F):
            key = tuple(sorted(v[list(f)])[:3])
            el_faces[e, j] = face_id.setdefault(key, len(face_id))
    ne, nf = len(edge_id), len(face_id)
    dofs = np.empty((n_el, 27), np.int64)
    dofs[:, :8] = conn
    dofs[:, 8:20] = n_vertices + el_edges
    dofs[:, 20:26] = n_vertices + ne + el_faces
    dofs[:, 26] = n_vertices + ne + nf + np.arange(n_el)
    return dofs, n_vertices + ne + nf + n_el


def mfem_h1_q2_element_dofs(conn: np.ndarray, n_vertices: int):
    return mfem_h1_element_dofs(conn, n_vertices, 2)


def nodes_1d(btype: int, p: int) -> np.ndarray:
    """The 1D nodes on [0, 1] of MFEM's nodal bases of order p."""
    if p == 0:
        return np.array([0.5])
    if btype == GAUSS_LEGENDRE:
        x, _ = np.polynomial.legendre.leggauss(p + 1)
        return 0.5 * (x + 1.0)
    if btype == GAUSS_LOBATTO:
        inner = np.polynomial.legendre.Legendre.basis(p).deriv().roots() if p > 1 else np.array([])
        return 0.5 * (np.concatenate([[-1.0], np.sort(inner.real), [1.0]]) + 1.0)
    raise ValueError(f"basis type {btype} has no nodes")


def nodal_to_bernstein(nodal: np.ndarray, p: int, nodes: np.ndarray) -> np.ndarray:
    """[..., (p+1)^3] values at the tensor nodes (lexicographic, i fastest) -> Bernstein coefficients: on each
    axis solve sum_j B_j^p(x_i) c_j = f_i (interpolation in the same space)."""
    m = np.array([[comb(p, j) * x ** j * (1.0 - x) ** (p - j) for j in range(p + 1)] for x in nodes])
    minv = np.linalg.inv(m)
    a = np.asarray(nodal, np.float64).reshape(nodal.shape[:-1] + (p + 1, p + 1, p + 1))   # [..., k, j, i]
    for axis in (-1, -2, -3):
        a = np.moveaxis(np.tensordot(a, minv, axes=([axis], [1])), -1, axis)
    return a.reshape(nodal.shape)


def lagrange_q2_to_bernstein(nodal: np.ndarray) -> np.ndarray:
    """[..., 27] values at the 3^3 lattice (i + 3 (j + 3 k), nodes 0, 1/2, 1) -> Bernstein coefficients:
    b0 = f(0), b2 = f(1), b1 = 2 f(1/2) - (f(0) + f(1)) / 2 per axis."""
    a = np.asarray(nodal, np.float64).reshape(nodal.shape[:-1] + (3, 3, 3))   # [..., k, j, i]
    for axis in (-1, -2, -3):
        f0 = np.take(a, 0, axis=axis)
        f1 = np.take(a, 1, axis=axis)
        f2 = np.take(a, 2, axis=axis)
        b1 = 2.0 * f1 - 0.5 * (f0 + f2)
        a = np.stack([f0, b1, f2], axis=axis)
    return a.reshape(nodal.shape)


def elevate_bernstein_hex(coef: np.ndarray, p: int, q: int) -> np.ndarray:
    """Degree elevation of tensor Bernstein coefficients [..., (p+1)^3] (lexicographic) from order p to q."""
    a = np.asarray(coef, np.float64).reshape(coef.shape[:-1] + (p + 1, p + 1, p + 1))
    for _ in range(p, q):
        for axis in (-1, -2, -3):
            n = a.shape[axis] - 1   # current degree along this axis
            parts = []
            for i in range(n + 2):
                lo = np.take(a, i - 1, axis=axis) if i >= 1 else 0.0
                hi = np.take(a, i, axis=axis) if i <= n else 0.0
                parts.append((i / (n + 1)) * lo + (1 - i / (n + 1)) * hi)
            a = np.stack(parts, axis=axis)
    return a.reshape(coef.shape[:-1] + ((q + 1) ** 3,))


def parse_basis(name: str, what: str = "field"):
    """MFEM collection name -> (kind "H1" | "L2", basis type, order)."""
    if name == "Quadratic":
        return "H1", GAUSS_LOBATTO, 2
    if name == "Linear":
        return "H1", GAUSS_LOBATTO, 1
    m = re.fullmatch(r"H1Pos_3D_P(\d+)", name)
    if m:
        return "H1", POSITIVE, int(m.group(1))
    m = re.fullmatch(r"(H1|L2)(?:_T(\d+))?_3D_P(\d+)", name)
    if not m:
        raise ValueError(f"{what}: basis '{name}' is not supported")
    kind = m.group(1)
    btype = int(m.group(2)) if m.group(2) is not None else (GAUSS_LOBATTO if kind == "H1" else GAUSS_LEGENDRE)
    return kind, btype, int(m.group(3))


def _to_bernstein(nodal: np.ndarray, btype: int, p: int) -> np.ndarray:
    if btype == POSITIVE:
        return np.asarray(nodal, np.float64)
    if btype == GAUSS_LOBATTO and p == 2:
        return lagrange_q2_to_bernstein(nodal)
    if btype == GAUSS_LOBATTO and p == 1:
        return np.asarray(nodal, np.float64)
    return nodal_to_bernstein(nodal, p, nodes_1d(btype, p))


class Domain:
    """One domain: a HexMesh in Bernstein form (per-element control points, no sharing), its scalar fields
    (float32 as dray's Float, plus the double coefficients) at the mesh's order, and the domain id."""

    def __init__(self, mesh: HexMesh, fields: dict, fields64: dict, domain_id: int = 0, topology: str = "",
                 ranges: Optional[dict] = None):
        self.mesh = mesh
        self.fields = fields
        self.fields64 = fields64
        # min/max of each field's float32 coefficients at its own order (degree elevation can pull interior
        # maxima in; dray's range is taken from the field as imported)
        self.ranges = ranges if ranges is not None else {
            k: (float(v.min()), float(v.max())) for k, v in fields.items()}
        self.domain_id = domain_id
        self.topology = topology

    def mesh_with_field(self, name: str) -> HexMesh:
        return HexMesh(n=self.mesh.n, order=self.mesh.order, ctrl_idx=self.mesh.ctrl_idx, values=self.mesh.values,
                       field=self.fields[name], field64=self.fields64[name])


def _components(h: H5File, path: str) -> list:
    """The component arrays under path/values (a group of x0 x1 x2, v0 v1 v2, x y z ...) in listed order."""
    if h.is_group(f"{path}/values"):
        return [np.asarray(h.read(f"{path}/values/{c}"), np.float64) for c in h.keys(f"{path}/values")]
    return [np.asarray(h.read(f"{path}/values"), np.float64)]


def load_domain(path: str) -> Domain:
    h = H5File(path)
    topos = h.keys("/topologies")
    if not topos:
        raise ValueError(f"{path}: no topologies")
    tname = topos[0]
    topo = f"/topologies/{tname}"
    if h.read(f"{topo}/type") != "unstructured" or h.read(f"{topo}/elements/shape") != "hex":
        raise ValueError(f"{path}: only unstructured hex topologies are supported")
    conn = h.read(f"{topo}/elements/connectivity").astype(np.int64).reshape(-1, 8)
    n_el = len(conn)
    cset = h.read(f"{topo}/coordset")
    verts = _components(h, f"/coordsets/{cset}")
    n_vertices = len(verts[0])
    tkeys = h.keys(topo)
    gf = h.read(f"{topo}/grid_function") if "grid_function" in tkeys else None

    # mesh nodes: the grid function (H1), else the vertices (order 1)
    if gf is not None:
        kind, btype, mp = parse_basis(h.read(f"/fields/{gf}/basis"), f"{path}:{gf}")
        if kind != "H1":
            raise ValueError(f"{path}: mesh nodes in an {kind} space are not supported")
        comps = _components(h, f"/fields/{gf}")
    else:
        kind, btype, mp, comps = "H1", GAUSS_LOBATTO, 1, verts
    if len(comps) != 3:
        raise ValueError(f"{path}: {len(comps)} node components (3 supported)")
    dofs, ndofs = mfem_h1_element_dofs(conn, n_vertices, mp)
    if any(len(c) != ndofs for c in comps):
        raise ValueError(f"{path}: {len(comps[0])} node dofs, the numbering gives {ndofs}")
    lat = _local_lattice(mp)
    npe = (mp + 1) ** 3
    nodes = np.empty((3, n_el, npe))
    for c in range(3):
        nodes[c][:, lat] = comps[c][dofs]
    mesh_b = [_to_bernstein(nodes[c], btype, mp) for c in range(3)]

    # scalar fields
    raw = {}
    for name in h.keys("/fields"):
        if name == gf or "_attribute" in name or not h.is_group(f"/fields/{name}"):
            continue
        fk = h.keys(f"/fields/{name}")
        if "basis" not in fk:
            continue
        comps_f = _components(h, f"/fields/{name}")
        if len(comps_f) != 1:
            continue   # vectors
        k2, b2, p2 = parse_basis(h.read(f"/fields/{name}/basis"), f"{path}:{name}")
        v = comps_f[0]
        if k2 == "L2":
            if len(v) != n_el * (p2 + 1) ** 3:
                raise ValueError(f"{path}:{name}: {len(v)} values for {n_el} order-{p2} elements")
            el = v.reshape(n_el, -1)
        else:
            d2, nd2 = mfem_h1_element_dofs(conn, n_vertices, p2)
            if len(v) != nd2:
                raise ValueError(f"{path}:{name}: {len(v)} dofs, the numbering gives {nd2}")
            el = np.empty((n_el, (p2 + 1) ** 3))
            el[:, _local_lattice(p2)] = v[d2]
        raw[name] = (_to_bernstein(el, b2, p2), p2)

    q = max([mp] + [p for _, p in raw.values()])
    values = np.stack([elevate_bernstein_hex(mesh_b[c], mp, q) for c in range(3)], axis=2).reshape(-1, 3)
    fields, fields64, ranges = {}, {}, {}
    for name, (c, p) in raw.items():
        c32 = c.astype(np.float32)
        ranges[name] = (float(c32.min()), float(c32.max()))
        fields64[name] = elevate_bernstein_hex(c, p, q).reshape(-1)
        fields[name] = fields64[name].astype(np.float32)
    npq = (q + 1) ** 3
    mesh = HexMesh(n=0, order=q, ctrl_idx=np.arange(n_el * npq, dtype=np.int32), values=values.astype(np.float32),
                   field=np.zeros(n_el * npq, np.float32))
    did = 0
    if "state" in h.keys("/") and "domain_id" in h.keys("/state"):
        did = int(np.asarray(h.read("/state/domain_id")).reshape(-1)[0])
    return Domain(mesh, fields, fields64, did, tname, ranges)


def load(root_path: str) -> list:
    """BlueprintReader::load: every domain file a root file names, in file order."""
    r = H5File(root_path)
    pattern = r.read("/file_pattern")
    n_files = int(np.asarray(r.read("/number_of_files")).reshape(-1)[0])
    base = os.path.dirname(os.path.abspath(root_path))
    return [load_domain(os.path.join(base, pattern % i)) for i in range(n_files)]

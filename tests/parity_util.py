"""Helpers shared by the parity tests (not collected by pytest)."""
from __future__ import annotations

import numpy as np

BAND = 1e-4     # dray's acceptance band: ref coordinates in [-1e-4, 1 + 1e-4] (element.hpp:481-484)
DELTA = 1e-6    # a located-id difference is explained only within this distance (plus the FP32 solution's own
                # error, at most FP32_REF_ERR) of the band edge
FP32_REF_ERR = 1e-5   # bound on an FP32 reference point's error against the FP64 map (checked on its own)


def dfs_leaves(inner: np.ndarray, pt: np.ndarray) -> list:
    """The leaves whose boxes contain pt, in DeviceMesh::locate's visiting order (device_mesh.hpp:262-320):
    depth-first, left child before right, a child entered when pt is inside its box (inclusive bounds).
    inner: dray's emitted nodes [n-1, 16] float32 (child codes: 4 x inner index, or -(leaf + 1))."""
    ch = np.ascontiguousarray(inner[:, 12:14]).view(np.int32)
    p = np.asarray(pt, np.float32)
    out = []

    def visit(k):
        for side in (0, 1):
            lo = inner[k, 6 * side:6 * side + 3]
            hi = inner[k, 6 * side + 3:6 * side + 6]
            if np.all(p >= lo) and np.all(p <= hi):
                c = int(ch[k, side])
                if c < 0:
                    out.append(-c - 1)
                else:
                    visit(c // 4)

    visit(0)
    return out


def band_edge_distance(ref64: np.ndarray) -> np.ndarray:
    """Distance of FP64 reference points [n, 3] to the nearest edge of the acceptance band (per point)."""
    r = np.asarray(ref64, np.float64)
    return np.minimum(np.abs(r + BAND), np.abs(r - (1 + BAND))).min(axis=1)


def explain_mismatches(orc, mesh, ob, pts, gc, gr, oc, orf, delta=DELTA):
    """Points where the GPU (gc, gr) and the FP32 oracle (oc, orf) located different elements.  Both walk
    the same candidates (BVH `ob`) in the same order, so one side accepted an element the other rejected.
    A difference is explained in one of two ways.

    Band edge: FP32 rounding decides acceptance only for a point whose exact (FP64) reference point lies
    at the acceptance band edge to within FP32 error: in an element one side accepted, the FP64 inverse
    map (iterated from that side's reference point) lies within delta + e of the band edge, where e, the
    FP32 error scale at the point, is the larger of the two sides' |r32 - r64|inf for the elements they
    accepted, capped at FP32_REF_ERR.

    Exact answer: the GPU's element is the one the FP64 oracle (the same walk in double) locates.  The FP32
    oracle evaluates absolute control-point coordinates, the GPU element-centred ones, so on small
    high-order elements the GPU's FP32 error is the smaller (the FP32 oracle can accept an element by a
    margin the exact map does not give).

    Returns (mismatch indices, explained mask, edge distance minus the allowance, mask of those explained
    by agreeing with the FP64 oracle)."""
    mism = np.nonzero(gc != oc)[0]
    if len(mism) == 0:
        return mism, np.zeros(0, bool), np.zeros(0), np.zeros(0, bool)
    edge = np.full(len(mism), np.inf)
    err = np.zeros(len(mism))
    for cells, refs in ((gc[mism], gr[mism]), (oc[mism], orf[mism])):
        ok = cells >= 0
        if ok.any():
            r32 = refs[ok].astype(np.float64)
            ex = orc.refine_f64(mesh, pts[mism][ok], cells[ok], r32)
            edge[ok] = np.minimum(edge[ok], band_edge_distance(ex))
            err[ok] = np.maximum(err[ok], np.minimum(np.abs(r32 - ex).max(axis=1), FP32_REF_ERR))
    excess = edge - (delta + err)
    oc64, _ = orc.locate(mesh, ob, pts[mism], dtype=np.float64)
    exact = gc[mism] == oc64
    return mism, (excess <= 0) | exact, excess, exact

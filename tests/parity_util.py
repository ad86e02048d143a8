"""Helpers shared by the parity tests (not collected by pytest)."""
from __future__ import annotations

import numpy as np

BAND = 1e-4     # dray's acceptance band: ref coordinates in [-1e-4, 1 + 1e-4] (element.hpp:481-484)
DELTA = 1e-6    # a located-id difference is explained only within this distance of the band edge


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


def explain_mismatches(orc, mesh, pts, gc, gr, oc, orf, delta=DELTA):
    """Points where the GPU (gc, gr) and the oracle (oc, orf) located different elements.  Each side walks
    the same candidates in the same order, so one side rejected the element the other accepted; the
    difference is explained when the point lies within `delta` of the acceptance band edge of one of the two
    elements (FP64 inverse map, iterated from the side's own reference point, or the centre for an element
    that side did not accept).  Returns (mismatch indices, explained mask, edge distances)."""
    mism = np.nonzero(gc != oc)[0]
    if len(mism) == 0:
        return mism, np.zeros(0, bool), np.zeros(0)
    edge = np.full(len(mism), np.inf)
    for cells, refs in ((gc[mism], gr[mism]), (oc[mism], orf[mism])):
        ok = cells >= 0
        if ok.any():
            ex = orc.refine_f64(mesh, pts[mism][ok], cells[ok], refs[ok].astype(np.float64))
            edge[ok] = np.minimum(edge[ok], band_edge_distance(ex))
    return mism, edge < delta, edge

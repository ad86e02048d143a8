"""CPU tests that pin the oracle (oracle/drx_oracle.cpp) before it is trusted as the parity checker.

Each check is chosen so that a plausible mistake (dropped term, swapped index, wrong sign, transposed
operand, wrong tie rule) fails: the reference's own known answers (t_dray_external_evals.cpp,
t_dray_matrix.cpp), closed forms, independent re-derivations and brute force on tiny inputs.
"""
import math

import numpy as np
import pytest

from devil_ray_amd import synthetic as syn
from parity_util import band_edge_distance, dfs_leaves


# ---------------------------------------------------------------- element evaluation (a13)
@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_eval_d_iota_known_answer(orc, p):
    """t_dray_external_evals.cpp:48 control values iota(100...): the element is the affine map
    100 + p u + p(p+1) v + p(p+1)^2 w (sum_i i B_i^p(t) = p t), so value and derivatives are exact."""
    npe = (p + 1) ** 3
    dofs = np.arange(100, 100 + npe, dtype=np.float64).reshape(-1, 1)
    g = np.linspace(0, 1, p + 1)
    pts = [(a, b, c) for a in g for b in g for c in g] + [(0.3, 0.4, 0.9), (0.123, 0.77, 0.5)]
    for r in pts:
        v, d = orc.eval_d(dofs, p, r)
        exp = 100 + p * r[0] + p * (p + 1) * r[1] + p * (p + 1) ** 2 * r[2]
        assert abs(v[0] - exp) < 1e-9
        assert np.allclose(d[:, 0], [p, p * (p + 1), p * (p + 1) ** 2], atol=1e-9)
        v32, d32 = orc.eval_d(dofs, p, r, dtype=np.float32)
        # reference tolerances 1e-4 (value) / 5e-4 (derivative), plus 8 float32 ulps of the magnitude
        assert abs(v32[0] - exp) < 1e-4 + 8 * np.spacing(np.float32(exp))
        assert np.allclose(d32[:, 0], d[:, 0], atol=5e-4)


def _bernstein(p, t):
    return np.array([math.comb(p, i) * t ** i * (1 - t) ** (p - i) for i in range(p + 1)])


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_eval_d_matches_direct_bernstein_sum(orc, p):
    """Independent tensor Bernstein sum with dof index i + (p+1)(j + (p+1)k) (u fastest) and
    central finite differences for the derivatives."""
    rng = np.random.default_rng(p)
    npe = (p + 1) ** 3
    dofs = rng.normal(size=(npe, 3))
    for _ in range(20):
        r = rng.random(3)
        v, d = orc.eval_d(dofs, p, r)
        bu, bv, bw = _bernstein(p, r[0]), _bernstein(p, r[1]), _bernstein(p, r[2])
        W = np.einsum("k,j,i->kji", bw, bv, bu).reshape(-1)
        assert np.allclose(v, W @ dofs, atol=1e-12)
        h = 1e-6
        for a in range(3):
            e = np.zeros(3)
            e[a] = h
            fd = (orc.eval_d(dofs, p, r + e)[0] - orc.eval_d(dofs, p, r - e)[0]) / (2 * h)
            assert np.allclose(d[a], fd, atol=1e-6)


# ---------------------------------------------------------------- LUP (a14)
def test_lup_known_answer(orc):
    """t_dray_matrix.cpp:31-41: inverse of [[1,0,5],[2,1,6],[3,4,0]] = [[-24,20,-5],[18,-15,4],[5,-4,1]]."""
    A = np.array([[1, 0, 5], [2, 1, 6], [3, 4, 0]], float)
    inv = np.array([[-24, 20, -5], [18, -15, 4], [5, -4, 1]], float)
    for dt, tol in ((np.float64, 1e-12), (np.float32, 1e-4)):
        cols = [orc.lup_solve(A, e, dtype=dt)[0] for e in np.eye(3)]
        assert np.allclose(np.stack(cols, axis=1), inv, atol=tol)


def test_lup_random_and_pivoting(orc):
    rng = np.random.default_rng(5)
    for _ in range(200):
        A = rng.normal(size=(3, 3)) + 3 * np.eye(3)
        b = rng.normal(size=3)
        x, valid = orc.lup_solve(A, b)
        assert valid and np.allclose(x, np.linalg.solve(A, b), atol=1e-10)
    A = np.array([[0, 2, 1], [1, 0, 0], [0, 0, 3.0]])   # zero leading pivot: needs the row swap
    x, valid = orc.lup_solve(A, [1, 2, 3])
    assert valid and np.allclose(A @ x, [1, 2, 3])
    x, valid = orc.lup_solve(np.zeros((3, 3)), [1, 2, 3])
    assert not valid                                     # pivot below epsilon<T>() flags the solve


# ---------------------------------------------------------------- Morton / clz (a4, a6)
def _interleave(x, y, z):
    c = 0
    for b in range(10):
        c |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
    return c


def test_morton_known_answers(orc):
    assert orc.lib().orc_morton_3d(0.0, 0.0, 0.0) == 0
    assert orc.lib().orc_morton_3d(1.0, 1.0, 1.0) == 0x3FFFFFFF          # clamps to 1023 per axis
    assert orc.lib().orc_morton_3d(2.0, -1.0, float("nan")) == _interleave(1023, 0, 0)
    rng = np.random.default_rng(0)
    for _ in range(500):
        x, y, z = rng.random(3).astype(np.float32)
        q = [min(max(int(np.float32(v) * np.float32(1024)), 0), 1023) for v in (x, y, z)]
        assert orc.lib().orc_morton_3d(float(x), float(y), float(z)) == _interleave(*q)


def test_clz(orc):
    rng = np.random.default_rng(1)
    for x in [0, 1, 2, 3, 0x80000000, 0xFFFFFFFF] + [int(v) for v in rng.integers(0, 2 ** 32, 200)]:
        assert orc.lib().orc_clz(x) == 32 - int(x).bit_length()


# ---------------------------------------------------------------- LBVH (a3-a9)
def _rand_boxes(n, seed, dup=True):
    rng = np.random.default_rng(seed)
    lo = rng.random((n, 3), dtype=np.float32)
    hi = lo + rng.random((n, 3), dtype=np.float32) * np.float32(0.05)
    a = np.zeros((n, 6), np.float32)
    a[:, 0::2], a[:, 1::2] = lo, hi
    if dup and n > 6:
        a[4] = a[2]
        a[6] = a[2]
    return a


def _karras_bruteforce(codes):
    """Recursive split of sorted keys at the highest differing bit of (code, index) -> set of
    (first, last, split) per inner node; Karras' construction yields the same partition."""
    keys = [(int(c) << 32) | i for i, c in enumerate(codes)]
    out = []

    def rec(a, b):
        if a == b:
            return
        hb = (keys[a] ^ keys[b]).bit_length()
        s = a
        for k in range(a, b):
            if (keys[k + 1] ^ keys[a]).bit_length() == hb:
                s = k
                break
        out.append((a, b, s))
        rec(a, s)
        rec(s + 1, b)

    rec(0, len(keys) - 1)
    return sorted(out)


@pytest.mark.parametrize("n", [2, 3, 5, 17, 64, 300])
def test_bvh_invariants_and_bruteforce(orc, n):
    a = _rand_boxes(n, n)
    prims = np.arange(n, dtype=np.int32) * 3
    b = orc.bvh_build(a, prims)
    codes = b.mcodes_unsorted
    # sort: non-decreasing, ties in index order (stable), permutation
    assert np.all(np.diff(b.mcodes_sorted.astype(np.int64)) >= 0)
    assert np.array_equal(b.aabb_ids, np.lexsort((np.arange(n), codes)))
    assert np.array_equal(b.leaf_nodes, prims[b.aabb_ids])
    # bounds
    assert np.array_equal(b.bounds[0::2], a[:, 0::2].min(axis=0)) and np.array_equal(b.bounds[1::2], a[:, 1::2].max(axis=0))
    # topology: every node but the root has exactly one parent, and the tree equals the brute force split
    inner = n - 1
    par = b.parent
    assert par[0] == -1 and np.all(par[1:] >= 0)
    kids = np.concatenate([b.left, b.right])
    assert sorted(kids.tolist()) == list(range(1, 2 * n - 1))

    def span(node):
        if node >= inner:
            return node - inner, node - inner
        return span(b.left[node])[0], span(b.right[node])[1]

    got = sorted((span(i)[0], span(i)[1], span(b.left[i])[1]) for i in range(inner))
    assert got == _karras_bruteforce(b.mcodes_sorted)
    # refit + emit: node boxes are the exact union of their leaves' boxes
    sa = a[b.aabb_ids]
    for i in range(inner):
        for side, child in ((0, b.left[i]), (6, b.right[i])):
            s0, s1 = span(child)
            box = b.inner[i, side:side + 6]
            assert np.array_equal(box[:3], sa[s0:s1 + 1, 0::2].min(axis=0))
            assert np.array_equal(box[3:], sa[s0:s1 + 1, 1::2].max(axis=0))
        ci = b.inner[i, 12:14].view(np.int32)
        for c, ch in zip(ci, (b.left[i], b.right[i])):
            assert c == (-(ch - inner + 1) if ch >= inner else 4 * ch)


def test_bvh_zero_and_one_box(orc):
    b0 = orc.bvh_build(np.zeros((0, 6), np.float32), np.zeros(0, np.int32))
    assert b0.inner.shape == (1, 16) and list(b0.leaf_nodes) == [0, 0]
    assert np.all(np.isinf(b0.bounds)) and b0.bounds[0] > 0        # empty bounds stay (inf, -inf)
    a = np.array([[0.1, 0.2, 0.3, 0.4, 0.5, 0.6]], np.float32)
    b1 = orc.bvh_build(a, np.array([7], np.int32))
    assert list(b1.leaf_nodes) == [7, 7]
    assert np.array_equal(b1.bounds, np.array([0.0, 0.2, 0.0, 0.4, 0.0, 0.6], np.float32))  # padded with the zero box


# ---------------------------------------------------------------- element split (a2)
@pytest.mark.parametrize("p,amp", [(1, 0.0), (2, 0.05), (3, 0.1), (4, 0.25)])
def test_split_partition_and_containment(orc, p, amp):
    mesh = syn.hex_mesh(2, p, amp, 0.02, 9)
    aabbs, prims, subrefs = orc.split_elements(mesh)
    assert aabbs.shape == (8 * 33, 6) and np.array_equal(prims, np.repeat(np.arange(8), 33))
    rng = np.random.default_rng(p)
    for e in range(8):
        sr = subrefs[33 * e:33 * (e + 1)].astype(np.float64)
        vol = np.prod(sr[:, 3:] - sr[:, :3], axis=1)
        assert abs(vol.sum() - 1.0) < 1e-12 and np.all(vol > 0)             # the 33 subrefs tile [0,1]^3
        dofs = mesh.values[mesh.elem_dofs(e)].astype(np.float64)
        for k in range(33):
            box = aabbs[33 * e + k]
            for _ in range(6):
                r = sr[k, :3] + rng.random(3) * (sr[k, 3:] - sr[k, :3])
                x = orc.eval_d(dofs, p, r)[0]
                assert np.all(x >= box[0::2] - 1e-6) and np.all(x <= box[1::2] + 1e-6)


def test_split_affine_identity_element(orc):
    """A lattice control net is the identity map: every sub-box is its subref (x1.000001 about the centre)."""
    mesh = syn.hex_mesh(1, 3, 0.0, 0.0, 0)
    aabbs, _, subrefs = orc.split_elements(mesh)
    lo, hi = subrefs[:, :3].astype(np.float64), subrefs[:, 3:].astype(np.float64)
    c, half = (lo + hi) / 2, (hi - lo) / 2 * 1.000001
    assert np.allclose(aabbs[:, 0::2], c - half, atol=1e-7) and np.allclose(aabbs[:, 1::2], c + half, atol=1e-7)


# The 33 subrefs [u0, v0, w0, u1, v1, w1] of construct_bvh's split (mesh_utils.cpp:586-639) worked out by hand
# for two exactly representable elements.  The rule: split the box of largest volume (surface area when the
# volume is 0), ties to the lowest index (strict >); cut its subref at 0.5 along its longest reference axis
# (ties to the lowest axis); the lower half stays in place, the upper half is appended.
# Identity map (a lattice net): every box is its subref, so the volumes halve in generations: split 1 cuts u,
# splits 2-3 cut v of boxes 0 and 1, splits 4-7 cut w of boxes 0-3, ...; after 31 splits the 32 boxes are
# [1/4 x 1/4 x 1/2] bricks and split 32 cuts box 0 along w, so box 0 ends as [0,1/4]^3 and box 32 as
# [0,1/4]^2 x [1/4,1/2].
_IDENTITY_SUBREFS = [
    (0, 0, 0, 1, 1, 1), (2, 0, 0, 3, 1, 2), (0, 2, 0, 1, 3, 2), (2, 2, 0, 3, 3, 2), (0, 0, 2, 1, 1, 4),
    (2, 0, 2, 3, 1, 4), (0, 2, 2, 1, 3, 4), (2, 2, 2, 3, 3, 4), (1, 0, 0, 2, 1, 2), (3, 0, 0, 4, 1, 2),
    (1, 2, 0, 2, 3, 2), (3, 2, 0, 4, 3, 2), (1, 0, 2, 2, 1, 4), (3, 0, 2, 4, 1, 4), (1, 2, 2, 2, 3, 4),
    (3, 2, 2, 4, 3, 4), (0, 1, 0, 1, 2, 2), (2, 1, 0, 3, 2, 2), (0, 3, 0, 1, 4, 2), (2, 3, 0, 3, 4, 2),
    (0, 1, 2, 1, 2, 4), (2, 1, 2, 3, 2, 4), (0, 3, 2, 1, 4, 4), (2, 3, 2, 3, 4, 4), (1, 1, 0, 2, 2, 2),
    (3, 1, 0, 4, 2, 2), (1, 3, 0, 2, 4, 2), (3, 3, 0, 4, 4, 2), (1, 1, 2, 2, 2, 4), (3, 1, 2, 4, 2, 4),
    (1, 3, 2, 2, 4, 4), (3, 3, 2, 4, 4, 4), (0, 0, 1, 1, 1, 2)]
# Flat element (identity in x, y; every control point at z = 0): every box has volume 0, so its surface area
# 2 |du| |dv| decides, and a cut along w leaves both halves with their parent's area.  After split 4 (box 0
# along w) all five boxes tie at area 1/2, so split 5 takes box 0 again and cuts it along u (box 5 starts as
# [1/4,1/2] x [0,1/2] x [0,1/2]), where the identity element takes box 1, the largest volume; the two orders
# differ from box 5 on.
_FLAT_SUBREFS = [
    (0, 0, 0, 1, 1, 1), (2, 0, 0, 3, 1, 2), (0, 2, 0, 1, 3, 2), (2, 2, 0, 3, 3, 2), (0, 0, 2, 1, 1, 4),
    (1, 0, 0, 2, 1, 2), (2, 0, 2, 3, 1, 4), (3, 0, 0, 4, 1, 2), (0, 2, 2, 1, 3, 4), (1, 2, 0, 2, 3, 2),
    (2, 2, 2, 3, 3, 4), (3, 2, 0, 4, 3, 2), (1, 0, 2, 2, 1, 4), (3, 0, 2, 4, 1, 4), (1, 2, 2, 2, 3, 4),
    (3, 2, 2, 4, 3, 4), (0, 1, 0, 1, 2, 2), (2, 1, 0, 3, 2, 2), (0, 3, 0, 1, 4, 2), (2, 3, 0, 3, 4, 2),
    (0, 1, 2, 1, 2, 4), (1, 1, 0, 2, 2, 2), (2, 1, 2, 3, 2, 4), (3, 1, 0, 4, 2, 2), (0, 3, 2, 1, 4, 4),
    (1, 3, 0, 2, 4, 2), (2, 3, 2, 3, 4, 4), (3, 3, 0, 4, 4, 2), (1, 1, 2, 2, 2, 4), (3, 1, 2, 4, 2, 4),
    (1, 3, 2, 2, 4, 4), (3, 3, 2, 4, 4, 4), (0, 0, 1, 1, 1, 2)]


def _quarters(t):
    return np.asarray(t, np.float32) * np.float32(0.25)


@pytest.mark.parametrize("p", [1, 2])
def test_split_identity_element_known_answer(orc, p):
    """The hand-derived 33 subrefs of the identity element, in order; each box is its subref x1.000001."""
    mesh = syn.elements_from_nets([syn.box_net((0, 0, 0), (1, 1, 1), p)], p)
    aabbs, prims, subrefs = orc.split_elements(mesh)
    exp = _quarters(_IDENTITY_SUBREFS)
    assert np.array_equal(subrefs, exp)
    assert np.array_equal(prims, np.zeros(33, np.int32))
    lo, hi = exp[:, :3].astype(np.float64), exp[:, 3:].astype(np.float64)
    c, half = (lo + hi) / 2, (hi - lo) / 2 * 1.000001
    assert np.abs(aabbs[:, 0::2] - (c - half)).max() < 1e-7 and np.abs(aabbs[:, 1::2] - (c + half)).max() < 1e-7


@pytest.mark.parametrize("p", [1, 2])
def test_split_flat_element_uses_surface_area(orc, p):
    """A flat element (zero volume everywhere) is split by surface area, ties to the lowest index: the
    hand-derived order differs from the identity element's from box 5 on, and the z extent stays 0."""
    net = syn.box_net((0, 0, 0), (1, 1, 1), p)
    net[:, 2] = 0.0
    mesh = syn.elements_from_nets([net], p)
    aabbs, _, subrefs = orc.split_elements(mesh)
    exp = _quarters(_FLAT_SUBREFS)
    assert np.array_equal(subrefs, exp)
    assert not np.array_equal(exp, _quarters(_IDENTITY_SUBREFS))
    assert np.all(aabbs[:, 4] == 0) and np.all(aabbs[:, 5] == 0)
    uv = exp[:, [0, 1, 3, 4]].astype(np.float64)
    c, half = (uv[:, :2] + uv[:, 2:]) / 2, (uv[:, 2:] - uv[:, :2]) / 2 * 1.000001
    assert np.abs(aabbs[:, [0, 2]] - (c - half)).max() < 1e-7 and np.abs(aabbs[:, [1, 3]] - (c + half)).max() < 1e-7


# ---------------------------------------------------------------- locate (a10-a12, a17)
def test_locate_affine_closed_form(orc):
    n = 4
    mesh = syn.hex_mesh(n, 2, 0.0, 0.0, 0)
    bvh = orc.build_mesh_bvh(mesh)
    rng = np.random.default_rng(3)
    pts = (0.002 + 0.996 * rng.random((3000, 3))).astype(np.float32)
    for dt in (np.float32, np.float64):
        cells, refs = orc.locate(mesh, bvh, pts, dtype=dt)
        q = pts.astype(np.float64) * n
        idx = np.floor(q).astype(int)
        frac = q - idx
        interior = np.all((frac > 2e-4) & (frac < 1 - 2e-4), axis=1)
        exp_cell = idx[:, 0] + n * (idx[:, 1] + n * idx[:, 2])
        assert np.array_equal(cells[interior], exp_cell[interior])
        assert np.abs(refs[interior] - frac[interior]).max() < (1e-5 if dt == np.float32 else 1e-9)
    out = np.array([[1.5, 0.5, 0.5], [-0.1, 0.2, 0.2]], np.float32)
    c, r = orc.locate(mesh, bvh, out)
    assert np.all(c == -1) and np.all(r == -1)


def test_locate_start_modes_on_affine_mesh(orc):
    """On a lattice (identity-map) mesh the corner-affine start is the answer itself, so each located
    point takes exactly one Newton step; the SubRef-centre start (subref_center) takes two: the first moves
    from the centre to the answer (more than the 1e-4 convergence step), the second confirms it.  Both
    give the closed-form element and reference point."""
    n = 4
    mesh = syn.hex_mesh(n, 2, 0.0, 0.0, 0)
    bvh = orc.build_mesh_bvh(mesh)
    rng = np.random.default_rng(4)
    pts = (0.002 + 0.996 * rng.random((3000, 3))).astype(np.float32)
    q = pts.astype(np.float64) * n
    idx = np.floor(q).astype(int)
    frac = q - idx
    interior = np.all((frac > 2e-4) & (frac < 1 - 2e-4), axis=1)
    exp_cell = idx[:, 0] + n * (idx[:, 1] + n * idx[:, 2])
    res = {}
    try:
        for mode in ("affine", "centre"):
            orc.set_newton_start(mode)
            cells, refs, steps, _ = orc.locate(mesh, bvh, pts, want_steps=True)
            assert np.array_equal(cells[interior], exp_cell[interior])
            assert np.abs(refs[interior] - frac[interior]).max() < 1e-5
            res[mode] = steps[interior]
    finally:
        orc.set_newton_start("affine")
    assert np.all(res["affine"] == 1)
    assert np.all((res["centre"] == 1) | (res["centre"] == 2)) and (res["centre"] == 2).mean() > 0.99
    with pytest.raises(ValueError):
        orc.set_newton_start("mlp")


@pytest.mark.parametrize("make", [lambda: syn.hex_mesh(3, 2, 0.05, 0.02, 4),
                                  lambda: syn.hex_mesh(2, 4, 0.25, 0.01, 2),
                                  lambda: syn.mixed_hex_mesh(3, 0.15, 0.05, 8)])
def test_locate_bvh_loses_no_hits(orc, make):
    """Brute force over every element (no BVH) finds a containing element exactly when the BVH
    traversal does, and both reference points map back to the query point."""
    mesh = make()
    bvh = orc.build_mesh_bvh(mesh)
    pts = syn.query_points(800, bvh.bounds, seed=5)
    cells, refs = orc.locate(mesh, bvh, pts, dtype=np.float64)
    bc, br = orc.locate_brute_f64(mesh, pts)
    assert np.array_equal(cells >= 0, bc >= 0)
    for i in np.nonzero(cells >= 0)[0]:
        e = int(cells[i])
        dofs = mesh.values[mesh.elem_dofs(e)].astype(np.float64)
        p = mesh.order if mesh.orders is None else int(mesh.orders[e])
        x = orc.eval_d(dofs, p, refs[i])[0]
        assert np.abs(x - pts[i]).max() < 1e-6


def _dfs_winners(bvh, boxes, pts):
    """The element of the first leaf, in depth-first left-first visiting order, whose axis-aligned affine
    element accepts the point (ref coordinates inside [-1e-4, 1 + 1e-4]); None where a ref coordinate is
    within 1e-6 of that band's edge in an element the walk reaches (rounding may decide there)."""
    out = []
    for p in pts:
        win = -1
        for leaf in dfs_leaves(bvh.inner, p):
            e = int(bvh.leaf_nodes[leaf])
            lo, hi = boxes[e]
            r = (p.astype(np.float64) - lo) / (hi - lo)
            if band_edge_distance(r[None, :])[0] < 1e-6:
                win = None
                break
            if np.all(r >= -1e-4) and np.all(r <= 1 + 1e-4):
                win = e
                break
        out.append(win)
    return out


@pytest.mark.parametrize("p", [1, 2])
def test_locate_first_accepting_candidate_wins(orc, p):
    """DeviceMesh::locate tries the candidate leaves depth-first, left child first, and the first element
    whose Newton solve is accepted wins (device_mesh.hpp:307-316).  Two overlapping affine elements (every
    point of the overlap lies inside both) and two elements sharing a face (points within 5e-5 of the face
    lie inside both acceptance bands): the oracle's winners equal an independent walk over its own emitted
    nodes, and both elements win somewhere, so a last-wins, right-first or lowest-id rule fails."""
    cases = [
        [((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)), ((0.5, 0.25, 0.0), (1.5, 1.25, 1.0))],
        [((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)), ((1.0, 0.0, 0.0), (2.0, 1.0, 1.0))],
    ]
    rng = np.random.default_rng(17)
    for boxes in cases:
        mesh = syn.elements_from_nets([syn.box_net(lo, hi, p) for lo, hi in boxes], p)
        bvh = orc.build_mesh_bvh(mesh)
        b = [(np.array(lo, np.float64), np.array(hi, np.float64)) for lo, hi in boxes]
        if boxes[1][0][0] == 0.5:
            pts = np.stack([0.5 + 0.5 * rng.random(1500), 0.25 + 0.75 * rng.random(1500), rng.random(1500)], 1)
        else:
            pts = np.stack([1.0 + 5e-5 * (2 * rng.random(1500) - 1), rng.random(1500), rng.random(1500)], 1)
        pts = pts.astype(np.float32)
        exp = _dfs_winners(bvh, b, pts)
        keep = np.array([e is not None for e in exp])
        exp = np.array([e for e in exp if e is not None], np.int32)
        assert keep.sum() > 1400 and (exp == 0).sum() > 50 and (exp == 1).sum() > 50, ((exp == 0).sum(), (exp == 1).sum())
        for dt in (np.float32, np.float64):
            cells, _ = orc.locate(mesh, bvh, pts, dtype=dt)
            assert np.array_equal(cells[keep], exp)


def test_refine_is_exact_inverse(orc):
    mesh = syn.hex_mesh(3, 3, 0.1, 0.01, 2)
    bvh = orc.build_mesh_bvh(mesh)
    pts = syn.query_points(500, bvh.bounds, seed=1)
    c, r = orc.locate(mesh, bvh, pts)
    ex = orc.refine_f64(mesh, pts, c, r.astype(np.float64))
    for i in np.nonzero(c >= 0)[0][:100]:
        dofs = mesh.values[mesh.elem_dofs(int(c[i]))].astype(np.float64)
        assert np.abs(orc.eval_d(dofs, 3, ex[i])[0] - pts[i]).max() < 1e-13


# ---------------------------------------------------------------- colour table, blend, volume (a18-a25)
def test_color_table_closed_form(orc):
    rgb = syn.grey_to_red_table()
    a = syn.table_alpha_for_corrected(0.02, 200)
    lut = orc.color_table_sample(rgb, [[0.0, a]], smooth=True, factor=10.0 / 200)
    s = np.arange(1024, dtype=np.float64) / 1023
    assert np.allclose(lut[:, 0], 0.5 + 0.5 * s, atol=1e-6)
    assert np.allclose(lut[:, 1], 0.5 - 0.5 * s, atol=1e-6) and np.allclose(lut[:, 2], lut[:, 1])
    assert np.allclose(lut[:, 3], 0.02, atol=1e-6)
    step = orc.color_table_sample(rgb, [[0.0, 0.5]], smooth=False)   # dray default: nearest of two
    assert np.array_equal(step[:512, 0], np.full(512, 0.5, np.float32)) and np.all(step[512:, 0] == 1.0)


def _cube_identity(n):
    return syn.hex_mesh(n, 1, 0.0, 0.0, 0)


def test_volume_constant_field_closed_form(orc):
    """Constant field: every sample has the same colour c, so alpha follows a_{k+1} = a_k + c_a (1 - a_k)
    in float32; early termination happens on the first sample whose blend reaches alpha >= 0.95."""
    mesh = _cube_identity(2)
    mesh.field[:] = 0.25
    bvh = orc.build_mesh_bvh(mesh)
    samples = 200
    a = syn.table_alpha_for_corrected(0.02, samples)
    lut = orc.color_table_sample(syn.grey_to_red_table(), [[0.0, a]], True, 10.0 / samples)
    rays = syn.orthographic_rays(4, 4)
    act = orc.remove_missed_rays(rays, bvh.bounds)
    parts, located, _ = orc.integrate(mesh, bvh, act, orc.sample_dist(bvh.bounds, samples), lut, 0.0, 1.0)
    parts = orc.compact_partials(parts)
    c = lut[int(np.float32(0.25) * np.float32(1023))]
    f = np.zeros(4, np.float32)
    for k in range(int(located[0])):
        alpha = np.float32(c[3] * np.float32(1 - f[3]))
        f[:3] = f[:3] + c[:3] * alpha
        f[3] = alpha + f[3]
    assert len(parts) == 16 and np.all(located == located[0])
    assert np.array_equal(parts["rgba"][0], f)
    # the cube is sqrt(3)/200 * k deep: the ray crosses 1.0 in depth with the step of the diagonal
    step = orc.sample_dist(bvh.bounds, samples)
    assert abs(int(located[0]) - 1.0 / step) <= 2


def test_volume_early_termination(orc):
    mesh = _cube_identity(2)
    mesh.field[:] = 0.5
    bvh = orc.build_mesh_bvh(mesh)
    samples = 2000          # small step: the ray would cross ~1155 samples without the alpha cut-off
    a = 0.95                 # table alpha, corrected to 1-(0.05)^(1/200) ~ 0.0149 per sample
    lut = orc.color_table_sample(syn.grey_to_red_table(), [[0.0, a]], True, 10.0 / samples)
    rays = syn.orthographic_rays(2, 2)
    act = orc.remove_missed_rays(rays, bvh.bounds)
    parts, located, _ = orc.integrate(mesh, bvh, act, orc.sample_dist(bvh.bounds, samples), lut, 0.0, 1.0)
    parts = orc.compact_partials(parts)
    ca = lut[511, 3]
    k = 0
    al = np.float32(0)
    while al < np.float32(0.95):
        al = np.float32(ca * np.float32(1 - al)) + al
        k += 1
    assert len(parts) == 4 and np.all(parts["rgba"][:, 3] >= 0.95)
    assert np.all(located == k + 1)          # one more locate happens after the last blend (volume.cpp:199-204)


def test_composite_known_answer(orc):
    p = np.zeros(3, dtype=syn.PARTIAL_DTYPE)
    p["pixel_id"] = [0, 0, 2]
    p["rgba"] = [[0.2, 0.1, 0.0, 0.5], [0.4, 0.4, 0.4, 0.5], [0.1, 0.1, 0.1, 0.25]]
    img = orc.composite(p, 2, 2)
    # pixel 0: front (0.2,0.1,0,.5) + back*(1-.5) -> (0.4,0.3,0.2,0.75), then white under it
    assert np.allclose(img[0], [0.4 + 0.25, 0.3 + 0.25, 0.2 + 0.25, 1.0])
    assert np.allclose(img[1], [1, 1, 1, 1]) and np.allclose(img[3], [1, 1, 1, 1])
    assert np.allclose(img[2], [0.1 + 0.75, 0.1 + 0.75, 0.1 + 0.75, 1.0])


def test_camera_rays(orc):
    r = orc.camera_perspective(64, 64, 45.0, (0.5, 0.5, 3.0), (0.5, 0.5, 0.5))
    d = r["dir"].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    assert np.array_equal(r["pixel_id"], np.arange(64 * 64))
    c = d[32 * 64 + 32]
    assert c[2] < -0.9999 and abs(c[0]) < 1e-6 and abs(c[1]) < 1e-6     # centre pixel looks down -z
    assert d[0, 0] < 0 and d[0, 1] < 0 and d[63, 0] > 0                # i -> +x (right), j -> +y
    half = math.degrees(math.atan2(-d[32 * 64, 0], -d[32 * 64, 2]))
    assert abs(half - 22.5) < 1.0


def test_remove_missed_rays(orc):
    rays = syn.orthographic_rays(4, 1)
    rays["orig"][:, 0] = [-0.5, 0.25, 0.75, 1.5]
    act = orc.remove_missed_rays(rays, [0, 1, 0, 1, 0, 1])
    assert list(act["pixel_id"]) == [1, 2]
    assert np.allclose(act["near"], 1.0 - 0.0005, atol=1e-6) and np.allclose(act["far"], 2.0 + 0.0005, atol=1e-6)


# ---- lit shading (VolumeShader::shaded_color) ------------------------------------------------
def test_default_light_closed_form(orc):
    # camera on +z looking at the centre, up y: the light sits 0.1 x distance above the camera
    assert np.allclose(orc.default_light((0.5, 0.5, 2.5), (0.5, 0.5, 0.5))[:3], (0.5, 0.7, 2.5), atol=1e-6)
    # camera on +x looking at the origin, up z: right = look x up = +y, up' = right x look = +z
    L = orc.default_light((3.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    assert np.allclose(L[:3], (3.0, 0.0, 0.3), atol=1e-6)
    assert np.allclose(L[3:], [0.4] * 3 + [0.75] * 3 + [0.8] * 3 + [15.0])


def _sheared_linear_mesh(a, c):
    """Order-2 mesh of 2^3 elements under a non-symmetric affine map x = A r + b; field f = a.x + c
    (exact in the Bernstein basis, so the physical gradient is exactly a)."""
    mesh = syn.hex_mesh(2, 2, 0.0, 0.0, 0)
    A = np.array([[1.0, 0.3, -0.2], [0.1, 0.8, 0.25], [-0.15, 0.05, 1.2]])
    mesh.values = (mesh.values.astype(np.float64) @ A.T + np.array([0.2, -0.1, 0.3])).astype(np.float32)
    mesh.field = (mesh.values.astype(np.float64) @ np.asarray(a) + c).astype(np.float32)
    return mesh


def _lit_setup(orc, a, c, cell, ref):
    mesh = _sheared_linear_mesh(a, c)
    lut = np.zeros((1024, 4), np.float32)
    lut[:, 0] = np.linspace(0.2, 0.9, 1024)
    lut[:, 1] = 0.5
    lut[:, 2] = np.linspace(0.9, 0.1, 1024)
    lut[:, 3] = 0.3
    dofs = mesh.values[mesh.elem_dofs(cell)].astype(np.float64)
    wp, _ = orc.eval_d(dofs, 2, ref)
    scalar = float(np.dot(a, wp) + c)
    lo, hi = float(mesh.field.min()), float(mesh.field.max())
    idx = int(np.float32((scalar - lo) / (hi - lo)) * np.float32(1023))
    return mesh, lut, wp, lut[idx], lo, hi


def _light(pos, amb=0.4, diff=0.75, spec=0.8, pw=15.0):
    return np.array(list(pos) + [amb] * 3 + [diff] * 3 + [spec] * 3 + [pw], np.float32)


def test_shade_headlight_closed_form(orc):
    """Light on the gradient line and the viewer looking down it: diffuse = specular = 1, so
    rgb = clamp((amb + diff + spec) * colour), alpha = colour alpha."""
    a = np.array([0.3, -0.2, 0.9])
    ah = a / np.linalg.norm(a)
    mesh, lut, wp, col, lo, hi = _lit_setup(orc, a, 0.1, 5, (0.3, 0.6, 0.2))
    out = orc.shade(mesh, 5, (0.3, 0.6, 0.2), -ah, lut, lo, hi, _light(wp + 4.0 * ah))
    exp = np.minimum(1.0, (0.4 + 0.75 + 0.8) * col[:3])
    assert np.allclose(out[:3], exp, atol=2e-5) and out[3] == col[3]


def test_shade_back_light_flips_gradient(orc):
    """Light behind the surface (ld . g = -1): the gradient flips, diffuse = 1; viewer perpendicular to
    the gradient: half vector at 45 deg, specular = (1/sqrt 2)^15."""
    a = np.array([0.3, -0.2, 0.9])
    ah = a / np.linalg.norm(a)
    perp = np.cross(ah, [1.0, 0.0, 0.0])
    perp /= np.linalg.norm(perp)
    mesh, lut, wp, col, lo, hi = _lit_setup(orc, a, 0.1, 2, (0.7, 0.2, 0.5))
    out = orc.shade(mesh, 2, (0.7, 0.2, 0.5), -perp, lut, lo, hi, _light(wp - 3.0 * ah))
    spec = (1.0 / np.sqrt(2.0)) ** 15
    exp = np.minimum(1.0, (0.4 + 0.75 + 0.8 * spec) * col[:3])
    assert np.allclose(out[:3], exp, atol=2e-5)


def test_shade_grazing_light_ambient_only(orc):
    """Light in the tangent plane and viewer along the light: diffuse = 0 and specular = 0^15."""
    a = np.array([0.3, -0.2, 0.9])
    ah = a / np.linalg.norm(a)
    t = np.cross(ah, [0.0, 1.0, 0.0])
    t /= np.linalg.norm(t)
    mesh, lut, wp, col, lo, hi = _lit_setup(orc, a, 0.1, 6, (0.4, 0.4, 0.6))
    out = orc.shade(mesh, 6, (0.4, 0.4, 0.6), -t, lut, lo, hi, _light(wp + 2.0 * t))
    assert np.allclose(out[:3], 0.4 * col[:3], atol=1e-4)


def test_shade_two_lights_and_transparent_samples(orc):
    a = np.array([0.3, -0.2, 0.9])
    ah = a / np.linalg.norm(a)
    mesh, lut, wp, col, lo, hi = _lit_setup(orc, a, 0.1, 5, (0.3, 0.6, 0.2))
    L = _light(wp + 4.0 * ah)
    two = orc.shade(mesh, 5, (0.3, 0.6, 0.2), -ah, lut, lo, hi, np.stack([L, L]))
    assert np.allclose(two[:3], np.minimum(1.0, 2 * (0.4 + 0.75 + 0.8) * col[:3]), atol=2e-5)
    assert np.isclose(two[3], 2 * col[3])          # each light adds the sample alpha (dray's accumulation)
    lut2 = lut.copy()
    lut2[:, 3] = 0.01                              # alpha <= 0.01: the sample contributes nothing
    assert np.all(orc.shade(mesh, 5, (0.3, 0.6, 0.2), -ah, lut2, lo, hi, L) == 0)


def test_volume_lit_matches_per_sample_shading(orc):
    """A lit render of one ray equals blending shade() at the ray's located samples."""
    a = np.array([0.3, -0.2, 0.9])
    mesh = _sheared_linear_mesh(a, 0.1)
    ob = orc.build_mesh_bvh(mesh)
    lut = np.zeros((1024, 4), np.float32)
    lut[:, 0] = np.linspace(0.2, 0.9, 1024)
    lut[:, 1] = 0.5
    lut[:, 3] = 0.05
    lo, hi = float(mesh.field.min()), float(mesh.field.max())
    rays = orc.camera_perspective(8, 8, 40.0, (0.7, 0.6, 3.0), (0.6, 0.4, 0.9))
    act = orc.remove_missed_rays(rays, ob.bounds)
    L = orc.default_light((0.7, 0.6, 3.0), (0.6, 0.4, 0.9))
    sd = orc.sample_dist(ob.bounds, 64)
    parts, _, _ = orc.integrate(mesh, ob, act[:1], sd, lut, lo, hi, dtype=np.float64, lights=L)
    r = act[0]
    d = np.float64(r["near"]) + sd
    f = np.zeros(4, np.float32)
    while d < r["far"] and f[3] < 0.95:
        pt = r["orig"].astype(np.float64) + d * r["dir"].astype(np.float64)
        c, rf = orc.locate(mesh, ob, pt[None, :], dtype=np.float64)
        if c[0] < 0:
            break
        s = orc.shade(mesh, int(c[0]), rf[0], r["dir"], lut, lo, hi, L)
        al = s[3] * (np.float32(1) - f[3])
        f[:3] = f[:3] + s[:3] * al
        f[3] = al + f[3]
        d += sd
    assert np.allclose(parts[0]["rgba"], f, atol=1e-5)


# ---- cross-domain compositing ------------------------------------------------------------------------------
def _partials(rows):
    from devil_ray_amd.synthetic import PARTIAL_DTYPE
    a = np.zeros(len(rows), PARTIAL_DTYPE)
    for i, (pid, depth, rgba) in enumerate(rows):
        a[i]["pixel_id"], a[i]["depth"], a[i]["rgba"] = pid, depth, rgba
    return a


def test_composite_domains_orders_by_depth(orc):
    """Two domains' partials for one pixel arrive back-to-front and interleaved with another pixel: the
    result is the front-to-back blend in depth order over white."""
    front, back = (0.2, 0.1, 0.0, 0.4), (0.0, 0.3, 0.5, 0.5)
    p = _partials([(1, 2.0, back), (0, 1.0, (0.1, 0.1, 0.1, 0.2)), (1, 1.0, front)])
    out = orc.composite_domains(p, 2, 1)
    a = 1 - front[3]
    blended = np.array([front[c] + back[c] * a for c in range(3)] + [front[3] + back[3] * a], np.float32)
    exp1 = blended[:3] + (1 - blended[3]) * 1.0          # white background under alpha < 1
    assert np.allclose(out[1, :3], exp1, atol=1e-6) and np.isclose(out[1, 3], 1.0)
    assert np.allclose(out[0, :3], np.array([0.1] * 3) + 0.8, atol=1e-6)


def test_composite_domains_matches_single_domain_and_is_order_free(orc):
    mesh = syn.hex_mesh(4, 2, 0.05, 0.01, 3)
    ob = orc.build_mesh_bvh(mesh)
    w = h = 24
    rays = orc.camera_perspective(w, h, 45.0, (0.5, 0.5, 2.5), (0.5, 0.5, 0.5))
    act = orc.remove_missed_rays(rays, ob.bounds)
    lut = np.zeros((1024, 4), np.float32)
    lut[:, 0] = np.linspace(0, 1, 1024)
    lut[:, 3] = 0.05
    parts, _, _ = orc.integrate(mesh, ob, act, orc.sample_dist(ob.bounds, 64), lut, float(mesh.field.min()),
                                float(mesh.field.max()))
    parts = orc.compact_partials(parts)
    ref = orc.composite(parts, w, h)
    assert np.array_equal(orc.composite_domains(parts, w, h), ref)            # already in (pixel, depth) order
    shuffled = parts[np.random.default_rng(0).permutation(len(parts))]
    assert np.array_equal(orc.composite_domains(shuffled, w, h), ref)         # any input order


# ---- framebuffer (Renderer volume compositing, composite_background) and camera orbits -------------------
def test_partials_to_framebuffer_replace_and_blend(orc):
    """Hand-computed: pixel 1 gets two partials (back first in the input), pixel 0 one, pixel 2 none.
    blend=False replaces pixels 0/1 and keeps pixel 2; blend=True puts the volume over the framebuffer
    colour; depths take the nearest partial."""
    front, back, solo = (0.2, 0.1, 0.0, 0.4), (0.0, 0.3, 0.5, 0.5), (0.1, 0.1, 0.1, 0.2)
    p = _partials([(1, 2.0, back), (0, 1.5, solo), (1, 1.0, front)])
    surf = np.array([[0.3, 0.3, 0.3, 1.0]] * 3, np.float32)
    a = 1 - front[3]
    vol1 = np.array([front[c] + back[c] * a for c in range(4)], np.float32)
    for blend in (False, True):
        col = surf.copy()
        dep = np.full(3, np.inf, np.float32)
        orc.partials_to_framebuffer(p, 3, 1, col, dep, blend=blend)
        if blend:
            exp1 = vol1 + (1 - vol1[3]) * surf[1]
            exp0 = np.array(solo, np.float32) + (1 - solo[3]) * surf[0]
        else:
            exp1, exp0 = vol1, np.array(solo, np.float32)
        assert np.allclose(col[1], exp1, atol=1e-7) and np.allclose(col[0], exp0, atol=1e-7)
        assert np.array_equal(col[2], surf[2])                               # no partials: untouched
        assert dep[0] == np.float32(1.5) and dep[1] == np.float32(1.0) and dep[2] == np.inf


def test_framebuffer_background_and_volume_save_equivalence(orc):
    """composite_background puts alpha < 1 pixels over a coloured background (closed form) and leaves
    opaque ones; clear + partials_to_framebuffer + white background equals composite_domains."""
    col = np.array([[0.2, 0.0, 0.1, 0.5], [0.3, 0.3, 0.3, 1.0], [0.0, 0.0, 0.0, 0.0]], np.float32)
    bg = (0.5, 0.25, 1.0, 1.0)
    ref = col.copy()
    orc.framebuffer_composite_background(col, bg)
    assert np.allclose(col[0], ref[0] + np.float32(0.5) * np.array(bg, np.float32), atol=1e-7)
    assert np.array_equal(col[1], ref[1]) and np.allclose(col[2], bg)
    rng = np.random.default_rng(5)
    n, w, h = 400, 9, 7
    rows = [(int(rng.integers(0, w * h)), float(rng.integers(0, 4)), tuple(rng.uniform(0, 0.3, 4))) for _ in range(n)]
    p = _partials(rows)
    col = np.zeros((w * h, 4), np.float32)
    dep = np.full(w * h, np.inf, np.float32)
    orc.partials_to_framebuffer(p, w, h, col, dep)
    orc.framebuffer_composite_background(col)
    assert np.array_equal(col, orc.composite_domains(p, w, h))


def test_camera_orbits_closed_forms(orc):
    """azimuth(90) about +y moves a camera at +z to +x; elevate(90) moves it to +y; 360 degrees return to
    the start; reset_to_bounds looks at the padded box centre from one diagonal away, fov 60.  A flipped
    rotation sign sends azimuth(90) to -x and elevate(90) to -y."""
    pos, look, up = (0.0, 0.0, 5.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)
    p, l, f = orc.camera_orbit(pos, look, up, 30.0, "azimuth", 90.0)
    assert np.allclose(p, (5, 0, 0), atol=1e-5) and np.array_equal(l, np.zeros(3, np.float32)) and f == 30.0
    p, _, _ = orc.camera_orbit(pos, look, up, 30.0, "elevate", 90.0)
    assert np.allclose(p, (0, 5, 0), atol=1e-5)
    for mode in ("azimuth", "elevate"):
        p, _, _ = orc.camera_orbit(pos, look, up, 30.0, mode, 360.0)
        assert np.allclose(p, pos, atol=1e-5)
    p, l, f = orc.camera_orbit((3.0, 0.5, 0.5), (0.0, 0.5, 0.5), up, 30.0, "reset",
                               bounds6=(0, 1, 0, 2, -1, 1), pads=(0.0, 0.0, 0.0))
    assert np.allclose(l, (0.5, 1.0, 0.0)) and f == 60.0
    assert np.allclose(p, np.array([0.5, 1.0, 0.0]) + np.array([1.0, 0, 0]) * np.sqrt(1 + 4 + 4), atol=1e-6)
    # pads grow each axis by pad x length on both sides
    p, l, _ = orc.camera_orbit((0.5, 0.5, 9.0), (0.5, 0.5, 0.5), up, 30.0, "reset",
                               bounds6=(0, 1, 0, 1, 0, 1), pads=(0.5, 0.0, 0.0))
    assert np.allclose(l, (0.5, 0.5, 0.5)) and np.allclose(p, (0.5, 0.5, 0.5 + np.sqrt(4 + 1 + 1)), atol=1e-6)


def test_explain_mismatches_rule(orc):
    """The parity bar for located ids: a difference is excused only at the acceptance band edge or where the
    GPU agrees with the FP64 oracle.  A planted wrong GPU id (the right element's neighbour, for a point deep
    inside its element) is flagged, and the same wrong id on the FP32 oracle's side is excused; a point the
    FP64 map puts 5e-7 outside the band in one element and inside its neighbour, reported by one side in
    each, is excused; the same point 1e-4 further out is not."""
    from parity_util import explain_mismatches
    mesh = syn.hex_mesh(3, 2, 0.0, 0.0, 1)        # affine 3^3 lattice: element size 1/3, ref = 3 x - cell
    ob = orc.build_mesh_bvh(mesh)
    pts = np.array([[0.5, 0.5, 0.5], [0.16, 0.5, 0.83]], np.float32)
    oc, orf = orc.locate(mesh, ob, pts)
    assert (oc >= 0).all()
    gc = oc.copy()
    gc[0] = oc[0] + 1                              # planted mistake
    mism, ok, _, by64 = explain_mismatches(orc, mesh, ob, pts, gc, orf, oc, orf)
    assert list(mism) == [0] and not ok.any() and not by64.any()
    # the other way round (the GPU right, an FP32 oracle answer the exact map rejects): excused because the
    # FP64 oracle agrees with the GPU
    mism, ok, _, by64 = explain_mismatches(orc, mesh, ob, pts, oc, orf, gc, orf)
    assert list(mism) == [0] and ok.all() and by64.all()
    # x = 1/3 + 1e-4/3 + 5e-7/3 lies 5e-7 past u = 1 + 1e-4 in element 0 (u = 3x) and at u = 1e-4 + 5e-7
    # inside element 1 (u = 3x - 1)
    for shift, want in ((5e-7, True), (1e-4, False)):
        x = (1 + 1e-4 + shift) / 3.0
        p = np.array([[x, 0.5, 0.5]], np.float64)
        p32 = p.astype(np.float32)
        e0, e1 = 0 + 3 * (1 + 3 * 1), 1 + 3 * (1 + 3 * 1)   # cells (0,1,1) and (1,1,1): i + 3 (j + 3 k)
        r0 = np.array([[3 * x, 1.5 - 1, 1.5 - 1]], np.float32)
        r1 = np.array([[3 * x - 1, 0.5, 0.5]], np.float32)
        mism, ok, excess, by64 = explain_mismatches(orc, mesh, ob, p32, np.array([e0], np.int32), r0,
                                                    np.array([e1], np.int32), r1)
        assert list(mism) == [0] and bool(ok[0]) == want and not by64.any(), (shift, excess)


def test_point_location_closed_form(orc):
    """PointLocation over two overlapping affine domains with linear fields, where the answer is known in
    closed form: an affine (lattice) net with a field sampled from a linear function reproduces it
    exactly, so a point in the second domain (x in [0.5, 1.5]) takes f + 10, a point only in the first
    takes f, and a point in neither keeps empty_val (the last domain that locates a point wins)."""
    a = syn.hex_mesh(3, 2, 0.0, 0.0, 1)
    b = syn.HexMesh(n=a.n, order=a.order, ctrl_idx=a.ctrl_idx,
                    values=(a.values + np.float32([0.5, 0.0, 0.0])).astype(np.float32), field=a.field)
    lin = lambda v: (v[:, 0] + 2 * v[:, 1] + 3 * v[:, 2]).astype(np.float32)
    fa, fb = lin(a.values), (lin(b.values) + np.float32(10)).astype(np.float32)
    rng = np.random.default_rng(8)
    pts = (rng.random((4000, 3)) * [2.0, 1.2, 1.0]).astype(np.float32)
    dom = [(a, orc.build_mesh_bvh(a), [fa]), (b, orc.build_mesh_bvh(b), [fb])]
    vals, cells = orc.point_location(dom, pts, 1, -7.0)
    x = pts.astype(np.float64)
    f = x[:, 0] + 2 * x[:, 1] + 3 * x[:, 2]
    # away from the domains' boundaries (the +-1e-4 acceptance band)
    inside_a = np.all((x > 1e-3) & (x < [1 - 1e-3, 1 - 1e-3, 1 - 1e-3]), axis=1)
    inside_b = np.all((x > [0.5 + 1e-3, 1e-3, 1e-3]) & (x < [1.5 - 1e-3, 1 - 1e-3, 1 - 1e-3]), axis=1)
    near_a = np.all((x > -1e-3) & (x < 1 + 1e-3), axis=1)
    near_b = np.all((x > [0.5 - 1e-3, -1e-3, -1e-3]) & (x < [1.5 + 1e-3, 1 + 1e-3, 1 + 1e-3]), axis=1)
    outside = ~near_a & ~near_b
    only_a = inside_a & (x[:, 0] < 0.5 - 1e-3)
    assert inside_b.sum() > 500 and only_a.sum() > 500 and outside.sum() > 500
    assert np.abs(vals[0][inside_b] - (f[inside_b] + 10)).max() < 1e-4
    assert np.abs(vals[0][only_a] - f[only_a]).max() < 1e-4
    assert np.all(vals[0][outside] == np.float32(-7.0))


def test_split_by_order_domains(orc):
    """The mixed-order mesh as one domain per order: the domains partition the elements, each domain's
    element nets are the mixed mesh's, and PointLocation over them gives, for points well inside an
    element, the element and value the mixed mesh's own locate + Field::eval give."""
    mesh = syn.mixed_hex_mesh(4, 0.15, 0.05, 5)
    doms = syn.split_by_order(mesh)
    assert sorted(int(p) for p, _, _ in doms) == sorted(np.unique(mesh.orders).tolist())
    ids = np.sort(np.concatenate([e for _, _, e in doms]))
    assert np.array_equal(ids, np.arange(mesh.n_el))
    for p, d, els in doms:
        assert d.n_el == len(els) and np.all(mesh.orders[els] == p)
        for k in (0, len(els) // 2, len(els) - 1):
            assert np.array_equal(d.values[d.elem_dofs(k)], mesh.values[mesh.elem_dofs(int(els[k]))])
    ob = orc.build_mesh_bvh(mesh)
    pts = syn.query_points(6000, ob.bounds, seed=4)
    cm, rm = orc.locate(mesh, ob, pts)
    vm = orc.field_eval(mesh, cm, rm)
    vals, cells = orc.point_location([(d, orc.build_mesh_bvh(d), [mesh.field]) for _, d, _ in doms], pts, 1, -1.0)
    # the domain and element each point took (the last domain that located it)
    glob = np.full(len(pts), -1)
    for (_, _, els), c in zip(doms, cells):
        glob = np.where(c >= 0, els[np.maximum(c, 0)], glob)
    deep = (cm >= 0) & (np.abs(rm - 0.5).max(axis=1) < 0.45)
    assert deep.sum() > 3000
    assert np.array_equal(glob[deep], cm[deep])
    assert np.abs(vals[0][deep] - vm[deep]).max() < 1e-6

"""The SD-tree recording filters in the oracle.

The reference's box filters (SpatialFilterBox, DirectionalFilterBox,
OPX7UtilsPathGuiding.h:31-43, 126-169, 429-443, 573-611, 649-655, 806-813) are
pinned by hand-computed overlap shares; the stochastic form that the device
runs (one uniform point of the same box per record) is then checked to equal
them in expectation."""
import numpy as np

LEAF = 0x80000000


def _state(stree, n_dtrees, b_child=None):
    b_child = np.zeros((n_dtrees, 4), np.uint16) if b_child is None else b_child
    nodes = b_child.shape[0]
    cnt = np.ones(n_dtrees, np.uint32) if nodes == n_dtrees else np.array([nodes], np.uint32)
    return {"stree": np.asarray(stree, np.uint32), "b_cnt": cnt, "s_cnt": cnt.copy(),
            "scal": np.zeros((6, n_dtrees), np.float32), "b_sums": np.zeros((nodes, 4), np.float32),
            "b_child": b_child, "s_sums": np.zeros((nodes, 4), np.float32), "s_child": b_child.copy()}


def _octants():
    # root splits x, its children y, the grandchildren z: leaf = 4 x + 2 y + z
    return [1, 3, 5, 7, 9, 11, 13] + [LEAF | k for k in range(8)]


def _box(orc, st, pos, canon, irr, sw, spatial, directional):
    return orc.sdtree_record_box(st, (0, 0, 0), 1.0, np.asarray(pos, np.float32).reshape(-1, 3),
                                 np.asarray(canon, np.float32).reshape(-1, 2), np.asarray(irr, np.float32),
                                 np.asarray(sw, np.float32), spatial, directional)


def test_spatial_box_shares_are_overlap_volumes(orc):
    st = _state(_octants(), 8)
    # centre of the cube, voxel 0.5^3: the box [.25, .75]^3 takes 1/8 of each octant
    out = _box(orc, st, [0.5, 0.5, 0.5], [0.5, 0.5], [0], [8], 1, 0)
    assert np.allclose(out["scal"][0], 1.0, rtol=1e-6)
    # box equal to an octant
    out = _box(orc, st, [0.25, 0.25, 0.25], [0.5, 0.5], [0], [8], 1, 0)
    assert np.allclose(out["scal"][0], [8, 0, 0, 0, 0, 0, 0, 0], rtol=1e-6)
    # box [-.15, .35] x [.05, .55] x [.05, .55]: overlaps along x 0.35 (and
    # nothing past x = 0.5), along y and z 0.45 below 0.5 and 0.05 above; the
    # part outside the cube is lost
    out = _box(orc, st, [0.1, 0.3, 0.3], [0.5, 0.5], [0], [1], 1, 0)
    lo, hi = 0.45, 0.05
    exp = np.array([0.35 * a * b / 0.125 for a in (lo, hi) for b in (lo, hi)] + [0] * 4)
    assert np.allclose(out["scal"][0], exp, rtol=1e-5, atol=1e-7)
    assert abs(out["scal"][0].sum() - 0.7) < 1e-6


def test_spatial_box_uses_the_voxel_of_the_vertex_leaf(orc):
    # root split along x only: voxel 0.5 x 1 x 1, so a vertex at the centre
    # spreads over both leaves equally
    st = _state([1, LEAF | 0, LEAF | 1], 2)
    out = _box(orc, st, [0.5, 0.5, 0.5], [0.5, 0.5], [0], [2], 1, 0)
    assert np.allclose(out["scal"][0], [1, 1], rtol=1e-6)
    out = _box(orc, st, [0.625, 0.5, 0.5], [0.5, 0.5], [0], [4], 1, 0)
    assert np.allclose(out["scal"][0], [1, 3], rtol=1e-6)


def test_directional_box_shares_are_overlap_areas(orc):
    st = _state([LEAF], 1)
    # root only: depth 1, square of edge 0.5
    out = _box(orc, st, [0.5, 0.5, 0.5], [0.5, 0.5], [4], [1], 0, 1)
    assert np.allclose(out["b_sums"][0], 1.0, rtol=1e-6) and out["scal"][0][0] == 1.0
    out = _box(orc, st, [0.5, 0.5, 0.5], [0.1, 0.1], [4], [1], 0, 1)
    assert np.allclose(out["b_sums"][0], [4 * 0.35 ** 2 / 0.25, 0, 0, 0], rtol=1e-5)
    # slot 0 of the root subdivided: depth 2 at (0.2, 0.2), square [.075, .325]^2
    ch = np.zeros((2, 4), np.uint16)
    ch[0, 0] = 1
    st = _state([LEAF], 1, ch)
    out = _box(orc, st, [0.5, 0.5, 0.5], [0.2, 0.2], [2], [1], 0, 1)
    a, b = 0.175, 0.075
    assert np.allclose(out["b_sums"][1], np.array([a * a, b * a, a * b, b * b]) * 2 / 0.0625, rtol=1e-5)
    assert np.all(out["b_sums"][0] == 0) and abs(out["b_sums"].sum() - 2.0) < 1e-5
    # nearest: the whole flux in the slot holding the direction
    out = _box(orc, st, [0.5, 0.5, 0.5], [0.2, 0.2], [2], [1], 0, 0)
    assert out["b_sums"][1, 0] == 2.0 and out["b_sums"].sum() == 2.0


def _binomial_close(x, m, n):
    # x: counts of n unit records; m: their expected values (box filter with weight n)
    x, m = np.asarray(x, np.float64).ravel(), np.asarray(m, np.float64).ravel()
    p = np.clip(m / n, 0.0, 1.0)
    return np.all(np.abs(x - m) <= 5.0 * np.sqrt(n * p * (1.0 - p)) + 1.0), (x, m)


def test_stochastic_filters_equal_the_box_filters_in_expectation(orc):
    # mixed-depth spatial tree (x split, then y on the left half) and DTree 0
    # subdivided twice; each point recorded N times with unit weight and flux
    # (counts, exact in float) against the box filter of one record of weight
    # N: per node the count is binomial with the box filter's share as mean
    stree = [1, 3, LEAF | 2, LEAF | 0, LEAF | 1]
    ch = np.zeros((5, 4), np.uint16)
    ch[0, 0], ch[1, 3] = 1, 2  # DTree 0: nodes 0..2; DTree 1: node 3; DTree 2: node 4
    st = _state(stree, 3, ch)
    st["b_cnt"] = np.array([3, 1, 1], np.uint32)
    st["s_cnt"] = st["b_cnt"].copy()
    pts = np.array([[0.3, 0.45, 0.5], [0.55, 0.2, 0.9], [0.05, 0.6, 0.02]], np.float32)
    dirs = np.array([[0.2, 0.3], [0.49, 0.52], [0.9, 0.05]], np.float32)
    n = 60000
    pos, canon = np.repeat(pts, n, axis=0), np.repeat(dirs, n, axis=0)
    ones = np.ones(pos.shape[0], np.float32)
    for k in range(3):
        for filt, sp, di in [(1, 1, 0), (2, 0, 1), (3, 1, 1)]:
            sl = slice(k * n, (k + 1) * n)
            sto = orc.sdtree_record(st, (0, 0, 0), 1.0, pos[sl], canon[sl], ones[sl], ones[sl], filter=filt, seed=7)
            det = _box(orc, st, pts[k], dirs[k], [1.0], [float(n)], sp, di)
            ok, info = _binomial_close(sto["scal"][0], det["scal"][0], n)
            assert ok, (k, filt, "weight", info)
            ok, info = _binomial_close(sto["b_sums"], det["b_sums"], n)
            assert ok, (k, filt, "flux", info)
    # the draws are a function of (seed, record index, batch) only
    a = orc.sdtree_record(st, (0, 0, 0), 1.0, pos, canon, ones, ones, filter=3, seed=7)
    b = orc.sdtree_record(st, (0, 0, 0), 1.0, pos, canon, ones, ones, filter=3, seed=7)
    c = orc.sdtree_record(st, (0, 0, 0), 1.0, pos, canon, ones, ones, filter=3, seed=7, batch=1)
    assert np.array_equal(a["b_sums"], b["b_sums"]) and not np.array_equal(a["b_sums"], c["b_sums"])


def test_stochastic_filter_on_deep_dtree(orc):
    # a DTree subdivided three levels towards the recorded direction: the
    # square shrinks with depth, and its expected shares match the box filter
    ch = np.zeros((4, 4), np.uint16)
    ch[0, 0], ch[1, 3], ch[2, 0] = 1, 2, 3
    st = _state([LEAF], 1, ch)
    n = 200000
    pos = np.full((n, 3), 0.5, np.float32)
    canon = np.tile(np.array([[0.26, 0.27]], np.float32), (n, 1))
    ones = np.ones(n, np.float32)
    sto = orc.sdtree_record(st, (0, 0, 0), 1.0, pos, canon, ones, ones, filter=2, seed=3)
    det = _box(orc, st, pos[:1], canon[:1], [1.0], [float(n)], 0, 1)
    m = det["b_sums"].astype(np.float64)
    assert abs(m.sum() - n) < 1e-4 * n  # square fully inside [0,1)^2
    assert np.count_nonzero(m) >= 3
    ok, info = _binomial_close(sto["b_sums"], m, n)
    assert ok, info

"""Device SD-tree against the oracle: record, build, reset (refine), lookup,
sample and pdf on identical inputs.  Integer-valued flux keeps the float
atomics of the device record exact, so every comparison is bit-exact."""
import numpy as np
import pytest

from rtlib_amd.tracer import RTSTreeWrapper

pytestmark = pytest.mark.gpu

LO = np.zeros(3, np.float32)
HI = np.array([4.0, 2.0, 3.0], np.float32)
SIZE = 4.0


def _tree(ctx):
    return RTSTreeWrapper(ctx, LO, HI, max_dtree_depth=20, fraction=0.3)


def _assert_state_equal(a, b, keys=("stree", "b_cnt", "s_cnt", "scal", "b_sums", "b_child", "s_sums", "s_child")):
    for k in keys:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape, (k, x.shape, y.shape)
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), k


def _records(rng, n, scale=SIZE):
    pos = rng.uniform(0, 1, (n, 3)).astype(np.float32) * np.array([4, 2, 3], np.float32)
    canon = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    irr = rng.integers(0, 8, n).astype(np.float32)
    sw = np.ones(n, np.float32)
    return pos, canon, irr, sw


def _pass(tree, orc, st, rng, n):
    pos, canon, irr, sw = _records(rng, n)
    tree.Record(pos, canon, irr, sw)
    st = orc.sdtree_record(st, LO, SIZE, pos, canon, irr, sw)
    _assert_state_equal(tree.Download(), st)
    tree.Build()
    st = orc.sdtree_build(st)
    _assert_state_equal(tree.Download(), st)
    return st


def test_lifecycle_record_build_reset(ctx, orc):
    rng = np.random.default_rng(0)
    tree = _tree(ctx)
    tree.SetSplit(mode=0, c=40.0)
    st = tree.Download()
    assert st["stree"].tolist() == [0x80000000] and st["b_cnt"].tolist() == [1]
    spp = 1
    for it in range(5):
        if it > 0:
            tree.Reset(it, spp)
            st = orc.sdtree_reset(st, it, spp, mode=0, c=40.0)
            _assert_state_equal(tree.Download(), st)
        st = _pass(tree, orc, st, rng, 4000 * (it + 1))
        spp *= 2
    sz = tree.Sizes()
    assert sz["dtrees"] > 8 and sz["sampling_nodes"] > 100


def test_octree_lifecycle_matches_oracle(ctx, orc):
    # the STree2 variant: 8-way spatial splits, the reference's fixed rule
    # (mode 0, c = 4000) scaled down by c so that a few thousand records split
    rng = np.random.default_rng(3)
    tree = _tree(ctx)
    tree.SetSplit(mode=0, c=60.0)
    tree.SetOctree(True)
    st = tree.Download()
    spp = 1
    for it in range(4):
        if it > 0:
            tree.Reset(it, spp)
            st = orc.sdtree_reset(st, it, spp, mode=0, c=60.0, octree=True)
            _assert_state_equal(tree.Download(), st)
        st = _pass(tree, orc, st, rng, 5000 * (it + 1))
        spp *= 2
    n = tree.Sizes()["dtrees"]
    assert n > 8 and (n - 1) % 7 == 0
    pos = rng.uniform(0, 1, (4000, 3)).astype(np.float32) * np.array([4, 2, 3], np.float32)
    assert np.array_equal(tree.Lookup(pos), orc.stree_lookup(st["stree"], LO, SIZE, pos))


@pytest.mark.parametrize("octree", [False, True])
def test_refine_many_levels_in_one_reset(ctx, orc, octree):
    # The first reset splits the root leaf 11 times (binary) or 4 times
    # (octree) in one refine, the next ones split some thousands of leaves a
    # few levels each: block offsets, walk order and DTree numbering of the
    # device refine against the oracle's sequential walk, bit for bit.
    rng = np.random.default_rng(11 if octree else 12)
    tree = _tree(ctx)
    tree.SetSplit(mode=1, c=40.0)
    tree.SetOctree(octree)
    st = tree.Download()
    for it in range(4):
        if it > 0:
            tree.Reset(it, 1)
            st = orc.sdtree_reset(st, it, 1, mode=1, c=40.0, octree=octree)
            _assert_state_equal(tree.Download(), st)
        st = _pass(tree, orc, st, rng, 60000 * (it + 1))
    sz = tree.Sizes()
    assert sz["dtrees"] > 2000, sz
    pos = rng.uniform(0, 1, (20000, 3)).astype(np.float32) * np.array([4, 2, 3], np.float32)
    assert np.array_equal(tree.Lookup(pos), orc.stree_lookup(st["stree"], LO, SIZE, pos))


def test_refine_respects_memory_cap(ctx, orc):
    # max_mb = 0: the footprint test (footprint / 10^6 >= max_mb) always
    # holds, so the reset keeps the spatial tree and only resets the DTrees.
    rng = np.random.default_rng(5)
    tree = _tree(ctx)
    tree.SetSplit(mode=0, c=40.0, max_mb=0)
    st = _pass(tree, orc, tree.Download(), rng, 8000)
    tree.Reset(1, 1)
    st = orc.sdtree_reset(st, 1, 1, mode=0, c=40.0, max_mb=0)
    _assert_state_equal(tree.Download(), st)
    assert st["stree"].tolist() == [0x80000000]


def test_lookup_sample_pdf_match_oracle(ctx, orc):
    rng = np.random.default_rng(1)
    tree = _tree(ctx)
    tree.SetSplit(mode=1, c=300.0)
    st = tree.Download()
    for it in range(3):
        if it:
            tree.Reset(it, 1)
            st = orc.sdtree_reset(st, it, 1, mode=1, c=300.0)
        st = _pass(tree, orc, st, rng, 6000)
    pos = rng.uniform(0, 1, (5000, 3)).astype(np.float32) * np.array([4, 2, 3], np.float32)
    cells = tree.Lookup(pos)
    assert np.array_equal(cells, orc.stree_lookup(st["stree"], LO, SIZE, pos))
    smp = tree.Sample(cells, 11)
    assert np.array_equal(smp, orc.sdtree_sample(st, cells, 11))
    canon = rng.uniform(0, 1, (5000, 2)).astype(np.float32)
    assert np.array_equal(tree.Pdf(cells, canon), orc.sdtree_pdf(st, cells, canon))


@pytest.mark.parametrize("filt", [1, 2, 3])
def test_box_filtered_record_matches_oracle(ctx, orc, filt):
    # stochastic box filters (bit 0 spatial, bit 1 directional): the device and
    # the oracle draw the same points from the filter stream, so the recorded
    # trees are identical; two calls check the per-call batch index
    rng = np.random.default_rng(5 + filt)
    tree = _tree(ctx)
    tree.SetSplit(mode=1, c=300.0)
    st = tree.Download()
    for it in range(3):
        if it:
            tree.Reset(it, 1)
            st = orc.sdtree_reset(st, it, 1, mode=1, c=300.0)
        st = _pass(tree, orc, st, rng, 6000)
    tree.Reset(3, 1)
    st = orc.sdtree_reset(st, 3, 1, mode=1, c=300.0)
    assert tree.Sizes()["dtrees"] > 8
    tree.SetFilter(spatial=filt & 1, directional=filt >> 1, seed=99)
    for batch in range(2):
        pos, canon, irr, sw = _records(rng, 20000)
        tree.Record(pos, canon, irr, sw)
        st = orc.sdtree_record(st, LO, SIZE, pos, canon, irr, sw, filter=filt, seed=99, batch=batch)
        _assert_state_equal(tree.Download(), st)


def test_import_export_round_trip_and_validation(ctx, orc):
    rng = np.random.default_rng(2)
    tree = _tree(ctx)
    tree.SetSplit(mode=1, c=200.0)
    st = tree.Download()
    st = _pass(tree, orc, st, rng, 3000)
    tree.Reset(1, 1)
    st = orc.sdtree_reset(st, 1, 1, mode=1, c=200.0)
    other = _tree(ctx)
    other.Upload(st)
    _assert_state_equal(other.Download(), st)
    bad = {k: np.array(v, copy=True) for k, v in st.items()}
    bad["b_child"][0, 0] = 60000
    with pytest.raises(Exception):
        other.Upload(bad)


def test_build_of_empty_tree_and_uniform_reset(ctx, orc):
    tree = _tree(ctx)
    tree.Build()
    tree.Reset(1, 1)
    st = tree.Download()
    # an empty sampling tree resets to the complete depth-4 tree (85 nodes)
    assert st["b_cnt"].tolist() == [85]


def test_lookup_deep_tree_and_box_edges(ctx, orc):
    # A spatial tree refined 27 levels deep towards one point (deeper than the
    # device's table of the first levels), looked up at random points, at
    # points on the power-of-two cell faces, near the refined point and
    # outside the box: identical leaves to the oracle's plain descent.
    depth = 27
    target = np.array([0.3712, 0.8123, 0.0917])
    words = np.zeros(1 + 2 * depth, np.uint32)
    cur, leaf = 0, 0
    for k in range(depth):
        first = 1 + 2 * k
        words[cur] = first
        bit = int(np.floor(target[k % 3] * 2.0 ** (k // 3 + 1))) & 1
        words[first + 1 - bit] = 0x80000000 | leaf
        leaf += 1
        cur = first + bit
    words[cur] = 0x80000000 | leaf
    leaf += 1
    st = {"stree": words, "b_cnt": np.ones(leaf, np.uint32), "s_cnt": np.ones(leaf, np.uint32),
          "scal": np.zeros((6, leaf), np.float32), "b_sums": np.zeros((leaf, 4), np.float32),
          "b_child": np.zeros((leaf, 4), np.uint16), "s_sums": np.zeros((leaf, 4), np.float32),
          "s_child": np.zeros((leaf, 4), np.uint16)}
    tree = _tree(ctx)
    tree.Upload(st)
    rng = np.random.default_rng(5)
    q = [rng.uniform(0, 1, (2000, 3)),
         target + rng.normal(0, 1, (4000, 3)) * 10.0 ** rng.uniform(-7, -0.5, (4000, 1)),
         rng.integers(0, 65, (2000, 3)) / 64.0,
         rng.integers(0, 33, (500, 3)) / 32.0,
         rng.uniform(-0.3, 1.3, (2000, 3)),
         np.array([[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [1, 0, 1]], np.float64)]
    pos = (np.concatenate(q) * SIZE + LO).astype(np.float32)
    got = tree.Lookup(pos)
    want = orc.stree_lookup(words, LO, SIZE, pos)
    assert np.array_equal(got, want)
    assert len(np.unique(want)) == depth + 1  # every leaf is reached


def test_context_outlives_its_guides(gpu):
    # A context with live guides refuses to be destroyed (they use its stream);
    # Context.close() destroys the guides it created first, so a guide object
    # that outlives close() holds no handle to a freed context.
    from rtlib_amd import _lib
    from rtlib_amd.tracer import Context, MLPGuide, RTSTreeWrapper

    c = Context(0)
    tree = _tree(c)
    mlp = MLPGuide(c, LO, HI, log2_table=12)
    with pytest.raises(_lib.RtpgError, match="still alive"):
        _lib.call("rtpg_context_destroy", c.h)
    c.close()
    assert c.h is None and tree.h is None and mlp.h is None
    tree.Destroy()
    mlp.close()
    # the other order works too
    c2 = Context(0)
    t2 = RTSTreeWrapper(c2, LO, HI)
    t2.Destroy()
    c2.close()
    assert c2.h is None

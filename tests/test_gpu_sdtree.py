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

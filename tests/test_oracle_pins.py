"""Pins the CPU oracle before it is trusted as the checker (no GPU needed).

Each test names what fixes the expected value: a published known answer, a
closed form, an invariant, or brute force on small inputs.
"""
import math

import numpy as np
import pytest


# ---------------------------------------------------------------- random numbers
def test_pcg32_known_answer(orc):
    # pcg32-demo of pcg-c-basic: pcg32_srandom_r(&rng, 42, 54), first six outputs.
    want = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
    assert list(orc.pcg32_raw(42, 54, 6)) == want


def test_xorshift32_known_answer(orc):
    # Marsaglia, "Xorshift RNGs" (2003): xor32 from y = 2463534242 yields 723471715 first.
    assert int(orc.xorshift32(2463534242, 1)[0]) == 723471715


def test_to_uniform_float_closed_form(orc):
    rng = np.random.default_rng(0)
    for v in list(rng.integers(0, 2**32, 1000, dtype=np.uint64)) + [0, 2**32 - 1, 511, 512]:
        v = int(v)
        assert orc.to_uniform_float(v) == (v >> 9) / float(1 << 23)


def test_path_streams_are_distinct(orc):
    a = orc.rng_stream(0, 5, 0, 64)
    b = orc.rng_stream(0, 6, 0, 64)
    c = orc.rng_stream(0, 5, 1, 64)
    d = orc.rng_stream(1, 5, 0, 64)
    for x in (b, c, d):
        assert np.count_nonzero(a == x) < 3
    f = orc.rng_floats(0, 5, 0, 20000)
    assert f.min() >= 0.0 and f.max() < 1.0 and abs(f.mean() - 0.5) < 0.01


def test_pcg1d_is_a_permutation_on_a_range(orc):
    # pcg1d is an LCG step followed by an invertible output permutation, so it
    # has no collisions; a dropped or altered step would very likely collide.
    vals = np.array([orc.pcg1d(i) for i in range(1 << 16)], np.uint64)
    assert np.unique(vals).size == vals.size


# ---------------------------------------------------------------- Morton codes
def _interleave2(x, y):
    return int("".join(f"{(y >> b) & 1}{(x >> b) & 1}" for b in range(15, -1, -1)), 2)


def test_morton2_brute_force(orc):
    for x in range(0, 256, 3):
        for y in range(0, 256, 5):
            assert orc.morton2(x, y) == _interleave2(x, y)
    assert orc.morton2(0xFFFF, 0) == 0x55555555 and orc.morton2(0, 0xFFFF) == 0xAAAAAAAA


def test_morton3_brute_force_and_range(orc):
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, 1024, (500, 3)):
        want = 0
        for b in range(10):
            want |= ((int(x) >> b) & 1) << (3 * b) | ((int(y) >> b) & 1) << (3 * b + 1) | ((int(z) >> b) & 1) << (3 * b + 2)
        assert orc.morton3(int(x), int(y), int(z)) == want
    assert orc.morton3(1023, 1023, 1023) == (1 << 30) - 1
    # order preservation along each axis at fixed others
    assert all(orc.morton3(x + 1, 7, 9) > orc.morton3(x, 7, 9) for x in range(1023))


# ---------------------------------------------------------------- hash grid
def _grid(orc, key_mode=0, slots=4096, bounds=(16, 8, 16)):
    return orc.grid((0.0, 0.0, 0.0), (16.0, 8.0, 16.0), bounds, slots, key_mode)


@pytest.mark.parametrize("key_mode", [0, 1])
def test_grid_insert_then_probe(orc, key_mode):
    g = _grid(orc, key_mode)
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 1, (3000, 3)).astype(np.float32) * np.array([16, 8, 16], np.float32)
    cur = np.zeros(4096, np.uint32)
    prv = np.zeros(4096, np.uint32)
    s_ins = orc.grid_find(g, cur, prv, pos, True)
    assert np.all(s_ins != 0xFFFFFFFF)
    # same cell -> same slot; checksum stored is the linear cell index + 1 (mode 0)
    cells = np.floor(pos / 1.0).astype(np.int64)
    lin = 1 + 16 * 8 * cells[:, 2] + 16 * cells[:, 1] + cells[:, 0]
    if key_mode == 0:
        assert np.all(cur[s_ins] == lin)
    for c in np.unique(lin)[:200]:
        assert np.unique(s_ins[lin == c]).size == 1
    s_prv = orc.grid_find(g, cur, cur.copy(), pos, False)
    assert np.array_equal(s_prv, s_ins)


def test_grid_full_block_returns_invalid(orc):
    # 33 distinct cells in one 32-slot bucket: the 33rd cannot be placed
    # (GetCurCellIndex returns UINT32_MAX, OPX7UtilsGrid.h:288-291).
    g = _grid(orc, 0, slots=64, bounds=(64, 64, 64))  # 2 buckets
    cur = np.zeros(64, np.uint32)
    prv = np.zeros(64, np.uint32)
    picked = []
    for x in range(64):
        for y in range(64):
            p = np.array([[x + 0.5, y + 0.5, 0.5]], np.float32) * np.array([16 / 64, 8 / 64, 16 / 64], np.float32)
            s = orc.grid_find(g, cur.copy(), prv, p, True)[0]
            if s < 32:
                picked.append(p[0])
            if len(picked) == 33:
                break
        if len(picked) == 33:
            break
    out = orc.grid_find(g, cur, prv, np.array(picked), True)
    assert np.all(out[:32] < 32) and out[32] == 0xFFFFFFFF


# ---------------------------------------------------------------- Morton quadtree
def _mq_pyramid_by_reshape(leaves, level):
    """Sums of 2x2 blocks in code order: level l nodes = sum of their 4 children."""
    out = [leaves]
    cur = leaves
    for _ in range(level):
        cur = cur.reshape(cur.shape[0], -1, 4).sum(-1)
        out.append(cur)
    return np.concatenate(out[::-1], axis=1)


@pytest.mark.parametrize("level", [1, 3, 6])
def test_mq_build_equals_block_sums(orc, level):
    rng = np.random.default_rng(3)
    per = ((1 << (2 * (level + 1))) - 1) // 3
    w = np.zeros((5, per), np.float32)
    off = ((1 << (2 * level)) - 1) // 3
    w[:, off:] = rng.integers(0, 100, (5, 4**level)).astype(np.float32)  # integers: exact sums
    want = _mq_pyramid_by_reshape(w[:, off:].astype(np.float64), level)
    orc.mq_build(w, level)
    assert np.array_equal(w.astype(np.float64), want)


def test_mq_record_bins_by_morton(orc):
    level = 3
    per = 85
    w = np.zeros((2, per), np.float32)
    canon = np.array([[0.0, 0.0], [0.999, 0.0], [0.0, 0.999], [0.5, 0.5], [1.0, 1.0]], np.float32)
    orc.mq_record(w, level, np.array([1, 1, 1, 1, 1]), canon, np.array([1, 2, 3, 4, 5], np.float32))
    leaf = w[1, 21:]
    assert leaf[0] == 1 and leaf[_interleave2(7, 0)] == 2 and leaf[_interleave2(0, 7)] == 3
    assert leaf[_interleave2(4, 4)] == 4 and leaf[_interleave2(7, 7)] == 5  # u = 1 clamps to the last bin


def test_mq_pdf_integrates_to_one_and_samples_follow_it(orc):
    level = 3
    rng = np.random.default_rng(4)
    w = np.zeros((1, 85), np.float32)
    w[0, 21:] = rng.gamma(0.5, 1.0, 64).astype(np.float32)
    w[0, 21 + 5] = 0.0
    orc.mq_build(w, level)
    n = 8
    g = (np.arange(n) + 0.5) / n
    uv = np.stack(np.meshgrid(g, g, indexing="xy"), -1).reshape(-1, 2).astype(np.float32)
    pdf = orc.mq_pdf(w, level, np.zeros(uv.shape[0], np.uint32), uv)
    assert abs(pdf.sum() / uv.shape[0] * 4 * math.pi - 1.0) < 1e-5
    s = orc.mq_sample(w, level, np.zeros(40000, np.uint32), 7)
    # sampled pdf equals the pdf at the sample; histogram matches the leaf weights
    pdf_at = orc.mq_pdf(w, level, np.zeros(s.shape[0], np.uint32), s[:, :2])
    assert np.allclose(pdf_at, s[:, 2], rtol=1e-6)
    ix = np.minimum((s[:, 0] * 8).astype(int), 7)
    iy = np.minimum((s[:, 1] * 8).astype(int), 7)
    counts = np.bincount(iy * 8 + ix, minlength=64).astype(np.float64)
    p = np.array([w[0, 21 + _interleave2(x, y)] for y in range(8) for x in range(8)], np.float64)
    p /= p.sum()
    exp = p * s.shape[0]
    assert counts[p == 0].sum() == 0
    m = exp > 5
    chi2 = ((counts[m] - exp[m]) ** 2 / exp[m]).sum()
    assert chi2 < m.sum() + 5 * math.sqrt(2 * m.sum())


# ---------------------------------------------------------------- SD-tree
def _single_tree(sums_rows, child_rows):
    return np.array(sums_rows, np.float32).reshape(-1, 4), np.array(child_rows, np.uint16).reshape(-1, 4)


def test_split_threshold_reference_formula(orc):
    # size_t(sqrt(2^iter * spp / 4) * 4000) (OPX7UtilsPathGuiding.h:1649)
    assert orc.lib().orc_split_threshold(0, 4000.0, 1, 2) == 4000
    assert orc.lib().orc_split_threshold(0, 4000.0, 3, 8) == 16000
    assert orc.lib().orc_split_threshold(0, 4000.0, 2, 1) == 4000
    assert orc.lib().orc_split_threshold(1, 12000.0, 4, 99) == 48000


def test_dtree_reset_uniform_when_empty(orc):
    # total <= 0: fraction = 0.25^depth > 0.01 for depth <= 3 -> complete tree, 1+4+16+64 nodes
    s, c = _single_tree([0, 0, 0, 0], [0, 0, 0, 0])
    out = np.zeros((65536, 4), np.uint16)
    import ctypes as C

    d = C.c_uint32()
    n = orc.lib().orc_dtree_reset(orc._p(s), orc._p(c, orc.u16p), C.c_float(0.0), 20, C.c_float(0.01),
                                  orc._p(out, orc.u16p), C.byref(d))
    assert n == 85 and d.value == 4
    assert np.count_nonzero(out[:n]) == 84


def test_dtree_reset_single_slot_chain(orc):
    # all mass in slot 0 of a leaf root: depth 1 slot 0 (fraction 1) splits; its
    # children carry 1/4 (depth 2, split), 1/16 (depth 3, split), 1/64 (depth 4,
    # split: 0.0156 > 0.01), 1/256 (depth 5, no) -> 1 + 1 + 4 + 16 + 64 nodes.
    s, c = _single_tree([8.0, 0, 0, 0], [0, 0, 0, 0])
    import ctypes as C

    out = np.zeros((65536, 4), np.uint16)
    d = C.c_uint32()
    n = orc.lib().orc_dtree_reset(orc._p(s), orc._p(c, orc.u16p), C.c_float(8.0), 20, C.c_float(0.01),
                                  orc._p(out, orc.u16p), C.byref(d))
    assert n == 86 and d.value == 5
    # with max depth 3 the chain stops after depth 2: 1 + 1 + 4 nodes
    n3 = orc.lib().orc_dtree_reset(orc._p(s), orc._p(c, orc.u16p), C.c_float(8.0), 3, C.c_float(0.01),
                                   orc._p(out, orc.u16p), C.byref(d))
    assert n3 == 6


def _canon(sums, child, k=0):
    """Numbering-free serialization of a DTree (children as nested tuples)."""
    return tuple((float(sums[k, i]), _canon(sums, child, child[k, i]) if child[k, i] else None) for i in range(4))


def _shape(child, k=0):
    return tuple(_shape(child, child[k, i]) if child[k, i] else None for i in range(4))


def _lifo_numbering(shape):
    """Node numbers the reference's LIFO DFS assigns to a tree of this shape."""
    order = {}
    nodes = [shape]
    count = 1
    stack = [(0, shape)]
    child_idx = {}
    while stack:
        idx, sh = stack.pop()
        kids = []
        for i in range(4):
            if sh[i] is not None:
                kids.append((i, count))
                stack.append((count, sh[i]))
                count += 1
        child_idx[idx] = kids
    return child_idx, count


def _random_dtree(rng, depth=5, p_split=0.55):
    """A random sampling tree in LIFO numbering with built (consistent) sums."""
    rows_c, rows_s = [[0, 0, 0, 0]], [[0, 0, 0, 0]]
    stack = [(0, 1)]
    while stack:
        k, d = stack.pop()
        for i in range(4):
            if d < depth and rng.random() < p_split:
                rows_c[k][i] = len(rows_c)
                rows_c.append([0, 0, 0, 0])
                rows_s.append([0, 0, 0, 0])
                stack.append((rows_c[k][i], d + 1))
            else:
                rows_s[k][i] = float(rng.integers(0, 50)) if rng.random() < 0.8 else 0.0
    s = np.array(rows_s, np.float32)
    c = np.array(rows_c, np.uint16)
    cnt = np.array([s.shape[0]], np.uint32)
    tot, area = np.zeros(1, np.float32), np.zeros(1, np.float32)
    orc_lib = None
    return s, c


def _spec_reset(sums, child, total, max_depth, th):
    """The subdivision rule of RTDTreeT::Reset written recursively: the shape
    (nested tuples) of the tree it must produce."""
    def rec(src, depth, val):
        out = []
        for i in range(4):
            si = val if src is None else float(sums[src, i])
            frac = (np.float32(si) / np.float32(total)) if total > 0 else 0.25**depth
            if depth < max_depth and frac > np.float32(th):
                nxt = int(child[src, i]) if src is not None and child[src, i] else None
                out.append(rec(nxt, depth + 1, None if nxt is not None else np.float32(si) / np.float32(4.0)))
            else:
                out.append(None)
        return tuple(out)

    return rec(0, 1, None)


@pytest.mark.parametrize("seed", range(6))
def test_dtree_reset_matches_rule_and_numbering(orc, seed):
    import ctypes as C

    rng = np.random.default_rng(10 + seed)
    s, c = _random_dtree(rng)
    cnt = np.array([s.shape[0]], np.uint32)
    tot, area = np.zeros(1, np.float32), np.zeros(1, np.float32)
    orc.lib().orc_dtree_build(1, orc._p(cnt, orc.u32p), orc._p(s), orc._p(c, orc.u16p), orc._p(tot), orc._p(area))
    out = np.zeros((65536, 4), np.uint16)
    d = C.c_uint32()
    th = 0.01 if seed % 2 == 0 else 0.05
    n = orc.lib().orc_dtree_reset(orc._p(s), orc._p(c, orc.u16p), C.c_float(float(tot[0])), 20, C.c_float(th),
                                  orc._p(out, orc.u16p), C.byref(d))
    want_shape = _spec_reset(s, c, float(tot[0]), 20, th)
    got = out[:n]
    assert _shape(got) == want_shape
    kids, count = _lifo_numbering(want_shape)
    assert count == n
    for k, lst in kids.items():
        for i, num in lst:
            assert got[k, i] == num


@pytest.mark.parametrize("seed", range(4))
def test_dtree_build_sums_and_area(orc, seed):
    rng = np.random.default_rng(20 + seed)
    s, c = _random_dtree(rng)
    s0 = s.copy()
    cnt = np.array([s.shape[0]], np.uint32)
    tot, area = np.zeros(1, np.float32), np.zeros(1, np.float32)
    orc.lib().orc_dtree_build(1, orc._p(cnt, orc.u32p), orc._p(s), orc._p(c, orc.u16p), orc._p(tot), orc._p(area))
    leaf_mass = sum(float(s0[k, i]) for k in range(s0.shape[0]) for i in range(4) if c[k, i] == 0)
    assert float(tot[0]) == leaf_mass  # integer-valued leaves: exact
    for k in range(s.shape[0]):
        for i in range(4):
            if c[k, i]:
                assert s[k, i] == s[c[k, i]].sum()

    def cover(k):  # fraction of the square under non-zero slots
        a = 0.0
        for i in range(4):
            if s[k, i] > 0:
                a += 0.25 if c[k, i] == 0 else cover(c[k, i]) / 4
        return a

    assert abs(float(area[0]) - cover(0)) < 1e-6


def test_dtree_pdf_integrates_to_one_and_sampling_matches(orc):
    rng = np.random.default_rng(30)
    s, c = _random_dtree(rng, depth=4)
    st = {"stree": np.array([0x80000000], np.uint32), "b_cnt": np.array([s.shape[0]], np.uint32),
          "s_cnt": np.array([s.shape[0]], np.uint32), "scal": np.array([[5.0], [0], [0], [0], [0], [0]], np.float32),
          "b_sums": s, "b_child": c, "s_sums": s, "s_child": c}
    st = orc.sdtree_build(st)
    n = 64  # finer than depth 4 leaves (16 x 16)
    g = (np.arange(n) + 0.5) / n
    uv = np.stack(np.meshgrid(g, g, indexing="xy"), -1).reshape(-1, 2).astype(np.float32)
    pdf = orc.sdtree_pdf(st, np.zeros(uv.shape[0], np.uint32), uv)
    assert abs(pdf.mean() * 4 * math.pi - 1.0) < 1e-5
    smp = orc.sdtree_sample(st, np.zeros(50000, np.uint32), 3)
    cell = np.minimum((smp * 16).astype(int), 15)
    counts = np.bincount(cell[:, 1] * 16 + cell[:, 0], minlength=256).astype(np.float64)
    g16 = (np.arange(16) + 0.5) / 16
    uv16 = np.stack(np.meshgrid(g16, g16, indexing="xy"), -1).reshape(-1, 2).astype(np.float32)
    p16 = orc.sdtree_pdf(st, np.zeros(256, np.uint32), uv16).astype(np.float64)
    exp = p16 / p16.sum() * smp.shape[0]
    assert counts[exp == 0].sum() == 0
    m = exp > 5
    chi2 = ((counts[m] - exp[m]) ** 2 / exp[m]).sum()
    assert chi2 < m.sum() + 5 * math.sqrt(2 * m.sum())


def test_stree_refine_split_counts(orc):
    # one leaf with weight w splits while w / 2^k > th (children inherit half):
    # w = 1000, th = 100 -> 1000, 500, 250, 125 > 100 split, 62.5 stops: 16 leaves.
    import ctypes as C

    for w, th, leaves in [(1000.0, 100, 16), (100.0, 100, 1), (100.5, 100, 2), (0.0, 0, 1)]:
        out = np.zeros(1024, np.uint32)
        mp = np.zeros(1024, np.uint32)
        nd = C.c_uint32()
        n = orc.lib().orc_stree_refine(orc._p(np.array([0x80000000], np.uint32), orc.u32p), 1,
                                       orc._p(np.array([w], np.float32)), orc._p(np.array([1], np.uint32), orc.u32p),
                                       orc._p(np.array([1], np.uint32), orc.u32p), C.c_uint64(th), 2000,
                                       orc._p(out, orc.u32p), 1024, orc._p(mp, orc.u32p), C.byref(nd), 0)
        assert nd.value == leaves and n == 2 * leaves - 1
        assert np.all(mp[:leaves] == 0)
        # every inner node's children are a consecutive pair after it
        for i in range(n):
            if not out[i] & 0x80000000:
                assert out[i] > i and out[i] + 1 < n


def test_stree_octree_refine_matches_uniform_grid(orc):
    # 8-way splits (the STree2 variant): w = 1000, th = 100 -> 8 leaves of 125
    # split again -> 64 leaves of 15.6: the 4 x 4 x 4 grid, one DTree per cell,
    # every leaf three binary levels per octree level (depth 6)
    import ctypes as C

    for w, th, leaves in [(1000.0, 100, 64), (100.5, 100, 8), (99.0, 100, 1)]:
        out = np.zeros(4096, np.uint32)
        mp = np.zeros(4096, np.uint32)
        nd = C.c_uint32()
        n = orc.lib().orc_stree_refine(orc._p(np.array([0x80000000], np.uint32), orc.u32p), 1,
                                       orc._p(np.array([w], np.float32)), orc._p(np.array([1], np.uint32), orc.u32p),
                                       orc._p(np.array([1], np.uint32), orc.u32p), C.c_uint64(th), 2000,
                                       orc._p(out, orc.u32p), 4096, orc._p(mp, orc.u32p), C.byref(nd), 1)
        assert nd.value == leaves and n == 1 + 14 * (leaves - 1) // 7
        k = round(leaves ** (1 / 3))
        c = (np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) / k
        got = orc.stree_lookup(out[:n], (0, 0, 0), 1.0, c.astype(np.float32))
        assert sorted(got.tolist()) == list(range(leaves))
        # jittered points stay in the DTree of their grid cell
        rng = np.random.default_rng(0)
        jit = (c + rng.uniform(-0.45, 0.45, c.shape) / k).astype(np.float32)
        assert np.array_equal(orc.stree_lookup(out[:n], (0, 0, 0), 1.0, jit), got)


def test_stree_lookup_axis_cycle(orc):
    # root splits x at .5, its children split y, grandchildren z
    stree = np.array([1, 3, 5, 0x80000000, 0x80000001, 0x80000002, 0x80000003], np.uint32)
    pts = np.array([[0.2, 0.2, 0.9], [0.2, 0.7, 0.1], [0.7, 0.2, 0.1], [0.7, 0.9, 0.9]], np.float32)
    got = orc.stree_lookup(stree, (0, 0, 0), 1.0, pts)
    assert list(got) == [0, 1, 2, 3]


def test_sdtree_record_descends_to_first_leaf(orc):
    s = np.zeros((2, 4), np.float32)
    c = np.array([[1, 0, 0, 0], [0, 0, 0, 0]], np.uint16)
    st = {"stree": np.array([0x80000000], np.uint32), "b_cnt": np.array([2], np.uint32),
          "s_cnt": np.array([2], np.uint32), "scal": np.zeros((6, 1), np.float32),
          "b_sums": s, "b_child": c, "s_sums": s, "s_child": c}
    pos = np.full((4, 3), 0.5, np.float32)
    canon = np.array([[0.1, 0.1], [0.4, 0.1], [0.9, 0.2], [0.3, 0.3]], np.float32)
    out = orc.sdtree_record(st, (0, 0, 0), 1.0, pos, canon, np.array([1, 2, 4, 0], np.float32),
                            np.array([1, 1, 2, 1], np.float32))
    assert out["scal"][0][0] == 5.0  # all positive weights counted, also for zero irradiance
    # (0.1,0.1) -> root slot 0 -> node 1 slot 0; (0.4,0.1) -> node 1 slot 1; (0.9,0.2) -> root slot 1
    assert out["b_sums"][1, 0] == 1 and out["b_sums"][1, 1] == 2 and out["b_sums"][0, 1] == 8


# ---------------------------------------------------------------- geometry
def test_brute_force_intersection_closed_form(orc):
    verts = np.array([[-1, -1, 5], [3, -1, 5], [-1, 3, 5], [-1, -1, 7], [3, -1, 7], [-1, 3, 7]], np.float32)
    idx = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    rays = np.array([[0, 0, 0, 1e-4, 0, 0, 1, 1e20], [0, 0, 10, 1e-4, 0, 0, -1, 1e20],
                     [5, 5, 0, 1e-4, 0, 0, 1, 1e20], [0, 0, 0, 1e-4, 0, 0, 1, 4.0]], np.float32)
    t, u, v, prim = orc.brute_closest(verts, idx, rays)
    assert t[0] == 5.0 and prim[0] == 0 and u[0] == 0.25 and v[0] == 0.25
    assert t[1] == 3.0 and prim[1] == 1
    assert np.isinf(t[2]) and prim[2] == 0xFFFFFFFF
    assert np.isinf(t[3])

"""BVH4 traversal on the GPU against brute-force intersection (oracle)."""
import numpy as np
import pytest

from rtlib_amd import scenes

pytestmark = pytest.mark.gpu


def _rays(rng, n, lo, hi, tmax=1e20):
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 3] = 1e-4
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def _check(ctx, orc, sc, rays):
    ctx.upload_scene(sc)
    t, u, v, prim = ctx.trace_closest(rays)
    bt, bu, bv, bprim = orc.brute_closest(sc["verts"], sc["idx"], rays)
    same = prim == bprim
    # Rays through a shared edge, or onto coplanar faces (a box standing on the
    # floor), can report either triangle at the same t.
    assert np.all(t[~same] == bt[~same])
    assert same.mean() > 0.99
    assert np.array_equal(t, bt)
    assert np.array_equal(u[same], bu[same]) and np.array_equal(v[same], bv[same])
    occ = ctx.trace_occluded(np.concatenate([rays[:, :7], np.minimum(bt, 1e20)[:, None] * 0.5], 1))
    assert np.array_equal(occ, np.zeros_like(occ))  # nothing before half the closest distance
    occ2 = ctx.trace_occluded(rays)
    assert np.array_equal(occ2, np.isfinite(bt))


def test_cornell_closest_and_any_hit(ctx, orc):
    rng = np.random.default_rng(0)
    sc = scenes.cornell(16, 16)
    _check(ctx, orc, sc, _rays(rng, 20000, (-0.99, 0.01, -0.99), (0.99, 1.99, 0.99)))


def test_triangle_soup(ctx, orc):
    rng = np.random.default_rng(1)
    nt = 3000
    c = rng.uniform(-5, 5, (nt, 1, 3))
    verts = (c + rng.normal(scale=0.4, size=(nt, 3, 3))).reshape(-1, 3).astype(np.float32)
    sc = {"verts": verts, "idx": np.arange(3 * nt, dtype=np.uint32).reshape(-1, 3),
          "albedo": np.full((nt, 4), 0.5, np.float32), "mat": np.zeros(nt, np.uint32),
          "emission": np.zeros((nt, 3), np.float32), "spec": None, "camera": None}
    _check(ctx, orc, sc, _rays(rng, 4000, (-6, -6, -6), (6, 6, 6)))


def _soup(rng, nt, center, spread, size):
    c = center + rng.uniform(-spread, spread, (nt, 1, 3))
    verts = (c + rng.normal(scale=size, size=(nt, 3, 3))).reshape(-1, 3).astype(np.float32)
    return {"verts": verts, "idx": np.arange(3 * nt, dtype=np.uint32).reshape(-1, 3),
            "albedo": np.full((nt, 4), 0.5, np.float32), "mat": np.zeros(nt, np.uint32),
            "emission": np.zeros((nt, 3), np.float32), "spec": None, "camera": None}


@pytest.mark.parametrize("center,spread,size", [(0.0, 2.0e5, 2.0e4), (1.0e3, 0.05, 0.004)])
def test_node_planes_at_extreme_scales(ctx, orc, center, spread, size):
    # the nodes hold binary16 plane offsets from their corner: offsets beyond
    # 65504 (first scene) become +inf upper planes, and tiny boxes far from
    # the origin (second) rely on the offsets being relative to the corner
    rng = np.random.default_rng(4)
    sc = _soup(rng, 3000, center, spread, size)
    r = 1.2 * spread
    _check(ctx, orc, sc, _rays(rng, 4000, (center - r,) * 3, (center + r,) * 3))


def test_heightfield_room_deep_bvh(ctx, orc):
    rng = np.random.default_rng(2)
    sc = scenes.guided_room(grid=(256, 128))
    sc["camera"] = None
    rays = _rays(rng, 1500, (0.2, 0.3, 0.2), (7.8, 2.8, 3.8))
    _check(ctx, orc, sc, rays)
    info = ctx.scene_info()
    assert info["tris"] == sc["idx"].shape[0] and info["stack_bound"] >= 1

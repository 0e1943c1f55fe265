"""rtpg_trace_closest_noalloc (device buffers, fixed scratch array) against
rtpg_trace_closest on the same rays: identical hits, on two calls in a row,
the second with fewer rays."""
import ctypes as C

import numpy as np
import pytest

from rtlib_amd import _lib, scenes

pytestmark = pytest.mark.gpu


def _rays(rng, n, lo, hi):
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    r[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 3] = 1e-4
    r[:, 7] = 1e20
    return r


def _noalloc(ctx, rays):
    import torch

    n = rays.shape[0]
    d_rays = torch.from_numpy(rays).to("cuda")
    d_hits = torch.full((n, 4), -1.0, dtype=torch.float32, device="cuda")
    # the library runs on its own stream: the copies above must have landed
    torch.cuda.synchronize()
    assert d_rays.is_contiguous() and d_rays.shape == (n, 8) and d_hits.shape == (n, 4)
    print(f"noalloc trace: {n} rays", flush=True)
    _lib.call("rtpg_trace_closest_noalloc", ctx.h, C.c_void_p(d_rays.data_ptr()), n, C.c_void_p(d_hits.data_ptr()))
    return d_hits.cpu().numpy()


def test_noalloc_trace_matches_allocating_trace(ctx):
    rng = np.random.default_rng(11)
    sc = scenes.guided_room(grid=(256, 128))
    sc["camera"] = None
    ctx.upload_scene(sc)
    info = ctx.scene_info()
    # the room's BVH is deep enough that rays use the spill stacks in scratch
    assert info["stack_bound"] > 8, info
    per_ray = 8 * (info["stack_bound"] - 8)  # bytes of spill stack per ray
    # first call: more rays than one launch's stacks hold in 32 MiB
    n1 = (32 << 20) // per_ray + 5000
    for n in (n1, 4099):
        rays = _rays(rng, n, (0.2, 0.3, 0.2), (7.8, 2.8, 3.8))
        got = _noalloc(ctx, rays)
        t, u, v, prim = ctx.trace_closest(rays)
        want = np.stack([t, u, v, prim.view(np.float32)], 1)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), n
        assert np.isfinite(t).mean() > 0.99

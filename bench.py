"""Benchmark: SD-tree-guided vs plain path tracing on one MI355X.

Workload (the guided-room configuration): room with a half-open dividing wall,
a small ceiling light (0.5% of the ceiling) hidden behind it, a 1024 x 512
fBm heightfield floor (~1.05M triangles); Lambertian albedo U[0.2,0.8] per
triangle with 20% GGX (roughness U[0.05,0.5]); 1280 x 720, path depth 10, NEE
with MIS and Russian roulette; SD-tree training passes of 1, 2, 4, ... spp
(split threshold c = 12000 * sqrt(2^k), flux threshold 0.01, DTree depth 20)
up to a 64-spp budget.

One step = one complete guided render of the budget (all training passes,
recording, refinement and the final pass).  `value` is Msamples/s (camera
paths per second) of the guided render.  The line also carries the plain path
tracer's rate, the relMSE of both images at equal wall-clock time against a
high-spp reference, the roofline of the dominant kernel and the CPU baseline.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Msamples/s + relMSE at equal time: guided (SD-tree/MLP) vs plain PT, 1 GPU"
HBM_PEAK_GBS = 8000.0
NODE_BYTES = 128      # one BVH4 node (four child boxes + child refs)
TRI_BYTES = 36        # three float3 vertices per tested triangle
RAY_IN_BYTES = 32     # origin + direction
HIT_OUT_BYTES = 16    # t, u, v, triangle
QUEUE_BYTES = 4       # queue index read


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_mse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref * ref + 1e-2)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--budget", type=int, default=64)
    ap.add_argument("--ref-spp", type=int, default=1024)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--small", action="store_true", help="reduced size for a quick functional check")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in the design notes "
                          "but not launched", "n_gpus": a.gpus}))
        return 2

    from rtlib_amd import scenes
    from rtlib_amd.tracer import (FLAG_MIS, FLAG_NEE, FLAG_RR, Context, RTSTreeController, RTSTreeWrapper,
                                  render_guided)

    W, H = (320, 180) if a.small else (1280, 720)
    grid = (256, 128) if a.small else (1024, 512)
    depth = 10
    base = FLAG_NEE | FLAG_MIS | FLAG_RR
    t0 = time.time()
    sc = scenes.guided_room(width=W, height=H, grid=grid)
    log(f"scene: {sc['idx'].shape[0]} triangles ({time.time() - t0:.1f} s)")
    ctx = Context(0)
    t0 = time.time()
    ctx.upload_scene(sc)
    info = ctx.scene_info()
    log(f"BVH build + upload {time.time() - t0:.1f} s: {info['nodes']} nodes, depth {info['bvh_depth']}, "
        f"stack bound {info['stack_bound']}")
    npix = W * H
    lo, hi = info["aabb_min"] - 0.005, info["aabb_max"] + 0.005

    tree = RTSTreeWrapper(ctx, lo, hi, 20, 0.3)
    tree.SetSplit(mode=1, c=12000.0, max_mb=2000, subdiv_th=0.01)

    def guided_step(seed, sample_offset=0):
        ctrl = RTSTreeController(tree, a.budget, 0, 0.5, 1)
        ctx.clear()
        n = render_guided(ctx, ctrl, tree, "tree", base, max_depth=depth, seed=seed)
        return n

    def timed(fn):
        ctx.synchronize()
        t = time.perf_counter()
        r = fn()
        ctx.synchronize()
        return time.perf_counter() - t, r

    # ---- guided: warmup, then K timed steps with kernel timing on the render stream
    for w in range(a.warmup):
        dt, _ = timed(lambda: guided_step(100 + w))
        log(f"warmup step {w}: {dt * 1e3:.1f} ms")
    ctx.set_profile(time_kernels=True, count_work=False)
    times, kms = [], np.zeros(5)
    rays = shadow = 0
    for k in range(a.steps):
        dt, n = timed(lambda: guided_step(k + 1))
        times.append(dt)
        log(f"step {k}: {dt * 1e3:.1f} ms ({n} spp)")
    # kernel times accumulate only over the last render call; re-time one step
    # pass by pass to collect per-kernel totals over a whole step
    ctx.set_profile(time_kernels=True, count_work=False)
    per_kernel = np.zeros(5)
    launches = 0
    ctrl = RTSTreeController(tree, a.budget, 0, 0.5, 1)
    ctx.clear()

    def acc_pass(state, spp, traced):
        nonlocal launches, rays, shadow
        ms, _ = ctx.kernel_stats()
        per_kernel[:] += np.array([ms[k] for k in ("extend", "shade", "shadow", "accumulate", "commit")])
        st = ctx.render_stats()
        rays += st["rays"]
        shadow += st["shadow_rays"]
        launches += 1

    t_prof0 = time.perf_counter()
    render_guided(ctx, ctrl, tree, "tree", base, max_depth=depth, seed=1, on_pass=acc_pass)
    ctx.synchronize()
    t_prof = time.perf_counter() - t_prof0
    guided_img = ctx.image()
    # ---- work counts of the same step (separate, untimed run)
    ctx.set_profile(time_kernels=False, count_work=True)
    work = np.zeros(4)
    rays_c = 0

    def acc_work(state, spp, traced):
        nonlocal rays_c
        _, wk = ctx.kernel_stats()
        work[:] += np.array([wk["extend_nodes"], wk["extend_tris"], wk["shadow_nodes"], wk["shadow_tris"]])
        rays_c += ctx.render_stats()["rays"]

    ctrl = RTSTreeController(tree, a.budget, 0, 0.5, 1)
    ctx.clear()
    render_guided(ctx, ctrl, tree, "tree", base, max_depth=depth, seed=1, on_pass=acc_work)
    ctx.set_profile(False, False)

    t_step = float(np.median(times))
    samples = a.budget * npix
    value = samples / t_step / 1e6

    # ---- plain path tracing on the same scene
    ctx.clear()
    dt_plain, _ = timed(lambda: ctx.render(spp=a.budget, flags=base, max_depth=depth, seed=77))
    plain_rate = a.budget * npix / dt_plain / 1e6
    st_plain = ctx.render_stats()
    # equal time: as many plain samples as the guided step's wall clock allows
    spp_eq = max(1, int(round(t_step * plain_rate * 1e6 / npix)))
    ctx.clear()
    ctx.render(spp=spp_eq, flags=base, max_depth=depth, seed=78)
    plain_img = ctx.image()
    # ---- reference image (plain, many samples, independent seed)
    t0 = time.time()
    ctx.clear()
    done = 0
    while done < a.ref_spp:
        k = min(256, a.ref_spp - done)
        ctx.render(spp=k, flags=base, max_depth=depth, seed=4242, sample_offset=done)
        done += k
    ref = ctx.image()
    log(f"reference {a.ref_spp} spp in {time.time() - t0:.1f} s")
    rm_guided = rel_mse(guided_img, ref)
    rm_plain = rel_mse(plain_img, ref)

    # ---- roofline of the dominant kernel
    names = ["extend", "shade", "shadow", "accumulate", "commit"]
    dom = int(np.argmax(per_kernel))
    ext_bytes = rays_c * (RAY_IN_BYTES + HIT_OUT_BYTES + QUEUE_BYTES) + work[0] * NODE_BYTES + work[1] * TRI_BYTES
    n_ext_launch = launches * depth
    achieved = ext_bytes / (per_kernel[0] * 1e-3) / 1e9 if per_kernel[0] > 0 else 0.0
    roofline = {"kernel": "k_extend (closest-hit BVH4 traversal)", "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None, "bytes_per_step": float(ext_bytes), "launches_per_step": n_ext_launch,
                "avg_launch_ms": round(per_kernel[0] / max(n_ext_launch, 1), 4),
                "dominant_by_time": names[dom],
                "kernel_ms_per_step": {n: round(float(x), 2) for n, x in zip(names, per_kernel)}}

    # ---- CPU baseline: the oracle integrator on the host cores, guided final pass
    cpu = cpu_baseline(sc, tree, lo, hi, W, H, depth, a.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_step * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (procedural scene, seeded)",
        "config": {"workload": "guided room: 1.05M-tri fBm heightfield, hidden light, 1280x720, depth 10, "
                               "SD-tree passes 1,2,4.. to 64 spp" if not a.small else "guided room (small)",
                   "width": W, "height": H, "triangles": int(sc["idx"].shape[0]), "budget_spp": a.budget,
                   "max_depth": depth, "parallelism": "single GPU"},
        "plain_msamples_per_s": round(plain_rate, 2),
        "mrays_per_s": {"guided": round((rays + shadow) / t_prof / 1e6, 1),
                        "plain": round((st_plain["rays"] + st_plain["shadow_rays"]) / dt_plain / 1e6, 1)},
        "relmse_equal_time": {"guided": round(rm_guided, 5), "plain": round(rm_plain, 5),
                              "plain_spp": spp_eq, "guided_spp": a.budget, "reference_spp": a.ref_spp},
        "sdtree": tree.Sizes(),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(sc, tree, lo, hi, W, H, depth, seconds):
    """Times the oracle's CPU integrator on a bounded pixel sample of the same
    guided render (final-pass state of the trained tree)."""
    try:
        from oracle import oracle
        from rtlib_amd import scenes
    except Exception as e:  # noqa: BLE001
        return {"error": str(e)}
    oracle.build()
    cam = sc["camera"]
    uvw = scenes.camera_uvw(cam["eye"], cam["lookat"], cam["up"], cam["fovy"], W / H)
    t0 = time.time()
    osc = oracle.Scene(sc)
    t_build = time.time() - t0
    state = tree.Download()
    guide = oracle.Guide(state, lo, float(np.max(hi - lo)))
    threads = min(os.cpu_count() or 1, 16)
    flags = 1 | 128 | 256 | 8 | 16 | 32  # NEE, MIS, RR, tree, guide sampling, final
    # calibrate on one row, then size the sample to the time budget
    acc = np.zeros((H, W, 3), np.float32)
    t0 = time.perf_counter()
    oracle.render(osc, uvw, W, H, 1, flags, max_depth=depth, seed=5, guide=guide, px_range=(0, W),
                  threads=threads, accum=acc)
    t_row = max(time.perf_counter() - t0, 1e-3)
    rows = int(max(1, min(H, seconds / t_row)))
    stride = max(1, H // rows)
    total_px, t_total, rays = 0, 0.0, 0
    for r in range(0, H, stride):
        t0 = time.perf_counter()
        _, nr = oracle.render(osc, uvw, W, H, 1, flags, max_depth=depth, seed=5, guide=guide,
                              px_range=(r * W, (r + 1) * W), threads=threads, accum=acc)
        t_total += time.perf_counter() - t0
        total_px += W
        rays += nr
        if t_total > seconds:
            break
    return {"value": round(total_px / t_total / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"{total_px} camera paths (1 spp on every {stride}th row of {W}x{H}), guided final pass "
                      f"with the GPU-trained SD-tree; BVH build {t_build:.1f} s excluded",
            "mrays_per_s": round(rays / t_total / 1e6, 3)}


if __name__ == "__main__":
    sys.exit(main())

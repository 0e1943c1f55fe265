"""Benchmark: guided (SD-tree, learned) vs plain path tracing on one MI355X.

Workload (the guided-room configuration): room with a half-open dividing wall,
a small ceiling light (0.5% of the ceiling) hidden behind it, a 1024 x 512
fBm heightfield floor (~1.05M triangles); Lambertian albedo U[0.2,0.8] per
triangle with 20% GGX (roughness U[0.05,0.5]); 1280 x 720, path depth 10, NEE
with MIS and Russian roulette; SD-tree training passes of 1, 2, 4, ..., 64 spp
(split threshold c = 12000 * sqrt(2^k), flux threshold 0.01, DTree depth 20)
and a final pass, 256 spp in all.

One step = one complete SD-tree-guided render of the budget (all training
passes, recording, refinement and the final pass).  `value` is Msamples/s
(camera paths per second) of that render.  The line also carries the plain
path tracer's rate, the relMSE at equal wall-clock time of the SD-tree-guided
and the learned-guide renders against plain path tracing (reference: a
high-spp plain render), the roofline of the dominant kernel and the CPU
baseline.

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
NODE_BYTES = 112      # what the traversal reads of a 128-byte BVH4 node: four child boxes + child entries
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
    ap.add_argument("--budget", type=int, default=256,
                    help="samples per pixel of one guided render: SD-tree passes 1, 2, 4, .., 64 and the final pass")
    ap.add_argument("--ref-spp", type=int, default=16384)
    ap.add_argument("--no-configs", action="store_true", help="skip the Cornell / many-light / adjoint / "
                    "learned-guide configurations")
    ap.add_argument("--mlp-passes", type=int, default=8)
    ap.add_argument("--mlp-steps", type=int, default=4)
    ap.add_argument("--mlp-cache", type=int, default=48, help="spatial-cache cells per axis of the learned guide "
                    "(0: network inference at every guided vertex)")
    ap.add_argument("--large-budget", type=int, default=1024, help="second budget for the equal-time comparison "
                    "of the SD-tree and the learned guide (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--small", action="store_true", help="reduced size for a quick functional check")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (once untimed, then 5 times) and exit; used to scale "
                         "the FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args(argv)
    if a.copy_calib:
        return copy_calibration(a.copy_calib)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in the design notes "
                          "but not launched", "n_gpus": a.gpus}))
        return 2

    from rtlib_amd import scenes
    from rtlib_amd.tracer import (FLAG_MIS, FLAG_NEE, FLAG_RR, Context, DoubleBufferedHashGrid3Buffer, MLPGuide,
                                  RTMortonQuadTreeController, RTMortonQuadTreeWrapper, RTSTreeController,
                                  RTSTreeWrapper, render_guided, render_mlp_guided)

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

    def guided_step(seed, on_pass=None):
        ctrl = RTSTreeController(tree, a.budget, 0, 0.5, 1)
        ctx.clear()
        return render_guided(ctx, ctrl, tree, "tree", base, max_depth=depth, seed=seed, on_pass=on_pass)

    def timed(fn):
        ctx.synchronize()
        t = time.perf_counter()
        r = fn()
        ctx.synchronize()
        return time.perf_counter() - t, r

    # ---- guided: warmup, then K timed steps.  Kernel classes are timed with
    # HIP events on the render stream during the timed steps themselves; the
    # window of the timed steps is logged (CLOCK_BOOTTIME and CLOCK_MONOTONIC
    # ns) so that a kernel trace can be cut to exactly these dispatches.
    for w in range(a.warmup):
        dt, _ = timed(lambda: guided_step(100 + w))
        log(f"warmup step {w}: {dt * 1e3:.1f} ms")
    names = ["extend", "shade", "shadow", "accumulate", "commit"]
    ctx.set_profile(time_kernels=True, count_work=False)
    times, step_ms, step_launches = [], [], []
    rays = shadow = 0

    def acc_pass(state, spp, traced):
        nonlocal rays, shadow
        ms, _, nl = ctx.kernel_stats()
        step_ms[-1] += np.array([ms[k] for k in names])
        step_launches[-1] += np.array([nl[k] for k in names])
        st = ctx.render_stats()
        rays += st["rays"]
        shadow += st["shadow_rays"]

    win = [time.clock_gettime_ns(time.CLOCK_BOOTTIME), time.monotonic_ns()]
    for k in range(a.steps):
        step_ms.append(np.zeros(5))
        step_launches.append(np.zeros(5, np.int64))
        dt, n = timed(lambda: guided_step(k + 1, on_pass=acc_pass))
        times.append(dt)
        log(f"step {k}: {dt * 1e3:.1f} ms ({n} spp)")
    win += [time.clock_gettime_ns(time.CLOCK_BOOTTIME), time.monotonic_ns()]
    log(f"timed window ns: boottime {win[0]} {win[2]} monotonic {win[1]} {win[3]}")
    guided_img = ctx.image()
    per_kernel = np.mean(step_ms, axis=0)
    launches = step_launches[0]
    # ---- work counts of step 0 (seed 1) in a separate, untimed run: same
    # random streams, so the same paths and the same traversal work
    ctx.set_profile(time_kernels=False, count_work=True)
    work = np.zeros(4)
    rays_c = 0

    def acc_work(state, spp, traced):
        nonlocal rays_c
        _, wk, _ = ctx.kernel_stats()
        work[:] += np.array([wk["extend_nodes"], wk["extend_tris"], wk["shadow_nodes"], wk["shadow_tris"]])
        rays_c += ctx.render_stats()["rays"]

    guided_step(1, on_pass=acc_work)
    ctx.set_profile(False, False)
    # the cost of the per-kernel events: step 0 again without them
    dt_noev, _ = timed(lambda: guided_step(1))
    log(f"step 0 without kernel events: {dt_noev * 1e3:.1f} ms")

    t_step = float(np.median(times))
    samples = a.budget * npix
    value = samples / t_step / 1e6
    t_rays = float(np.sum(times))

    # ---- plain path tracing on the same scene
    ctx.clear()
    dt_plain, _ = timed(lambda: ctx.render(spp=a.budget, flags=base, max_depth=depth, seed=77))
    plain_rate = a.budget * npix / dt_plain / 1e6
    st_plain = ctx.render_stats()
    plain_budget_img = ctx.image()
    # kernel classes of the same plain render, timed in a separate run
    ctx.set_profile(time_kernels=True, count_work=False)
    ctx.clear()
    ctx.render(spp=a.budget, flags=base, max_depth=depth, seed=77)
    plain_kernel_ms = ctx.kernel_stats()[0]
    ctx.set_profile(False, False)
    quality = None
    if a.ref_spp > 0:
        # learned guide on the same budget: a.mlp_passes recording passes of 1
        # spp, each followed by a.mlp_steps Adam steps, then one final guided
        # pass that reads the spatial cache of the network (a.mlp_cache cells
        # per axis); and the same with inference at every guided vertex
        mlp = MLPGuide(ctx, lo, hi, seed=1, cache_res=a.mlp_cache)
        mlp_pv = MLPGuide(ctx, lo, hi, seed=1)

        def mlp_step(g=mlp, budget=a.budget):
            g.reset(1)
            ctx.clear()
            return render_mlp_guided(ctx, g, budget, base, max_depth=depth, seed=11,
                                     train_passes=a.mlp_passes, steps_per_pass=a.mlp_steps)

        timed(mlp_step)
        t_mlp, _ = timed(mlp_step)
        mlp_img = ctx.image()
        log(f"learned guide step: {t_mlp * 1e3:.1f} ms ({a.budget} spp, cache {a.mlp_cache})")
        timed(lambda: mlp_step(mlp_pv))
        t_pv, _ = timed(lambda: mlp_step(mlp_pv))
        pv_img = ctx.image()
        log(f"learned guide step, per-vertex inference: {t_pv * 1e3:.1f} ms ({a.budget} spp)")
        mlp_pv.close()

        def plain_equal_time(t, seed):
            spp = max(1, int(round(t * plain_rate * 1e6 / npix)))
            ctx.clear()
            ctx.render(spp=spp, flags=base, max_depth=depth, seed=seed)
            return spp, ctx.image()

        # hash-grid variant: 30-bit Morton cell keys, 2^20 slots, 8 x 8 direction
        # bins per cell, the reference's Morton controller (Locate first, trees
        # built from iteration 3)
        grid = DoubleBufferedHashGrid3Buffer(ctx, lo, hi, bounds=(256, 64, 256), slots=1 << 20, key_mode=1)
        mq = RTMortonQuadTreeWrapper(ctx, 1 << 20, 3, 0.3)

        def grid_step():
            grid.Clear()
            mq.Allocate()
            ctrl = RTMortonQuadTreeController(mq, grid, a.budget, 3, 0, 0.5, 1)
            ctx.clear()
            return render_guided(ctx, ctrl, mq, "grid", base, max_depth=depth, seed=13, grid=grid)

        timed(grid_step)
        t_grid, _ = timed(grid_step)
        grid_img = ctx.image()
        log(f"hash-grid guided step: {t_grid * 1e3:.1f} ms ({a.budget} spp)")

        # SD-tree with the stochastic box recording filters (spatial + directional)
        tree.SetFilter(1, 1, 17)
        timed(lambda: guided_step(90))
        t_box, _ = timed(lambda: guided_step(91))
        box_img = ctx.image()
        box_sizes = tree.Sizes()
        tree.SetFilter(0, 0)
        log(f"SD-tree guided step, box filters: {t_box * 1e3:.1f} ms ({a.budget} spp)")

        spp_tree, plain_tree = plain_equal_time(t_step, 78)
        spp_box, plain_box = plain_equal_time(t_box, 81)
        spp_mlp, plain_mlp = plain_equal_time(t_mlp, 79)
        spp_pv, plain_pv = plain_equal_time(t_pv, 82)
        spp_grid, plain_grid = plain_equal_time(t_grid, 80)
        # the SD-tree and the learned guide on a larger budget, where their
        # training passes are a smaller share of the time
        if a.large_budget > 0:
            def tree_large():
                ctrl = RTSTreeController(tree, a.large_budget, 0, 0.5, 1)
                ctx.clear()
                return render_guided(ctx, ctrl, tree, "tree", base, max_depth=depth, seed=92)

            timed(tree_large)
            t_tl, _ = timed(tree_large)
            tl_img = ctx.image()
            timed(lambda: mlp_step(mlp, a.large_budget))
            t_ml, _ = timed(lambda: mlp_step(mlp, a.large_budget))
            ml_img = ctx.image()
            log(f"budget {a.large_budget}: SD-tree {t_tl * 1e3:.1f} ms, learned guide {t_ml * 1e3:.1f} ms")
            spp_tl, plain_tl = plain_equal_time(t_tl, 83)
            spp_ml, plain_ml = plain_equal_time(t_ml, 84)
        # reference image (plain, many samples, independent seed)
        t0 = time.time()
        ctx.clear()
        done = 0
        while done < a.ref_spp:
            k = min(256, a.ref_spp - done)
            ctx.render(spp=k, flags=base, max_depth=depth, seed=4242, sample_offset=done)
            done += k
        ref = ctx.image()
        log(f"reference {a.ref_spp} spp in {time.time() - t0:.1f} s")
        quality = {
            "reference_spp": a.ref_spp,
            "plain_at_budget": {"spp": a.budget, "relmse": round(rel_mse(plain_budget_img, ref), 5)},
            "sdtree": {"ms": round(t_step * 1e3, 1), "relmse": round(rel_mse(guided_img, ref), 5),
                       "plain_spp": spp_tree, "plain_relmse": round(rel_mse(plain_tree, ref), 5)},
            "sdtree_box_filters": {"ms": round(t_box * 1e3, 1), "relmse": round(rel_mse(box_img, ref), 5),
                                   "plain_spp": spp_box, "plain_relmse": round(rel_mse(plain_box, ref), 5),
                                   "dtrees": box_sizes["dtrees"]},
            "hashgrid": {"ms": round(t_grid * 1e3, 1), "relmse": round(rel_mse(grid_img, ref), 5),
                         "plain_spp": spp_grid, "plain_relmse": round(rel_mse(plain_grid, ref), 5)},
            "mlp": {"ms": round(t_mlp * 1e3, 1), "relmse": round(rel_mse(mlp_img, ref), 5),
                    "plain_spp": spp_mlp, "plain_relmse": round(rel_mse(plain_mlp, ref), 5),
                    "train_passes": a.mlp_passes, "steps_per_pass": a.mlp_steps, "cache_res": a.mlp_cache},
            "mlp_per_vertex_inference": {"ms": round(t_pv * 1e3, 1), "relmse": round(rel_mse(pv_img, ref), 5),
                                         "plain_spp": spp_pv, "plain_relmse": round(rel_mse(plain_pv, ref), 5)},
        }
        if a.large_budget > 0:
            quality[f"budget_{a.large_budget}"] = {
                "sdtree": {"ms": round(t_tl * 1e3, 1), "relmse": round(rel_mse(tl_img, ref), 5),
                           "plain_spp": spp_tl, "plain_relmse": round(rel_mse(plain_tl, ref), 5)},
                "mlp": {"ms": round(t_ml * 1e3, 1), "relmse": round(rel_mse(ml_img, ref), 5),
                        "plain_spp": spp_ml, "plain_relmse": round(rel_mse(plain_ml, ref), 5)},
            }
        mlp.close()

    # ---- roofline of the dominant kernel (k_extend): algorithmic bytes of
    # step 0 over its event-timed k_extend time
    dom = int(np.argmax(per_kernel))
    ext_bytes = rays_c * (RAY_IN_BYTES + HIT_OUT_BYTES + QUEUE_BYTES) + work[0] * NODE_BYTES + work[1] * TRI_BYTES
    n_ext = int(launches[0])
    ext_ms0 = float(step_ms[0][0])
    achieved = ext_bytes / (ext_ms0 * 1e-3) / 1e9 if ext_ms0 > 0 else 0.0
    traffic, traffic_src = _traffic(W, H)
    roofline = {"kernel": "k_extend (closest-hit BVH4 traversal)", "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "traffic_source": traffic_src,
                "algorithmic_bytes_per_launch": round(ext_bytes / max(n_ext, 1)),
                "launches_per_step": n_ext,
                "avg_launch_ms": round(float(per_kernel[0]) / max(n_ext, 1), 4),
                "dominant_by_time": names[dom],
                "kernel_ms_per_step": {n: round(float(x), 2) for n, x in zip(names, per_kernel)},
                "plain_kernel_ms_per_step": {n: round(float(plain_kernel_ms[n]), 2) for n in names},
                "kernel_launches_per_step": {n: int(x) for n, x in zip(names, launches)}}

    # ---- the other configurations, measured on the same context
    configs = None if (a.small or a.no_configs) else other_configs(ctx, timed)

    # ---- CPU baseline: the oracle integrator on the host cores, guided final pass
    cpu = cpu_baseline(sc, tree, lo, hi, W, H, depth, a.cpu_seconds) if a.cpu_seconds > 0 else None

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_step * 1e3, 2),
        "ms_per_step_without_kernel_events": round(dt_noev * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (procedural scene, seeded)",
        "config": {"workload": "guided room: 1.05M-tri fBm heightfield, hidden light, 1280x720, depth 10, "
                               f"SD-tree passes 1,2,4..64 spp + final pass, {a.budget} spp budget" if not a.small
                   else "guided room (small)",
                   "width": W, "height": H, "triangles": int(sc["idx"].shape[0]), "budget_spp": a.budget,
                   "max_depth": depth, "parallelism": "single GPU"},
        "plain_msamples_per_s": round(plain_rate, 2),
        "mrays_per_s": {"guided": round((rays + shadow) / t_rays / 1e6, 1),
                        "plain": round((st_plain["rays"] + st_plain["shadow_rays"]) / dt_plain / 1e6, 1)},
        "relmse_equal_time": quality,
        "sdtree": tree.Sizes(),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "configs": configs,
    }
    print(json.dumps(out), flush=True)
    return 0


TRAFFIC_FILE = ROOT / "profiles" / "k_extend_guided_room_traffic.json"


def other_configs(ctx, timed):
    """Throughput (and, for the learned guide, equal-time relMSE) of the other
    benchmark configurations: Cornell box plain path tracing (256^2, 64 spp,
    depth 8, RR after 3); many-light RIS (2000 icospheres, 2.56M tris, 4096
    emissive triangles, 1920x1080, 16 spp, depth 6, M = 32); path-replay
    adjoint (8 orbit views 45 degrees apart at 512^2, 16 spp, depth 6, on the
    Cornell box and on the 1.05M-triangle guided room);
    learned guide on the Cornell box against SD-tree guiding and plain path
    tracing at equal time (64-spp budgets, 16k-spp reference)."""
    from rtlib_amd import scenes
    from rtlib_amd.tracer import (FLAG_MIS, FLAG_NEE, FLAG_RIS, FLAG_RR, MLPGuide, RTSTreeController,
                                  RTSTreeWrapper, render_guided, render_mlp_guided)
    out = {}
    nee_mis_rr = FLAG_NEE | FLAG_MIS | FLAG_RR
    # Cornell box, plain
    sc = scenes.cornell(256, 256)
    ctx.upload_scene(sc)
    ctx.clear()
    ctx.render(spp=64, flags=nee_mis_rr, max_depth=8, rr_depth=3, seed=1)
    ts = []
    for k in range(3):
        ctx.clear()
        dt, _ = timed(lambda: ctx.render(spp=64, flags=nee_mis_rr, max_depth=8, rr_depth=3, seed=2 + k))
        ts.append(dt)
    st = ctx.render_stats()
    t = float(np.median(ts))
    out["cornell_plain"] = {"msamples_per_s": round(256 * 256 * 64 / t / 1e6, 1), "ms": round(t * 1e3, 2),
                            "mrays_per_s": round((st["rays"] + st["shadow_rays"]) / t / 1e6, 1)}
    log(f"cornell: {out['cornell_plain']}")
    # learned guide vs SD-tree vs plain on the Cornell box at equal time
    ctx.clear()
    for k in range(64):
        ctx.render(spp=256, flags=nee_mis_rr, max_depth=8, seed=900, sample_offset=256 * k)
    ref = ctx.image()
    info = ctx.scene_info()
    lo, hi = info["aabb_min"] - 0.005, info["aabb_max"] + 0.005
    plain_rate = 256 * 256 * 64 / t
    mlp = MLPGuide(ctx, lo, hi, seed=1)

    def mlp_step():
        mlp.reset(1)
        ctx.clear()
        return render_mlp_guided(ctx, mlp, 64, nee_mis_rr, max_depth=8, seed=21, train_passes=8, steps_per_pass=4)

    timed(mlp_step)
    t_mlp, _ = timed(mlp_step)
    img_mlp = ctx.image()
    mlp.close()
    tree = RTSTreeWrapper(ctx, lo, hi, 20, 0.3)
    tree.SetSplit(mode=1, c=12000.0)

    def tree_step():
        tree.Clear()
        ctrl = RTSTreeController(tree, 64, 0, 0.5, 1)
        ctx.clear()
        return render_guided(ctx, ctrl, tree, "tree", nee_mis_rr, max_depth=8, seed=22)

    timed(tree_step)
    t_tree, _ = timed(tree_step)
    img_tree = ctx.image()
    res = {}
    for name, tt, img in (("mlp", t_mlp, img_mlp), ("sdtree", t_tree, img_tree)):
        spp = max(1, int(round(tt * plain_rate / (256 * 256))))
        ctx.clear()
        ctx.render(spp=spp, flags=nee_mis_rr, max_depth=8, seed=23)
        res[name] = {"ms": round(tt * 1e3, 2), "relmse": round(rel_mse(img, ref), 6), "plain_spp": spp,
                     "plain_relmse": round(rel_mse(ctx.image(), ref), 6)}
    ctx.clear()
    ctx.render(spp=64, flags=nee_mis_rr, max_depth=8, seed=24)
    res["plain_at_budget"] = {"spp": 64, "relmse": round(rel_mse(ctx.image(), ref), 6)}
    out["cornell_guided_equal_time"] = res
    log(f"cornell guided: {res}")
    # many-light RIS
    sc = scenes.many_lights()
    ctx.upload_scene(sc)
    ris = FLAG_NEE | FLAG_RIS
    ctx.clear()
    ctx.render(spp=4, flags=ris, max_depth=6, num_candidates=32, seed=1)
    ctx.clear()
    dt, _ = timed(lambda: ctx.render(spp=16, flags=ris, max_depth=6, num_candidates=32, seed=2))
    st = ctx.render_stats()
    out["many_light_ris"] = {"msamples_per_s": round(1920 * 1080 * 16 / dt / 1e6, 1), "ms": round(dt * 1e3, 1),
                             "mrays_per_s": round((st["rays"] + st["shadow_rays"]) / dt / 1e6, 1),
                             "triangles": int(sc["idx"].shape[0])}
    log(f"many-light RIS: {out['many_light_ris']}")
    # path-replay adjoint, 8 orbit views
    sc = scenes.cornell(512, 512)
    ctx.upload_scene(sc)
    rng = np.random.default_rng(0)
    dldi = rng.uniform(-1, 1, (512, 512, 3)).astype(np.float32)
    views = []
    for k in range(8):
        ang = np.radians(45.0 * k)
        views.append(((3.8 * np.sin(ang), 1.0, 3.8 * np.cos(ang)), (0.0, 1.0, 0.0)))

    def adjoint_batch():
        for eye, look in views:
            ctx.set_camera_look(eye, look, (0, 1, 0), 40.0, 512, 512)
            ctx.adjoint(dldi, spp=16, flags=FLAG_NEE | FLAG_MIS, max_depth=6, seed=3)

    adjoint_batch()
    dt, _ = timed(adjoint_batch)
    out["adjoint"] = {"ms_per_batch": round(dt * 1e3, 1),
                      "msamples_per_s": round(8 * 512 * 512 * 16 / dt / 1e6, 1),
                      "note": "primal + replay per view, gradients downloaded and permuted on the host"}
    log(f"adjoint: {out['adjoint']}")

    # the same adjoint batch on the 1.05M-triangle guided room (about 1M x 3
    # albedo parameters): 8 views 45 degrees apart orbiting the left part of
    # the room from inside
    sc = scenes.guided_room(512, 512)
    ctx.upload_scene(sc)
    views = []
    for k in range(8):
        ang = np.radians(45.0 * k)
        views.append(((2.5 + 1.5 * np.sin(ang), 1.5, 2.0 + 1.5 * np.cos(ang)), (2.5, 1.0, 2.0)))
    adjoint_batch()
    dt, _ = timed(adjoint_batch)
    out["adjoint_room"] = {"ms_per_batch": round(dt * 1e3, 1),
                           "msamples_per_s": round(8 * 512 * 512 * 16 / dt / 1e6, 1),
                           "parameters": int(sc["idx"].shape[0]) * 3 + 3}
    log(f"adjoint (room): {out['adjoint_room']}")
    return out


def _traffic(W, H):
    """HBM bytes per k_extend launch from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by the counter's reading for a copy kernel of known size)."""
    try:
        t = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None, None
    if t.get("width") != W or t.get("height") != H:
        return None, None
    b = t["fetch_kib_per_dispatch"] * 1024 * t["fetch_scale"] + t["write_kib_per_dispatch"] * 1024 * t["write_scale"]
    return round(b), TRAFFIC_FILE.name


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 4
    x = torch.arange(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "dispatches": 5}), flush=True)
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

"""Benchmark: guided (SD-tree, hash grid, learned) vs plain path tracing on one MI355X.

Workload (the guided-room configuration): room with a half-open dividing wall,
a small ceiling light (0.5% of the ceiling) hidden behind it, a 1024 x 512
fBm heightfield floor (~1.05M triangles); Lambertian albedo U[0.2,0.8] per
triangle with 20% GGX (roughness U[0.05,0.5]); 1280 x 720, path depth 10, NEE
with MIS and Russian roulette; SD-tree training passes of 1, 2, 4, ..., 64 spp
(split threshold c = 12000 * sqrt(2^k), flux threshold 0.01, DTree depth 20)
and a final pass, 256 spp in all; guide fraction 0.7 and the spatial box
recording filter (--fraction, --spatial-filter; the library defaults are the
reference's 0.3 and Nearest, run as a separate leg).

One step = one complete SD-tree-guided render of the budget (all training
passes, recording, refinement and the final pass).  `value` is Msamples/s
(camera paths per second) of that render.  The line also carries:
- equal-time quality legs: each guided render (SD-tree, hash grid, both with
  RIS, learned guide) against plain path tracing given the same wall-clock
  time, as relMSE, SMAPE and rMAE against a high-spp plain reference, mean and
  standard deviation over --seeds seeds;
- the roofline of the dominant kernel (k_extend, closest-hit traversal): its
  VALU issue rate from the committed counter summary of this workload over
  the launch time measured here, beside its HBM traffic;
- the other configurations and the CPU baseline.

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
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every 4
# cycles (16 lanes), at the 2.4 GHz peak engine clock
VALU_PEAK_GWIS = 256 * 4 * 2.4 / 4  # 614.4 G wave-instructions/s
NODE_BYTES = 80       # a BVH4 node: four child boxes (binary16 planes), child entries, corner
TRI_BYTES = 36        # three float3 vertices per tested triangle
RAY_IN_BYTES = 32     # origin + direction
HIT_OUT_BYTES = 16    # t, u, v, triangle
QUEUE_BYTES = 4       # queue index read
SHADOW_REC_BYTES = 48  # shadow ray record: origin + tmax, direction + path, contribution
SHADE_VTX_BYTES = 200  # hit 16 + indices 12 + vertices 36 + material 16 + state in 44 + out 44 + next ray 32
GUIDE_VTX_BYTES = 48   # recorded guide vertex (three float4), written by shade and read by commit
COMMIT_ATOMIC_BYTES = 8  # the DTree weight and one leaf slot (the record descent's nodes are not counted)
ACCUM_PIX_BYTES = 24   # accumulation buffer read-modify-write per pixel and launch
PATH_L_BYTES = 16      # per-sample radiance read by the accumulation
PMC_FILE = ROOT / "profiles" / "guided_room_counters.json"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_mse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref * ref + 1e-2)))


def metrics(img, ref):
    """relMSE, and the reference comparison tool's SMAPE and rMAE."""
    from rtlib_amd import results
    rmae, smape = results.compare(ref.reshape(-1, 3), img.reshape(-1, 3))
    return {"relmse": rel_mse(img, ref), "smape": smape, "rmae": rmae}


def summarize(rows):
    """rows: per seed {ms, plain_spp, guided: metrics, plain: metrics} ->
    mean and standard deviation of each figure and of the guided / plain ratios."""
    out = {"seeds": len(rows)}

    def ms(vals):
        a = np.asarray(vals, np.float64)
        return [round(float(a.mean()), 6), round(float(a.std(ddof=1)) if a.size > 1 else 0.0, 6)]

    out["ms"] = ms([r["ms"] for r in rows])
    out["plain_spp"] = ms([r["plain_spp"] for r in rows])
    for k in ("relmse", "smape", "rmae"):
        out[k] = ms([r["guided"][k] for r in rows])
        out["plain_" + k] = ms([r["plain"][k] for r in rows])
        out["ratio_" + k] = ms([r["guided"][k] / r["plain"][k] for r in rows])
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--budget", type=int, default=256,
                    help="samples per pixel of one guided render: SD-tree passes 1, 2, 4, .., 64 and the final pass")
    ap.add_argument("--fraction", type=float, default=0.7,
                    help="probability of sampling the guide instead of the BSDF (SD-tree and hash grid)")
    ap.add_argument("--spatial-filter", type=int, default=1, help="SD-tree spatial recording filter: 0 Nearest, "
                    "1 Box (stochastic form)")
    ap.add_argument("--directional-filter", type=int, default=0, help="SD-tree directional recording filter: "
                    "0 Nearest, 1 Box (stochastic form)")
    ap.add_argument("--grid-bounds", default="64,32,64", help="hash-grid cells per axis over the scene box")
    ap.add_argument("--seeds", type=int, default=4, help="seeds per equal-time quality leg")
    ap.add_argument("--ref-spp", type=int, default=16384, help="reference image (0: no quality legs)")
    ap.add_argument("--no-configs", action="store_true", help="skip the Cornell / many-light / adjoint / "
                    "learned-guide configurations")
    ap.add_argument("--mlp-passes", type=int, default=8)
    ap.add_argument("--mlp-steps", type=int, default=4)
    ap.add_argument("--mlp-cache", type=int, default=32, help="spatial-cache cells per axis of the learned guide "
                    "(0: network inference at every guided vertex)")
    ap.add_argument("--large-budget", type=int, default=1024, help="second budget for the equal-time comparison "
                    "of the SD-tree and the learned guide (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--small", action="store_true", help="reduced size for a quick functional check")
    ap.add_argument("--pmc-step", action="store_true", help="run exactly one guided step of the workload (seed 1, "
                    "the first timed step's) and exit: the command that the counter summaries under profiles/ "
                    "are collected from")
    ap.add_argument("--pmc-grid", action="store_true", help="run one hash-grid-guided render of the workload "
                    "(seed 1) and exit: the command the hash-grid kernel trace in profiles/ is taken from")
    ap.add_argument("--pmc-mlp", action="store_true", help="run one learned-guide render of the workload and exit "
                    "(for the matrix-core counters of its training and inference kernels)")
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
    from rtlib_amd.tracer import (FLAG_MIS, FLAG_NEE, FLAG_RIS, FLAG_RR, Context, MLPGuide, RTSTreeController,
                                  RTSTreeWrapper, render_guided, render_mlp_guided)

    W, H = (320, 180) if a.small else (1280, 720)
    grid_res = (256, 128) if a.small else (1024, 512)
    depth = 10
    frac = a.fraction
    bounds = tuple(int(x) for x in a.grid_bounds.split(","))
    base = FLAG_NEE | FLAG_MIS | FLAG_RR
    ris = FLAG_NEE | FLAG_RIS | FLAG_RR
    t0 = time.time()
    sc = scenes.guided_room(width=W, height=H, grid=grid_res)
    log(f"scene: {sc['idx'].shape[0]} triangles ({time.time() - t0:.1f} s)")
    ctx = Context(0)
    t0 = time.time()
    ctx.upload_scene(sc)
    info = ctx.scene_info()
    log(f"BVH build + upload {time.time() - t0:.1f} s: {info['nodes']} nodes, depth {info['bvh_depth']}, "
        f"stack bound {info['stack_bound']}")
    npix = W * H
    lo, hi = info["aabb_min"] - 0.005, info["aabb_max"] + 0.005

    tree = RTSTreeWrapper(ctx, lo, hi, 20, frac)
    tree.SetSplit(mode=1, c=12000.0, max_mb=2000, subdiv_th=0.01)
    tree.SetFilter(a.spatial_filter, a.directional_filter, 17)

    def guided_step(seed, on_pass=None, flags=base):
        ctrl = RTSTreeController(tree, a.budget, 0, 0.5, 1)
        ctx.clear()
        return render_guided(ctx, ctrl, tree, "tree", flags, max_depth=depth, seed=seed, fraction=frac,
                             on_pass=on_pass)

    if a.pmc_grid:
        from rtlib_amd.tracer import (DoubleBufferedHashGrid3Buffer, RTMortonQuadTreeController,
                                      RTMortonQuadTreeWrapper)
        grid = DoubleBufferedHashGrid3Buffer(ctx, lo, hi, bounds=bounds, slots=1 << 20, key_mode=1)
        mq = RTMortonQuadTreeWrapper(ctx, 1 << 20, 3, frac)
        ctrl = RTMortonQuadTreeController(mq, grid, a.budget, 3, 0, 0.5, 1)
        ctx.clear()
        render_guided(ctx, ctrl, mq, "grid", base, max_depth=depth, seed=1, fraction=frac, grid=grid)
        ctx.synchronize()
        log("one hash-grid-guided render done")
        return 0
    if a.pmc_step:
        guided_step(1)
        ctx.synchronize()
        log("one guided step done")
        return 0
    if a.pmc_mlp:
        mlp = MLPGuide(ctx, lo, hi, seed=1, cache_res=a.mlp_cache)
        render_mlp_guided(ctx, mlp, a.budget, base, max_depth=depth, seed=11, train_passes=a.mlp_passes,
                          steps_per_pass=a.mlp_steps)
        ctx.synchronize()
        mlp.close()
        log("one learned-guide render done")
        return 0

    def timed(fn):
        ctx.synchronize()
        t = time.perf_counter()
        r = fn()
        ctx.synchronize()
        return time.perf_counter() - t, r

    # ---- guided: warmup, then K timed steps.  Kernel classes are timed with
    # HIP events on the render stream during the timed steps themselves.
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

    for k in range(a.steps):
        step_ms.append(np.zeros(5))
        step_launches.append(np.zeros(5, np.int64))
        dt, n = timed(lambda: guided_step(k + 1, on_pass=acc_pass))
        times.append(dt)
        log(f"step {k}: {dt * 1e3:.1f} ms ({n} spp)")
    per_kernel = np.mean(step_ms, axis=0)
    launches = step_launches[0]
    # ---- work counts of step 0 (seed 1) in a separate, untimed run: same
    # random streams, so the same paths and the same traversal work
    ctx.set_profile(time_kernels=False, count_work=True)
    work = np.zeros(4)
    rays_c = shadow_c = rec_rays = 0

    def acc_work(state, spp, traced):
        nonlocal rays_c, shadow_c, rec_rays
        _, wk, _ = ctx.kernel_stats()
        work[:] += np.array([wk["extend_nodes"], wk["extend_tris"], wk["shadow_nodes"], wk["shadow_tris"]])
        st = ctx.render_stats()
        rays_c += st["rays"]
        shadow_c += st["shadow_rays"]
        if state != 2:  # recording passes (every state but the final Sample one)
            rec_rays += st["rays"]

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
    ctx.set_profile(time_kernels=True, count_work=False)
    ctx.clear()
    ctx.render(spp=a.budget, flags=base, max_depth=depth, seed=77)
    plain_kernel_ms = ctx.kernel_stats()[0]
    ctx.set_profile(False, False)
    ctx.clear()
    dt_plain_ris, _ = timed(lambda: ctx.render(spp=a.budget, flags=ris, max_depth=depth, seed=78))
    plain_ris_rate = a.budget * npix / dt_plain_ris / 1e6

    quality = None
    if a.ref_spp > 0:
        quality = quality_legs(a, ctx, sc, tree, lo, hi, bounds, guided_step, timed, plain_rate, plain_ris_rate,
                               npix, depth, base, ris)

    # ---- roofline of the dominant kernel (k_extend)
    ext_bytes = rays_c * (RAY_IN_BYTES + HIT_OUT_BYTES + QUEUE_BYTES) + work[0] * NODE_BYTES + work[1] * TRI_BYTES
    roofline = extend_roofline(W, H, per_kernel, launches, ext_bytes, names, plain_kernel_ms)
    algo = {"extend": ext_bytes,
            "shade": rays_c * SHADE_VTX_BYTES + shadow_c * SHADOW_REC_BYTES + rec_rays * GUIDE_VTX_BYTES,
            "shadow": shadow_c * SHADOW_REC_BYTES + work[2] * NODE_BYTES + work[3] * TRI_BYTES,
            "accumulate": a.budget * npix * PATH_L_BYTES + int(launches[3]) * npix * ACCUM_PIX_BYTES,
            "commit": rec_rays * (GUIDE_VTX_BYTES + COMMIT_ATOMIC_BYTES)}
    roofline.update(per_kernel_roofline(W, H, names, per_kernel, launches, algo, t_step))

    # ---- the other configurations, measured on the same context
    configs = None if (a.small or a.no_configs) else other_configs(ctx, timed, frac)

    # ---- CPU baseline: the oracle integrator on the host cores, guided final pass
    cpu = cpu_baseline(sc, tree, lo, hi, W, H, depth, a.cpu_seconds, frac) if a.cpu_seconds > 0 else None

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
                   "max_depth": depth, "guide_fraction": frac,
                   "sdtree_filters": {"spatial": ["nearest", "box"][a.spatial_filter],
                                      "directional": ["nearest", "box"][a.directional_filter]},
                   "hash_grid": {"cells": list(bounds), "slots": 1 << 20,
                                                                             "directional_bins": "8x8"},
                   "parallelism": "single GPU"},
        "plain_msamples_per_s": round(plain_rate, 2),
        "plain_ris_msamples_per_s": round(plain_ris_rate, 2),
        "mrays_per_s": {"guided": round((rays + shadow) / t_rays / 1e6, 1),
                        "plain": round((st_plain["rays"] + st_plain["shadow_rays"]) / dt_plain / 1e6, 1)},
        "equal_time": quality,
        "sdtree": tree.Sizes(),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "configs": configs,
    }
    print(json.dumps(out), flush=True)
    return 0


def quality_legs(a, ctx, sc, tree, lo, hi, bounds, guided_step, timed, plain_rate, plain_ris_rate, npix, depth, base,
                 ris):
    """Equal-time legs: every guided render against plain path tracing (plain
    RIS for the RIS legs) given the same wall-clock time, over a.seeds seeds."""
    from rtlib_amd.tracer import (DoubleBufferedHashGrid3Buffer, MLPGuide, RTMortonQuadTreeController,
                                  RTMortonQuadTreeWrapper, RTSTreeController, RTSTreeWrapper, render_guided,
                                  render_mlp_guided)
    frac = a.fraction
    t0 = time.time()
    ctx.clear()
    done = 0
    while done < a.ref_spp:
        k = min(256, a.ref_spp - done)
        ctx.render(spp=k, flags=base, max_depth=depth, seed=4242, sample_offset=done)
        done += k
    ref = ctx.image()
    log(f"reference {a.ref_spp} spp in {time.time() - t0:.1f} s")
    grid = DoubleBufferedHashGrid3Buffer(ctx, lo, hi, bounds=bounds, slots=1 << 20, key_mode=1)
    mq = RTMortonQuadTreeWrapper(ctx, 1 << 20, 3, frac)

    def grid_step(seed, flags=base):
        grid.Clear()
        mq.Allocate()
        ctrl = RTMortonQuadTreeController(mq, grid, a.budget, 3, 0, 0.5, 1)
        ctx.clear()
        return render_guided(ctx, ctrl, mq, "grid", flags, max_depth=depth, seed=seed, fraction=frac, grid=grid)

    mlp = MLPGuide(ctx, lo, hi, seed=1, cache_res=a.mlp_cache)

    def mlp_step(seed, g=None, budget=None):
        g = g or mlp
        g.reset(1)
        ctx.clear()
        return render_mlp_guided(ctx, g, budget or a.budget, base, max_depth=depth, seed=seed,
                                 train_passes=a.mlp_passes, steps_per_pass=a.mlp_steps)

    # the learned guide with its fallback SD-tree (recorded in the same
    # passes) and the judge that picks the estimator for the rest of the budget
    fb_tree = RTSTreeWrapper(ctx, lo, hi, 20, frac)
    fb_tree.SetSplit(mode=1, c=12000.0, max_mb=2000, subdiv_th=0.01)
    fb_tree.SetFilter(spatial=a.spatial_filter, directional=a.directional_filter, seed=11)
    choices = []

    def mlp_judge_step(seed):
        j = {}
        mlp.reset(1)
        ctx.clear()
        n = render_mlp_guided(ctx, mlp, a.budget, base, max_depth=depth, seed=seed, train_passes=a.mlp_passes,
                              steps_per_pass=a.mlp_steps, fallback=fb_tree, fraction=frac, judge=j)
        choices.append(j["choice"])
        return n

    def leg(name, fn, seeds, plain_flags=base, rate=plain_rate, seed0=1000):
        fn(seed0 - 1)  # warm
        rows = []
        for s in range(seeds):
            t, _ = timed(lambda: fn(seed0 + s))
            img = ctx.image()
            spp = max(1, int(round(t * rate * 1e6 / npix)))
            ctx.clear()
            done = 0
            while done < spp:
                k = min(256, spp - done)
                ctx.render(spp=k, flags=plain_flags, max_depth=depth, seed=seed0 + 500 + s, sample_offset=done)
                done += k
            rows.append({"ms": t * 1e3, "plain_spp": spp, "guided": metrics(img, ref), "plain": metrics(ctx.image(),
                                                                                                        ref)})
        r = summarize(rows)
        log(f"{name}: {r['ms'][0]:.1f} ms, relMSE {r['relmse'][0]:.5f} vs plain {r['plain_relmse'][0]:.5f} "
            f"(ratio {r['ratio_relmse'][0]:.3f} +- {r['ratio_relmse'][1]:.3f}), SMAPE ratio {r['ratio_smape'][0]:.3f}")
        return r

    out = {"reference_spp": a.ref_spp, "budget_spp": a.budget}
    ctx.clear()
    ctx.render(spp=a.budget, flags=base, max_depth=depth, seed=77)
    out["plain_at_budget"] = metrics(ctx.image(), ref)
    out["sdtree"] = leg("SD-tree", lambda s: guided_step(s), a.seeds, seed0=1000)
    out["hashgrid"] = leg("hash grid", lambda s: grid_step(s), a.seeds, seed0=2000)
    out["sdtree_ris"] = leg("SD-tree + RIS", lambda s: guided_step(s, flags=ris), a.seeds, plain_flags=ris,
                            rate=plain_ris_rate, seed0=3000)
    out["hashgrid_ris"] = leg("hash grid + RIS", lambda s: grid_step(s, flags=ris), a.seeds, plain_flags=ris,
                              rate=plain_ris_rate, seed0=4000)
    out["mlp"] = leg("learned guide (spatial cache)", lambda s: mlp_step(s), a.seeds, seed0=5000)
    info = mlp.info()
    out["mlp"].update({"train_passes": a.mlp_passes, "steps_per_pass": a.mlp_steps, "cache_res": a.mlp_cache,
                       "samples_per_step": 1 << 18, "max_res": float(info["res"][-1])})
    out["mlp_judge"] = leg("learned guide, judge + SD-tree fallback", mlp_judge_step, a.seeds, seed0=5500)
    out["mlp_judge"]["choices"] = choices[1:]
    fb_tree.Destroy()
    pv = MLPGuide(ctx, lo, hi, seed=1)
    out["mlp_per_vertex_inference"] = leg("learned guide (per-vertex inference)", lambda s: mlp_step(s, g=pv), 1,
                                          seed0=6000)
    pv.close()
    # the SD-tree with the reference's defaults: fraction 0.3, Nearest filters
    ref_tree = RTSTreeWrapper(ctx, lo, hi, 20, 0.3)
    ref_tree.SetSplit(mode=1, c=12000.0, max_mb=2000, subdiv_th=0.01)

    def ref_tree_step(s):
        ctrl = RTSTreeController(ref_tree, a.budget, 0, 0.5, 1)
        ctx.clear()
        return render_guided(ctx, ctrl, ref_tree, "tree", base, max_depth=depth, seed=s, fraction=0.3)

    out["sdtree_reference_defaults"] = leg("SD-tree, fraction 0.3, Nearest", ref_tree_step, 2, seed0=7000)
    out["sdtree_reference_defaults"].update({"fraction": 0.3, "filters": "nearest"})
    ref_tree.Destroy()
    if a.large_budget > 0:
        def tree_large(s):
            ctrl = RTSTreeController(tree, a.large_budget, 0, 0.5, 1)
            ctx.clear()
            return render_guided(ctx, ctrl, tree, "tree", base, max_depth=depth, seed=s, fraction=frac)

        out[f"budget_{a.large_budget}"] = {
            "sdtree": leg(f"SD-tree {a.large_budget} spp", tree_large, 1, seed0=8000),
            "mlp": leg(f"learned guide {a.large_budget} spp", lambda s: mlp_step(s, budget=a.large_budget), 1,
                       seed0=9000)}
    mlp.close()
    return out


def load_counters(W, H):
    """Per-dispatch counter averages of one guided step of this workload,
    from the committed rocprofv3 --pmc summary (or None)."""
    try:
        t = json.loads(PMC_FILE.read_text())
    except (OSError, ValueError):
        return None
    if t.get("width") != W or t.get("height") != H:
        return None
    return t


def extend_roofline(W, H, per_kernel, launches, ext_bytes, names, plain_kernel_ms):
    """k_extend (closest-hit BVH4 traversal) against the VALU issue roofline:
    SQ_INSTS_VALU per launch (counter summary of this workload) over the
    launch time measured in this run.  Beside it, its HBM traffic (FETCH_SIZE
    and WRITE_SIZE per launch, each scaled by its reading for a copy kernel of
    known size) and the algorithmic bytes it touches (rays, visited nodes,
    tested triangles), most of which hit L2 and the MALL."""
    n_ext = int(launches[0])
    avg_s = float(per_kernel[0]) / max(n_ext, 1) * 1e-3
    pmc = load_counters(W, H)
    r = {"kernel": "k_extend (closest-hit BVH4 traversal)", "bound": "valu", "unit": "G wave-instructions/s",
         "peak": VALU_PEAK_GWIS, "achieved": None, "frac": None, "traffic": None,
         "launches_per_step": n_ext, "avg_launch_ms": round(avg_s * 1e3, 4),
         "kernel_ms_per_step": {n: round(float(x), 2) for n, x in zip(names, per_kernel)},
         "plain_kernel_ms_per_step": {n: round(float(plain_kernel_ms[n]), 2) for n in names},
         "kernel_launches_per_step": {n: int(x) for n, x in zip(names, launches)},
         "touched_bytes_per_launch": round(ext_bytes / max(n_ext, 1)),
         "touched_gbs": round(ext_bytes / max(n_ext, 1) / avg_s / 1e9, 1) if avg_s > 0 else None}
    if pmc is not None and avg_s > 0:
        e = pmc["kernels"]["k_extend"]
        valu = e["SQ_INSTS_VALU"]
        r["achieved"] = round(valu / avg_s / 1e9, 1)
        r["frac"] = round(valu / avg_s / 1e9 / VALU_PEAK_GWIS, 4)
        r["valu_per_launch"] = valu
        r["valu_per_wave"] = round(valu / e["SQ_WAVES"], 1)
        if "ta_busy_frac" in e:
            # the texture addresser that issues its node and triangle loads:
            # busy fraction over the counter run's dispatches (co-bound)
            r["ta_busy_frac"] = e["ta_busy_frac"]
        traffic = e["FETCH_SIZE"] * 1024 * pmc["fetch_scale"] + e["WRITE_SIZE"] * 1024 * pmc["write_scale"]
        r["traffic"] = round(traffic)
        r["hbm"] = {"gbs": round(traffic / avg_s / 1e9, 1), "peak_gbs": HBM_PEAK_GBS,
                    "frac": round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4),
                    "traffic_over_touched": round(traffic / max(ext_bytes / max(n_ext, 1), 1), 3)}
        r["counters_source"] = f"profiles/{PMC_FILE.name}"
    return r


def per_kernel_roofline(W, H, names, per_kernel, launches, algo, t_step):
    """Every kernel class of the guided step: the rate of the algorithmic
    bytes it touches (from the work-count pass: rays, shadow rays, visited
    nodes, tested triangles, recorded vertices; most hit L2 and the MALL, so
    this rate is no HBM fraction) over its event-timed ms per step, and, where
    the counter summary holds the kernel, its HBM traffic per step
    (FETCH_SIZE and WRITE_SIZE per dispatch x dispatches) as a fraction of
    the HBM peak; the frame aggregate divides the sums by the step's wall
    time."""
    pmc = load_counters(W, H)
    key = {"extend": "k_extend", "shade": "k_shade", "shadow": "k_shadow", "commit": "k_commit"}
    out, tot_algo, tot_hbm = {}, 0.0, 0.0
    for i, n in enumerate(names):
        ms = float(per_kernel[i])
        b = float(algo[n])
        e = {"ms_per_step": round(ms, 2), "algorithmic_bytes_per_step": round(b),
             "algorithmic_gbs": round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
        tot_algo += b
        if pmc is not None and key.get(n) in pmc["kernels"]:
            k = pmc["kernels"][key[n]]
            per_disp = k["FETCH_SIZE"] * 1024 * pmc["fetch_scale"] + k["WRITE_SIZE"] * 1024 * pmc["write_scale"]
            hbm = per_disp * int(launches[i])
            tot_hbm += hbm
            e["hbm_bytes_per_step"] = round(hbm)
            e["hbm_frac"] = round(hbm / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if ms > 0 else None
            if "active_lanes_per_valu" in k:
                e["active_lanes_per_valu"] = k["active_lanes_per_valu"]
            if "ta_busy_frac" in k:
                e["ta_busy_frac"] = k["ta_busy_frac"]
        out[n] = e
    frame = {"ms_per_step": round(t_step * 1e3, 2), "algorithmic_bytes_per_step": round(tot_algo),
             "algorithmic_gbs": round(tot_algo / t_step / 1e9, 1),
             "hbm_bytes_per_step": round(tot_hbm) if tot_hbm else None,
             "hbm_frac": round(tot_hbm / t_step / 1e9 / HBM_PEAK_GBS, 4) if tot_hbm else None}
    return {"per_kernel": out, "frame": frame}


def other_configs(ctx, timed, frac):
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
    fb = RTSTreeWrapper(ctx, lo, hi, 20, frac)
    fb.SetSplit(mode=1, c=12000.0)
    judge = {}

    def mlp_judge_step():
        judge.clear()
        mlp.reset(1)
        ctx.clear()
        return render_mlp_guided(ctx, mlp, 64, nee_mis_rr, max_depth=8, seed=25, train_passes=8, steps_per_pass=4,
                                 fallback=fb, fraction=frac, judge=judge)

    timed(mlp_judge_step)
    t_judge, _ = timed(mlp_judge_step)
    img_judge = ctx.image()
    fb.Destroy()
    mlp.close()
    tree = RTSTreeWrapper(ctx, lo, hi, 20, frac)
    tree.SetSplit(mode=1, c=12000.0)

    def tree_step():
        tree.Clear()
        ctrl = RTSTreeController(tree, 64, 0, 0.5, 1)
        ctx.clear()
        return render_guided(ctx, ctrl, tree, "tree", nee_mis_rr, max_depth=8, seed=22, fraction=frac)

    timed(tree_step)
    t_tree, _ = timed(tree_step)
    img_tree = ctx.image()
    res = {}
    for name, tt, img in (("mlp", t_mlp, img_mlp), ("mlp_judge", t_judge, img_judge), ("sdtree", t_tree, img_tree)):
        spp = max(1, int(round(tt * plain_rate / (256 * 256))))
        ctx.clear()
        ctx.render(spp=spp, flags=nee_mis_rr, max_depth=8, seed=23)
        res[name] = {"ms": round(tt * 1e3, 2), "relmse": round(rel_mse(img, ref), 6), "plain_spp": spp,
                     "plain_relmse": round(rel_mse(ctx.image(), ref), 6)}
    ctx.clear()
    ctx.render(spp=64, flags=nee_mis_rr, max_depth=8, seed=24)
    res["plain_at_budget"] = {"spp": 64, "ms": round(t * 1e3, 2), "relmse": round(rel_mse(ctx.image(), ref), 6)}
    res["mlp_judge"]["choice"] = judge.get("choice")
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


def cpu_baseline(sc, tree, lo, hi, W, H, depth, seconds, frac):
    """Times the oracle's CPU integrator on a bounded pixel sample of the same
    guided render (final-pass state of the trained tree), with one thread per
    CPU this process may run on, at most the 16 CPUs a one-GPU box allots."""
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
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = min(avail, 16)
    flags = 1 | 128 | 256 | 8 | 16 | 32  # NEE, MIS, RR, tree, guide sampling, final
    # calibrate on one row, then size the sample to the time budget
    acc = np.zeros((H, W, 3), np.float32)
    t0 = time.perf_counter()
    oracle.render(osc, uvw, W, H, 1, flags, max_depth=depth, seed=5, guide=guide, px_range=(0, W),
                  threads=threads, accum=acc, fraction=frac)
    t_row = max(time.perf_counter() - t0, 1e-3)
    rows = int(max(1, min(H, seconds / t_row)))
    stride = max(1, H // rows)
    total_px, t_total, rays = 0, 0.0, 0
    for r in range(0, H, stride):
        t0 = time.perf_counter()
        _, nr = oracle.render(osc, uvw, W, H, 1, flags, max_depth=depth, seed=5, guide=guide,
                              px_range=(r * W, (r + 1) * W), threads=threads, accum=acc, fraction=frac)
        t_total += time.perf_counter() - t0
        total_px += W
        rays += nr
        if t_total > seconds:
            break
    return {"value": round(total_px / t_total / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(), "cpus_available": avail,
            "sample": f"{total_px} camera paths (1 spp on every {stride}th row of {W}x{H}), guided final pass "
                      f"with the GPU-trained SD-tree; BVH build {t_build:.1f} s excluded",
            "mrays_per_s": round(rays / t_total / 1e6, 3)}


if __name__ == "__main__":
    sys.exit(main())

"""Benchmark: Mrays/s of the plain HIP path tracer on the full-shading workload.

Workload (one "step"): the 256^3 terrain+boxes scene with 70/20/10
Lambert/metal/emissive IDs, the 1024x512 FP32 sky+sun environment
(importance sampled), a 1024x1024 thin-lens camera, 16 spp and 8 bounces,
rendered from a cleared accumulator.  Inputs are resident in HBM before the
timed region; every step repeats the same 16 passes, so the ray count of a
step is the count of the (untimed) statistics launch.

Rays = every camera sample + every shadow and continuation traversal.
The roofline object prices the render kernel at its algorithmic HBM bytes
(1 B per DDA voxel visit, 32 B per surface interaction, 32 B per sample
accumulator read-modify-write, binary-search and texel bytes per image
environment lookup) over the kernel's HIP-event time.  The CPU baseline is
the C oracle (OpenMP over the host cores) on the same frame for as many of
the step's passes as fit a ~12 s budget.

The same line carries the metric's other half (leave out with --no-extras):
"batch", 64 mixed scenes rendered as one batch against the same renders one
after another, and "time_to_rmse", the 1920x1080 scene at one sample per
pixel per frame, 256 frames of the plain tracer against the radiance cache
and plain path truncation given the same time, scored against a
16384-sample FP64 reference.  It also carries "voxelize", the 200k-triangle
mesh voxelized into 512^3 and rendered at 1920x1080 (10 steps; the same
workload as --workload voxelize).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s; time-to-RMSE 1e-2 vs FP64 ref, plain vs batched+radiance-cache"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the dominant kernel on the full workload (one
# stream), from rocprofv3 FETCH_SIZE (x2.000) + WRITE_SIZE (x1.000), each
# factor being what the counter reads for a copy kernel of known size
# (profiles/wf_traverse_kernel_full_1024x1024_16spp.md).
TRAFFIC: dict = {("wf_traverse_kernel", 1): 481225141}


def env_lookup_bytes(cu_w: int, cv_n: int) -> tuple[int, int]:
    """Bytes of one importance-sampled environment draw (two binary searches
    of 4-byte CDF entries, 4 bracket reads, 4 RGB32F texels) and of one
    lookup by direction (4 RGB32F texels)."""
    search = 4 * (math.ceil(math.log2(cv_n)) + math.ceil(math.log2(cu_w)) + 4)
    return search + 48, 48


def algorithmic_bytes(st: dict, cu_w: int, cv_n: int) -> int:
    sample_b, eval_b = env_lookup_bytes(cu_w, cv_n)
    return (st["visits"] * 1 + st["interactions"] * 32 + st["samples"] * 32 + st["env_samples"] * sample_b
            + st["env_evals"] * eval_b)


def cpu_baseline(wl, budget_s: float = 12.0):
    """Oracle on the same workload, all host cores: the whole frame for as
    many of the step's passes as fit the time budget (at least one)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.build()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    sc = wl.scene
    o = oracle.OracleScene(sc.offsets(), sc.mat, sc.emissive)
    o.set_camera(**wl.camera_arrays())
    cu, cv, integ = wl.cdf
    o.set_environment(wl.env_top, wl.env_bottom, wl.env_image, cu, cv, integ, wl.env_rotation)
    o.set_settings(max_bounces=wl.bounces)
    t = time.perf_counter()
    acc, cnt = o.render(wl.width, wl.height, wl.seed, 0, 1, threads=cores)
    dt1 = time.perf_counter() - t
    passes = int(max(1, min(wl.spp - 1, budget_s / max(dt1, 1e-3))))
    t = time.perf_counter()
    _, cnt2 = o.render(wl.width, wl.height, wl.seed, 1, passes, accum=acc, threads=cores)
    dt = time.perf_counter() - t + dt1
    rays = cnt["rays"] + cnt2["rays"]
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"full {wl.width}x{wl.height} frame, passes 0-{passes} of the step's {wl.spp}, {rays} rays in "
                      f"{dt:.2f} s (C oracle, OpenMP)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--mode", type=int, default=1, help="0 megakernel, 1 wavefront (default)")
    ap.add_argument("--streams", type=int, default=1,
                    help="wavefront: 1 (default) runs the step on one stream, 2 runs its pass halves on two streams")
    ap.add_argument("--calibrate", action="store_true",
                    help="run only the streaming copy kernel (--copy-bytes read and written, --copy-iters times), "
                         "the known-size reference for the FETCH_SIZE / WRITE_SIZE counter passes")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--copy-iters", type=int, default=3)
    ap.add_argument("--workload", default="full", choices=["full", "batch", "cache", "voxelize"],
                    help="full: the 256^3 full-shading scene (default); batch: 64 mixed scenes, one batched "
                         "graph launch against 64 sequential renders; cache: time to RMSE 1e-2 against an FP64 "
                         "reference, plain tracer against the radiance cache; voxelize: 200k triangles into 512^3, "
                         "then 1920x1080, 8 spp, 6 bounces")
    ap.add_argument("--ref-spp", type=int, default=16384, help="time-to-RMSE race: passes of the FP64 reference")
    ap.add_argument("--frames", type=int, default=256,
                    help="time-to-RMSE race: frames of the plain tracer, whose time every contender gets")
    ap.add_argument("--only", choices=["plain", "cache", "truncation"], default=None,
                    help="cache workload: run --frames frames of this contender alone (for a kernel trace)")
    ap.add_argument("--no-extras", action="store_true",
                    help="full workload only: leave out the batch, time-to-RMSE and voxelize objects")
    ap.add_argument("--cache-bounce", type=int, default=2, help="cache workload: bounce at which paths end")
    ap.add_argument("--train-paths", type=int, default=65536,
                    help="cache workload: training paths per frame (65536 give about 58k samples per frame)")
    ap.add_argument("--train-extra", type=int, default=1,
                    help="cache workload: bounces training paths trace past the cache bounce (-1: to the end)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (asked for {args.gpus}); the path shards by "
              "image tiles and scenes, which is not exercised here", file=sys.stderr)
        sys.exit(2)

    from voxeltoy_amd import api, workloads

    if args.calibrate:
        with api.Context(0) as ctx:
            ms = ctx.copy_probe(args.copy_bytes, args.copy_iters)
        print(json.dumps({"calibrate": "copy", "bytes": args.copy_bytes, "iters": args.copy_iters,
                          "kernel_ms": ms, "gbps_read_plus_write": 2 * args.copy_bytes / (ms / 1e3) / 1e9}))
        return None
    if args.workload == "batch":
        return bench_batch(args, api, workloads)
    if args.workload == "cache":
        return bench_cache(args, api, workloads)
    if args.workload == "voxelize":
        return bench_voxelize(args, api, workloads)

    wl = workloads.full()
    ctx = api.Context(0)
    workloads.load(ctx, wl)
    ctx.set_render_mode(args.mode)
    ctx.set_overlap(args.streams)
    cu_w, cv_n = wl.cdf[0].shape[1], len(wl.cdf[1])

    # statistics launch (untimed): exact work of one step
    ctx.set_stats(True)
    ctx.reset_accum()
    ctx.render(wl.seed, 0, wl.spp)
    st = ctx.stats()
    ctx.set_stats(False)

    for _ in range(args.warmup):
        ctx.reset_accum()
        ctx.render(wl.seed, 0, wl.spp)
    ctx.synchronize()

    # traversal launches are bracketed by HIP events on the render stream
    ctx.set_launch_timing(args.mode == 1)
    kernel_ms, trav_ms, trav_launches, trav_union = [], [], [], []
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.reset_accum()
        ctx.render(wl.seed, 0, wl.spp)
        kernel_ms.append(ctx.stats()["kernel_ms"])
        lt = ctx.launch_timing()
        trav_ms.append(lt["traverse_ms"])
        trav_launches.append(lt["launches"])
        trav_union.append(lt["union_ms"])
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps
    k_ms = sum(kernel_ms) / len(kernel_ms)

    rays = st["rays"]
    value = rays / (ms_per_step / 1000.0) / 1e6
    bytes_step = algorithmic_bytes(st, cu_w, cv_n)
    render_gbs = bytes_step / (k_ms / 1000.0) / 1e9
    if args.mode == 1 and sum(trav_launches) > 0:
        # dominant kernel: the traversal (1 B per DDA voxel visit)
        n_launch = sum(trav_launches) / len(trav_launches)
        launch_ms = sum(trav_ms) / sum(trav_launches)
        launch_bytes = st["visits"] / n_launch
        kernel_name, achieved = "wf_traverse_kernel", launch_bytes / (launch_ms / 1000.0) / 1e9
        busy_ms = sum(trav_union) / len(trav_union)
        busy_gbs = st["visits"] / (busy_ms / 1000.0) / 1e9
    else:
        n_launch, launch_ms, launch_bytes = 1, k_ms, bytes_step
        kernel_name, achieved = "render_kernel", render_gbs
        busy_ms, busy_gbs = k_ms, render_gbs
    cpu = None if args.no_cpu_baseline else cpu_baseline(wl, args.cpu_budget)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "full-shading 256^3 terrain+boxes, 70/20/10 lambert/metal/emissive, "
                               "1024x512 fp32 sky+sun env (importance sampled), 1024x1024 thin lens, "
                               "16 spp, 8 bounces (plain tracer)",
                   "grid": 256, "width": wl.width, "height": wl.height, "spp": wl.spp, "bounces": wl.bounces,
                   "emissive_voxels": int(len(wl.scene.emissive)), "rays_per_step": int(rays),
                   "render_mode": ["megakernel", "wavefront"][args.mode], "streams": args.streams,
                   "kernel_ms": round(k_ms, 3), "counters": {k: int(st[k]) for k in
                                                             ("visits", "interactions", "samples", "env_samples",
                                                              "env_evals")}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": TRAFFIC.get((kernel_name, args.streams)) if args.mode == 1 else None,
                     "kernel": kernel_name, "launches_per_step": round(n_launch, 2),
                     "avg_launch_ms": round(launch_ms, 4), "algorithmic_bytes_per_launch": int(launch_bytes),
                     "busy_ms_per_step": round(busy_ms, 3), "busy_achieved_gbs": round(busy_gbs, 1),
                     "render_algorithmic_bytes_per_step": int(bytes_step),
                     "render_achieved_gbs": round(render_gbs, 1)},
        "cpu_baseline": cpu,
    }
    ctx.close()
    if not args.no_extras:
        # the metric's other half: batched against sequential scenes, and time
        # to RMSE of the plain tracer against the radiance cache
        out["batch"] = batch_race(api, workloads, steps=3, warmup=1)
        out["time_to_rmse"] = rmse_race(api, workloads, frames=args.frames, ref_spp=args.ref_spp,
                                        cache_bounce=args.cache_bounce, train_paths=args.train_paths,
                                        train_extra=args.train_extra)
        # the voxelize-then-render workload (200k triangles into 512^3)
        out["voxelize"] = voxelize_race(api, workloads, steps=10, warmup=3)
    print(json.dumps(out))


def voxelize_race(api, workloads, steps: int = 10, warmup: int = 3, mode: int = 1, streams: int = 1) -> dict:
    """A seeded displaced icosphere of 200 000 triangles (5 % radial noise)
    voxelized THIN into an empty 512^3 grid (computeMeshTransform placement),
    then rendered at 1920x1080, 8 spp, 6 bounces under a gradient sky.  A
    step is one voxelization plus one 8-pass render; the voxelizer kernel and
    the render are also timed on their own.  Returns the measurements."""
    import math as _m
    import numpy as np
    from voxeltoy_amd import scenes
    verts, tris = scenes.displaced_icosphere(seed=4, freq=100)
    res, W, H, spp, bounces = 512, 1920, 1080, 8, 6
    M = api.mesh_transform(verts.min(0), verts.max(0), [res] * 3)
    mat = np.asarray(scenes.lambert((0.0, 0.0, 0.0), (0.7, 0.65, 0.6)), np.float32)
    id_offset = np.full(256, -1, np.int32)
    ctx = api.Context(0)
    ctx.set_volume_u8(np.zeros((res, res, res), np.uint8), id_offset, mat)
    bmin, bmax = workloads.reference_bounds((res, res, res))
    cam = scenes.orbit_camera(bmin, bmax, _m.radians(70), _m.radians(35), W, H)
    wl = workloads.Workload("c3", None, cam, W, H, spp=spp, bounces=bounces, seed=77, bmin=bmin, bmax=bmax,
                            env_top=(1.6, 1.7, 1.9), env_bottom=(0.35, 0.3, 0.25))
    ctx.set_camera(**wl.camera_arrays())
    ctx.set_environment(wl.env_top, wl.env_bottom)
    ctx.set_settings(max_bounces=bounces)
    ctx.set_frame(W, H)
    ctx.set_render_mode(mode)
    ctx.set_overlap(streams)

    # first voxelization (untimed) and the statistics launch
    occ = ctx.voxelize(verts, tris, M, mode=0, material_offset=0, want_occupancy=True)
    covered = int(occ.sum())
    del occ
    ctx.set_stats(True)
    ctx.reset_accum()
    ctx.render(wl.seed, 0, spp)
    st = ctx.stats()
    ctx.set_stats(False)
    for _ in range(warmup):
        ctx.voxelize(verts, tris, M, mode=0, material_offset=0, want_occupancy=False)
        ctx.reset_accum()
        ctx.render(wl.seed, 0, spp)
    ctx.synchronize()
    vox_ms, ren_ms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.voxelize(verts, tris, M, mode=0, material_offset=0, want_occupancy=False)
        vox_ms.append(ctx.stats()["kernel_ms"])
        ctx.reset_accum()
        ctx.render(wl.seed, 0, spp)
        ren_ms.append(ctx.stats()["kernel_ms"])
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / steps
    v_ms = sum(vox_ms) / len(vox_ms)
    r_ms = sum(ren_ms) / len(ren_ms)
    rays = st["rays"]
    # voxelizer: 3 uint32 indices + 3 float3 vertices per triangle read, 1 B per covered voxel written
    vox_bytes = len(tris) * (12 + 36) + covered
    vox_gbs = vox_bytes / (v_ms / 1000.0) / 1e9
    # the reference's mesh fixture (resources/bunny.obj, 4968 triangles) at
    # 512^3: large triangles, a wave each
    from voxeltoy_amd import renderer as vr
    bv, bt, _, _ = vr.load_obj(os.path.join(ROOT, "tests", "golden", "bunny.obj"))
    bM = api.mesh_transform(bv.min(0), bv.max(0), [res] * 3)
    bctx = api.Context(0)
    bctx.set_volume_u8(np.zeros((res, res, res), np.uint8), id_offset, mat)
    b_cov = int(bctx.voxelize(bv, bt, bM, mode=0, material_offset=0, want_occupancy=True).sum())
    b_ms = []
    for _ in range(max(3, steps)):
        bctx.voxelize(bv, bt, bM, mode=0, material_offset=0, want_occupancy=False)
        b_ms.append(bctx.stats()["kernel_ms"])
    bctx.close()
    b_ms = sum(b_ms) / len(b_ms)
    ctx.close()
    return {"workload": "200k-triangle displaced icosphere voxelized (THIN) into 512^3, then 1920x1080, "
                        "8 spp, 6 bounces, gradient sky; a step = voxelization (host arrays uploaded) + render",
            "steps": steps, "ms_per_step": round(ms_per_step, 3),
            "mrays_per_s": round(rays / (ms_per_step / 1000.0) / 1e6, 2),
            "triangles": int(len(tris)), "covered_voxels": covered, "rays_per_step": int(rays),
            "voxelize_kernel_ms": round(v_ms, 4),
            "voxelize_mtri_per_s": round(len(tris) / (v_ms / 1000.0) / 1e6, 1),
            "voxelize_algorithmic_bytes": int(vox_bytes), "voxelize_achieved_gbs": round(vox_gbs, 1),
            "render_kernel_ms": round(r_ms, 3),
            "render_mrays_per_s": round(rays / (r_ms / 1000.0) / 1e6, 2), "streams": streams,
            "bunny_512": {"triangles": int(len(bt)), "covered_voxels": b_cov,
                          "voxelize_kernel_ms": round(b_ms, 4),
                          "voxelize_mtri_per_s": round(len(bt) / (b_ms / 1000.0) / 1e6, 2)}}


def bench_voxelize(args, api, workloads):
    """The voxelize-then-render workload (voxelize_race) as the bench line."""
    r = voxelize_race(api, workloads, steps=args.steps, warmup=args.warmup, mode=args.mode, streams=args.streams)
    cfg = {k: v for k, v in r.items() if k not in ("steps", "ms_per_step", "mrays_per_s", "voxelize_algorithmic_bytes",
                                                    "voxelize_achieved_gbs")}
    out = {
        "metric": METRIC,
        "value": r["mrays_per_s"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": cfg,
        "roofline": {"bound": "hbm", "kernel": "voxelize_kernel", "achieved": r["voxelize_achieved_gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["voxelize_achieved_gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": r["voxelize_algorithmic_bytes"],
                     "avg_launch_ms": r["voxelize_kernel_ms"]},
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def batch_race(api, workloads, steps: int = 5, warmup: int = 1) -> dict:
    """64 scenes (grid 32^3-512^3, image 128^2-1024^2, 1-64 spp, 5-60%
    occupancy, 4 bounces): one batched render (HIP graph replay, launch
    shapes chosen by vt_autotune_batch on this batch) against the same 64
    renders launched one scene after another, each on its own context
    stream; wall times with the device synchronised."""
    wls = workloads.batch(64)
    ctxs = []
    for wl in wls:
        c = api.Context(0)
        workloads.load(c, wl)
        ctxs.append(c)
    seeds = [wl.seed for wl in wls]
    counts = [wl.spp for wl in wls]
    b = api.Batch(ctxs)
    b.render(seeds, counts, stats=True)
    st = b.stats()
    rays = st["rays"]
    # the batch's launch shapes (solo threshold, path capacity, merged
    # traversal tuning) chosen by timing graph replays of this very batch
    t_tune = time.perf_counter()
    tuning = b.autotune(seeds, counts)
    t_tune = time.perf_counter() - t_tune

    def batched():
        for c in ctxs:
            c.reset_accum()
        b.render(seeds, counts, graph=True)

    def sequential():
        for c, wl in zip(ctxs, wls):
            c.reset_accum()
            c.render(wl.seed, 0, wl.spp)
        for c in ctxs:
            c.synchronize()

    for _ in range(warmup):
        batched()
        sequential()
    t0 = time.perf_counter()
    for _ in range(steps):
        batched()
    t_b = (time.perf_counter() - t0) / steps
    k_ms = b.stats()["kernel_ms"]
    t0 = time.perf_counter()
    for _ in range(steps):
        sequential()
    t_s = (time.perf_counter() - t0) / steps
    b.close()
    for c in ctxs:
        c.close()
    return {"workload": "batch of 64 scenes: grid 32^3-512^3, image 128^2-1024^2, 1-64 spp, 5-60% occupancy, "
                        "70/20/10 lambert/metal/emissive, gradient env, 4 bounces",
            "scenes": len(wls), "steps": steps, "rays_per_batch": int(rays), "samples_per_batch": int(st["samples"]),
            "batched_ms": round(t_b * 1000.0, 3), "batched_kernel_ms": round(k_ms, 3),
            "batched_mrays_per_s": round(rays / t_b / 1e6, 2),
            "sequential_ms": round(t_s * 1000.0, 3), "sequential_mrays_per_s": round(rays / t_s / 1e6, 2),
            "speedup_vs_sequential": round(t_s / t_b, 3),
            "autotuned": {"solo_paths": tuning["solo_paths"], "max_paths": tuning["max_paths"],
                          "dda_budget": tuning["dda_budget"], "refill_min": tuning["refill_min"],
                          "refill_chunk": tuning["refill_chunk"], "candidates": tuning["candidates"],
                          "best_replay_ms": round(tuning["ms"], 3), "seconds": round(t_tune, 2)}}


def bench_batch(args, api, workloads):
    res = batch_race(api, workloads, args.steps, args.warmup)
    out = {
        "metric": METRIC,
        "value": res["batched_mrays_per_s"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["batched_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": res["workload"]},
        "batch": res,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def profile_contender(api, workloads, which: str, frames: int, cache_bounce: int = 2, train_paths: int = 65536,
                      train_extra: int = 1) -> dict:
    """`frames` frames of one time-to-RMSE contender alone (no reference, no
    scoring), for a kernel trace of its frames: wall ms per frame."""
    import numpy as np
    wl = workloads.full(width=1920, height=1080, spp=1, bounces=8)
    ctx = api.Context(0)
    workloads.load(ctx, wl)
    cache = None
    if which != "plain":
        cache = api.RadianceCache(ctx, seed=1)
        if which == "truncation":
            cache.set_params(np.zeros(cache.n_params, np.float32))
    ms = []
    for n in range(frames):
        t0 = time.perf_counter()
        if which == "plain":
            ctx.render(wl.seed, n, 1)
        elif which == "cache":
            cache.render(wl.seed, n, 1, cache_bounce=cache_bounce, train_paths=train_paths, lr=3e-3,
                         train_extra=train_extra)
        else:
            cache.render(wl.seed, n, 1, cache_bounce=cache_bounce, train_paths=0)
        ctx.synchronize()
        ms.append((time.perf_counter() - t0) * 1000.0)
    if cache:
        cache.close()
    ctx.close()
    return {"contender": which, "frames": frames, "ms_per_frame": round(sum(ms[1:]) / max(1, frames - 1), 3)}


def rmse_race(api, workloads, frames: int = 256, ref_spp: int = 16384, cache_bounce: int = 2,
              train_paths: int = 65536, train_extra: int = 1, levels=(1e-2, 3e-3)) -> dict:
    """Time to RMSE against an FP64 reference on the 1920x1080 full-shading
    scene, one sample per pixel per frame: the plain tracer (8 bounces) for
    `frames` frames, whose wall time is the budget the other contenders get:
    radiance-cache rendering (paths end into the cache at bounce
    `cache_bounce`; `train_paths` training paths per frame trace
    `train_extra` more bounces and train it) and, as the control, the same
    paths ending into an untrained cache that predicts 0 (plain truncation).
    RMSE is over pixels x RGB of the tonemapped running mean (the reference
    output convention), on one pixel mask for all contenders: the pixels
    finite in the reference and in every contender's last image (a non-finite
    sample stays in a running mean, so such a pixel is non-finite from then on
    in every later image).  Frame times are wall times with the device
    synchronised; reading the images for the curve is not timed."""
    import numpy as np
    wl = workloads.full(width=1920, height=1080, spp=1, bounces=8)
    t_ref = time.perf_counter()
    ref_ctx = api.Context(0)
    workloads.load(ref_ctx, wl, fp64=True)
    done, bias_spp, ref_part = 0, min(4096, ref_spp), None
    while done < ref_spp:
        k = min(512, ref_spp - done, bias_spp - done if done < bias_spp else ref_spp)
        ref_ctx.render(wl.seed + 1, done, k)
        done += k
        if done == bias_spp:
            ref_part = ref_ctx.read_sum64() / bias_spp
    ref = ref_ctx.read_sum64() / ref_spp
    ref_ctx.close()
    t_ref = time.perf_counter() - t_ref
    ref_finite = np.isfinite(ref).all(axis=-1)

    # what ending the paths at the cache bounce leaves out: the first
    # `bias_spp` passes of the reference against the same passes (same
    # random streams, so the first bounces are identical) with every path
    # ended there into a cache that predicts 0
    t_bias = time.perf_counter()
    trunc_ctx = api.Context(0)
    workloads.load(trunc_ctx, wl, fp64=True)
    trunc0 = api.RadianceCache(trunc_ctx, seed=1)
    trunc0.set_params(np.zeros(trunc0.n_params, np.float32))
    for n0 in range(0, bias_spp, 512):
        trunc0.render(wl.seed + 1, n0, min(512, bias_spp - n0), cache_bounce=cache_bounce, train_paths=0)
    trunc_part = trunc_ctx.read_sum64() / bias_spp
    trunc0.close()
    trunc_ctx.close()
    t_bias = time.perf_counter() - t_bias

    def sq_err(img):
        """Per-pixel squared error summed over RGB (float64)."""
        d = img[..., :3].astype(np.float64) - ref
        return np.einsum("hwc,hwc->hw", d, d)

    def run(step_fn, read_fn, stop):
        """Frames while `stop(frames, elapsed_ms, mean_frame_ms)` is false;
        images are read at frames 1, 2, 4, ... and every 32 frames and at
        the end."""
        t, n, marks = 0.0, 0, []
        while not stop(n, t, t / n if n else 0.0):
            t0 = time.perf_counter()
            step_fn(n)
            t += (time.perf_counter() - t0) * 1000.0
            n += 1
            if (n & (n - 1)) == 0 or n % 32 == 0:
                marks.append((n, t, sq_err(read_fn())))
        if not marks or marks[-1][0] != n:
            marks.append((n, t, sq_err(read_fn())))
        return {"frames": n, "ms": t, "marks": marks, "nonfinite": int((~np.isfinite(marks[-1][2])).sum())}

    plain_ctx = api.Context(0)
    workloads.load(plain_ctx, wl)

    def plain_step(n):
        plain_ctx.render(wl.seed, n, 1)
        plain_ctx.synchronize()

    plain = run(plain_step, plain_ctx.read_accum, lambda n, t, f: n >= frames)
    plain_ctx.close()
    budget_ms = plain["ms"]

    def within_budget(n, t, f):
        return n > 0 and t + f > budget_ms

    cache_ctx = api.Context(0)
    workloads.load(cache_ctx, wl)
    cache = api.RadianceCache(cache_ctx, seed=1)
    cstats = []

    def cache_step(n):
        cstats.append(cache.render(wl.seed, n, 1, cache_bounce=cache_bounce, train_paths=train_paths, lr=3e-3,
                                   train_extra=train_extra))

    cached = run(cache_step, cache_ctx.read_accum, within_budget)
    cache.close()
    cache_ctx.close()

    cut_ctx = api.Context(0)
    workloads.load(cut_ctx, wl)
    zero = api.RadianceCache(cut_ctx, seed=1)
    zero.set_params(np.zeros(zero.n_params, np.float32))

    def cut_step(n):
        zero.render(wl.seed, n, 1, cache_bounce=cache_bounce, train_paths=0)

    truncated = run(cut_step, cut_ctx.read_accum, within_budget)
    zero.close()
    cut_ctx.close()

    runs = {"plain": plain, "cache": cached, "truncation": truncated}
    mask = ref_finite.copy()
    for r in runs.values():
        mask &= np.isfinite(r["marks"][-1][2])
    n_mask = int(mask.sum())

    def summary(r):
        curve = [(round(t / 1000.0, 4), float(np.sqrt(e[mask].sum() / (3 * n_mask)))) for _, t, e in r["marks"]]
        out = {"frames": r["frames"], "seconds": round(r["ms"] / 1000.0, 4),
               "ms_per_frame": round(r["ms"] / max(1, r["frames"]), 3),
               "rmse_at_equal_time": round(curve[-1][1], 6), "nonfinite_pixels": r["nonfinite"]}
        for level in levels:
            hit, prev = None, None
            for t, e in curve:
                if e <= level:
                    hit = t if prev is None or prev[1] <= level else prev[0] + (prev[1] - level) / (prev[1] - e) * (
                        t - prev[0])
                    break
                prev = (t, e)
            out[f"seconds_to_{level:g}"] = None if hit is None else round(hit, 4)
        out["curve"] = [(t, round(e, 6)) for t, e in curve]
        return out

    res = {name: summary(r) for name, r in runs.items()}
    bmask = mask & np.isfinite(ref_part).all(axis=-1) & np.isfinite(trunc_part).all(axis=-1)
    dtail = (ref_part - trunc_part)[bmask]
    res["truncation_bias"] = {
        "passes": bias_spp, "seconds": round(t_bias, 2), "pixels": int(bmask.sum()),
        "rmse": round(float(np.sqrt(np.mean(dtail * dtail))), 7),
        "mean": round(float(dtail.mean()), 7),
        "note": f"tonemapped mean of {bias_spp} plain passes minus the same passes ended at bounce {cache_bounce} "
                "into a zero cache (identical random streams): what truncation leaves out; compare with "
                "rmse_at_equal_time"}
    res["plain"]["note"] = f"{frames} frames; its time is every contender's budget"
    for lv in levels:
        key = f"seconds_to_{lv:g}"
        tp, tc, tt = res["plain"][key], res["cache"][key], res["truncation"][key]
        res[f"cache_speedup_to_{lv:g}_vs_plain"] = None if tp is None or tc is None else round(tp / tc, 3)
        res[f"cache_speedup_to_{lv:g}_vs_truncation"] = None if tt is None or tc is None else round(tt / tc, 3)
    res.update({
        "workload": "full-shading 256^3 scene at 1920x1080, 1 spp per frame, 8 bounces; FP64 reference of "
                    f"{ref_spp} spp (own seed); RMSE over pixels x RGB of the tonemapped mean",
        "reference_spp": ref_spp, "reference_seconds": round(t_ref, 2),
        "reference_nonfinite_pixels": int((~ref_finite).sum()), "scored_pixels": n_mask,
        "cache_bounce": cache_bounce, "train_paths_per_frame": train_paths, "train_extra_bounces": train_extra,
        "cache_mean_samples_per_frame": round(sum(x["samples"] for x in cstats) / max(1, len(cstats)), 1),
        "cache_mean_queries_per_frame": round(sum(x["queries"] for x in cstats) / max(1, len(cstats)), 1),
    })
    return res


def bench_cache(args, api, workloads):
    """The time-to-RMSE race on its own line (value: the cache's time to 1e-2);
    with --only, frames of one contender alone, for a kernel trace."""
    if args.only:
        print(json.dumps(profile_contender(api, workloads, args.only, args.frames, args.cache_bounce,
                                           args.train_paths, args.train_extra)))
        return
    res = rmse_race(api, workloads, frames=args.frames, ref_spp=args.ref_spp, cache_bounce=args.cache_bounce,
                    train_paths=args.train_paths, train_extra=args.train_extra)
    out = {
        "metric": METRIC,
        "value": res["cache"]["seconds_to_0.01"],
        "unit": "s",
        "n_gpus": 1,
        "steps": res["plain"]["frames"] + res["cache"]["frames"] + res["truncation"]["frames"],
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 tracer, fp16 mlp",
        "data": "synthetic",
        "config": {"workload": res["workload"]},
        "time_to_rmse": res,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

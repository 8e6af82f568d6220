"""Benchmark of the MI355X path tracer on one GPU: Mrays/s of the plain renderer
on the procedural 1.31M-triangle scene (the `value`), the batched renderer vs
sequential plain renders on the 64-scene batch, time-to-relMSE 1e-3 on the
Cornell box, and the path-replay adjoint's cost relative to the forward pass.

Workloads (seeded, synthetic; no files):
  procedural   64 Perlin-displaced level-5 icospheres (1,310,720 triangles) +
               ground + 4 rectangle lights, GGX conductors, 1920x1080, 64 spp,
               max depth 8, NEE + MIS, stratified sampler, box filter. A step is
               one sample pass over the frame; --steps 64 is the full render.
  batch        64 procedural-sphere scenes, triangles log-uniform 1e4..2e6
               (26.5M total), square resolutions log-uniform 128..1024, spp
               1..64, depth 4..16: one batched render (autotuned launch shape)
               vs mpt_render of each scene in turn, both with the pass-by-pass
               renderer (each sample pass bounces in lock step) and the default
               streaming one; equal relMSE because every sample is the same path
               in all three (films compared).
  cornell      36 triangles, 256x256, depth 8: progressive spp until relMSE
               <= 1e-3 against a 16384-spp render with another seed.
  textured     204,802 triangles, 256^2 albedo + roughness textures, 512x512,
               16 spp, depth 4: forward render vs path-replay adjoint (GPU
               events), then 200 Adam steps of texture / radiance recovery.

Reported (one JSON line):
  value          all ray queries (camera, continuation, MIS-BSDF, shadow) / s of the
                 procedural render, whole job, inputs resident in HBM
  roofline       the traversal kernel (one mixed launch per wavefront step: the
                 continuation and MIS-BSDF rays closest hit, the shadow rays any
                 hit): algorithmic bytes per query (32 B ray + 16 B hit + 24 B per
                 slab test + 36 B per primitive test, counted by the kernel) / its
                 average launch time (HIP events on the render stream)
  cpu_baseline   the compiled reference (oracle/_ref, built from the reference's
                 sources) on this host's cores, on a bounded sample of the same
                 frame at 1 spp; the oracle port when the reference build is absent
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s + time-to-relMSE 1e-3 (batched vs plain GPU); adjoint/forward time"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(spec, seconds_target=15.0, seed=7):
    """Rays/s of the reference's CPU path on a bounded pixel sample (1 spp)."""
    from oracle.binding import OracleScene, RefScene, ref_available
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 64))
    npix = spec.width * spec.height
    kind = "reference" if ref_available() else "port"
    if kind == "reference":
        side = tempfile.mkdtemp(prefix="mpt_bench_")
        S = RefScene(spec.to_crsd(sidecar_dir=side))
    else:
        S = OracleScene(spec)
    # calibrate on a small window around the image centre, then size the sample
    # (pixels around the centre, then whole frames of more samples) to ~seconds_target
    begin = npix // 2 - 2000
    secs, rays, smp = S.time_render(begin, begin + 2000, 1, threads, seed)
    want = 2000 * seconds_target / max(secs, 1e-3)
    n = int(min(npix, max(4000, want)))
    spp = int(max(1, min(spec.spp, round(want / npix)))) if n == npix else 1
    begin = max(0, npix // 2 - n // 2)
    secs, rays, smp = S.time_render(begin, begin + n, spp, threads, seed)
    total = int(rays.sum())
    return {"value": round(total / secs / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": kind,
            "sample": f"{n} pixels around the frame centre x {spp} spp, depth 8 ({total} ray queries, {secs:.1f} s)",
            "samples_per_s": round(smp / secs, 1)}


def all_rays(st):
    return st["camera_rays"] + st["continuation_rays"] + st["mis_rays"] + st["shadow_rays"]


def relmse(img, ref):
    return float(np.mean((img.astype(np.float64) - ref) ** 2 / (ref.astype(np.float64) ** 2 + 0.01)))


def bench_batch(ctx):
    """64-scene batch: sequential plain renders vs one autotuned batched render."""
    from cadise_amd import scenes
    from cadise_amd.api import Batch, Scene
    t0 = time.time()
    specs = scenes.batch_scenes()
    gen_s = time.time() - t0
    t0 = time.time()
    with cf.ThreadPoolExecutor(16) as ex:
        scs = list(ex.map(lambda s: Scene.from_spec(ctx, s), specs))
    build_s = time.time() - t0
    seeds = [0xB000 + i for i in range(len(scs))]
    work = sum(s.width * s.height * s.spp for s in specs)
    log(f"batch: {len(specs)} scenes, {sum(s.num_prims() for s in specs)} primitives, {work} samples "
        f"(generated {gen_s:.1f} s, built {build_s:.1f} s)")
    scs[0].render(seed=1, spp_begin=0, spp_end=1, readback=False)
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prays = 0
    for sc, sd in zip(scs, seeds):
        sc.render(seed=sd, readback=False, passwise=True)
        prays += all_rays(sc.last_stats())
    passwise_s = time.perf_counter() - t0
    log(f"batch: plain pass-by-pass sequential {passwise_s:.2f} s ({prays / passwise_s / 1e6:.0f} Mrays/s)")
    t0 = time.perf_counter()
    rays = 0
    for sc, sd in zip(scs, seeds):
        sc.render(seed=sd, readback=False)
        rays += all_rays(sc.last_stats())
    plain_s = time.perf_counter() - t0
    log(f"batch: plain streaming sequential {plain_s:.2f} s ({rays / plain_s / 1e6:.0f} Mrays/s)")
    b = Batch(ctx, scs)
    t0 = time.perf_counter()
    best, trials = b.autotune(seeds, trial_spp=1)
    tune_s = time.perf_counter() - t0
    log(f"batch: autotune {tune_s:.1f} s, {len(trials)} trials -> {best}")
    b.render(seeds)   # the graph for the full work is recorded on the first render
    t0 = time.perf_counter()
    st = b.render(seeds)
    batch_s = time.perf_counter() - t0
    brays = all_rays(st)
    log(f"batch: batched {batch_s:.2f} s ({brays / batch_s / 1e6:.0f} Mrays/s), speedup {passwise_s / batch_s:.2f}x "
        f"over pass-by-pass, {plain_s / batch_s:.2f}x over streaming single-scene renders")
    # equal relMSE: the films agree up to float summation order
    worst = 0.0
    for i, sc in enumerate(scs):
        f = sc.render(seed=seeds[i])
        g = b.film(i)
        worst = max(worst, float(np.abs(g - f).max() / max(1e-12, np.abs(f).max())))
    return {"scenes": len(specs), "primitives": int(sum(s.num_prims() for s in specs)), "samples": int(work),
            "plain_passwise_s": round(passwise_s, 3), "plain_streaming_s": round(plain_s, 3),
            "batched_s": round(batch_s, 3), "speedup_vs_passwise": round(passwise_s / batch_s, 2),
            "speedup_vs_streaming": round(plain_s / batch_s, 2),
            "plain_passwise_mrays_s": round(prays / passwise_s / 1e6, 1),
            "plain_streaming_mrays_s": round(rays / plain_s / 1e6, 1), "batched_mrays_s": round(brays / batch_s / 1e6, 1),
            "rays_equal": bool(brays == rays == prays), "films_max_rel_diff": worst, "autotune_s": round(tune_s, 2),
            "autotuned": best, "autotune_trials": len(trials)}


def bench_time_to_relmse(ctx, target=1e-3):
    """Cornell box: progressive spp until relMSE <= target against 16k spp."""
    from cadise_amd import scenes
    from cadise_amd.api import Scene, film_mean
    ref_spp = 16384
    rs = Scene.from_spec(ctx, scenes.cornell_box(spp=ref_spp))
    t0 = time.time()
    ref = film_mean(rs.render(seed=0x5EF)).astype(np.float64)
    ref_s = time.time() - t0
    del rs
    spp_max = 4096
    sc = Scene.from_spec(ctx, scenes.cornell_box(spp=spp_max))
    sc.render(seed=3, spp_begin=0, spp_end=1, readback=False)
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done, chunk, reached, err = 0, 4, None, None
    while done < spp_max:
        nxt = min(spp_max, done + chunk)
        film = sc.render(seed=0x1234, spp_begin=done, spp_end=nxt, accumulate=done > 0)
        done = nxt
        err = relmse(film_mean(film), ref)
        if err <= target:
            reached = done
            break
        chunk = max(4, done // 4)
    t = time.perf_counter() - t0
    log(f"cornell: relMSE {err:.2e} at {done} spp in {t:.3f} s (reference {ref_spp} spp in {ref_s:.1f} s)")
    return {"workload": "cornell 256x256 depth 8", "target": target, "seconds": round(t, 4), "spp": reached,
            "relmse": err, "reference_spp": ref_spp,
            "note": "wall time of progressive renders incl. readback and relMSE checks after each chunk"}


def bench_adjoint(ctx, adam_steps):
    from cadise_amd import scenes
    from cadise_amd.api import Scene, film_mean
    spec = scenes.textured_spheres()
    sc = Scene.from_spec(ctx, spec)
    G = np.random.default_rng(1).normal(size=(spec.height, spec.width, 3)).astype(np.float32) * 1e-6
    fw, aw = [], []
    for rep in range(4):
        sc.render(seed=100 + rep, keep_samples=True, readback=False)
        fw.append(sc.last_stats()["ms_gpu"])
        sc.adjoint(G, seed=100 + rep)
        aw.append(sc.last_stats()["ms_adjoint"])
    fms, ams = float(np.median(fw[1:])), float(np.median(aw[1:]))
    log(f"textured: forward {fms:.2f} ms, adjoint {ams:.2f} ms (ratio {ams / fms:.2f})")
    out = {"workload": "textured spheres 204802 tris, 512x512, 16 spp, depth 4, 256^2 albedo + roughness",
           "forward_ms": round(fms, 3), "adjoint_ms": round(ams, 3), "ratio": round(ams / fms, 3)}
    if adam_steps > 0:
        from cadise_amd.inverse import differentiation_workload
        _, r = differentiation_workload(ctx, steps=adam_steps, log=None)
        out["adam"] = {"steps": adam_steps, "seconds": round(r.seconds, 2), "loss_first": r.losses[0],
                       "loss_last10": float(np.mean(r.losses[-10:])), "tex_rmse_seen": r.tex_rmse,
                       "tex_rmse_better_observed_half": r.tex_rmse_well, "radiance": r.radiance}
        log(f"textured: {adam_steps} Adam steps in {r.seconds:.1f} s, loss {r.losses[0]:.3e} -> "
            f"{np.mean(r.losses[-10:]):.3e}, albedo RMSE {r.tex_rmse['albedo']}, roughness RMSE {r.tex_rmse['rough']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=4, help="BVH width (4 or 8)")
    ap.add_argument("--leaf-max", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batch")
    ap.add_argument("--no-relmse", action="store_true", help="skip the Cornell time-to-relMSE")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint / Adam workload")
    ap.add_argument("--no-passwise", action="store_true", help="skip the pass-by-pass render of the procedural frame")
    ap.add_argument("--adam-steps", type=int, default=200)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 3 device copies of 1 GiB (scale reference for rocprof FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("copy calibration: 3 x 1 GiB device copies (expect 3145728 KiB read, 3145728 KiB written)", flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; sharding pixels and samples over GPUs "
                                   "(one film reduction at the end) is not launched"}))
        sys.exit(2)

    import torch

    from cadise_amd import scenes
    from cadise_amd.api import Context, Scene, bvh_opts

    spec = scenes.procedural_spheres(spp=max(64, args.steps))
    t0 = time.time()
    ctx = Context(0)
    scene = Scene.from_spec(ctx, spec, bvh_opts(width=args.width, leaf_max=args.leaf_max))
    setup_s = time.time() - t0
    st_scene = scene.stats()
    log(f"procedural: {st_scene['nprims']} prims, {st_scene['wide_nodes']} BVH{st_scene['width']} nodes, "
        f"host build {st_scene['build_ms']:.0f} ms + collapse {st_scene['collapse_ms']:.0f} ms")
    seed = 0x5EED
    if args.warmup > 0:
        scene.render(seed=seed, spp_begin=0, spp_end=args.warmup, readback=False)
        log(f"procedural: warmup {args.warmup} passes done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    scene.render(seed=seed, spp_begin=0, spp_end=args.steps, readback=False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = scene.last_stats()
    closest = st["camera_rays"] + st["continuation_rays"] + st["mis_rays"]
    rays = closest + st["shadow_rays"]
    mrays = rays / dt / 1e6
    # roofline of the traversal kernel: one mixed launch per step traces every query
    boxes = st["box_tests"] + st["box_tests_any"]
    prims = st["tri_tests"] + st["tri_tests_any"]
    bytes_trace = 48.0 * rays + 24.0 * boxes + 36.0 * prims
    launches = max(1, st["closest_launches"])
    per_launch_bytes = bytes_trace / launches
    avg_launch_ms = st["ms_closest"] / launches
    achieved = per_launch_bytes / (avg_launch_ms * 1e-3) / 1e9
    log(f"procedural: {args.steps} passes in {dt:.3f} s: {rays} rays ({mrays:.1f} Mrays/s), closest "
        f"{st['ms_closest']:.1f} ms (closest + shadow) shade {st['ms_shade']:.1f} ms regen "
        f"{st['ms_camera']:.1f} ms over {launches} steps")
    kernel = f"k_trace<{st_scene['width']},{st_scene['stack_lds']},mixed>"
    # the same frame pass by pass (every pass's paths bounce in lock step)
    dt_pw = None
    if not args.no_passwise:
        scene.render(seed=seed, spp_begin=0, spp_end=1, readback=False, passwise=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        scene.render(seed=seed, spp_begin=0, spp_end=args.steps, readback=False, passwise=True)
        torch.cuda.synchronize()
        dt_pw = time.perf_counter() - t0
        log(f"procedural: pass by pass {dt_pw:.3f} s ({all_rays(scene.last_stats()) / dt_pw / 1e6:.1f} Mrays/s)")
    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scenes)",
        "config": {"workload": "procedural spheres: 64 displaced level-5 icospheres, 1310726 primitives, "
                               "1920x1080, 64 spp, depth 8, GGX, 4 quad lights, NEE+MIS",
                   "primitives": st_scene["nprims"], "width": spec.width, "height": spec.height,
                   "spp_per_step": 1, "max_depth": spec.max_depth, "bvh_width": st_scene["width"],
                   "leaf_max": st_scene["leaf_max"], "renderer": "streaming wavefront, 2^24 path slots"},
        "samples_per_s": round(st["samples"] / dt, 1),
        "passwise_mrays_s": round(rays / dt_pw / 1e6, 1) if dt_pw else None,
        "rays": {"camera": st["camera_rays"], "continuation": st["continuation_rays"], "mis_bsdf": st["mis_rays"],
                 "shadow": st["shadow_rays"]},
        "kernel_ms": {"trace": round(st["ms_closest"], 3),
                      "shade": round(st["ms_shade"], 3), "regenerate": round(st["ms_camera"], 3),
                      "gpu_total": round(st["ms_gpu"], 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": f"{kernel} (closest hit + shadow any hit, one launch per step)",
                     "launches": launches, "per_launch_bytes": round(per_launch_bytes),
                     "avg_launch_ms": round(avg_launch_ms, 4),
                     "box_tests_per_ray": round(boxes / max(rays, 1), 2),
                     "prim_tests_per_ray": round(prims / max(rays, 1), 2)},
        "setup_s": round(setup_s, 2),
    }
    # HBM bytes per launch of this kernel on this workload, measured with
    # rocprofv3 (FETCH_SIZE, WRITE_SIZE, copy-calibrated): profiles/traffic.json
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            meas = json.load(f).get(kernel)
        if meas and st_scene["leaf_max"] == meas.get("leaf_max", 2) and meas.get("renderer") == "streaming":
            out["roofline"]["traffic"] = meas["hbm_bytes_per_launch"]
            out["roofline"]["traffic_source"] = meas["source"]
            out["roofline"]["l2_hit_rate"] = meas.get("l2_hit_rate")
            out["roofline"]["valu_busy_pct"] = meas.get("valu_busy_pct")
    except (OSError, ValueError):
        pass
    del scene
    if not args.no_batch:
        out["batch"] = bench_batch(ctx)
    if not args.no_relmse:
        out["time_to_relmse"] = bench_time_to_relmse(ctx)
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint(ctx, args.adam_steps)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(spec, args.cpu_seconds)
        except Exception as e:   # the GPU line stands on its own
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

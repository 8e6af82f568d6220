"""Benchmark: Mrays/s of the wavefront path tracer on one MI355X.

Workload (the plain-GPU configuration): 64 Perlin-displaced level-5
icospheres (1,310,720 triangles) + ground + 4 rectangle lights, GGX
conductors, 1920x1080, 64 spp, max depth 8, NEE + MIS, stratified sampler,
box filter. A step is one sample pass over the whole frame (one path per
pixel, every bounce); --steps 64 is the full 64-spp render.

Reported (one JSON line):
  value          all ray queries (camera, continuation, MIS-BSDF, shadow) / s, whole job
  roofline       the closest-hit traversal kernel: algorithmic bytes per query
                 (32 B ray + 16 B hit + 24 B per slab test + 36 B per primitive
                 test, counted by the kernel) / its average launch time, HIP events
  cpu_baseline   the compiled reference (oracle/_ref, built from the reference's
                 sources) on this host's cores, on a bounded sample of the same
                 frame at 1 spp; the oracle port when the reference build is absent
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s + time-to-relMSE 1e-3 (batched vs plain GPU); adjoint/forward time"
HBM_PEAK_GBS = 8000.0


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
    return {"value": total / secs / 1e6, "unit": "Mrays/s", "cores": threads, "kind": kind,
            "sample": f"{n} pixels around the frame centre x {spp} spp, depth 8 ({total} ray queries, {secs:.1f} s)",
            "samples_per_s": smp / secs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=4, help="BVH width (4 or 8)")
    ap.add_argument("--leaf-max", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
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
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU design (pixel/sample sharding "
                                   "with one film reduction) is described in DESIGN.md and not launched"}))
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
    print(f"[bench] scene {st_scene['nprims']} prims, {st_scene['wide_nodes']} BVH{st_scene['width']} nodes, "
          f"host build {st_scene['build_ms']:.0f} ms + collapse {st_scene['collapse_ms']:.0f} ms", file=sys.stderr,
          flush=True)
    seed = 0x5EED
    if args.warmup > 0:
        scene.render(seed=seed, spp_begin=0, spp_end=args.warmup, readback=False)
        print(f"[bench] warmup {args.warmup} passes done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    scene.render(seed=seed, spp_begin=0, spp_end=args.steps, readback=False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = scene.last_stats()
    closest = st["camera_rays"] + st["continuation_rays"] + st["mis_rays"]
    rays = closest + st["shadow_rays"]
    mrays = rays / dt / 1e6
    # roofline of the closest-hit traversal kernel
    bytes_closest = 48.0 * closest + 24.0 * st["box_tests"] + 36.0 * st["tri_tests"]
    launches = max(1, st["closest_launches"])
    per_launch_bytes = bytes_closest / launches
    avg_launch_ms = st["ms_closest"] / launches
    achieved = per_launch_bytes / (avg_launch_ms * 1e-3) / 1e9
    print(f"[bench] {args.steps} passes in {dt:.3f} s: {rays} rays ({mrays:.1f} Mrays/s), "
          f"closest {st['ms_closest']:.1f} ms any {st['ms_anyhit']:.1f} ms shade {st['ms_shade']:.1f} ms "
          f"film {st['ms_film']:.1f} ms camera {st['ms_camera']:.1f} ms", file=sys.stderr, flush=True)
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
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "procedural spheres: 64 displaced level-5 icospheres, 1310726 primitives, "
                               "1920x1080, 64 spp, depth 8, GGX, 4 quad lights, NEE+MIS",
                   "primitives": st_scene["nprims"], "width": spec.width, "height": spec.height,
                   "spp_per_step": 1, "max_depth": spec.max_depth, "bvh_width": st_scene["width"],
                   "leaf_max": st_scene["leaf_max"]},
        "samples_per_s": round(st["samples"] / dt, 1),
        "rays": {"camera": st["camera_rays"], "continuation": st["continuation_rays"], "mis_bsdf": st["mis_rays"],
                 "shadow": st["shadow_rays"]},
        "kernel_ms": {"closest": round(st["ms_closest"], 3), "anyhit": round(st["ms_anyhit"], 3),
                      "shade": round(st["ms_shade"], 3), "camera": round(st["ms_camera"], 3),
                      "film": round(st["ms_film"], 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_trace<4,24,false> (closest hit)",
                     "per_launch_bytes": round(per_launch_bytes), "avg_launch_ms": round(avg_launch_ms, 4),
                     "box_tests_per_ray": round(st["box_tests"] / max(closest, 1), 2),
                     "prim_tests_per_ray": round(st["tri_tests"] / max(closest, 1), 2)},
        "setup_s": round(setup_s, 2),
    }
    # HBM bytes per launch of this kernel, measured with rocprofv3 (FETCH_SIZE,
    # WRITE_SIZE, copy-calibrated) on this workload: profiles/traffic.json
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            meas = json.load(f).get(f"k_trace<{st_scene['width']},24,false>")
        if meas and st_scene["leaf_max"] == 2:
            out["roofline"]["traffic"] = meas["hbm_bytes_per_launch"]
            out["roofline"]["traffic_source"] = meas["source"]
            out["roofline"]["l2_hit_rate"] = meas["l2_hit_rate"]
            out["roofline"]["valu_busy_pct"] = meas["valu_busy_pct"]
    except (OSError, ValueError):
        pass
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(spec, args.cpu_seconds)
        except Exception as e:   # the GPU line stands on its own
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark of the MI355X path tracer on one GPU: Mrays/s of the plain renderer
on the procedural 1.31M-triangle scene (the `value`), the batched renderer vs
sequential plain renders on the 64-scene batch, time-to-relMSE 1e-3 on the
Cornell box (GPU and the reference's CPU path against the reference's own
16384-spp image), and the path-replay adjoint's cost relative to the forward
pass.

Workloads (seeded, synthetic; no files except the committed Cornell-box image R):
  procedural   64 Perlin-displaced level-5 icospheres (1,310,720 triangles) +
               ground + 4 rectangle lights, GGX conductors, 1920x1080, 64 spp,
               max depth 8, NEE + MIS, stratified sampler, box filter. A step is
               one sample pass over the frame; --steps 64 is the full render.
  batch        64 procedural-sphere scenes, triangles log-uniform 1e4..2e6
               (26.5M total), square resolutions log-uniform 128..1024, spp
               1..64, depth 4..16: one batched render (autotuned launch shape)
               vs mpt_render of each scene in turn, both with the pass-by-pass
               renderer (each sample pass bounces in lock step) and the default
               streaming one, at equal spp (every sample is the same path in all
               three, so relMSE is equal); with --batch-relmse also at the spp
               each scene needs for relMSE 1e-3.
  cornell      36 triangles, 256x256, depth 8: progressive spp until relMSE
               <= 1e-3 against the compiled reference's 16384-spp image
               (tests/golden/cornell_r16k.npz), on the GPU and on the reference.
  textured     204,802 triangles, 256^2 albedo + roughness textures, 512x512,
               16 spp, depth 4: forward render vs path-replay adjoint (GPU
               events), then 200 Adam steps of texture / radiance recovery.

Reported (one JSON line):
  value          all ray queries of the reference's algorithm (camera, continuation,
                 MIS-BSDF, shadow) answered per second on the procedural render,
                 whole job, inputs resident in HBM. A MIS-BSDF query whose ray
                 misses the picked area light's own primitive is answered in
                 shading by that one primitive test (its closest hit cannot be the
                 light; the reference traces it and finds the same zero term), and
                 so is a shadow query whose NEE term is exactly zero (occluded or
                 not, it adds +0): both count as queries; traced_mrays_s counts
                 traced rays only
  roofline       the traversal kernel (one mixed launch per wavefront step: the
                 continuation and MIS-BSDF rays closest hit, the shadow rays any
                 hit): algorithmic bytes per query (32 B ray + 16 B hit + 24 B per
                 slab test + 36 B per primitive test, counted by the kernel) / its
                 average launch time (HIP events on the render stream), against
                 the 8 TB/s HBM peak; beside it the same launches' VALU work (16
                 lane-ops per slab test, 32 per primitive test) against 39.3 T
                 lane-ops/s, and their HBM traffic and issued VALU instructions
                 measured by rocprofv3 at this exact launch shape and device
                 source digest (profiles/traffic.json) when it was profiled
  cpu_baseline   the compiled reference (oracle/_ref, built from the reference's
                 sources) on this host's cores, on a bounded sample of the same
                 frame at 1 spp; the oracle port when the reference build is absent
"""
from __future__ import annotations

import argparse
import hashlib
import concurrent.futures as cf
import ctypes as C
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
VALU_PEAK_TOPS = 39.3     # 256 CUs x 64 lanes x 2.4 GHz, non-packed FP32 lane-ops
C1_REFERENCE = os.path.join(ROOT, "tests", "golden", "cornell_r16k.npz")

_hip = None


def sync():
    """hipDeviceSynchronize of the HIP runtime the library runs on."""
    global _hip
    if _hip is None:
        from cadise_amd import native
        native.load_library()
        _hip = C.CDLL("libamdhip64.so.7")
    if _hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(n or 1, 64))


def cpu_baseline(spec, seconds_target=15.0, seed=7):
    """Rays/s of the reference's CPU path on a bounded pixel sample (1 spp)."""
    from oracle.binding import OracleScene, RefScene, ref_available
    threads = host_threads()
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
            "samples_per_s": round(smp / secs, 1), "cpu_model": cpu_model()}


def all_rays(st):
    """Ray queries of the reference's algorithm: traced ones, the MIS-BSDF
    queries answered by the picked light's own record and the shadow queries
    whose NEE term is exactly zero."""
    return (st["camera_rays"] + st["continuation_rays"] + st["mis_rays"] + st["mis_resolved"] + st["shadow_rays"]
            + st["shadow_resolved"])


def relmse(img, ref):
    return float(np.mean((img.astype(np.float64) - ref) ** 2 / (ref.astype(np.float64) ** 2 + 0.01)))


def bench_batch(ctx, relmse_target=None):
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
    sync()

    def sequential(spps, passwise):
        t0 = time.perf_counter()
        nr = 0
        for sc, sd, n in zip(scs, seeds, spps):
            sc.render(seed=sd, spp_begin=0, spp_end=n, readback=False, passwise=passwise)
            nr += all_rays(sc.last_stats())
        sync()
        return time.perf_counter() - t0, nr

    spp0 = [s.spp for s in specs]
    passwise_s, prays = sequential(spp0, True)
    log(f"batch: plain pass-by-pass sequential {passwise_s:.2f} s ({prays / passwise_s / 1e6:.0f} Mrays/s)")
    plain_s, rays = sequential(spp0, False)
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
    out = {"scenes": len(specs), "primitives": int(sum(s.num_prims() for s in specs)), "samples": int(work),
           "plain_passwise_s": round(passwise_s, 3), "plain_streaming_s": round(plain_s, 3),
           "batched_s": round(batch_s, 3), "speedup_vs_passwise": round(passwise_s / batch_s, 2),
           "speedup_vs_streaming": round(plain_s / batch_s, 2),
           "plain_passwise_mrays_s": round(prays / passwise_s / 1e6, 1),
           "plain_streaming_mrays_s": round(rays / plain_s / 1e6, 1), "batched_mrays_s": round(brays / batch_s / 1e6, 1),
           "rays_equal": bool(brays == rays == prays), "films_max_rel_diff": worst, "autotune_s": round(tune_s, 2),
           "autotuned": best, "autotune_trials": len(trials),
           "note": "equal spp per scene (the scenes' own sample counts): every sample is the same path in all three "
                   "renderers, so each scene's relMSE is equal across them"}
    del b
    if relmse_target:
        ctx.release_stream_buffers()
        out["time_to_relmse"] = batch_time_to_relmse(ctx, specs, scs, seeds, best, relmse_target)
    return out


def batch_time_to_relmse(ctx, specs, scs, seeds, best, target):
    """Per scene, the spp that reaches relMSE <= target, then the scenes at those
    spp rendered one after another with the default streaming renderer and as
    one batch (the autotuned launch shape), and one after another pass by pass.

    relMSE against the unknown converged image is estimated without a
    reference render from two independent renders A, B of equal spp (the
    estimator is unbiased, so E[(A - B)^2] = 2 Var):
        relMSE ~ mean((A - B)^2 / 2 / (((A + B) / 2)^2 + 0.01)).
    The needed spp follows from the 1/spp law on the scenes' own spp (x1.15);
    at those spp the batch (seeds s) and the sequential renders (seeds s') are
    two such independent renders, which checks the result."""
    import dataclasses

    from cadise_amd.api import Batch, Scene, batch_opts, film_mean

    def est(a, b):
        m = 0.5 * (a + b)
        return float(np.mean((a - b) ** 2 / 2.0 / (m * m + 0.01)))

    t0 = time.time()
    need, base = [], []
    for i, (s, sc) in enumerate(zip(specs, scs)):
        a = film_mean(sc.render(seed=seeds[i])).astype(np.float64)
        b = film_mean(sc.render(seed=0xE000 + i)).astype(np.float64)
        e = est(a, b)
        base.append(e)
        need.append(int(np.ceil(s.spp * e / target * 1.15)))
        if i % 16 == 15:
            log(f"batch relMSE: {i + 1} scenes estimated ({time.time() - t0:.1f} s)")
    prep_s = time.time() - t0
    specs2 = [dataclasses.replace(s, spp=n) for s, n in zip(specs, need)]
    with cf.ThreadPoolExecutor(16) as ex:
        scs2 = list(ex.map(lambda s: Scene.from_spec(ctx, s), specs2))
    samples = int(sum(n * s.width * s.height for n, s in zip(need, specs)))
    log(f"batch relMSE: estimates {prep_s:.1f} s; spp for relMSE {target:g}: {min(need)}..{max(need)}, "
        f"{samples / 1e9:.2f} G samples (configured: {sum(s.spp * s.width * s.height for s in specs) / 1e9:.2f} G)")
    seeds2 = [0xF000 + i for i in range(len(scs2))]
    scs2[0].render(seed=1, spp_begin=0, spp_end=1, readback=False)
    sync()
    t0 = time.perf_counter()
    rays = 0
    for k, (sc, sd) in enumerate(zip(scs2, seeds2)):
        sc.render(seed=sd, readback=False)
        rays += all_rays(sc.last_stats())
        if k % 16 == 15:
            sync()
            log(f"batch relMSE: {k + 1} sequential renders ({time.perf_counter() - t0:.1f} s)")
    sync()
    plain_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    prays = 0
    for k, (sc, sd) in enumerate(zip(scs2, seeds2)):
        sc.render(seed=sd, readback=False, passwise=True)
        prays += all_rays(sc.last_stats())
        if k % 16 == 15:
            sync()
            log(f"batch relMSE: {k + 1} sequential pass-by-pass renders ({time.perf_counter() - t0:.1f} s)")
    sync()
    passwise_s = time.perf_counter() - t0
    bb = Batch(ctx, scs2, batch_opts(**best))
    # the launch shape for this work: the autotuner on trial renders of 8 spp per scene
    t0 = time.perf_counter()
    best2, trials2 = bb.autotune(seeds, trial_spp=8)
    log(f"batch relMSE: autotune {time.perf_counter() - t0:.1f} s, {len(trials2)} trials -> {best2}")
    bb.render(seeds)        # records the graph for this work
    log("batch relMSE: graph recorded")
    t0 = time.perf_counter()
    stats = bb.render(seeds)
    batch_s = time.perf_counter() - t0
    errs = [est(film_mean(bb.film(i)).astype(np.float64), film_mean(sc.render(seed=sd)).astype(np.float64))
            for i, (sc, sd) in enumerate(zip(scs2, seeds2))]
    log(f"batch relMSE: sequential pass-by-pass {passwise_s:.2f} s, sequential streaming {plain_s:.2f} s, "
        f"batched {batch_s:.2f} s ({passwise_s / batch_s:.2f}x, {plain_s / batch_s:.2f}x); "
        f"estimated relMSE at those spp: max {max(errs):.2e}, mean {np.mean(errs):.2e}")
    del bb, scs2
    ctx.release_stream_buffers()
    return {"target": target,
            "estimate": "relMSE to the converged image from two independent renders of equal spp: "
                        "mean((A - B)^2 / 2 / (((A + B) / 2)^2 + 0.01))",
            "spp_min": int(min(need)), "spp_max": int(max(need)), "samples": samples,
            "autotuned": best2, "plain_passwise_s": round(passwise_s, 3), "plain_streaming_s": round(plain_s, 3),
            "batched_s": round(batch_s, 3), "speedup_vs_passwise": round(passwise_s / batch_s, 2),
            "speedup_vs_streaming": round(plain_s / batch_s, 2), "relmse_at_spp_max": max(errs),
            "relmse_at_spp_mean": float(np.mean(errs)), "batched_rays": all_rays(stats), "plain_rays": rays,
            "plain_passwise_rays": prays,
            "relmse_at_configured_spp_mean": float(np.mean(base))}


def bench_time_to_relmse(ctx, target=1e-3, cpu=True):
    """Cornell box: progressive spp until relMSE <= target against the compiled
    reference's 16384-spp image; the reference's CPU path at the spp the GPU
    needed (raised by 25% until it reaches the target)."""
    from cadise_amd import scenes
    from cadise_amd.api import Scene, film_mean
    if os.path.exists(C1_REFERENCE):
        g = np.load(C1_REFERENCE)
        ref = g["image"].astype(np.float64)
        ref_desc = (f"the compiled reference's own renderer at {int(g['spp'])} spp (tests/golden/cornell_r16k.npz, "
                    f"{float(g['seconds']):.0f} s on {int(g['threads'])} threads)")
    else:   # the fixture is committed; this keeps the line whole if it is not
        rs = Scene.from_spec(ctx, scenes.cornell_box(spp=16384))
        ref = film_mean(rs.render(seed=0x5EF)).astype(np.float64)
        ref_desc = "GPU render at 16384 spp (another seed): the reference image fixture is missing"
        del rs
    spp_max = 4096
    sc = Scene.from_spec(ctx, scenes.cornell_box(spp=spp_max))
    sc.render(seed=3, spp_begin=0, spp_end=1, readback=False)
    sync()
    t0 = time.perf_counter()
    # progressive: after each chunk the film is read back and its relMSE checked;
    # relMSE falls as 1/spp, so the next chunk ends where that law (with a 10%
    # margin) puts the target, at least 4 spp further
    done, nxt, reached, err = 0, 32, None, None
    while done < spp_max:
        film = sc.render(seed=0x1234, spp_begin=done, spp_end=nxt, accumulate=done > 0)
        done = nxt
        err = relmse(film_mean(film), ref)
        if err <= target:
            reached = done
            break
        nxt = min(spp_max, max(done + 4, int(np.ceil(done * err / target * 1.1))))
    t = time.perf_counter() - t0
    log(f"cornell: GPU relMSE {err:.2e} at {done} spp in {t:.3f} s against {ref_desc}")
    out = {"workload": "cornell 256x256 depth 8 (36 triangles, box filter, stratified)", "target": target,
           "seconds": round(t, 4), "spp": reached, "relmse": err, "reference_image": ref_desc,
           "note": "GPU: wall time of progressive renders incl. readback and relMSE checks after each chunk"}
    if cpu and reached:
        from oracle.binding import RefScene, ref_available
        if ref_available():
            threads = host_threads()
            spp = reached
            for _ in range(4):
                R = RefScene(scenes.cornell_box(spp=spp).to_crsd())
                t0 = time.perf_counter()
                f = R.render_linear(seed=0x77, threads=threads)
                secs = time.perf_counter() - t0
                e = relmse(film_mean(f), ref)
                if e <= target:
                    break
                spp = int(np.ceil(spp * 1.25))
            log(f"cornell: reference CPU path relMSE {e:.2e} at {spp} spp in {secs:.2f} s on {threads} threads")
            out["cpu_reference"] = {"seconds": round(secs, 3), "spp": spp, "relmse": e, "cores": threads,
                                    "reached": bool(e <= target), "cpu_model": cpu_model(),
                                    "note": "one render of the compiled reference at that spp (its sampling "
                                            "renderer, seeded random streams), wall time"}
            out["gpu_vs_cpu"] = round(secs / t, 1) if e <= target else None
    return out


def bench_adjoint(ctx, adam_steps):
    from cadise_amd import scenes
    from cadise_amd.api import Scene
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


def device_source_digest():
    """First 12 hex digits of the SHA-256 over the device sources (sorted file
    names and contents): a traffic record applies only to the kernels it measured."""
    d = os.path.join(ROOT, "cadise_amd", "csrc", "device")
    h = hashlib.sha256()
    for f in sorted(os.listdir(d)):
        h.update(f.encode())
        with open(os.path.join(d, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:12]


def measured_traffic(kernel, shape):
    """rocprofv3's HBM bytes per launch for `kernel` at this launch shape
    (profiles/traffic.json), or None when that shape was not profiled."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or any(rec.get(k) != v for k, v in shape.items()):
        return None
    return rec


def calibrate_copy():
    """3 device-to-device copies of 1 GiB (scale of rocprofv3's FETCH_SIZE / WRITE_SIZE)."""
    from cadise_amd import native
    native.load_library()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    n = 1 << 30
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), n) == 0 and hip.hipMalloc(C.byref(b), n) == 0
    assert hip.hipMemset(a, 1, n) == 0
    hip.hipDeviceSynchronize()
    for _ in range(3):
        assert hip.hipMemcpy(b, a, n, 3) == 0   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    print("copy calibration: 3 x 1 GiB device copies (expect 3145728 KiB read, 3145728 KiB written)", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=4, help="BVH width (4 or 8)")
    ap.add_argument("--leaf-max", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the CPU time-to-relMSE")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batch")
    ap.add_argument("--batch-relmse", action="store_true",
                    help="also time the batch and sequential renders at the spp each scene needs for relMSE 1e-3")
    ap.add_argument("--no-relmse", action="store_true", help="skip the Cornell time-to-relMSE")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint / Adam workload")
    ap.add_argument("--no-passwise", action="store_true", help="skip the pass-by-pass render of the procedural frame")
    ap.add_argument("--adam-steps", type=int, default=200)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 3 device copies of 1 GiB (scale reference for rocprof FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; sharding pixels and samples over GPUs "
                                   "(one film reduction at the end) is not launched"}))
        sys.exit(2)

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
    sync()
    t0 = time.perf_counter()
    scene.render(seed=seed, spp_begin=0, spp_end=args.steps, readback=False)
    sync()
    dt = time.perf_counter() - t0
    st = scene.last_stats()
    rays = all_rays(st)
    traced = rays - st["mis_resolved"] - st["shadow_resolved"]
    mrays = rays / dt / 1e6
    # roofline of the traversal kernel: one mixed launch per step traces every traced query
    boxes = st["box_tests"] + st["box_tests_any"]
    prims = st["tri_tests"] + st["tri_tests_any"]
    bytes_trace = 48.0 * traced + 24.0 * boxes + 36.0 * prims
    ops_trace = 16.0 * boxes + 32.0 * prims
    launches = max(1, st["closest_launches"])
    per_launch_bytes = bytes_trace / launches
    avg_launch_ms = st["ms_closest"] / launches
    achieved = per_launch_bytes / (avg_launch_ms * 1e-3) / 1e9
    valu_tops = ops_trace / launches / (avg_launch_ms * 1e-3) / 1e12
    slots = st["path_slots"]
    log(f"procedural: {args.steps} passes in {dt:.3f} s: {rays} rays ({mrays:.1f} Mrays/s), trace "
        f"{st['ms_closest']:.1f} ms (closest + shadow) shade {st['ms_shade']:.1f} ms regen "
        f"{st['ms_camera']:.1f} ms over {launches} steps, {slots} path slots")
    kernel = f"k_trace<{st_scene['width']},{st_scene['stack_lds']},mixed>"
    shape = {"steps": args.steps, "warmup": args.warmup, "path_slots": slots, "leaf_max": st_scene["leaf_max"],
             "block": st_scene["trace_block"], "blocks_per_cu": st_scene["blocks_per_cu"],
             "device_src": device_source_digest()}
    # the same frame pass by pass (every pass's paths bounce in lock step)
    dt_pw = None
    if not args.no_passwise:
        scene.render(seed=seed, spp_begin=0, spp_end=1, readback=False, passwise=True)
        sync()
        t0 = time.perf_counter()
        scene.render(seed=seed, spp_begin=0, spp_end=args.steps, readback=False, passwise=True)
        sync()
        dt_pw = time.perf_counter() - t0
        log(f"procedural: pass by pass {dt_pw:.3f} s ({all_rays(scene.last_stats()) / dt_pw / 1e6:.1f} Mrays/s)")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": f"{kernel} (closest hit + shadow any hit, one launch per step)",
            "launches": launches, "per_launch_bytes": round(per_launch_bytes),
            "avg_launch_ms": round(avg_launch_ms, 4),
            "box_tests_per_ray": round(boxes / max(traced, 1), 2),
            "prim_tests_per_ray": round(prims / max(traced, 1), 2),
            "valu_ops_per_launch": round(ops_trace / launches),
            "valu_achieved_tops": round(valu_tops, 3), "valu_peak_tops": VALU_PEAK_TOPS,
            "valu_frac": round(valu_tops / VALU_PEAK_TOPS, 4)}
    meas = measured_traffic(kernel, shape)
    if meas:
        traffic = meas["hbm_bytes_per_launch"]
        roof["traffic"] = traffic
        roof["traffic_source"] = meas["source"]
        roof["hbm_measured_gbs"] = round(traffic / (avg_launch_ms * 1e-3) / 1e9, 1)
        roof["hbm_measured_frac"] = round(traffic / (avg_launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        roof["traffic_per_algorithmic_byte"] = round(traffic / per_launch_bytes, 3)
        for k in ("l2_hit_rate", "valu_busy_pct", "valu_issued_frac", "avg_duration_us_rocprof"):
            if k in meas:
                roof[k] = meas[k]
    else:
        roof["traffic_note"] = f"no rocprofv3 traffic record for {kernel} at {shape} in profiles/traffic.json"
    shade_launches = launches
    vertices = st["camera_rays"] + st["continuation_rays"]
    shade = {"kernel": "k_shade<streaming>", "launches": shade_launches,
             "avg_launch_ms": round(st["ms_shade"] / shade_launches, 4), "path_vertices": vertices,
             "ns_per_vertex": round(st["ms_shade"] * 1e6 / max(vertices, 1), 3)}
    meas_s = measured_traffic("k_shade<streaming>", shape)
    if meas_s:
        shade["traffic_per_launch"] = meas_s["hbm_bytes_per_launch"]
        shade["bytes_per_vertex"] = round(meas_s["hbm_bytes_per_launch"] * shade_launches / max(vertices, 1), 1)
        shade["traffic_source"] = meas_s["source"]
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
                   "leaf_max": st_scene["leaf_max"], "renderer": f"streaming wavefront, {slots} path slots",
                   "trace_launch": f"{st_scene['blocks_per_cu']} blocks of {st_scene['trace_block']} per CU, "
                                   f"{st_scene['stack_lds']} LDS stack entries per lane"},
        "samples_per_s": round(st["samples"] / dt, 1),
        "traced_mrays_s": round(traced / dt / 1e6, 2),
        "passwise_mrays_s": round(rays / dt_pw / 1e6, 1) if dt_pw else None,
        "rays": {"camera": st["camera_rays"], "continuation": st["continuation_rays"], "mis_bsdf": st["mis_rays"],
                 "mis_bsdf_resolved_by_light_record": st["mis_resolved"], "shadow": st["shadow_rays"],
                 "shadow_resolved_zero_term": st["shadow_resolved"]},
        "kernel_ms": {"trace": round(st["ms_closest"], 3),
                      "shade": round(st["ms_shade"], 3), "regenerate": round(st["ms_camera"], 3),
                      "gpu_total": round(st["ms_gpu"], 3)},
        "roofline": roof,
        "shade": shade,
        "setup_s": round(setup_s, 2),
    }
    del scene
    ctx.release_stream_buffers()
    if not args.no_batch:
        out["batch"] = bench_batch(ctx, 1e-3 if args.batch_relmse else None)
    if not args.no_relmse:
        out["time_to_relmse"] = bench_time_to_relmse(ctx, cpu=not args.no_cpu)
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

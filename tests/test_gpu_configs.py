"""The plain-GPU and batching workloads at full size.

Procedural scene (64 displaced level-5 icospheres, 1,310,726 primitives,
1920 x 1080, GGX, four quad lights):
  * 1,000,000 seeded rays (half camera primaries, half secondary rays leaving
    surface points with cosine directions): closest-hit ids and t bit-exact
    against the oracle's binary-BVH traversal on every ray whose closest hit is
    unique, exact-t ties counted and listed; any-hit flags identical on the same
    rays with bounded maxT (bvhAccelerator.cpp:40-156);
  * the full 1920 x 1080 frame at 1 spp and depth 8, the streaming renderer
    against the oracle with the GPU's random layout: per-pixel linear radiance
    within 1e-4 relative (+1e-6) on >= 98% of pixels, image mean within 1e-3,
    filter weights identical (box filter).
64-scene batch (26.5M primitives, resolutions 128^2..1024^2, 1..64 spp, depth
4..16): the batched renderer's ray counts equal the sums over sequential
plain renders exactly, and every scene's film equals its plain film up to the
order of float additions (2e-5 relative + 2e-6).
"""
import time

import numpy as np
import pytest

from cadise_amd import scenes
from oracle.binding import OR_RNG_PHILOX, OracleScene
from tests.helpers import stress_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from cadise_amd.api import Context
    return Context(0)


@pytest.fixture(scope="module")
def procedural():
    t0 = time.time()
    spec = scenes.procedural_spheres()
    orc = OracleScene(spec)
    rays = stress_rays(spec, orc, 1_000_000, seed=0xC2)
    oprim, ot = orc.trace(rays)
    ties = orc.ties(rays)
    short = rays.copy()
    short[:, 7] = np.random.default_rng(0xC5).uniform(0.05, 12.0, len(rays)).astype(np.float32)
    oocc, _ = orc.trace(short, any_hit=True)
    print(f"[procedural] {spec.num_prims()} prims, {len(rays)} rays, oracle {time.time() - t0:.1f} s, "
          f"hits {int((oprim >= 0).sum())}, tied rays {int((ties > 1).sum())}", flush=True)
    return spec, orc, rays, short, oprim, ot, ties, oocc.astype(bool)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("width", [4, 8])
def test_procedural_trace_bit_exact(ctx, procedural, width):
    from cadise_amd.api import Scene, bvh_opts
    spec, orc, rays, short, oprim, ot, ties, oocc = procedural
    sc = Scene.from_spec(ctx, spec, bvh_opts(width=width, leaf_max=2 if width == 4 else 4))
    prim, t = sc.trace(rays)
    unique = ties <= 1
    bad = np.flatnonzero(unique & ((prim != oprim) | (t.view(np.int32) != ot.view(np.int32))))
    print(f"[procedural bvh{width}] {len(rays)} rays, {int((~unique).sum())} tied, {len(bad)} mismatches", flush=True)
    assert len(bad) == 0, f"first {bad[:8].tolist()}: gpu {prim[bad[:8]].tolist()} oracle {oprim[bad[:8]].tolist()}"
    for i in np.flatnonzero(~unique):
        print(f"  tie ray {i}: gpu prim {prim[i]} oracle prim {oprim[i]} t {ot[i]!r} ({ties[i]} primitives)")
    assert np.array_equal(t[~unique].view(np.int32), ot[~unique].view(np.int32))
    occ = sc.is_occluded(short)
    assert np.array_equal(occ, oocc), f"{int((occ != oocc).sum())} any-hit mismatches"


@pytest.mark.timeout(600)
def test_procedural_frame_vs_oracle(ctx, procedural):
    from cadise_amd.api import Scene, film_mean
    spec, orc, *_ = procedural
    sc = Scene.from_spec(ctx, spec)
    film = sc.render(seed=0xC2, spp_begin=0, spp_end=1)
    st = sc.last_stats()
    t0 = time.time()
    orf = orc.render_linear(0xC2, rng_mode=OR_RNG_PHILOX, spp_begin=0, spp_end=1, threads=16)
    g, o = film_mean(film), film_mean(orf)
    close = np.all(np.abs(g - o) <= 1e-4 * np.abs(o) + 1e-6, axis=-1)
    rel = abs(g.mean() - o.mean()) / max(o.mean(), 1e-12)
    print(f"[procedural frame] {spec.width}x{spec.height} 1 spp depth {spec.max_depth}: {st['samples']} samples, "
          f"GPU {st['ms_gpu']:.1f} ms, oracle {time.time() - t0:.1f} s; close={close.mean():.5f} "
          f"mean gpu={g.mean():.6f} oracle={o.mean():.6f} rel={rel:.2e}", flush=True)
    assert st["samples"] == spec.width * spec.height
    assert close.mean() >= 0.98 and rel < 1e-3
    assert np.array_equal(film[..., 3], orf[..., 3])


@pytest.mark.timeout(900)
def test_batch_64_scenes_matches_plain(ctx):
    from cadise_amd.api import Batch, Scene
    t0 = time.time()
    specs = scenes.batch_scenes()
    scs = [Scene.from_spec(ctx, s) for s in specs]
    seeds = [7000 + 13 * i for i in range(len(specs))]
    t1 = time.time()
    plain, stats = [], []
    for sc, sd in zip(scs, seeds):
        plain.append(sc.render(seed=sd))
        stats.append(sc.last_stats())
    b = Batch(ctx, scs)
    bst = b.render(seeds)
    print(f"[batch 64] {sum(s.num_prims() for s in specs)} prims, {bst['samples']} samples; built in {t1 - t0:.1f} s; "
          f"plain {sum(s['ms_gpu'] for s in stats):.0f} ms GPU, batched {bst['ms_gpu']:.0f} ms GPU", flush=True)
    for k in ("camera_rays", "continuation_rays", "mis_rays", "mis_resolved", "shadow_rays", "shadow_resolved",
              "samples"):
        assert bst[k] == sum(s[k] for s in stats), (k, bst[k], sum(s[k] for s in stats))
    for i in range(len(scs)):
        np.testing.assert_allclose(b.film(i), plain[i], rtol=2e-5, atol=2e-6, err_msg=f"scene {i}")

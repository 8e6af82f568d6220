"""Batched streaming renderer (mpt_batch_*) against the plain per-scene
renderer: every sample of the batch is the same path mpt_render draws for that
scene and seed, so each scene's film equals the plain film up to the order of
the float additions (the batch splats with atomics), and the ray counts of the
batch equal the sums of the plain renders' counts exactly. Checked with path
slots far fewer than the work (slots recycled many times), both BVH widths
(a width-8 batch over scenes built with width 4 re-collapses their binary trees
and must equal width-8 scenes node for node, so tie-breaking matches), all LDS
stack depths and block sizes, and after the autotuner."""
import os

import numpy as np
import pytest

from cadise_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from cadise_amd.api import Context
    return Context(0)


@pytest.fixture(scope="module")
def batch_set(ctx):
    from cadise_amd.api import Scene, bvh_opts
    specs = scenes.batch_scenes(seed=99, n_scenes=6, tris=(1e3, 5e4), res=(24, 64), spp=(1, 8), depth=(2, 8))
    specs.append(scenes.cornell_box(width=40, height=32, spp=4, filter="gaussian"))
    specs.append(scenes.textured_spheres(n_spheres=3, level=2, tex_res=16, width=32, height=32, spp=3, max_depth=5))
    specs.append(scenes.sky_spheres(n_spheres=4, level=2, width=40, height=24, spp=3, max_depth=6, env_size=(64, 32),
                                    sun_diameter_deg=6.0))
    specs.append(scenes.material_zoo(width=40, height=30, spp=3, max_depth=9, textured=True))
    specs.append(scenes.material_zoo(width=24, height=20, spp=5, max_depth=7, estimator="vpt"))
    # picture and checkerboard textures after other scenes' textures: the merge offsets
    # the checkerboards' odd / even indices
    here = os.path.dirname(os.path.abspath(__file__))
    pics = np.load(os.path.join(here, "golden", "images.npz"))
    specs.append(scenes.checker_cornell(os.path.join(here, "golden", "images", "rgb8_filters.png"),
                                        pics["rgb8_filters.png"], width=32, height=32, spp=3, max_depth=5))
    big = scenes.procedural_spheres(seed=5, n_spheres=6, level=3, width=256, height=200, spp=6, max_depth=9)
    specs.append(big)
    seeds = [1000 + 17 * i for i in range(len(specs))]
    out = {}
    for width in (4, 8):
        scs = [Scene.from_spec(ctx, s, bvh_opts(width=width)) for s in specs]
        plain, stats = [], []
        for sc, sd in zip(scs, seeds):
            plain.append(sc.render(seed=sd))
            stats.append(sc.last_stats())
        out[width] = (scs, plain, stats)
    work = sum(s.width * s.height * s.spp for s in specs)
    print(f"[batch] {len(specs)} scenes, {sum(s.num_prims() for s in specs)} primitives, {work} samples", flush=True)
    return specs, seeds, out


def _check(batch, batch_set, bst, label, width):
    specs, seeds, out = batch_set
    scs, plain, stats = out[width]
    for i in range(len(scs)):
        f = batch.film(i)
        np.testing.assert_allclose(f, plain[i], rtol=2e-5, atol=2e-6, err_msg=f"{label}: scene {i}")
    for k in ("camera_rays", "continuation_rays", "mis_rays", "mis_resolved", "shadow_rays", "shadow_resolved",
              "samples"):
        assert bst[k] == sum(s[k] for s in stats), (label, k, bst[k], sum(s[k] for s in stats))
    print(f"[batch {label}] {bst['samples']} samples, {bst['ms_gpu']:.2f} ms GPU, films match", flush=True)


@pytest.mark.parametrize("scene_width,opts", [
    (4, dict()),
    (4, dict(capacity_log2=16, graph_steps=4)),
    (8, dict(width=8, stack_lds=16, block=128, capacity_log2=17)),
    (4, dict(width=8, stack_lds=24, capacity_log2=16)),
    (4, dict(width=4, stack_lds=32, capacity_log2=18, graph_steps=32)),
])
def test_batch_matches_plain(ctx, batch_set, scene_width, opts):
    from cadise_amd.api import Batch, batch_opts
    specs, seeds, out = batch_set
    b = Batch(ctx, out[scene_width][0], batch_opts(**opts))
    bst = b.render(seeds)
    width = opts.get("width", 4)
    _check(b, batch_set, bst, str(opts), width)
    # a second render on the recorded graph gives the same films
    bst2 = b.render(seeds)
    _check(b, batch_set, bst2, str(opts) + " again", width)


def test_batch_autotune(ctx, batch_set):
    from cadise_amd.api import Batch
    specs, seeds, out = batch_set
    b = Batch(ctx, out[4][0])
    best, trials = b.autotune(seeds, trial_spp=1)
    print(f"[autotune] best {best} from {len(trials)} trials: " +
          ", ".join(f"{t[0]['width']}/{t[0]['stack_lds']}/{t[0]['block']}/2^{t[0]['capacity_log2']}: {t[1]:.2f} ms"
                    for t in trials), flush=True)
    assert len(trials) >= 6 and b.opts == best
    bst = b.render(seeds)
    _check(b, batch_set, bst, "autotuned", best["width"])


def test_batch_films_vs_oracle(ctx, batch_set):
    """Each film of a batched render against the oracle's render of that scene
    with the GPU's random layout (Philox per pixel / sample / bounce / slot): per
    pixel within 1e-4 relative (+1e-6) on >= 98% of pixels, image mean within 1e-3."""
    from cadise_amd.api import Batch, film_mean
    from oracle.binding import OR_RNG_PHILOX, OracleScene
    specs, seeds, out = batch_set
    b = Batch(ctx, out[4][0])
    b.render(seeds)
    for i, (spec, sd) in enumerate(zip(specs, seeds)):
        g = film_mean(b.film(i))
        o = film_mean(OracleScene(spec).render_linear(sd, rng_mode=OR_RNG_PHILOX))
        close = np.all(np.abs(g - o) <= 1e-4 * np.abs(o) + 1e-6, axis=-1)
        rel = abs(g.mean() - o.mean()) / max(o.mean(), 1e-12)
        print(f"[batch vs oracle] scene {i} ({spec.width}x{spec.height}, {spec.spp} spp, depth {spec.max_depth}): "
              f"close={close.mean():.5f} rel={rel:.2e}", flush=True)
        assert close.mean() >= 0.98 and rel < 1e-3, i

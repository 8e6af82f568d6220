"""GPU path-replay adjoint (mpt_render_adjoint) against

* the oracle's FP64 host gradient (or_adjoint, pinned in test_adjoint_oracle.py):
  the same paths, the same detached-sampling derivative;
* central finite differences (eps = 1e-3) of the GPU forward renderer with common
  random numbers, on 256 random albedo texels and the emitted radiance of the
  light. Albedo and Le change no sampled direction, so at max-depth 4 (Russian
  roulette inert) the difference quotient and the adjoint estimate the same
  per-sample derivative: bar 2% relative per parameter.
* roughness texels change the sampled GGX directions, so there the difference
  quotient of one sample set differs from the adjoint by zero-mean noise (both
  estimate the derivative of the expected image). They are checked against the
  FP64 host gradient, and in aggregate against central differences over
  independent seeds: per 16 x 16 texel block the mean difference quotient over
  the seeds against the mean adjoint, by the slope of their regression through
  the origin (test_roughness_gradient_vs_finite_differences).
"""
import time

import numpy as np
import pytest

from cadise_amd import scenes
from cadise_amd.api import film_mean
from oracle.binding import OR_RNG_PHILOX, OracleScene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from cadise_amd.api import Context
    return Context(0)


def _loss(img, target):
    return float(((img.astype(np.float64) - target) ** 2).mean())


def _dloss(img, target):
    return (2.0 * (img.astype(np.float64) - target) / img.size).astype(np.float32)


@pytest.mark.parametrize("max_depth", [4, 6])
def test_adjoint_matches_fp64_host_gradient(ctx, max_depth):
    from cadise_amd.api import Scene
    spec = scenes.textured_spheres(n_spheres=4, level=3, tex_res=16, width=48, height=40, spp=4, max_depth=max_depth)
    sc = Scene.from_spec(ctx, spec)
    O = OracleScene(spec)
    seed = 1234
    rng = np.random.default_rng(3)
    G = rng.normal(size=(spec.height, spec.width, 3)).astype(np.float32)
    film = sc.render(seed=seed, keep_samples=True)
    d_tex, d_light = sc.adjoint(G, seed=seed)
    views = sc.texture_views(d_tex)
    ofilm = O.render_linear(seed, OR_RNG_PHILOX, threads=8)
    close = np.isclose(film, ofilm, rtol=1e-4, atol=1e-6).all(-1).mean()
    print(f"[adjoint depth {max_depth}] film pixels matching the oracle: {close:.4f}", flush=True)
    og, olight = O.adjoint(seed, ofilm, G, threads=8)
    for name in ("albedo", "rough"):
        g = views[sc.tex_index[name]].astype(np.float64).reshape(og[name].shape)
        o = og[name]
        err = np.linalg.norm(g - o) / np.linalg.norm(o)
        big = np.abs(o) > 1e-2 * np.abs(o).max()
        rel = np.abs(g[big] - o[big]) / np.abs(o[big])
        print(f"[adjoint depth {max_depth}] {name}: rel L2 {err:.2e}, {big.sum()} large texels, "
              f"median rel {np.median(rel):.2e}, p99 {np.quantile(rel, 0.99):.2e}", flush=True)
        assert err < 1e-3, (name, err)
        assert np.quantile(rel, 0.99) < 1e-2, name
    np.testing.assert_allclose(d_light, olight, rtol=1e-3)


def test_streaming_adjoint_matches_passwise(ctx):
    """The streaming adjoint (default) and the pass-by-pass adjoint replay the
    same paths: gradients equal up to float summation order."""
    from cadise_amd.api import Scene
    spec = scenes.textured_spheres(n_spheres=4, level=3, tex_res=16, width=64, height=48, spp=4, max_depth=6)
    sc = Scene.from_spec(ctx, spec)
    G = np.random.default_rng(8).normal(size=(spec.height, spec.width, 3)).astype(np.float32)
    sc.render(seed=31, keep_samples=True)
    a_tex, a_light = sc.adjoint(G, seed=31)
    from cadise_amd.native import MPT_RENDER_PASSWISE, RenderParams
    import ctypes as C
    st = sc.stats()
    b_tex = np.zeros(st["tex_elements"], np.float32)
    b_light = np.zeros((st["nlights"], 3), np.float32)
    p = RenderParams(31, 0, 0, 0, 0, MPT_RENDER_PASSWISE)
    sc.ctx.check(sc.L.mpt_render_adjoint(sc.h, C.byref(p), G.ctypes.data, b_tex.ctypes.data, b_light.ctypes.data))
    assert np.linalg.norm(a_tex - b_tex) <= 1e-5 * np.linalg.norm(b_tex)
    np.testing.assert_allclose(a_light, b_light, rtol=1e-5)


@pytest.mark.timeout(600)
def test_adjoint_matches_finite_differences(ctx):
    """Differentiation workload at full size: 204,802 triangles, 256^2 albedo and
    roughness textures, 512^2, 16 spp, max depth 4, L2 loss to a target rendered
    with other textures and another seed."""
    from cadise_amd.api import Scene
    spec_t = scenes.textured_spheres()
    res = spec_t.textures[0].data.shape[0]
    spec = scenes.textured_spheres(albedo=scenes.smooth_noise_texture(res, 3, 0.2, 0.8, 99),
                                   roughness=scenes.smooth_noise_texture(res, 1, 0.2, 0.8, 98))
    sc_t = Scene.from_spec(ctx, spec_t)
    target = film_mean(sc_t.render(seed=777)).astype(np.float64)
    del sc_t
    sc = Scene.from_spec(ctx, spec)
    seed = 4242
    img = film_mean(sc.render(seed=seed, keep_samples=True))
    fwd_ms = sc.last_stats()["ms_total"]
    G = _dloss(img, target)
    d_tex, d_light = sc.adjoint(G, seed=seed)
    st = sc.last_stats()
    print(f"[fd] forward {fwd_ms:.1f} ms, adjoint {st['ms_adjoint']:.1f} ms GPU ({st['ms_total']:.1f} ms wall)",
          flush=True)
    views = sc.texture_views(d_tex)
    ia, ir = sc.tex_index["albedo"], sc.tex_index["rough"]
    ga, gr = views[ia], views[ir]
    alb = spec.textures[0].data.copy()
    rough = spec.textures[1].data.copy()
    eps = 1e-3

    def loss_with(a=None, r=None, le=None):
        if a is not None:
            sc.set_texture(ia, a)
        if r is not None:
            sc.set_texture(ir, r)
        if le is not None:
            sc.set_light_radiance(0, le)
        # pass by pass: its film sums in a fixed order, so the two renders of a
        # difference quotient differ only where the parameter acts
        v = _loss(film_mean(sc.render(seed=seed, passwise=True)), target)
        return v

    rng = np.random.default_rng(11)
    cand = np.argwhere(np.abs(ga) > 1e-3 * np.abs(ga).max())
    pick = cand[rng.choice(len(cand), 256, replace=False)]
    worst = 0.0
    t0 = time.time()
    for k, (y, x, c) in enumerate(pick):
        a = alb.copy()
        a[y, x, c] += eps
        lp = loss_with(a=a)
        a[y, x, c] -= 2 * eps
        lm = loss_with(a=a)
        fd = (lp - lm) / (2 * eps)
        rel = abs(fd - ga[y, x, c]) / abs(fd)
        worst = max(worst, rel)
        assert rel < 0.02, (k, (y, x, c), fd, float(ga[y, x, c]))
    sc.set_texture(ia, alb)
    print(f"[fd] 256 albedo texels: worst relative difference {worst:.2e} ({time.time() - t0:.1f} s)", flush=True)
    le = sc.light_radiance()[0].copy()
    le = np.array([10.0, 10.0, 10.0], np.float32)
    for c in range(3):
        x = le.copy()
        x[c] += eps
        lp = loss_with(le=x)
        x[c] -= 2 * eps
        lm = loss_with(le=x)
        fd = (lp - lm) / (2 * eps)
        rel = abs(fd - d_light[0, c]) / abs(fd)
        print(f"[fd] Le[{c}]: fd {fd:.6e} adjoint {d_light[0, c]:.6e} rel {rel:.2e}", flush=True)
        assert rel < 0.02
    sc.set_light_radiance(0, le)


@pytest.mark.timeout(600)
def test_adam_recovers_textures(ctx):
    """Inverse rendering at reduced size (256^2 image, 64^2 textures, 8 spp, 200
    Adam steps): the loss falls, the textures approach the truth and the
    emitted radiance approaches (10, 10, 10) from (7, 12, 8.5)."""
    from cadise_amd.inverse import differentiation_workload
    _, r = differentiation_workload(ctx, steps=200, log_every=50, n_spheres=6, level=4, tex_res=64, width=256,
                                    height=256, spp=8)
    print(f"[adam] loss {r.losses[0]:.4e} -> {np.mean(r.losses[-10:]):.4e}; texture RMSE seen {r.tex_rmse} "
          f"better-observed half {r.tex_rmse_well}; radiance {r.radiance}; {r.seconds:.1f} s, "
          f"forward {r.ms_forward:.1f} ms adjoint {r.ms_adjoint:.1f} ms", flush=True)
    assert np.mean(r.losses[-10:]) < 0.3 * np.mean(r.losses[:3])
    assert r.tex_rmse["albedo"][1] < 0.5 * r.tex_rmse["albedo"][0]
    assert r.tex_rmse["rough"][1] < 0.9 * r.tex_rmse["rough"][0]
    assert np.abs(np.array(r.radiance[1]) - 10.0).max() < 1.0


def test_adjoint_through_checkerboards_matches_finite_differences(ctx):
    """A picture texture read through checkerboards (the back wall: a checkerboard
    of the picture and a constant; the tall box: a checkerboard of a checkerboard
    that holds the picture): the adjoint sends each vertex's gradient to the
    picture texel its checkerboards resolve to. Checked by central differences of
    a fixed linear functional sum(G * image) on the pass-by-pass renderer (fixed
    summation order, common random numbers), max depth 4 (Russian roulette inert):
    bar 2% relative per texel."""
    import os
    from cadise_amd.api import Scene
    here = os.path.dirname(os.path.abspath(__file__))
    pics = np.load(os.path.join(here, "golden", "images.npz"))
    spec = scenes.checker_cornell(os.path.join(here, "golden", "images", "rgb8_filters.png"),
                                  pics["rgb8_filters.png"], width=96, height=96, spp=16, max_depth=4)
    sc = Scene.from_spec(ctx, spec)
    ip = sc.tex_index["pic"]
    pic = sc.get_texture(ip)
    G = np.random.default_rng(5).normal(size=(spec.height, spec.width, 3)).astype(np.float32)
    seed = 99
    sc.render(seed=seed, keep_samples=True)
    d_tex, _ = sc.adjoint(G, seed=seed)
    g = sc.texture_views(d_tex)[ip].reshape(pic.shape)

    def functional(p):
        sc.set_texture(ip, p)
        img = film_mean(sc.render(seed=seed, passwise=True)).astype(np.float64)
        return float((img * G).sum())

    rng = np.random.default_rng(12)
    cand = np.argwhere(np.abs(g) > 0.05 * np.abs(g).max())
    assert len(cand) >= 24
    eps = 1e-3
    worst = 0.0
    for (y, x, c) in cand[rng.choice(len(cand), 24, replace=False)]:
        p = pic.copy()
        p[y, x, c] += eps
        lp = functional(p)
        p[y, x, c] -= 2 * eps
        lm = functional(p)
        fd = (lp - lm) / (2 * eps)
        rel = abs(fd - g[y, x, c]) / abs(fd)
        worst = max(worst, rel)
        assert rel < 0.02, ((y, x, c), fd, float(g[y, x, c]))
    sc.set_texture(ip, pic)
    print(f"[fd checkerboard] 24 picture texels: worst relative difference {worst:.2e}", flush=True)


@pytest.mark.timeout(600)
def test_roughness_gradient_vs_finite_differences(ctx):
    """Roughness texels move the sampled GGX directions, so one sample set's
    central difference (eps 1e-3, common random numbers) and the adjoint differ
    by zero-mean noise; both estimate d E[F] / d theta for the linear functional
    F = sum(G * image) (fixed G). Over M = 16 independent seeds, per block k of
    16 x 16 roughness texels (the 32 blocks with the largest adjoint at a pilot
    seed), the seed means FD_k and AD_k are regressed through the origin:
    slope beta = sum FD AD / sum AD^2 with its standard error from the residuals.
    Bars: |beta - 1| <= 3 SE(beta) and SE(beta) <= 0.2 (the test pins the
    slope to within 0.6 at worst), and every block whose mean adjoint exceeds 3
    standard errors of the difference quotient's mean agrees with it in sign."""
    from cadise_amd.api import Scene
    spec = scenes.textured_spheres(max_depth=4)
    sc = Scene.from_spec(ctx, spec)
    ir = sc.tex_index["rough"]
    rough = sc.get_texture(ir).copy()
    res = rough.shape[0]
    G = np.random.default_rng(21).normal(size=(spec.height, spec.width, 3)).astype(np.float32)
    eps, B, K, M = 1e-3, 16, 32, 16

    def adjoint_blocks(seed):
        sc.render(seed=seed, keep_samples=True, readback=False)
        d_tex, _ = sc.adjoint(G, seed=seed)
        g = sc.texture_views(d_tex)[ir].reshape(res // B, B, res // B, B)
        return g.sum(axis=(1, 3)).astype(np.float64)

    def functional(r, seed):
        sc.set_texture(ir, r)
        img = film_mean(sc.render(seed=seed, passwise=True)).astype(np.float64)
        return float((img * G).sum())

    pilot = adjoint_blocks(5000)
    order = np.argsort(-np.abs(pilot).ravel())[:K]
    blocks = [divmod(int(i), res // B) for i in order]
    ad = np.zeros((M, K))
    fd = np.zeros((M, K))
    t0 = time.time()
    for m in range(M):
        seed = 6000 + 17 * m
        a = adjoint_blocks(seed)
        for k, (by, bx) in enumerate(blocks):
            ad[m, k] = a[by, bx]
            r = rough.copy()
            r[B * by:B * by + B, B * bx:B * bx + B] += eps
            lp = functional(r, seed)
            r[B * by:B * by + B, B * bx:B * bx + B] -= 2 * eps
            lm = functional(r, seed)
            fd[m, k] = (lp - lm) / (2 * eps)
    sc.set_texture(ir, rough)
    fdm, adm = fd.mean(0), ad.mean(0)
    fd_se = fd.std(0, ddof=1) / np.sqrt(M)
    beta = float((fdm * adm).sum() / (adm * adm).sum())
    resid = fdm - beta * adm
    se_beta = float(np.sqrt((resid ** 2).sum() / (K - 1) / (adm * adm).sum()))
    strong = np.abs(adm) > 3 * fd_se
    agree = np.sign(fdm[strong]) == np.sign(adm[strong])
    print(f"[roughness fd] {M} seeds x {K} blocks in {time.time() - t0:.1f} s: slope {beta:.3f} +- {se_beta:.3f}; "
          f"{int(strong.sum())} blocks with |adjoint| > 3 SE(fd), sign agreement {int(agree.sum())}/{int(strong.sum())}; "
          f"median |fd - ad| / SE(fd) {np.median(np.abs(fdm - adm) / fd_se):.2f}", flush=True)
    for k in range(K):
        print(f"  block {blocks[k]}: fd {fdm[k]:+.4e} +- {fd_se[k]:.2e}, adjoint {adm[k]:+.4e}")
    assert se_beta <= 0.2, se_beta
    assert abs(beta - 1.0) <= 3 * se_beta, (beta, se_beta)
    assert agree.all()

"""Inverse rendering with the path-replay adjoint: recover the albedo and
roughness textures and the emitted radiance of a scene from a target image by
Adam on the L2 image loss (gradients from mpt_render_adjoint through
torch_render.RenderFunction). Each step renders with a fresh seed (stochastic
gradients); parameters are projected onto their valid ranges after each step."""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import torch_render
from .api import Scene, film_mean


@dataclass
class Recovery:
    losses: list = field(default_factory=list)
    tex_rmse: dict = field(default_factory=dict)          # name -> [initial, final] over seen texels
    tex_rmse_all: dict = field(default_factory=dict)      # name -> [initial, final] over all texels
    tex_rmse_well: dict = field(default_factory=dict)     # name -> [initial, final] over the better-observed half
    radiance: list = field(default_factory=list)          # [initial, final, true]
    seconds: float = 0.0
    ms_forward: float = 0.0                               # mean GPU-side forward wall time per step
    ms_adjoint: float = 0.0                               # mean adjoint replay time per step (events)
    seen_fraction: dict = field(default_factory=dict)


def recover(scene: Scene, target: np.ndarray, init: dict, truth: dict, radiance_init, radiance_true, steps: int = 200,
            lr: dict | None = None, lr_radiance: float = 0.2, seed0: int = 1, log_every: int = 20,
            log=print, smooth: dict | None = None) -> Recovery:
    """init / truth: texture name -> H x W x C float32; lr: texture name -> Adam
    step size. Returns the loss curve and the texture errors (RMSE) before and
    after: over the texels that received gradient at any step ("seen"), over the
    better-observed half of those (accumulated |gradient| above the median) and
    over all texels. smooth: texture name -> weight of a squared-difference
    smoothness prior between neighbouring texels (added to the loss)."""
    names = sorted(init)
    lr = {**{"albedo": 0.02, "rough": 0.004}, **(lr or {})}
    params = [torch.tensor(init[n], dtype=torch.float32, requires_grad=True) for n in names]
    rad = torch.tensor(np.asarray(radiance_init, np.float32).reshape(1, 3), requires_grad=True)
    opt = torch.optim.Adam([{"params": [p], "lr": lr.get(n, 0.02)} for n, p in zip(names, params)] +
                           [{"params": [rad], "lr": lr_radiance}])
    tgt = torch.from_numpy(np.ascontiguousarray(target, np.float32))
    seen = {n: np.zeros(init[n].shape[:2], np.float64) for n in names}
    lo = {"albedo": 0.01, "rough": 0.05}
    out = Recovery()
    t0 = time.time()
    fwd_ms = adj_ms = 0.0
    for it in range(steps):
        opt.zero_grad()
        img = torch_render.render(scene, seed0 + it, names, rad, params)
        fwd_ms += scene.last_stats()["ms_total"]
        loss = ((img - tgt) ** 2).mean()
        total = loss
        for n, p in zip(names, params):
            lam = (smooth or {}).get(n, 0.0)
            if lam > 0:
                total = total + lam * (((p[1:] - p[:-1]) ** 2).mean() + ((p[:, 1:] - p[:, :-1]) ** 2).mean())
        total.backward()
        adj_ms += scene.last_stats()["ms_adjoint"]
        for n, p in zip(names, params):
            seen[n] += p.grad.abs().sum(-1).numpy()
        opt.step()
        with torch.no_grad():
            for n, p in zip(names, params):
                p.clamp_(lo.get(n, 0.01), 1.0)
            rad.clamp_(min=0.1)
        out.losses.append(float(loss.detach()))
        if log and (it % log_every == 0 or it == steps - 1):
            log(f"[adam] step {it}: loss {float(loss.detach()):.4e}")
    out.seconds = time.time() - t0
    out.ms_forward = fwd_ms / steps
    out.ms_adjoint = adj_ms / steps
    for n, p in zip(names, params):
        final = p.detach().numpy()
        m = seen[n] > 0
        well = seen[n] >= (np.median(seen[n][m]) if m.any() else np.inf)

        def rmse(a, mask):
            return float(np.sqrt(((a - truth[n]) ** 2)[mask].mean())) if mask.any() else float("nan")

        out.tex_rmse[n] = [rmse(init[n], m), rmse(final, m)]
        out.tex_rmse_well[n] = [rmse(init[n], well), rmse(final, well)]
        out.tex_rmse_all[n] = [rmse(init[n], np.ones_like(m)), rmse(final, np.ones_like(m))]
        out.seen_fraction[n] = float(m.mean())
    out.radiance = [list(map(float, np.asarray(radiance_init).ravel())), list(map(float, rad.detach().numpy().ravel())),
                    list(map(float, np.asarray(radiance_true).ravel()))]
    return out


def differentiation_workload(ctx, steps: int = 200, log=print, log_every: int = 20, lr: dict | None = None,
                             smooth: dict | None = None, **kw):
    """The differentiation workload: 204,802-triangle textured scene, 256^2 albedo
    and roughness textures, 512^2, 16 spp, max depth 4. Target rendered with the
    true parameters at 64 spp (another seed); Adam from other smooth-noise
    textures and a perturbed radiance, with a weak smoothness prior on both
    textures (the true ones are smooth noise; without it the weakly observed
    roughness texels drift under the stochastic gradients)."""
    from . import scenes
    spec_true = scenes.textured_spheres(**kw)
    res = spec_true.textures[0].data.shape[0]
    init = {"albedo": scenes.smooth_noise_texture(res, 3, 0.2, 0.8, 99),
            "rough": scenes.smooth_noise_texture(res, 1, 0.2, 0.8, 98)}
    truth = {"albedo": spec_true.textures[0].data, "rough": spec_true.textures[1].data.reshape(res, res, 1)}
    le_true = np.array([10.0, 10.0, 10.0], np.float32)
    le_init = np.array([7.0, 12.0, 8.5], np.float32)
    # target: the true parameters at 4x the samples (a scene of its own, so that
    # the stratified sampler's strata match the sample count)
    kt = dict(kw)
    kt["spp"] = 4 * kw.get("spp", spec_true.spp)
    tscene = Scene.from_spec(ctx, scenes.textured_spheres(**kt))
    target = film_mean(tscene.render(seed=0x7A5))
    del tscene
    scene = Scene.from_spec(ctx, scenes.textured_spheres(albedo=init["albedo"], roughness=init["rough"],
                                                         light_radiance=le_init, **kw))
    smooth = {"albedo": 0.1, "rough": 0.1} if smooth is None else smooth
    res_ = recover(scene, target, init, truth, le_init, le_true, steps=steps, log=log, log_every=log_every, lr=lr,
                   smooth=smooth)
    return scene, res_

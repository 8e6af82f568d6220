"""torch.autograd.Function around mpt_render / mpt_render_adjoint, for gradient
checks and texture optimisation. The forward pass renders on the GPU with the
current parameters (keeping each sample's radiance); the backward pass runs
the path-replay adjoint. Parameters: named textures (H x W x C tensors, the
scene's own layout) and the lights' emitted radiance (nlights x 3)."""
from __future__ import annotations

import numpy as np
import torch

from .api import Scene, film_mean


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, scene: Scene, seed: int, names: tuple, radiance: torch.Tensor, *textures: torch.Tensor):
        for name, t in zip(names, textures):
            scene.set_texture(scene.tex_index[name], t.detach().cpu().numpy().astype(np.float32))
        rad = radiance.detach().cpu().numpy().astype(np.float32)
        for li in range(rad.shape[0]):
            scene.set_light_radiance(li, rad[li])
        film = scene.render(seed=seed, keep_samples=True)
        ctx.scene, ctx.seed, ctx.names = scene, seed, names
        ctx.shapes = [t.shape for t in textures]
        return torch.from_numpy(film_mean(film))

    @staticmethod
    def backward(ctx, grad_image: torch.Tensor):
        scene = ctx.scene
        d_tex, d_light = scene.adjoint(grad_image.detach().cpu().numpy().astype(np.float32), seed=ctx.seed)
        views = scene.texture_views(d_tex)
        grads = [torch.from_numpy(views[scene.tex_index[n]].reshape(shape).copy())
                 for n, shape in zip(ctx.names, ctx.shapes)]
        return (None, None, None, torch.from_numpy(d_light.copy()), *grads)


def render(scene: Scene, seed: int, names, radiance: torch.Tensor, textures) -> torch.Tensor:
    """Linear image (H x W x 3) as a differentiable function of the textures
    named in `names` (tensors in `textures`) and the lights' radiance."""
    return RenderFunction.apply(scene, seed, tuple(names), radiance, *textures)

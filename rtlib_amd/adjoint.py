"""torch.autograd wrapper of the path-replay adjoint (rtpg_adjoint).

    img = render_differentiable(ctx, albedo, emission, spp=16, seed=0)
    loss = ((img - target) ** 2).sum(); loss.backward()

albedo: (n_triangles, 3) tensor of per-triangle Lambertian albedo; emission:
(3,) radiance shared by every light.  The forward pass uploads both and
renders on the GPU; the backward pass replays the same random streams and
returns d loss / d albedo and d loss / d emission.  Tensors may live on any
device: the data crosses to the library as host arrays.
"""
from __future__ import annotations

import numpy as np
import torch

from .tracer import FLAG_MIS, FLAG_NEE, Context


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(fctx, albedo, emission, rctx: Context, settings: dict):
        nl = rctx.scene_info()["lights"]
        rctx.set_albedo(albedo.detach().to("cpu", torch.float32).numpy())
        le = emission.detach().to("cpu", torch.float32).numpy().reshape(1, 3)
        rctx.set_emission(np.repeat(le, nl, axis=0))
        rctx.clear()
        rctx.render(**settings)
        img = rctx.image()
        fctx.rctx = rctx
        fctx.settings = settings
        fctx.dev = albedo.device
        return torch.from_numpy(img).to(albedo.device)

    @staticmethod
    def backward(fctx, grad_img):
        s = fctx.settings
        ga, gl = fctx.rctx.adjoint(grad_img.detach().to("cpu", torch.float32).numpy(), spp=s["spp"],
                                   flags=s["flags"], max_depth=s["max_depth"], seed=s["seed"],
                                   sample_offset=s.get("sample_offset", 0), bg=s.get("bg", (0.0, 0.0, 0.0)))
        return (torch.from_numpy(ga).to(fctx.dev), torch.from_numpy(gl.sum(0)).to(fctx.dev), None, None)


def render_differentiable(rctx: Context, albedo, emission, spp=16, flags=FLAG_NEE | FLAG_MIS, max_depth=6, seed=0,
                          sample_offset=0, bg=(0.0, 0.0, 0.0)):
    """Differentiable render; Russian roulette and RIS are off (see rtpg_adjoint)."""
    settings = dict(spp=spp, flags=flags, max_depth=max_depth, seed=seed, sample_offset=sample_offset, bg=bg)
    return _Render.apply(albedo, emission, rctx, settings)

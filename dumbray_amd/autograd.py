"""torch.autograd wrapper of the renderer: image = render(albedo, emission).

Forward runs dr_render_blocks (HIP); backward runs dr_backward (Renderer::
backward, HIP path replay) with the incoming image gradient.  The image is the
block-major frame (blocks*128, 4); use BlockBasedFrameBuffer.to_image for an
(H, W, 4) view.
"""
from __future__ import annotations

import torch

from .render import BlockBasedFrameBuffer, DumbRenderer


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, emission, renderer: DumbRenderer, iteration: int = 1):
        with torch.no_grad():
            renderer.albedo.copy_(albedo.detach().to(torch.float32))
            if renderer.nlights:
                renderer.emission.copy_(emission.detach().to(torch.float32))
            frame = BlockBasedFrameBuffer(renderer.width, renderer.height, albedo.device)
            renderer.render(frame, iteration)
        ctx.renderer = renderer
        ctx.iteration = iteration
        ctx.dtypes = (albedo.dtype, emission.dtype)
        return frame.data.clone()

    @staticmethod
    def backward(ctx, grad_frame):
        ga, ge = ctx.renderer.backward(grad_frame, ctx.iteration)
        return ga.to(ctx.dtypes[0]), ge.to(ctx.dtypes[1]), None, None


def render(albedo, emission, renderer: DumbRenderer, iteration: int = 1):
    return RenderFunction.apply(albedo, emission, renderer, iteration)

"""torch.autograd wrapper of the textured render and its path-replay adjoint.

The forward pass uploads the texture set (rt_scene_set_textures) and renders
the beauty film (rt_render); the backward pass runs rt_adjoint with the
incoming dLoss/dBeauty.  Both cross the C ABI with host arrays; torch only
carries the tensors.  The scene's texture set is left at the forward
texture, which the backward pass differentiates at.
"""
import numpy as np
import torch


class RenderTextures(torch.autograd.Function):
    """beauty (3, H, W) = render(texture (5, th, tw)); backward = rt_adjoint."""

    @staticmethod
    def forward(ctx, texture, renderer, width, height):
        tex = np.ascontiguousarray(texture.detach().cpu().numpy(), np.float32)
        renderer.scene.set_textures(tex)
        img = renderer.render(width, height).buffers.beauty
        ctx.renderer, ctx.size = renderer, (width, height)
        return torch.from_numpy(np.ascontiguousarray(img)).to(texture.device)

    @staticmethod
    def backward(ctx, grad_beauty):
        w, h = ctx.size
        g, _ = ctx.renderer.adjoint(w, h, grad_beauty.detach().cpu().numpy())
        return torch.from_numpy(g).to(grad_beauty.device), None, None, None


def render_textures(texture, renderer, width, height):
    return RenderTextures.apply(texture, renderer, width, height)

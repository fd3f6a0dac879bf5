"""Density gradients of the AD renderer behind ``torch.autograd``.

``render_density(density, scene, options, shape)`` renders the image (channel 0 of
``VolPathTracerAD.render``) with the given density grid of one heterogeneous medium.
Its backward runs ``dtrt_render_adjoint``, the path-replay adjoint. That is the
reverse mode of the same interior estimator (emission, surface and volume terms;
heterogeneous.cpp:40-141 weights) that the forward-mode lanes differentiate. So the
gradient over all N grid nodes costs one adjoint pass, instead of N/nder forward
passes. The backward pass uses its own random stream (``adjoint_seed``), so the
image and the gradient estimates are independent.
"""
import torch

from . import dtrt as G


class RenderDensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density, scene, options, shape, medium, adjoint_seed):
        G.set_density(scene, medium, density)
        out = torch.zeros(shape, dtype=torch.float32, device=density.device if density.is_cuda else "cuda")
        G.render_device(0, scene, options, out.data_ptr())
        ctx.scene, ctx.options, ctx.medium, ctx.adjoint_seed = scene, options, medium, adjoint_seed
        ctx.density_shape, ctx.density_device = density.shape, density.device
        return out[0].clone()

    @staticmethod
    def backward(ctx, grad_image):
        o = ctx.options
        opts = G.RenderOptions(ctx.adjoint_seed, o.spp, o.max_bounces, o.sppe, 0, True, o.mode, 0.0)
        g_img = grad_image.detach().float().contiguous()
        if not g_img.is_cuda:
            g_img = g_img.cuda()
        grad = torch.zeros(ctx.density_shape, dtype=torch.float32, device=g_img.device)
        G.render_adjoint(ctx.scene, opts, ctx.medium, g_img.data_ptr(), grad.data_ptr())
        torch.cuda.synchronize()
        return grad.to(ctx.density_device), None, None, None, None, None


def render_density(density, scene, options, shape, medium=0, adjoint_seed=None):
    """Image [H][W][3] as a differentiable function of ``density`` [rz][ry][rx]."""
    seed = options.seed + 0x9E3779B9 if adjoint_seed is None else adjoint_seed
    return RenderDensity.apply(density, scene, options, tuple(shape), int(medium), int(seed))

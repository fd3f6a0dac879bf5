"""Density gradients of the AD renderer behind ``torch.autograd``.

``render_density(density, scene, options, shape)`` renders the image (channel 0 of
``VolPathTracerAD.render``) with the given density grid of one heterogeneous medium.
Its backward runs ``dtrt_render_adjoint``, the path-replay adjoint. That is the
reverse mode of the same interior estimator (emission, surface and volume terms;
heterogeneous.cpp:40-141 weights) that the forward-mode lanes differentiate. So the
gradient over all N grid nodes costs one adjoint pass, instead of N/nder forward
passes. The backward pass uses its own random stream (``adjoint_seed``), so the
image and the gradient estimates are independent.

The reference's loss (``ADLossFunc.backward``, auto_grad.py:46-74) does not take the
plain image gradient: it sums ``pyramid_scale**i * <D^i dI/dp, clamp(D^i dL/dI)>``
over ``pyramid_level`` levels of a 2x2 box pyramid D. ``pyramid_weights`` turns
dL/dI into the one full-resolution image W with ``<dI/dp, W>`` equal to that sum
(W = sum_i s^i (D^T)^i clamp(D^i dL/dI)), so the adjoint returns, for every grid
node, what ``ADLossFunc`` would report for it as a forward-mode parameter.
"""
import math

import torch

from . import dtrt as G


def downsample(x):
    """2x2 box average of an [H][W][C] image; an odd last row / column is repeated
    first (auto_grad.py:10-16)."""
    if x.shape[0] % 2 == 1:
        x = torch.cat((x, x[-1:]), dim=0)
    if x.shape[1] % 2 == 1:
        x = torch.cat((x, x[:, -1:]), dim=1)
    return (x[0::2, 0::2] + x[1::2, 0::2] + x[0::2, 1::2] + x[1::2, 1::2]) * 0.25


def downsample_adjoint(y, h, w):
    """Transpose of ``downsample`` for an [h][w][C] input: each coarse pixel spreads a
    quarter of its value over its 2x2 block, and a repeated row / column folds back
    onto the last one."""
    up = 0.25 * y.repeat_interleave(2, dim=0).repeat_interleave(2, dim=1)
    if up.shape[0] > h:
        up = torch.cat((up[:h - 1], up[h - 1:h] + up[h:h + 1]), dim=0)
    if up.shape[1] > w:
        up = torch.cat((up[:, :w - 1], up[:, w - 1:w] + up[:, w:w + 1]), dim=1)
    return up


def pyramid_levels(h, w, pyramid_level):
    """Levels ADLossFunc.backward actually uses for an h x w image (auto_grad.py:49)."""
    return int(min(math.log(h, 2) + 1, math.log(w, 2) + 1, pyramid_level))


def pyramid_weights(grad_image, pyramid_level=1, pyramid_scale=4.0, clamping=0):
    """Full-resolution weight image W of ADLossFunc's multi-scale loss gradient:
    sum_i pyramid_scale**i (D^T)^i c(D^i grad_image), c = zeroing of entries beyond
    +-2 when ``clamping`` (auto_grad.py:58-63), else identity."""
    g = grad_image
    h, w = g.shape[0], g.shape[1]
    lvl = pyramid_levels(h, w, pyramid_level)
    chain = [g]
    for _ in range(1, lvl):
        chain.append(downsample(chain[-1]))
    acc = None
    for i in reversed(range(lvl)):
        gi = chain[i]
        if clamping:
            gi = torch.where((gi > 2.0) | (gi < -2.0), torch.zeros_like(gi), gi)
        term = (pyramid_scale ** i) * gi
        acc = term if acc is None else acc + term
        if i > 0:
            acc = downsample_adjoint(acc, chain[i - 1].shape[0], chain[i - 1].shape[1])
    return acc


class RenderDensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density, scene, options, shape, medium, adjoint_seed, pyramid):
        G.set_density(scene, medium, density)
        out = torch.zeros(shape, dtype=torch.float32, device=density.device if density.is_cuda else "cuda")
        G.render_device(0, scene, options, out.data_ptr())
        ctx.scene, ctx.options, ctx.medium, ctx.adjoint_seed = scene, options, medium, adjoint_seed
        ctx.density_shape, ctx.density_device, ctx.pyramid = density.shape, density.device, pyramid
        return out[0].clone()

    @staticmethod
    def backward(ctx, grad_image):
        o = ctx.options
        opts = G.RenderOptions(ctx.adjoint_seed, o.spp, o.max_bounces, o.sppe, 0, True, o.mode, 0.0)
        g_img = grad_image.detach().float()
        if not g_img.is_cuda:
            g_img = g_img.cuda()
        g_img = pyramid_weights(g_img, *ctx.pyramid).contiguous()
        grad = torch.zeros(ctx.density_shape, dtype=torch.float32, device=g_img.device)
        G.render_adjoint(ctx.scene, opts, ctx.medium, g_img.data_ptr(), grad.data_ptr())
        torch.cuda.synchronize()
        return grad.to(ctx.density_device), None, None, None, None, None, None


def render_density(density, scene, options, shape, medium=0, adjoint_seed=None, pyramid_level=1, pyramid_scale=4.0,
                   clamping=0):
    """Image [H][W][3] as a differentiable function of ``density`` [rz][ry][rx].

    ``pyramid_level``, ``pyramid_scale`` and ``clamping`` have ADLossFunc's meaning
    (auto_grad.py:25-30); the defaults give the plain gradient.
    """
    seed = options.seed + 0x9E3779B9 if adjoint_seed is None else adjoint_seed
    pyr = (int(pyramid_level), float(pyramid_scale), int(clamping))
    return RenderDensity.apply(density, scene, options, tuple(shape), int(medium), int(seed), pyr)

"""The reference's example scenes on the GPU, loaded by dtrt_amd.loaders from the
committed XML/OBJ files: 3-sigma parity with the oracle, and the glass_1 optimisation
(examples/glass_1/optimize.py) through SceneManager + ADLossFunc against the trace
the reference recorded (examples/glass_1/cvg_results/{param,loss}.txt)."""
import os

import numpy as np
import pytest

from dtrt_amd import loaders, scenes
from dtrt_amd.harness import SceneManager, SceneTransform
from tests.helpers import assert_3sigma, block_z, assert_parity

pytestmark = pytest.mark.gpu

SCENES = os.path.join(os.path.dirname(__file__), "golden", "scenes")
GLASS = os.path.join(SCENES, "glass_1", "scene.xml")
CBOX = os.path.join(SCENES, "shadow_1", "cbox.xml")


def _translated(path, shape, res, crop, vec):
    d, integ = loaders.load_mitsuba(path, nder=1)
    d.width = d.height = res
    d.crop = crop
    sm = SceneManager(d, [[SceneTransform("SHAPE_TRANSLATE", vec, shape)]])
    return sm.desc, integ


@pytest.mark.parametrize("name", ["glass_1", "shadow_1"])
def test_example_scene_ad_parity(gpu, oracle, name):
    if name == "glass_1":
        d, _ = _translated(GLASS, 0, 48, (12, 12, 24, 24), [1.0, 0, 0])
    else:
        d, _ = _translated(CBOX, 3, 48, (10, 18, 24, 24), [100.0, 0, 0])
    kw = dict(spp=16, sppe=512, sppse=4, max_bounces=4)
    sg, so = scenes.instantiate(gpu, d), scenes.instantiate(oracle, d)
    a = np.stack([scenes.render(gpu, d, seed=s, scene=sg, **kw) for s in range(16)]).astype(np.float64)
    b = np.stack([scenes.render(oracle, d, seed=100 + s, scene=so, **kw) for s in range(16)]).astype(np.float64)
    assert np.abs(a[:, 1]).mean() > 0
    assert_parity(a, b, 4, what=name, max_mean_z2=1.8)


def test_glass_1_primal_parity(gpu, oracle):
    d, integ = _translated(GLASS, 0, 48, (12, 12, 24, 24), [1.0, 0, 0])
    assert integ == "volpath"
    sg, so = scenes.instantiate(gpu, d), scenes.instantiate(oracle, d)
    a = np.stack([scenes.render_primal(gpu, d, seed=s, scene=sg, spp=32, max_bounces=4)[None] for s in range(16)])
    b = np.stack([scenes.render_primal(oracle, d, seed=50 + s, scene=so, spp=32, max_bounces=4)[None]
                  for s in range(16)])
    assert_parity(a.astype(np.float64), b.astype(np.float64), 4, what="glass_1 primal")


def test_glass_1_optimisation_follows_the_reference_trace(gpu):
    """First 6 Adam steps of examples/glass_1/optimize.py at its own settings (256^2,
    spp 1024, sppe 32768, sppse 10, 4 bounces, 10-level pyramid): the parameters and
    the loss stay within Monte Carlo error of the reference's recorded trace."""
    import torch
    from dtrt_amd.harness import ADLossFunc
    d, integ = loaders.load_mitsuba(GLASS, nder=4)
    opts = gpu.RenderOptions(0, 1024, 4, 32768, 10, True)
    sm = SceneManager(d, [[SceneTransform("SHAPE_TRANSLATE", [1.0, 0, 0], 0)],
                          [SceneTransform("SHAPE_TRANSLATE", [0, 1.0, 0], 0)],
                          [SceneTransform("SHAPE_TRANSLATE", [0, 0, 1.0], 0)],
                          [SceneTransform("BSDF_VARY", "roughdielectric", 0, 0.1, 0.0)]])
    sm.set_arguments([-5.0, -3.0, -4.0, -2.5])
    tgt = torch.zeros(d.out_shape(), device="cuda")
    gpu.render_device(loaders.integrator_for(gpu, integ).kind, sm.scene(gpu), opts, tgt.data_ptr())
    target = tgt.view(-1)[: d.height * d.width * 3].view(d.height, d.width, 3).clone()   # VolPathTracer image
    sm.reset()
    var = torch.zeros(4, requires_grad=True)
    optim = torch.optim.Adam([var], lr=2e-1)
    oor = torch.zeros(4)
    ref_p = np.loadtxt(os.path.join(SCENES, "glass_1", "param_trace.txt"), delimiter=",")
    ref_l = np.loadtxt(os.path.join(SCENES, "glass_1", "loss_trace.txt"), delimiter=",")
    for t in range(6):
        optim.zero_grad()
        opts.seed = t + 1
        img = ADLossFunc.apply(sm, gpu.VolPathTracerAD(), opts, var, oor, torch.tensor([10000.0] * 4), 10)
        loss = (img - target).pow(2).sum()
        loss.backward()
        optim.step()
        oor = torch.as_tensor(sm.set_arguments(var.detach().numpy()))
        v = var.detach().numpy()
        print("iter %d: %s loss %.3f (reference %s loss %.3f)" % (t, v[:3], loss.item(), ref_p[t, 1:], ref_l[t, 1]))
        assert abs(loss.item() / ref_l[t, 1] - 1) < 0.02
        assert np.abs(v[:3] - ref_p[t, 1:]).max() < 0.02


def test_adlossfunc_backward_equals_the_oracle_loop(gpu):
    """ADLossFunc.backward (pyramid weight image on the device) against the oracle's
    restatement of auto_grad.py:46-74 on the same tangent images, with clamping and one
    out-of-range variable."""
    import torch
    from dtrt_amd.harness import ADLossFunc
    from oracle import ad_loss
    d, _ = loaders.load_mitsuba(GLASS, nder=3)
    d.width, d.height = 40, 36
    sm = SceneManager(d, [[SceneTransform("SHAPE_TRANSLATE", [1.0, 0, 0], 0)],
                          [SceneTransform("SHAPE_TRANSLATE", [0, 1.0, 0], 0)],
                          [SceneTransform("BSDF_VARY", "roughdielectric", 0, 0.1, 0.0)]])
    var = torch.zeros(3, requires_grad=True)
    opts = gpu.RenderOptions(4, 64, 4, 2048, 4, True)
    oor, pen = torch.tensor([0.0, 0.0, 0.3]), torch.tensor([1.0, 1.0, 7.0])
    img = ADLossFunc.apply(sm, gpu.VolPathTracerAD(), opts, var, oor, pen, 4, 3.0, -1, 1)
    g = torch.linspace(-3, 3, img.numel()).reshape(img.shape)
    (img * g.cuda()).sum().backward()
    derivs = sm.last_image[1:].cpu().numpy().astype(np.float64)
    want = ad_loss.backward(derivs, g.numpy().astype(np.float64), 4, 3.0, oor.numpy(), pen.numpy(), clamping=1)
    assert np.abs(want[:2]).min() > 0 and want[2] == pytest.approx(2.1)
    assert np.allclose(var.grad.numpy(), want, rtol=1e-4, atol=1e-6 * np.abs(want).max())

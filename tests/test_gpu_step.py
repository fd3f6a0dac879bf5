"""The batched inverse-rendering step (dtrt_scene_step / capture_step / replay_step):
forward views, the L2 loss gradient, its contraction with every tangent lane, and the
density adjoint, run eagerly and as one replayed HIP graph."""
import numpy as np
import pytest

from dtrt_amd import scenes

pytestmark = pytest.mark.gpu
INTERIOR = 1 | 2 | 8


def _setup(gpu, nviews=3):
    import torch
    d = scenes.dielectric_medium(nder=8, res=48, spp=4, freq=10, grid_res=16)
    sc = scenes.instantiate(gpu, d)
    views = scenes.orbit_views(nviews, seed=3, res_choices=(32, 40, 48), spp_choices=(2, 4), radius=3.2)
    cams = [gpu.Camera(w, h, c2w.astype(np.float32), d.cam_to_ndc.astype(np.float32), d.clip_near, -1, None)
            for c2w, w, h, _ in views]
    fo = [gpu.RenderOptions(10 + i, spp, 6, 64, d.sppse, True, -1, d.ddist) for i, (_, _, _, spp) in enumerate(views)]
    ao = [gpu.RenderOptions(50 + i, spp, 6, 0, 0, True, INTERIOR, 0.0) for i, (_, _, _, spp) in enumerate(views)]
    rng = np.random.default_rng(4)
    tgts = [torch.tensor(rng.random((h, w, 3)) * 0.3, dtype=torch.float32, device="cuda") for _, w, h, _ in views]
    ims = [torch.zeros((d.nder + 1, h, w, 3), device="cuda") for _, w, h, _ in views]
    return d, sc, cams, fo, ao, tgts, ims


def _outs(d):
    import torch
    return (torch.zeros(d.media[0].density.shape, dtype=torch.float32, device="cuda"),
            torch.zeros(d.nder, dtype=torch.float64, device="cuda"), torch.zeros(1, dtype=torch.float64, device="cuda"))


def test_step_equals_its_parts(gpu):
    """Eager step = render_views, then (torch) the L2 gradient and the tangent
    contraction, then render_views_adjoint of that gradient, with the same seeds."""
    import torch
    d, sc, cams, fo, ao, tgts, ims = _setup(gpu)
    grad, pg, loss = _outs(d)
    gpu.step(sc, cams, fo, ims, tgts, 0, ao, grad, pg, loss)
    ref = [torch.zeros_like(x) for x in ims]
    gpu.render_views(sc, cams, fo, ref)
    dimgs = [2 * (r[0] - t) for r, t in zip(ref, tgts)]
    want_loss = sum(float(((r[0] - t).double() ** 2).sum()) for r, t in zip(ref, tgts))
    want_pg = sum((g.double()[None] * r[1:].double()).sum(dim=(1, 2, 3)) for g, r in zip(dimgs, ref))
    want_grad = torch.zeros_like(grad)
    gpu.render_views_adjoint(sc, cams, ao, 0, dimgs, want_grad)
    for a, b in zip(ims, ref):
        assert np.allclose(a.cpu().numpy(), b.cpu().numpy(), rtol=1e-5, atol=1e-6 * float(b.abs().max()))
    assert abs(loss.item() / want_loss - 1) < 1e-6
    assert torch.abs(want_pg).max() > 0
    assert np.allclose(pg.cpu().numpy(), want_pg.cpu().numpy(), rtol=1e-4, atol=1e-6 * float(want_pg.abs().max()))
    g, w = grad.cpu().numpy(), want_grad.cpu().numpy()
    assert np.abs(w).max() > 0
    assert np.allclose(g, w, rtol=1e-3, atol=1e-5 * np.abs(w).max())


def test_step_graph_replays_equal_eager_steps(gpu):
    """Three replays with new seeds, each followed by an eager step with the same seeds
    into separate outputs: images, loss, lane gradients and density gradient agree."""
    import torch
    d, sc, cams, fo, ao, tgts, ims = _setup(gpu)
    grad, pg, loss = _outs(d)
    gpu.capture_step(sc, cams, fo, ims, tgts, 0, ao, grad, pg, loss)
    for k in range(3):
        seeds = [1000 + 10 * k + i for i in range(2 * len(cams))]
        rc = gpu.replay_step(sc, seeds)
        assert rc in (0, 1)
        got = ([x.clone() for x in ims], grad.clone(), pg.clone(), loss.clone())
        fo2 = [gpu.RenderOptions(seeds[2 * i], o.spp, o.max_bounces, o.sppe, o.sppse, True, o.mode, o.ddistCoeff)
               for i, o in enumerate(fo)]
        ao2 = [gpu.RenderOptions(seeds[2 * i + 1], o.spp, o.max_bounces, 0, 0, True, INTERIOR, 0.0) for i, o in enumerate(ao)]
        ims2 = [torch.zeros_like(x) for x in ims]
        g2, pg2, l2 = _outs(d)
        gpu.step(sc, cams, fo2, ims2, tgts, 0, ao2, g2, pg2, l2)
        torch.cuda.synchronize()
        for a, b in zip(got[0], ims2):
            assert np.allclose(a.cpu().numpy(), b.cpu().numpy(), rtol=1e-5, atol=1e-6 * float(b.abs().max()))
        assert abs(got[3].item() / l2.item() - 1) < 1e-6
        assert np.allclose(got[2].cpu().numpy(), pg2.cpu().numpy(), rtol=1e-4, atol=1e-6 * float(pg2.abs().max()))
        w = g2.cpu().numpy()
        assert np.allclose(got[1].cpu().numpy(), w, rtol=1e-3, atol=1e-5 * np.abs(w).max())


@pytest.mark.parametrize("lane", [0, 1])
def test_shape_lane_gradient_matches_crn_finite_differences(gpu, lane):
    """The step's shape-parameter gradient is its loss gradient contracted with the
    tangent lanes (test_step_equals_its_parts). Here the lane itself is checked: for a
    fixed weight image w (the L2 gradient 2 (I - T) of a 2048-spp render against a
    target moved by 0.05 along the lane's field), sum(w * lane) agrees with central
    differences of sum(w * I), the vertices moved by +-h along the field through
    dtrt_scene_update_shape and rendered with the same seeds (common random numbers),
    within the joint sampling error of eight seeds per side. Lane 0 translates the
    sphere along x, lane 1 is a deformation field. With fixed weights the check carries
    no bias from the image's own noise.

    The sphere is black, an occluder inside the medium: its silhouette (primary edges),
    its shadows in the medium (volume edges) and the medium boundary it forms (Lins) carry
    the whole derivative. A lit sphere is not used: the estimator, as the reference's,
    has no boundary term for radiance steps inside a mesh's outline (shading-normal
    terminators where the view direction crosses the interpolated normal, facet edges of
    flat shading), so its shape derivative differs from finite differences there."""
    import torch
    d = scenes.dielectric_medium(nder=8, res=48, spp=256, freq=10, grid_res=16)
    d.bsdfs[2] = ("diffuse", (0.0, 0.0, 0.0))
    d.shapes[0].velocities[0] = np.array([1.0, 0.0, 0.0], np.float32)
    sc = scenes.instantiate(gpu, d)
    cam = [gpu.Camera(d.width, d.height, d.cam_to_world.astype(np.float32), d.cam_to_ndc.astype(np.float32),
                      d.clip_near, -1, None)]
    v0 = d.shapes[0].vertices.astype(np.float64)
    vel = d.shapes[0].velocities[lane].astype(np.float64)

    def render(seed, spp):
        img = torch.zeros((d.nder + 1, d.height, d.width, 3), device="cuda")
        gpu.render_views(sc, cam, [gpu.RenderOptions(seed, spp, 6, d.sppe, d.sppse, True, -1, d.ddist)], [img])
        return img.double()

    gpu.update_shape(sc, 0, v0 + 0.05 * vel)
    tgt = render(99, 2048)[0]
    gpu.update_shape(sc, 0, v0)
    w = 2 * (render(98, 2048)[0] - tgt)
    seeds = [7 + 10 * i for i in range(8)]
    ad = np.array([float((w * render(s, d.spp)[1 + lane]).sum()) for s in seeds])
    h = 0.02
    fds = []
    for s in seeds:
        gpu.update_shape(sc, 0, v0 + h * vel)
        up = float((w * render(s, d.spp)[0]).sum())
        gpu.update_shape(sc, 0, v0 - h * vel)
        dn = float((w * render(s, d.spp)[0]).sum())
        fds.append((up - dn) / (2 * h))
    gpu.update_shape(sc, 0, v0)
    fd, se_fd = np.mean(fds), np.std(fds, ddof=1) / np.sqrt(len(fds))
    g, se_g = ad.mean(), ad.std(ddof=1) / np.sqrt(len(ad))
    print("lane", lane, "AD", g, "+-", se_g, "FD", fd, "+-", se_fd)
    assert g < -4 * se_g, (g, se_g)
    assert abs(fd - g) < 4 * np.hypot(se_fd, se_g) + 0.05 * abs(g), (fd, se_fd, g, se_g)

"""Batched multi-view rendering (dtrt_render_views): one pass over views of uneven
resolution, crop, spp, sppe and seed gives, per view, the image of rendering that view
alone (same random streams, so equal up to float summation order)."""
import numpy as np
import pytest

from dtrt_amd import scenes

pytestmark = pytest.mark.gpu


def _camera(gpu, d, eye, res, crop=None, fov=45.0):
    w, h = res
    cam = gpu.Camera(w, h, scenes.look_at(eye, (0, 0.5, 0), (0, 1, 0)).astype(np.float32),
                     scenes.cam_to_ndc(fov).astype(np.float32), d.clip_near, -1, None)
    if crop:
        cam.set_rect(*crop)
    return cam


@pytest.mark.parametrize("nder", [1, 8])
def test_views_equal_single_renders(gpu, nder):
    """At nder 8 the AD samples run in the per-bounce refill kernel, whose waves mix
    samples of several pixels and views."""
    import torch
    d = scenes.ico_surface(nder=nder, res=32, spp=16)
    sc = scenes.instantiate(gpu, d)
    cams = [None, _camera(gpu, d, (1.5, 1.2, 3.0), (48, 40)), _camera(gpu, d, (-2.0, 2.0, 2.0), (64, 64), crop=(16, 8, 24, 32))]
    shapes = [(nder + 1, 32, 32, 3), (nder + 1, 40, 48, 3), (nder + 1, 32, 24, 3)]
    opts = [gpu.RenderOptions(1, 16, 5, 512, 4, True), gpu.RenderOptions(2, 40, 5, 256, 4, True),
            gpu.RenderOptions(3, 8, 5, -1, 4, True)]
    outs = [torch.zeros(s, device="cuda") for s in shapes]
    gpu.render_views(sc, cams, opts, outs)
    st = gpu.render_stats(sc)
    assert st["ad_samples"] == 32 * 32 * 16 + 48 * 40 * 40 + 24 * 32 * 8
    for v in range(3):
        single = torch.zeros(shapes[v], device="cuda")
        gpu.render_views(sc, [cams[v]], [opts[v]], [single])
        a, b = outs[v].cpu().numpy(), single.cpu().numpy()
        assert np.abs(a[0]).sum() > 0
        for lane in range(nder + 1):
            scale = np.abs(b[lane]).max() + 1e-12
            assert np.allclose(a[lane], b[lane], rtol=1e-4, atol=1e-5 * scale), (v, lane)
    # view 0 is the scene's own camera: dtrt_render gives the same image
    host = scenes.render(gpu, d, seed=1, scene=sc, sppe=512, sppse=4)
    assert np.allclose(host, outs[0].cpu().numpy(), rtol=1e-4, atol=1e-5 * np.abs(host).max())


def test_views_reject_mixed_settings(gpu):
    import torch
    d = scenes.ico_surface(nder=1, res=16, spp=4)
    sc = scenes.instantiate(gpu, d)
    outs = [torch.zeros((2, 16, 16, 3), device="cuda") for _ in range(2)]
    with pytest.raises(RuntimeError, match="share"):
        gpu.render_views(sc, None, [gpu.RenderOptions(0, 4, 5, 0, 4, True), gpu.RenderOptions(0, 4, 6, 0, 4, True)], outs)


def test_graph_replay_equals_eager(gpu):
    """Replays of the captured views with new seeds, with eager renders in between,
    each equal to an eager render with those seeds."""
    import torch
    d = scenes.ico_surface(nder=1, res=32, spp=16)
    sc = scenes.instantiate(gpu, d)
    cams = [None, _camera(gpu, d, (1.5, 1.2, 3.0), (48, 40))]
    shapes = [(2, 32, 32, 3), (2, 40, 48, 3)]
    opts = [gpu.RenderOptions(1, 16, 5, 512, 4, True), gpu.RenderOptions(2, 24, 5, 256, 4, True)]
    imgs = [torch.zeros(sh, device="cuda") for sh in shapes]
    gpu.capture_views(sc, cams, opts, imgs)
    for k, seeds in enumerate([(11, 12), (21, 22), (11, 12)]):
        assert gpu.replay_views(sc, seeds) == 0
        got = [im.cpu().numpy() for im in imgs]
        ref = [torch.zeros(sh, device="cuda") for sh in shapes]
        eo = [gpu.RenderOptions(sd, o.spp, o.max_bounces, o.sppe, o.sppse, True) for sd, o in zip(seeds, opts)]
        gpu.render_views(sc, cams, eo, ref)
        for v in range(2):
            r = ref[v].cpu().numpy()
            assert np.abs(r[0]).sum() > 0
            for lane in range(2):
                scale = np.abs(r[lane]).max() + 1e-12
                assert np.allclose(got[v][lane], r[lane], rtol=1e-4, atol=1e-5 * scale), (k, v, lane)


def test_graph_capture_refuses_overflowing_schedule(gpu):
    import torch
    d = scenes.ico_surface(nder=1, res=32, spp=16)
    sc = scenes.instantiate(gpu, d)
    gpu.set_queue_budget(sc, 64 * 128)
    with pytest.raises(RuntimeError, match="overflowed"):
        gpu.capture_views(sc, None, [gpu.RenderOptions(1, 16, 5, 0, 4, True)], [torch.zeros((2, 32, 32, 3), device="cuda")])
    gpu.set_queue_budget(sc, 0)


def test_scenes_render_together_as_alone(gpu):
    """dtrt_render_scenes over three different scenes (surfaces only, a medium, a
    dielectric in a medium), each on its own stream and host thread, gives every view
    the image of that scene's render_views alone; a scene listed twice is refused."""
    import torch
    ds = [scenes.ico_surface(nder=4, res=32, spp=16), scenes.cube_medium(nder=4, res=40, spp=8, grid_res=16),
          scenes.dielectric_medium(nder=4, res=32, spp=8, freq=8, grid_res=16)]
    scs = [scenes.instantiate(gpu, d) for d in ds]
    cams = [[None], [None, _camera(gpu, ds[1], (1.5, 1.2, 3.0), (48, 40))], [None]]
    opts = [[gpu.RenderOptions(1, 16, 5, 256, 4, True)],
            [gpu.RenderOptions(2, 8, 6, 256, 4, True), gpu.RenderOptions(3, 12, 6, 256, 4, True)],
            [gpu.RenderOptions(4, 8, 6, 256, 4, True)]]
    shapes = [[(5, 32, 32, 3)], [(5, 40, 40, 3), (5, 40, 48, 3)], [(5, 32, 32, 3)]]
    outs = [[torch.zeros(s, device="cuda") for s in sl] for sl in shapes]
    gpu.render_scenes(scs, cams, opts, outs)
    for i in range(3):
        alone = [torch.zeros(s, device="cuda") for s in shapes[i]]
        gpu.render_views(scs[i], cams[i], opts[i], alone)
        for a, b in zip(outs[i], alone):
            a, b = a.cpu().numpy(), b.cpu().numpy()
            assert np.abs(b[0]).sum() > 0
            for lane in range(5):
                scale = np.abs(b[lane]).max() + 1e-12
                assert np.allclose(a[lane], b[lane], rtol=1e-4, atol=1e-5 * scale), (i, lane)
    with pytest.raises(RuntimeError, match="twice"):
        gpu.render_scenes([scs[0], scs[0]], [cams[0], cams[0]], [opts[0], opts[0]], [outs[0], outs[0]])

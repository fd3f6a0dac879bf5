"""Parity at the benchmark configurations' own scales: the full 100,820-face dielectric
mesh in its 128^3 medium at depth 12 with 8 shape lanes (per-term 3-sigma against the
oracle), BVH closest hits on the 100K- and 1M-face meshes, the density adjoint at the
full 128^3 / 512^2 / 64 spp size, and the batched views of the 1M-face scene."""
import numpy as np
import pytest

from dtrt_amd import scenes
from tests.helpers import assert_parity, renders
from tests.test_gpu_adjoint import INTERIOR, _density_scene, _support_voxels
from tests.test_gpu_kernels import _global_face, _rays

pytestmark = pytest.mark.gpu

EMISSION, SURF_MAIN, SURF_BOUNDARY, VOL_MAIN, VOL_BOUNDARY1, VOL_BOUNDARY2 = 1, 2, 4, 8, 16, 32


@pytest.fixture(scope="module")
def dielectric_medium():
    """The dielectric configuration at full size: 100,820-face displaced sphere with
    shading normals, 128^3 fBm grid, depth 12, nder 8; a 24^2 crop across the sphere's
    right silhouette (the sphere spans x ~ 145..367 at y = 256)."""
    return scenes.dielectric_medium(nder=8, res=512, spp=8, crop=(355, 244, 24, 24))


@pytest.fixture(scope="module")
def displaced_views():
    return scenes.displaced_views()


@pytest.mark.parametrize("mode", [-1, EMISSION | SURF_MAIN | VOL_MAIN, VOL_BOUNDARY1, VOL_BOUNDARY2, 0])
def test_dielectric_medium_full_mesh_parity(gpu, oracle, dielectric_medium, mode):
    """Per EMode term (0: the primary-edge pass alone). The oracle side renders 4x the
    GPU's spp per image, so its estimate is the closer to converged of the two."""
    d = dielectric_medium
    assert sum(len(s.indices) for s in d.shapes) >= 100820 and d.media[0].density.shape == (128, 128, 128)
    assert d.max_bounces == 12 and d.nder == 8
    # the primary-edge pass alone is sparse (silhouette pixels only): more samples, 6x6 blocks
    sppe, block = (2048, 6) if mode == 0 else (256, 4)
    a = renders(gpu, d, range(16), mode=mode, sppe=sppe)
    b = renders(oracle, d, range(1000, 1012), mode=mode, sppe=4 * sppe, spp=4 * d.spp)
    assert np.abs(b[:, 1:].mean(0)).max() > 0   # the term has tangents in the crop
    assert_parity(a, b, block, what="dielectric_medium full mesh, mode %d" % mode, max_mean_z2=1.8)


@pytest.mark.parametrize("which", ["dielectric_medium", "displaced_views"])
def test_bvh_parity_on_config_meshes(gpu, oracle, dielectric_medium, displaced_views, which):
    """100K rays from around the displaced sphere towards it: the GPU's fp32 closest hit
    names the same face as the oracle's double-precision BVH for >= 99.99% of rays,
    and matched hits agree on t within 1e-5 relative."""
    d = dielectric_medium if which == "dielectric_medium" else displaced_views
    rng = np.random.default_rng(11)
    rays = _rays(100000, rng, -0.95, 0.95, -0.3, 0.3)   # from inside the medium cube
    sg = scenes.instantiate(gpu, d)
    so = scenes.instantiate(oracle, d)
    r32 = rays.astype(np.float32)
    fg, tg = gpu.trace(sg, r32)
    sh, tr, to = oracle.intersect(so, r32.astype(np.float64), brute=False)
    fo = _global_face(d, sh, tr)
    assert (sh == 0).mean() > 0.3   # a large share of the rays ends on the displaced sphere
    assert (fg == fo).mean() >= 0.9999, (fg != fo).sum()
    both = (fg == fo) & (fo >= 0)
    rel = np.abs(tg[both] - to[both]) / np.maximum(to[both], 1e-3)
    assert np.median(rel) <= 1e-6 and rel.max() <= 1e-5


def test_adjoint_full_size_equals_forward_voxel_lanes(gpu):
    """The density adjoint at the adjoint configuration's full size (128^3 grid,
    512^2, 64 spp, depth 10): per node, the reverse-mode gradient equals the forward-mode
    tangent of that node contracted with the loss gradient, for 16 nodes, same seed."""
    d = _density_scene(nder=16, res=512, crop=None, spp=64, grid_res=128)
    sc = scenes.instantiate(gpu, d)
    vox = _support_voxels(d.media[0].density, 16, seed=3)
    gpu.set_voxel_lanes(sc, 0, vox)
    img = scenes.render(gpu, d, seed=21, scene=sc, mode=INTERIOR, sppe=0, sppse=0).astype(np.float64)
    dimg = np.random.default_rng(0).normal(size=img.shape[1:]).astype(np.float32)
    expect = np.array([(dimg * img[1 + j]).sum() for j in range(16)])
    import torch
    grad = torch.zeros(d.media[0].density.shape, dtype=torch.float32, device="cuda")
    g = torch.as_tensor(dimg, device="cuda")
    gpu.render_adjoint(sc, gpu.RenderOptions(21, d.spp, d.max_bounces, 0, 0, True, INTERIOR), 0, g.data_ptr(),
                       grad.data_ptr())
    got = grad.cpu().numpy().astype(np.float64).ravel()[vox]
    assert np.count_nonzero(np.abs(expect) > 1e-3) >= 8
    assert np.allclose(got, expect, rtol=2e-3, atol=2e-4 * np.abs(expect).max()), np.c_[got, expect]
    assert np.count_nonzero(grad.cpu().numpy()) > 1000000   # most of the 2.1M nodes receive gradient


def test_adjoint_full_size_matches_crn_finite_differences(gpu):
    """CRN central differences of an image loss on 3 nodes of the full-size grid agree
    with the adjoint within their joint sampling error."""
    import torch
    d = _density_scene(nder=1, res=512, crop=None, spp=64, grid_res=128)
    sc = scenes.instantiate(gpu, d)
    dens = d.media[0].density.astype(np.float32)
    dimg = np.random.default_rng(1).normal(size=(d.height, d.width, 3)).astype(np.float32)
    g_t = torch.as_tensor(dimg, device="cuda")

    def adjoint(seed):
        grad = torch.zeros(dens.shape, dtype=torch.float32, device="cuda")
        gpu.render_adjoint(sc, gpu.RenderOptions(seed, d.spp, d.max_bounces, 0, 0, True, INTERIOR), 0,
                           g_t.data_ptr(), grad.data_ptr())
        return grad.cpu().numpy().astype(np.float64).ravel()

    grads = np.stack([adjoint(100 + s) for s in range(3)])
    g = grads.mean(0)
    cand = _support_voxels(dens, 2000, seed=5)
    vox = cand[np.argsort(-np.abs(g[cand]))[:3]]
    h = 0.05 * float(dens.max())

    def loss(rho, seed):
        gpu.set_density(sc, 0, rho)
        img = scenes.render(gpu, d, seed=seed, scene=sc, mode=INTERIOR, sppe=0, sppse=0)
        return float((dimg.astype(np.float64) * img[0]).sum())

    for v in vox:
        fds = []
        for s in range(3):
            up, dn = dens.copy(), dens.copy()
            up.ravel()[v] += h
            dn.ravel()[v] -= h
            fds.append((loss(up, 200 + s) - loss(dn, 200 + s)) / (2 * h))
        fd, se_fd = np.mean(fds), np.std(fds, ddof=1) / np.sqrt(3)
        se_g = grads[:, v].std(ddof=1) / np.sqrt(3)
        assert abs(fd - g[v]) < 4 * np.hypot(se_fd, se_g) + 0.05 * abs(g[v]), (v, fd, se_fd, g[v], se_g)
    gpu.set_density(sc, 0, dens)


def test_displaced_views_views_equal_single_renders(gpu, displaced_views):
    """The batched inverse-rendering scene (1,003,520 faces, 256^3 grid, nder 8): one
    dtrt_render_views call over three uneven views gives each view's single render."""
    import torch
    d = displaced_views
    sc = scenes.instantiate(gpu, d)
    views = scenes.orbit_views(32, seed=5, spp_choices=(16, 32, 64, 128, 256), radius=3.2)[:3]
    cams = [gpu.Camera(w, h, c2w.astype(np.float32), d.cam_to_ndc.astype(np.float32), d.clip_near, -1, None)
            for c2w, w, h, _ in views]
    opts = [gpu.RenderOptions(70 + i, 2, d.max_bounces, 64, d.sppse, True, -1, d.ddist) for i in range(3)]
    outs = [torch.zeros((d.nder + 1, h, w, 3), device="cuda") for _, w, h, _ in views]
    gpu.render_views(sc, cams, opts, outs)
    for i in range(3):
        one = torch.zeros_like(outs[i])
        gpu.render_views(sc, [cams[i]], [opts[i]], [one])
        a, b = outs[i].cpu().numpy(), one.cpu().numpy()
        assert np.abs(b[0]).max() > 0 and np.abs(b[1:]).max() > 0
        assert np.allclose(a, b, rtol=1e-4, atol=1e-5 * np.abs(b).max()), i

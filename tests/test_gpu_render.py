"""Render-level parity of the HIP path: closed forms on the GPU and per-term 3-sigma
agreement with the oracle on small versions of the benchmark scenes."""
import math

import numpy as np
import pytest

from dtrt_amd import scenes
from tests import scene_fixtures as F
from tests.helpers import assert_3sigma, block_z, renders, sum_z

pytestmark = pytest.mark.gpu

# EMode bits (integratorAD.h:12-24)
EMISSION, SURF_MAIN, SURF_BOUNDARY, VOL_MAIN, VOL_BOUNDARY1, VOL_BOUNDARY2 = 1, 2, 4, 8, 16, 32


@pytest.mark.parametrize("hetero", [False, True])
def test_furnace(gpu, hetero):
    d = F.furnace(nder=1, heterogeneous=hetero, res=32, spp=64)
    img = scenes.render(gpu, d, seed=3).astype(np.float64)
    v, g = img[0, ..., 0], img[1, ..., 0]
    assert abs(v.mean() - 1.0) < 4 * v.std() / math.sqrt(v.size) + 1e-3
    assert abs(g.mean()) < 4 * g.std() / math.sqrt(g.size) + 1e-3


def test_direct_lighting_matches_lambert_polygon(gpu):
    d, lv, rho, Le = F.direct_lighting(res=32, spp=1024)
    img = scenes.render(gpu, d, seed=1)
    ref = F.direct_expected(d, lv, rho, Le)
    assert abs(img[0, ..., 0].mean() / ref.mean() - 1) < 3e-3
    assert np.abs(img[0, ..., 0] - ref).max() < 0.1 * ref.max()


def test_absorbing_slab(gpu):
    d, L, dL = F.absorbing_slab(heterogeneous=False)
    img = scenes.render(gpu, d, seed=0)
    assert np.allclose(img[0, ..., 0], L, rtol=1e-4, atol=1e-5)
    assert np.allclose(img[1, ..., 0], dL, rtol=1e-3, atol=1e-4)
    d, L, dL = F.absorbing_slab(heterogeneous=True, spp=2048)
    img = scenes.render(gpu, d, seed=0)
    assert abs(img[0, ..., 0].mean() / L.mean() - 1) < 3e-3
    assert abs(img[1, ..., 0].mean() / dL.mean() - 1) < 1e-2


def test_primary_edges_translating_emitter(gpu):
    d, half = F.translating_emitter(res=64)
    g = scenes.render(gpu, d, seed=0)[1, ..., 0].astype(np.float64)
    w = d.width // 2
    assert abs(g[:, :w].sum() / -half - 1) < 5e-3
    assert abs(g[:, w:].sum() / half - 1) < 5e-3


def test_interface_edges_match_lambert_derivative(gpu):
    d, dE = F.translating_light(res=16, spp=4096)
    g = scenes.render(gpu, d, seed=0)[1, ..., 0]
    assert np.corrcoef(g.ravel(), dE.ravel())[0, 1] > 0.999
    assert np.abs(g - dE).max() < 0.03 * np.abs(dE).max()


def _crn_fd(gpu, make, h, seed=5):
    base = scenes.render(gpu, make(0.0), seed=seed).astype(np.float64)
    up = scenes.render(gpu, make(+h), seed=seed).astype(np.float64)
    dn = scenes.render(gpu, make(-h), seed=seed).astype(np.float64)
    return base[1], (up[0] - dn[0]) / (2 * h)


def test_reflectance_tangent_equals_crn_fd(gpu):
    ad, fd = _crn_fd(gpu, lambda e: F.ico_surface_reflectance_lane(res=32, spp=16, rho=0.8 + e), 1e-2)
    assert np.abs(fd).max() > 1e-2
    assert np.allclose(ad, fd, rtol=2e-3, atol=2e-3)


def test_albedo_tangent_equals_crn_fd(gpu):
    ad, fd = _crn_fd(gpu, lambda e: F.cube_medium_albedo_lane(res=24, spp=16, albedo=0.8 + e), 1e-2)
    assert np.abs(fd).max() > 1e-2
    assert np.allclose(ad, fd, rtol=5e-3, atol=5e-3)


def _parity(gpu, oracle, d, mode=-1, seeds=16, block=4, **kw):
    """Block z-scores GPU vs oracle; the image sums of every channel must agree too."""
    a = renders(gpu, d, range(seeds), mode=mode, **kw)
    b = renders(oracle, d, range(1000, 1000 + seeds), mode=mode, **kw)
    zs = sum_z(a, b)
    assert (np.abs(zs) <= 4).all(), "%s mode %d: image-sum z-scores %s" % (d.name, mode, np.array2string(zs, precision=2))
    return block_z(a, b, block)


@pytest.mark.parametrize("mode", [-1, EMISSION | SURF_MAIN, SURF_BOUNDARY])
def test_ico_surface_parity(gpu, oracle, mode):
    d = scenes.ico_surface(nder=1, res=24, spp=32)
    z = _parity(gpu, oracle, d, mode=mode, sppe=1024)
    assert_3sigma(z, what="ico_surface mode %d" % mode)


@pytest.mark.parametrize("mode", [-1, EMISSION | VOL_MAIN, VOL_BOUNDARY1, VOL_BOUNDARY2])
def test_cube_medium_parity(gpu, oracle, mode):
    d = scenes.cube_medium(nder=3, res=64, spp=32, grid_res=32, crop=(20, 20, 24, 24))
    z = _parity(gpu, oracle, d, mode=mode, sppe=256)
    assert_3sigma(z, what="cube_medium mode %d" % mode)


@pytest.mark.parametrize("move", ["light", "box"])
def test_medium_box_parity(gpu, oracle, move):
    d = F.medium_box(move=move, spp=64, res=12)
    if move == "box":
        d.bsdfs = [("null",), ("null",)]
    z = _parity(gpu, oracle, d, seeds=16, block=3)
    assert_3sigma(z, what="medium box " + move)


def test_dielectric_medium_small_parity(gpu, oracle):
    """Rough dielectric displaced sphere with shading normals in a medium, 8 lanes."""
    d = scenes.dielectric_medium(nder=8, res=48, spp=16, freq=10, grid_res=16, crop=(14, 14, 20, 20))
    z = _parity(gpu, oracle, d, seeds=12, block=4, sppe=128, max_bounces=6)
    assert_3sigma(z, what="dielectric_medium small", max_mean_z2=1.8)


def test_render_into_device_tensor(gpu):
    import torch
    d = scenes.ico_surface(nder=1, res=16, spp=8)
    sc = scenes.instantiate(gpu, d)
    host = scenes.render(gpu, d, seed=2, scene=sc)
    out = torch.zeros(d.out_shape(), dtype=torch.float32, device="cuda")
    gpu.render_device(0, sc, gpu.RenderOptions(2, d.spp, d.max_bounces, d.sppe, d.sppse, True), out.data_ptr())
    torch.cuda.synchronize()
    dev = out.cpu().numpy()
    assert np.allclose(host[0], dev[0], rtol=1e-5, atol=1e-6)
    st = gpu.render_stats(sc)
    assert st["ad_samples"] == 16 * 16 * 8 and st["children"] > 0


@pytest.mark.parametrize("nder", [3, 8])
def test_queue_overflow_replay_is_exact(gpu, nder):
    """A queue far smaller than one chunk's children forces many overflow/replay rounds;
    every child and every sample must still be counted exactly once. At nder 8 the AD
    samples run in the per-bounce refill kernel (k_ad_refill), at nder 3 in k_ad_paths."""
    d = scenes.cube_medium(nder=nder, res=64, spp=64, grid_res=16, crop=(16, 16, 32, 32))
    sc = scenes.instantiate(gpu, d)
    ref = scenes.render(gpu, d, seed=4, scene=sc, sppe=64).astype(np.float64)
    st_ref = gpu.render_stats(sc)
    assert st_ref["replayed"] == 0
    gpu.set_queue_budget(sc, 4096 * 128)
    try:
        got = scenes.render(gpu, d, seed=4, scene=sc, sppe=64).astype(np.float64)
        st = gpu.render_stats(sc)
    finally:
        gpu.set_queue_budget(sc, 0)
    assert st["replayed"] > 0 and st["children"] == st_ref["children"]
    scale = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
    for lane in range(ref.shape[0]):
        assert np.allclose(got[lane], ref[lane], rtol=1e-3, atol=1e-4 * scale[lane]), lane


def test_value_lane_independent_of_lane_count(gpu):
    """Channel 0 does not depend on how many tangent lanes ride along: with no motion
    tangents every lane count makes the same sampling decisions, so the same seed
    gives the same image at every compiled lane count (all kernel instantiations)."""
    ref = None
    for n in (1, 2, 3, 4, 8, 16):
        d = scenes.cube_medium(nder=n, res=64, spp=8, grid_res=24, crop=(16, 16, 32, 32))
        for s in d.shapes:
            if s.velocities is not None:
                s.velocities = np.zeros_like(s.velocities)
        d.media[0].trans_vel = np.zeros_like(d.media[0].trans_vel)
        img = scenes.render(gpu, d, seed=11, sppe=256)[0].astype(np.float64)
        assert img.sum() > 0
        if ref is None:
            ref = img
        else:
            assert np.allclose(img, ref, rtol=1e-5, atol=1e-6), n


@pytest.mark.parametrize("mode", [-1, EMISSION | SURF_MAIN | VOL_MAIN, SURF_BOUNDARY | VOL_BOUNDARY1 | VOL_BOUNDARY2])
def test_feature_sweep_parity(gpu, oracle, mode):
    """Phong, rough dielectric, AreaLightEx, homogeneous HG medium, camera motion and
    Lins with ddistCoeff > 0, each with its own tangents."""
    d = F.feature_sweep()
    z = _parity(gpu, oracle, d, mode=mode, seeds=16, block=4)
    assert_3sigma(z, what="feature sweep mode %d" % mode)


def test_rotating_grid_parity(gpu, oracle):
    d = F.rotating_grid()
    z = _parity(gpu, oracle, d, seeds=16, block=4)
    assert_3sigma(z, what="rotating grid")


def test_autotune_keeps_results_and_picks_listed_settings(gpu):
    """dtrt_scene_autotune only changes launch settings: the same seed renders the same
    image (up to float summation order) before and after tuning."""
    d = scenes.cube_medium(nder=3, res=64, spp=32, grid_res=16)
    sc = scenes.instantiate(gpu, d)
    before = scenes.render(gpu, d, seed=9, scene=sc).astype(np.float64)
    opts = gpu.RenderOptions(3, d.spp, d.max_bounces, d.sppe, d.sppse, True, -1, d.ddist)
    ms, st = gpu.autotune(sc, opts)
    assert ms > 0
    assert st["child_blocks_per_cu"] in (16, 24, 32, 48) and st["fill_percent"] in (50, 75, 90)
    assert st["max_chunk"] in (1 << 21, 1 << 22, 1 << 23)
    assert st["queue_budget_bytes"] in (1 << 30, 2 << 30, 4 << 30, 6 << 30)
    assert st["ad_workgroup"] in (64, 128, 256) and st["child_workgroup"] in (64, 128, 256)
    after = scenes.render(gpu, d, seed=9, scene=sc).astype(np.float64)
    scale = np.abs(before).reshape(before.shape[0], -1).max(axis=1)
    for lane in range(before.shape[0]):
        assert np.allclose(after[lane], before[lane], rtol=1e-3, atol=1e-4 * scale[lane]), lane


@pytest.mark.parametrize("groups", [(128, 256), (256, 128), (256, 64)])
def test_workgroup_sizes_render_the_same_image(gpu, groups):
    """The AD-path and child kernels at other work-group sizes (the traversal stack in
    LDS follows the block size) and a small queue budget: same seed, same image."""
    d = scenes.cube_medium(nder=3, res=48, spp=32, grid_res=16)
    sc = scenes.instantiate(gpu, d)
    ref = scenes.render(gpu, d, seed=5, scene=sc).astype(np.float64)
    gpu.set_launch_settings(sc, ad_workgroup=groups[0], child_workgroup=groups[1], queue_budget_bytes=1 << 24)
    got = scenes.render(gpu, d, seed=5, scene=sc).astype(np.float64)
    assert gpu.render_stats(sc)["replayed"] > 0   # the small queue overflowed and was replayed
    scale = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
    for lane in range(ref.shape[0]):
        assert np.allclose(got[lane], ref[lane], rtol=1e-3, atol=1e-4 * scale[lane]), lane

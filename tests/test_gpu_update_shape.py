"""dtrt_scene_update_shape: new vertex positions for one shape, with the BVH refitted on
the GPU instead of rebuilt. The updated scene must render like a scene built from scratch
with the new geometry (same seeds, so the same paths), trace like the oracle on it, and
carry the new geometry into the light and edge samplers."""
import copy

import numpy as np
import pytest

from dtrt_amd import scenes
from tests.test_gpu_kernels import _global_face, _rays

pytestmark = pytest.mark.gpu


def _moved(d, si, scale, shift):
    m = copy.deepcopy(d)
    v = m.shapes[si].vertices.astype(np.float64)
    c = v.mean(0)
    m.shapes[si].vertices = ((v - c) * scale + c + np.asarray(shift)).astype(np.float32)
    return m


def _same_images(a, b):
    """Equal paths: every value agrees to float summation order, except where a ray grazes
    a shared edge and the two trees (refitted, rebuilt) pick different faces."""
    tol = 1e-4 * np.abs(b).max() + 1e-6
    return (np.abs(a - b) > tol).mean() < 1e-3


@pytest.mark.parametrize("which", ["sphere", "light"])
def test_updated_scene_renders_like_a_rebuilt_one(gpu, which):
    d = scenes.ico_surface(nder=1, res=32, spp=16)
    si = 1 if which == "sphere" else 2   # ground, icosphere, light quad
    new = _moved(d, si, 1.15, (0.12, 0.05, -0.08))
    sc = scenes.instantiate(gpu, d)
    scenes.render(gpu, d, seed=3, scene=sc)   # renders once with the old geometry
    s = new.shapes[si]
    gpu.update_shape(sc, si, s.vertices, s.normals,
                     None if s.velocities is None else np.asarray(s.velocities, np.float32).reshape(-1, 3))
    a = scenes.render(gpu, new, seed=7, scene=sc)
    b = scenes.render(gpu, new, seed=7)
    assert np.abs(b[0]).max() > 0 and np.abs(b[1]).max() > 0
    assert _same_images(a, b)
    old = scenes.render(gpu, d, seed=7)
    assert not _same_images(old, b)   # the change is visible


def test_refitted_bvh_traces_like_the_oracle(gpu, oracle):
    d = scenes.dielectric_medium(nder=1, res=16, spp=1, freq=30, grid_res=8)
    new = _moved(d, 0, 0.9, (0.05, -0.1, 0.02))
    sc = scenes.instantiate(gpu, d)
    gpu.update_shape(sc, 0, new.shapes[0].vertices, new.shapes[0].normals)
    so = scenes.instantiate(oracle, new)
    rays = _rays(100000, np.random.default_rng(9), -0.95, 0.95, -0.3, 0.3).astype(np.float32)
    fg, tg = gpu.trace(sc, rays)
    sh, tr, to = oracle.intersect(so, rays.astype(np.float64), brute=False)
    fo = _global_face(new, sh, tr)
    assert (sh == 0).mean() > 0.3
    assert (fg == fo).mean() >= 0.9999
    both = (fg == fo) & (fo >= 0)
    rel = np.abs(tg[both] - to[both]) / np.maximum(to[both], 1e-3)
    assert rel.max() <= 1e-5


def test_update_releases_captured_graphs(gpu):
    import torch
    d = scenes.ico_surface(nder=1, res=16, spp=4)
    sc = scenes.instantiate(gpu, d)
    out = [torch.zeros(d.out_shape(), device="cuda")]
    opts = [gpu.RenderOptions(1, 4, d.max_bounces, d.sppe, d.sppse, True, -1, d.ddist)]
    gpu.capture_views(sc, None, opts, out)
    assert gpu.replay_views(sc, [2]) == 0
    gpu.update_shape(sc, 1, d.shapes[1].vertices)
    with pytest.raises(RuntimeError, match="no captured"):
        gpu.replay_views(sc, [3])

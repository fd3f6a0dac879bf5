"""The BVH in LDS. A launch copies the whole tree (small scenes) or its breadth-first
top levels (the N >= 8 kernels on large meshes) into LDS after the traversal stacks,
which are sized to the tree's depth. Traversal visits the same nodes in the same order
wherever they are read from, so an image rendered with the tree kept in HBM only
(dtrt_test_set_tree_in_hbm) is the same up to the order of the accumulator sums."""
import numpy as np
import pytest

from dtrt_amd import scenes
from tests.test_gpu_precision import precision

pytestmark = pytest.mark.gpu


def _same_image(gpu, d, rtol, seed=3, **kw):
    sc = scenes.instantiate(gpu, d)
    lds = scenes.render(gpu, d, seed=seed, scene=sc, **kw).astype(np.float64)
    gpu.set_tree_in_hbm(sc, True)
    hbm = scenes.render(gpu, d, seed=seed, scene=sc, **kw).astype(np.float64)
    gpu.set_tree_in_hbm(sc, False)
    again = scenes.render(gpu, d, seed=seed, scene=sc, **kw).astype(np.float64)
    assert np.abs(lds[0]).max() > 0
    scale = np.abs(lds).reshape(lds.shape[0], -1).max(axis=1)
    for lane in range(lds.shape[0]):
        assert np.allclose(hbm[lane], lds[lane], rtol=rtol, atol=rtol * 1e-2 * scale[lane]), lane
        assert np.allclose(again[lane], lds[lane], rtol=rtol, atol=rtol * 1e-2 * scale[lane]), lane


@pytest.mark.parametrize("bits", [64, 32])
def test_whole_tree_in_lds_renders_as_from_hbm(gpu, bits):
    """The benchmark scene's 194-node tree fits whole: the AD-path kernel (64-lane
    groups) and the child kernel (256-lane groups) read nodes and triangles from LDS."""
    with precision(gpu, bits):
        d = scenes.cube_medium(nder=3, res=48, spp=32, grid_res=16)
        assert gpu.scene_info(scenes.instantiate(gpu, d))["nodes"] == 194
        _same_image(gpu, d, rtol=1e-5 if bits == 64 else 1e-3)


def test_tree_top_in_lds_renders_as_from_hbm(gpu):
    """The 100,820-face dielectric mesh at nder 8: the refill and child kernels keep
    the first few hundred nodes (the top of the tree) in LDS and read the rest from HBM."""
    d = scenes.dielectric_medium(nder=8, res=512, spp=4, crop=(355, 244, 24, 24))
    assert sum(len(s.indices) for s in d.shapes) >= 100820 and d.nder == 8
    _same_image(gpu, d, rtol=1e-5)


def test_deep_tree_renders_as_from_hbm(gpu):
    """A comb of 2,000 slivers spaced geometrically along x, in front of the icosphere
    scene, makes a 24-level tree: the traversal stacks, sized to the tree's depth, then
    hold 24 entries per lane, and the top of the tree follows them in LDS (nder 8). A
    stack deeper than its size would overwrite the tree's copy."""
    base = scenes.ico_surface(nder=8, res=48, spp=16)
    n = 2000
    x = (np.cumsum(1.02 ** np.arange(n)) * 1e-3).astype(np.float32) - 1.2
    v = np.zeros((3 * n, 3), np.float32)
    v[0::3, 0], v[1::3, 0], v[2::3, 0] = x, x + 1e-3, x
    v[:, 1] = 0.1
    v[2::3, 1] = 0.6
    v[:, 2] = 0.8
    comb = scenes.ShapeDesc(v, np.arange(3 * n, dtype=np.int32).reshape(-1, 3), bsdf_id=1)
    base.shapes = list(base.shapes) + [comb]
    assert gpu.scene_info(scenes.instantiate(gpu, base))["nodes"] == 1931
    _same_image(gpu, base, rtol=1e-5)

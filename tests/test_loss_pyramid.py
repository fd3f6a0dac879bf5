"""The adjoint's loss-side weighting against ADLossFunc.backward (auto_grad.py:46-74).

The oracle restates the reference loop (downsample both images level by level);
the product folds the pyramid into one weight image with the transpose of the
downsampling. Their agreement for every parameter is the identity
<D^i a, c(D^i g)> = <a, (D^T)^i c(D^i g)>, so it also checks the transpose.
"""
import numpy as np
import pytest
import torch

from dtrt_amd.adjoint import downsample, downsample_adjoint, pyramid_levels, pyramid_weights
from oracle import ad_loss


def test_downsample_known_values():
    x = np.arange(9, dtype=np.float64).reshape(3, 3, 1)
    want = np.array([[2.0, 3.5], [6.5, 8.0]]).reshape(2, 2, 1)   # last row / column repeated, then 2x2 means
    assert np.array_equal(ad_loss.downsample(x), want)
    assert np.array_equal(downsample(torch.tensor(x)).numpy(), want)
    c = np.full((10, 7, 3), 1.25)
    assert np.array_equal(ad_loss.downsample(c), np.full((5, 4, 3), 1.25))


@pytest.mark.parametrize("h,w", [(8, 8), (7, 5), (13, 10), (1, 6)])
def test_downsample_adjoint_is_the_transpose(h, w):
    rng = np.random.default_rng(h * 31 + w)
    x = torch.tensor(rng.normal(size=(h, w, 3)))
    y = torch.tensor(rng.normal(size=((h + 1) // 2, (w + 1) // 2, 3)))
    lhs = (downsample(x) * y).sum().item()
    rhs = (x * downsample_adjoint(y, h, w)).sum().item()
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_level_count_follows_the_image_size():
    assert pyramid_levels(512, 512, 20) == 10
    assert pyramid_levels(24, 40, 9) == 5
    assert pyramid_levels(24, 40, 2) == 2
    assert pyramid_levels(1, 40, 3) == 1


@pytest.mark.parametrize("h,w", [(16, 16), (13, 10), (24, 9)])
@pytest.mark.parametrize("level,scale,clamping", [(1, 4.0, 0), (3, 4.0, 0), (4, 2.0, 1), (9, 4.0, 1)])
def test_weight_image_reproduces_adlossfunc_backward(h, w, level, scale, clamping):
    rng = np.random.default_rng(h + 7 * w + 101 * level)
    nder = 4
    derivs = rng.normal(size=(nder, h, w, 3))
    g = 1.6 * rng.normal(size=(h, w, 3))          # some entries beyond +-2 (clamped when asked)
    want = ad_loss.backward(derivs, g, level, scale, clamping=clamping)
    W = pyramid_weights(torch.tensor(g), level, scale, clamping).numpy()
    got = np.array([(derivs[j] * W).sum() for j in range(nder)])
    assert np.allclose(got, want, rtol=1e-10, atol=1e-10)
    if level == 1 and not clamping:
        assert np.array_equal(W, g)


def test_out_of_range_parameter_gets_the_penalty():
    rng = np.random.default_rng(0)
    derivs, g = rng.normal(size=(3, 8, 8, 3)), rng.normal(size=(8, 8, 3))
    r = ad_loss.backward(derivs, g, 3, 4.0, out_of_range=[0.0, 0.5, 0.0], penalty_scale=[1.0, 3.0, 1.0])
    free = ad_loss.backward(derivs, g, 3, 4.0)
    assert r[1] == 1.5 and r[0] == free[0] and r[2] == free[2]

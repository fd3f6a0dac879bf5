"""Shared test helpers: multi-seed renders and the per-pixel-block 3-sigma test."""
import numpy as np

from dtrt_amd import scenes


def renders(mod, desc, seeds, scene=None, **kw):
    """float64 stack [len(seeds), nder+1, H, W, 3] of independent renders."""
    scene = scene if scene is not None else scenes.instantiate(mod, desc)
    return np.stack([scenes.render(mod, desc, seed=s, scene=scene, **kw).astype(np.float64) for s in seeds])


def block_means(stack, b):
    k, c, h, w, _ = stack.shape
    hb, wb = h // b, w // b
    x = stack[:, :, :hb * b, :wb * b].reshape(k, c, hb, b, wb, b, 3)
    return x.mean(axis=(3, 5))


def block_z(a, b, block):
    """z-scores of block means between two stacks of independent renders.

    The standard error of each side comes from the spread over its renders.
    Blocks whose pooled error is zero on both sides must agree exactly.
    """
    ma, mb = block_means(a, block), block_means(b, block)
    ea = ma.std(axis=0, ddof=1) / np.sqrt(len(a))
    eb = mb.std(axis=0, ddof=1) / np.sqrt(len(b))
    se = np.sqrt(ea ** 2 + eb ** 2)
    diff = ma.mean(axis=0) - mb.mean(axis=0)
    z = np.where(se > 0, diff / np.where(se > 0, se, 1), np.where(np.abs(diff) > 1e-6, np.inf, 0.0))
    return z


def assert_3sigma(z, max_outlier_frac=0.02, max_mean_z2=1.6, what=""):
    """Per-block |z| <= 3 for all but a small fraction, and no systematic excess."""
    zf = z[np.isfinite(z)]
    assert np.isfinite(z).all(), "%s: %d blocks differ with zero variance" % (what, (~np.isfinite(z)).sum())
    frac = (np.abs(zf) > 3).mean()
    mz2 = (zf ** 2).mean()
    assert frac <= max_outlier_frac, "%s: %.3f of blocks beyond 3 sigma (mean z^2 %.2f)" % (what, frac, mz2)
    assert mz2 <= max_mean_z2, "%s: mean z^2 = %.2f" % (what, mz2)

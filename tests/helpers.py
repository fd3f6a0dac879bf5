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


def sum_z(a, b):
    """z-scores of the image sums (every pixel of a render added up), per channel and
    colour, between two stacks of independent renders. A bias too small to show in one
    block adds up over the image: 0.5 sigma per block over 36 blocks is 3 sigma here."""
    sa, sb = a.sum(axis=(2, 3)), b.sum(axis=(2, 3))
    se = np.sqrt(sa.var(axis=0, ddof=1) / len(a) + sb.var(axis=0, ddof=1) / len(b))
    diff = sa.mean(axis=0) - sb.mean(axis=0)
    scale = np.maximum(np.abs(sa.mean(axis=0)), np.abs(sb.mean(axis=0)))
    return np.where(se > 1e-12 * np.maximum(scale, 1e-300), diff / np.where(se > 0, se, 1),
                    np.where(np.abs(diff) > 1e-9 * np.maximum(scale, 1e-30), np.inf, 0.0))


def assert_parity(a, b, block, what="", max_outlier_frac=0.02, max_mean_z2=1.6, max_sum_z=4.0):
    """The per-block 3-sigma test, plus |z| <= max_sum_z for every channel's image sum."""
    assert_3sigma(block_z(a, b, block), max_outlier_frac=max_outlier_frac, max_mean_z2=max_mean_z2, what=what)
    zs = sum_z(a, b)
    assert (np.abs(zs) <= max_sum_z).all(), "%s: image-sum z-scores per channel %s" % (what, np.array2string(zs, precision=2))

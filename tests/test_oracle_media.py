"""Pins the oracle's heterogeneous medium: trilinear lookups with rigid-motion
tangents (gridvolume.h:144-326), the VOL loader (:88-132), delta tracking and
ratio tracking (heterogeneous.cpp:9-110)."""
import math
import os

import numpy as np

from dtrt_amd import scenes


def _grid_medium(o, dens, nder, bmin=(-1, -1, -1), bmax=(1, 1, 1), trans=None, rot=None, scale=1.0):
    o.set_nder(nder)
    keep = [np.ascontiguousarray(dens, np.float32), np.eye(4, dtype=np.float32),
            None if trans is None else np.ascontiguousarray(trans, np.float32),
            None if rot is None else np.ascontiguousarray(rot, np.float32)]
    rz, ry, rx = dens.shape
    m = o.HeterogeneousGrid(keep[0], (rx, ry, rz), bmin, bmax, keep[1], scale, (0.8, 0.8, 0.8), 0, keep[2], keep[3],
                            None, None)
    m._keep = keep
    return m


def test_trilinear_exact_on_affine_field_and_motion_tangents(oracle):
    res = 9
    a, b, c, d0 = 0.3, -0.2, 0.5, 4.0
    z, y, x = np.meshgrid(np.arange(res), np.arange(res), np.arange(res), indexing="ij")
    dens = a * x + b * y + c * z + d0
    trans = np.array([[1.0, 0.0, 0.0], [0.0, 0.5, -1.0]])
    rot = np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 0.1]])
    m = _grid_medium(oracle, dens, 2, trans=trans, rot=rot)
    rng = np.random.default_rng(0)
    p = rng.uniform(-0.95, 0.95, (300, 3))
    pts = np.concatenate([p, np.zeros((300, 6))], 1)
    out = oracle.grid_lookup(m, pts, 2)
    gs = (res - 1) / 2.0
    grad = np.array([a, b, c]) * gs                   # world-space gradient of the density
    g = (p + 1) * gs
    assert np.allclose(out[:, 0], a * g[:, 0] + b * g[:, 1] + c * g[:, 2] + d0, rtol=1e-6)   # float32 grid data
    # lane 0: translation e_x  -> d/dtheta rho(p - theta v) = -grad . v
    assert np.allclose(out[:, 1], -(grad @ trans[0]), rtol=1e-6, atol=1e-5)
    # lane 1: translation + rotation about the volume centre (origin)
    v = trans[1] + np.cross(rot[1], p)
    assert np.allclose(out[:, 2], -(v @ grad), rtol=1e-6, atol=1e-5)


def test_vol_roundtrip(oracle, tmp_path):
    dens = scenes.fbm_grid(12, seed=3)
    fn = str(tmp_path / "g.vol")
    scenes.write_vol(fn, dens, (-1, -1, -1), (1, 1, 1))
    oracle.set_nder(1)
    mf = oracle.Heterogeneous(fn, np.eye(4, dtype=np.float32), 1.0, (0.8, 0.8, 0.8), 0)
    mg = _grid_medium(oracle, dens, 1)
    rng = np.random.default_rng(1)
    pts = np.concatenate([rng.uniform(-1.1, 1.1, (500, 3)), np.zeros((500, 3))], 1)
    assert np.array_equal(oracle.grid_lookup(mf, pts, 1), oracle.grid_lookup(mg, pts, 1))


def test_fbm_grid_is_deterministic_and_in_range():
    a = scenes.fbm_grid(16, seed=7)
    b = scenes.fbm_grid(16, seed=7)
    assert np.array_equal(a, b) and abs(a.min()) < 1e-6 and abs(a.max() - 2.0) < 1e-6
    assert not np.array_equal(a, scenes.fbm_grid(16, seed=8))


def test_tracking_estimators_on_constant_density(oracle):
    sigma = 1.7
    dens = np.full((5, 5, 5), sigma, np.float32)
    m = _grid_medium(oracle, dens, 1)
    n = 40000
    for L in [0.3, 1.0, 1.8]:
        seg = np.tile([[-0.9, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, L]], (n, 1))
        T = oracle.medium_estimate(m, 0, seg, seed=1)
        assert abs(T.mean() - math.exp(-sigma * L)) < 4 * T.std() / math.sqrt(n) + 1e-4
        hit = oracle.medium_estimate(m, 1, seg, seed=2)
        p = 1 - math.exp(-sigma * L)
        assert abs(hit.mean() - p) < 4 * math.sqrt(p * (1 - p) / n)

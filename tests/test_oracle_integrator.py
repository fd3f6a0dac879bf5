"""Integrator-level pins of the oracle (volpathAD.cpp, integratorAD.cpp) against
closed forms and common-random-number finite differences."""
import math

import numpy as np
import pytest

from dtrt_amd import scenes
from tests import scene_fixtures as F


@pytest.mark.parametrize("hetero", [False, True])
def test_furnace(oracle, hetero):
    d = F.furnace(nder=1, heterogeneous=hetero, res=12, spp=32)
    img = scenes.render(oracle, d, seed=3)
    v = img[0, ..., 0]
    assert abs(v.mean() - 1.0) < 4 * v.std() / math.sqrt(v.size) + 1e-3
    g = img[1, ..., 0]                   # d/d(density scale) vanishes in a furnace
    assert abs(g.mean()) < 4 * g.std() / math.sqrt(g.size) + 1e-3


def test_direct_lighting_matches_lambert_polygon(oracle):
    d, lv, rho, Le = F.direct_lighting(res=16, spp=256)
    img = scenes.render(oracle, d, seed=1)
    ref = F.direct_expected(d, lv, rho, Le)
    rel = (img[0, ..., 0].mean() - ref.mean()) / ref.mean()
    assert abs(rel) < 5e-3
    assert np.abs(img[0, ..., 0] - ref).max() < 0.25 * ref.max()


def test_absorbing_slab_homogeneous_exact(oracle):
    d, L, dL = F.absorbing_slab(heterogeneous=False)
    img = scenes.render(oracle, d, seed=0)
    assert np.allclose(img[0, ..., 0], L, rtol=1e-5, atol=1e-6)
    assert np.allclose(img[1, ..., 0], dL, rtol=1e-4, atol=1e-5)


def test_absorbing_slab_heterogeneous_statistical(oracle):
    d, L, dL = F.absorbing_slab(heterogeneous=True, spp=512)
    img = scenes.render(oracle, d, seed=0)
    assert abs(img[0, ..., 0].mean() / L.mean() - 1) < 5e-3
    assert abs(img[1, ..., 0].mean() / dL.mean() - 1) < 2e-2


def _fd_crn(mod, make, h, seed=5):
    base = scenes.render(mod, make(0.0), seed=seed)
    up = scenes.render(mod, make(+h), seed=seed)
    dn = scenes.render(mod, make(-h), seed=seed)
    return base[1], (up[0] - dn[0]) / (2 * h)


def test_reflectance_tangent_equals_crn_finite_difference(oracle):
    ad, fd = _fd_crn(oracle, lambda e: F.ico_surface_reflectance_lane(res=16, spp=4, rho=0.8 + e), 1e-3)
    assert np.abs(fd).max() > 1e-2
    assert np.allclose(ad, fd, rtol=2e-3, atol=2e-4)


def test_medium_albedo_tangent_equals_crn_finite_difference(oracle):
    ad, fd = _fd_crn(oracle, lambda e: F.cube_medium_albedo_lane(res=12, spp=4, albedo=0.8 + e, grid_res=12), 1e-3)
    assert np.abs(fd).max() > 1e-2
    assert np.allclose(ad, fd, rtol=2e-3, atol=2e-3)


# ---------------------------------------------------------------- VolPathTracer (volpath.cpp)
@pytest.mark.parametrize("hetero", [False, True])
def test_primal_furnace(oracle, hetero):
    d = F.furnace(nder=1, heterogeneous=hetero, res=12, spp=32)
    v = scenes.render_primal(oracle, d, seed=3)[..., 0].astype(np.float64)
    assert abs(v.mean() - 1.0) < 4 * v.std() / math.sqrt(v.size) + 1e-3


def test_primal_direct_lighting_matches_lambert_polygon(oracle):
    d, lv, rho, Le = F.direct_lighting(res=16, spp=256)
    img = scenes.render_primal(oracle, d, seed=1)[..., 0]
    ref = F.direct_expected(d, lv, rho, Le)
    assert abs(img.mean() / ref.mean() - 1) < 5e-3
    assert np.abs(img - ref).max() < 0.25 * ref.max()


def test_primal_equals_ad_value_channel(oracle):
    """VolPathTracer and the value channel of VolPathTracerAD's interior terms estimate
    the same radiance (volpath.cpp:10-103 vs volpathAD.cpp:145-620 on the value lane)."""
    d = scenes.cube_medium(nder=1, res=32, spp=16, grid_res=16, crop=(10, 10, 12, 12))
    sc = scenes.instantiate(oracle, d)
    a = np.stack([scenes.render_primal(oracle, d, seed=s, scene=sc)[None] for s in range(48)]).astype(np.float64)
    b = np.stack([scenes.render(oracle, d, seed=1000 + s, scene=sc, mode=1 | 2 | 8, sppe=0, sppse=0)[:1]
                  for s in range(48)]).astype(np.float64)
    from tests.helpers import assert_3sigma, block_z
    assert_3sigma(block_z(a, b, 3), what="primal vs AD value")


@pytest.mark.parametrize("hetero", [False, True])
def test_primal_simple_furnace(oracle, hetero):
    """VolPathTracerSimple (volpath_simple.cpp): NEE alone also gives L = 1 in a furnace."""
    d = F.furnace(nder=1, heterogeneous=hetero, res=12, spp=32)
    v = scenes.render_primal(oracle, d, seed=3, simple=True)[..., 0].astype(np.float64)
    assert abs(v.mean() - 1.0) < 4 * v.std() / math.sqrt(v.size) + 1e-3


def test_primal_simple_direct_lighting_matches_lambert_polygon(oracle):
    d, lv, rho, Le = F.direct_lighting(res=16, spp=256)
    img = scenes.render_primal(oracle, d, seed=1, simple=True)[..., 0]
    ref = F.direct_expected(d, lv, rho, Le)
    assert abs(img.mean() / ref.mean() - 1) < 5e-3


def test_primal_simple_equals_mis_estimator(oracle):
    """NEE-only and NEE+MIS primal estimators have the same expectation."""
    d = scenes.ico_surface(nder=1, res=24, spp=8)
    sc = scenes.instantiate(oracle, d)
    a = np.stack([scenes.render_primal(oracle, d, seed=s, scene=sc, simple=True)[None] for s in range(32)])
    b = np.stack([scenes.render_primal(oracle, d, seed=500 + s, scene=sc)[None] for s in range(32)])
    from tests.helpers import assert_3sigma, block_z
    assert_3sigma(block_z(a.astype(np.float64), b.astype(np.float64), 3), what="simple vs MIS")


@pytest.mark.parametrize("kind,depth", [(3, 5), (4, 5), (3, 1)])
def test_surface_integrators_diffuse_furnace(oracle, kind, depth):
    """PathTracer (path.cpp) and DirectIntegrator (direct.cpp, one bounce) inside an
    emitting diffuse sphere: L = Le * sum_{k<=depth} a^k."""
    a = 0.5
    d = F.diffuse_furnace(albedo=a, Le=2.0, res=12, spp=32, depth=depth)
    v = scenes.render_primal(oracle, d, seed=2, kind=kind)[..., 0].astype(np.float64)
    eff = 1 if kind == 4 else depth
    want = 2.0 * sum(a ** k for k in range(eff + 1))
    assert abs(v.mean() / want - 1) < 4 * v.std() / math.sqrt(v.size) / want + 1e-3


@pytest.mark.parametrize("kind", [3, 4])
def test_surface_integrators_direct_lighting_matches_lambert_polygon(oracle, kind):
    d, lv, rho, Le = F.direct_lighting(res=16, spp=256)
    img = scenes.render_primal(oracle, d, seed=1, kind=kind)[..., 0]
    ref = F.direct_expected(d, lv, rho, Le)
    assert abs(img.mean() / ref.mean() - 1) < 5e-3


def test_particle_tracer_closed_forms(oracle):
    """ParticleTracer (ptracer.cpp): light paths splatted to the camera reproduce the
    Lambert polygon (one bounce) and the emitting diffuse sphere (L = Le sum a^k)."""
    d, lv, rho, Le = F.direct_lighting(res=16, spp=64)
    img = scenes.render_primal(oracle, d, seed=1, kind=5, spp=20)[..., 0]
    ref = F.direct_expected(d, lv, rho, Le)
    assert abs(img.mean() / ref.mean() - 1) < 5e-3
    assert np.abs(img - ref).max() < 0.1 * ref.max()
    d = F.diffuse_furnace(albedo=0.5, Le=2.0, res=12, spp=32, depth=5)
    v = scenes.render_primal(oracle, d, seed=2, kind=5, spp=10)[..., 0]
    assert abs(v.mean() / (2.0 * sum(0.5 ** k for k in range(6))) - 1) < 3e-3

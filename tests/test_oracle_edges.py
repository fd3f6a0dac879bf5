"""Pins the oracle's boundary terms: the primary-edge pass (integratorAD.cpp:100-165,
edge_manager.cpp), secondary interface and volume edge terms (volpathAD.cpp:207-305,
513-616) and the Lins boundary term (:321-467), against analytic derivatives and
common-random-number finite differences."""
import numpy as np

from dtrt_amd import scenes
from tests import scene_fixtures as F


def test_primary_edges_translating_emitter_exact(oracle):
    d, half = F.translating_emitter()
    g = scenes.render(oracle, d, seed=0)[1, ..., 0]
    w = d.width // 2
    assert abs(g[:, :w].sum() / -half - 1) < 5e-3
    assert abs(g[:, w:].sum() / half - 1) < 5e-3


def test_interface_edges_match_lambert_derivative(oracle):
    d, dE = F.translating_light(res=10, spp=512)
    g = scenes.render(oracle, d, seed=0)[1, ..., 0]
    assert np.corrcoef(g.ravel(), dE.ravel())[0, 1] > 0.999
    assert np.abs(g - dE).max() < 0.05 * np.abs(dE).max()


def _ad_fd(mod, d, seeds=6, h=0.01):
    ad = np.array([scenes.render(mod, d, seed=s)[1, ..., 0].sum() for s in range(seeds)])
    fd = np.array([(scenes.render(mod, F.shifted(d, h), seed=50 + s)[0, ..., 0].sum() -
                    scenes.render(mod, F.shifted(d, -h), seed=50 + s)[0, ..., 0].sum()) / (2 * h) for s in range(seeds)])
    se = np.sqrt(ad.var(ddof=1) / seeds + fd.var(ddof=1) / seeds)
    return ad.mean(), fd.mean(), se


def test_volume_edges_match_crn_fd(oracle):
    ad, fd, se = _ad_fd(oracle, F.medium_box(move="light", spp=256))
    assert abs(ad) > 1.0 and abs(ad - fd) < 4 * se + 0.02 * abs(fd)


def test_lins_matches_crn_fd_with_transmissive_light(oracle):
    d = F.medium_box(move="box", spp=256)
    d.bsdfs = [("null",), ("null",)]
    lins = np.mean([scenes.render(oracle, d, seed=s, mode=16)[1, ..., 0].sum() for s in range(3)])
    ad, fd, se = _ad_fd(oracle, d)
    assert lins > 1.0 and abs(ad - fd) < 4 * se + 0.02 * abs(fd)


def test_lins_nee_gated_by_light_bsdf(oracle):
    # scene.cpp:471-472 tests the *light's* BSDF for transmission, so NEE from the
    # inner side of a null boundary toward an opaque light is dropped; the Lins
    # estimate then keeps only its phase-sampled part (reference behaviour).
    d = F.medium_box(move="box", spp=128)
    opaque = np.mean([scenes.render(oracle, d, seed=s, mode=16)[1, ..., 0].sum() for s in range(3)])
    d.bsdfs = [("null",), ("null",)]
    transmissive = np.mean([scenes.render(oracle, d, seed=s, mode=16)[1, ..., 0].sum() for s in range(3)])
    assert abs(opaque) < 0.05 * abs(transmissive)


# ------------------------------------------- surface-only AD integrators (pathAD, directAD)
def _surface_ad(mod, name):
    return mod.PathTracerAD() if name == "path" else mod.DirectAD()


def test_surface_ad_primary_edges_translating_emitter(oracle):
    """integratorAD.cpp's primary-edge pass through PathTracerAD / DirectAD's pixelColor."""
    for name in ("path", "direct"):
        d, half = F.translating_emitter()
        g = scenes.render(oracle, d, seed=0, integrator=_surface_ad(oracle, name))[1, ..., 0]
        w = d.width // 2
        assert abs(g[:, :w].sum() / -half - 1) < 5e-3, name
        assert abs(g[:, w:].sum() / half - 1) < 5e-3, name


def test_surface_ad_edges_match_lambert_derivative(oracle):
    """pathAD.cpp:112-210 (angle-uniform points, primal sub-paths) and directAD.cpp:55-95
    (points linear in t, emission only) against the analytic derivative of the
    irradiance of a translating quad light."""
    for name in ("path", "direct"):   # unstratified edge samples: more spp than the volpathAD test
        d, dE = F.translating_light(res=10, spp=2048)
        img = scenes.render(oracle, d, seed=0, integrator=_surface_ad(oracle, name))
        g = img[1, ..., 0]
        assert np.corrcoef(g.ravel(), dE.ravel())[0, 1] > 0.998, name
        assert np.abs(g - dE).max() < 0.06 * np.abs(dE).max(), name


def test_surface_ad_direct_lighting_value(oracle):
    d, lv, rho, Le = F.direct_lighting(res=16, spp=256)
    ref = F.direct_expected(d, lv, rho, Le)
    for name in ("path", "direct"):
        v = scenes.render(oracle, d, seed=1, integrator=_surface_ad(oracle, name))[0, ..., 0]
        assert abs(v.mean() / ref.mean() - 1) < 5e-3, name


def test_path_ad_agrees_with_volpath_ad_on_surfaces(oracle):
    """Without media or null surfaces PathTracerAD and VolPathTracerAD estimate the same
    image and tangents (all EMode terms, both edge passes)."""
    from tests.helpers import assert_3sigma, block_z
    d = scenes.ico_surface(nder=1, res=16, spp=8)
    sc = scenes.instantiate(oracle, d)
    a = np.stack([scenes.render(oracle, d, seed=s, scene=sc, sppe=256) for s in range(24)]).astype(np.float64)
    b = np.stack([scenes.render(oracle, d, seed=100 + s, scene=sc, sppe=256, integrator=oracle.PathTracerAD())
                  for s in range(24)]).astype(np.float64)
    assert_3sigma(block_z(a, b, 4), what="pathAD vs volpathAD")

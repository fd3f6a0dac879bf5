"""The surface-only forward-mode integrators PathTracerAD (pathAD.cpp) and DirectAD
(directAD.cpp) on the GPU: closed forms and 3-sigma agreement with the oracle."""
import numpy as np
import pytest

from dtrt_amd import scenes
from tests import scene_fixtures as F
from tests.helpers import assert_3sigma, block_z, assert_parity

pytestmark = pytest.mark.gpu
NAMES = ["path", "direct"]


def _integ(mod, name):
    return mod.PathTracerAD() if name == "path" else mod.DirectAD()


@pytest.mark.parametrize("name", NAMES)
def test_primary_edges_translating_emitter(gpu, name):
    d, half = F.translating_emitter(res=64)
    g = scenes.render(gpu, d, seed=0, integrator=_integ(gpu, name))[1, ..., 0].astype(np.float64)
    w = d.width // 2
    assert abs(g[:, :w].sum() / -half - 1) < 5e-3
    assert abs(g[:, w:].sum() / half - 1) < 5e-3


@pytest.mark.parametrize("name", NAMES)
def test_edges_match_lambert_derivative(gpu, name):
    d, dE = F.translating_light(res=16, spp=8192)
    g = scenes.render(gpu, d, seed=0, integrator=_integ(gpu, name))[1, ..., 0]
    assert np.corrcoef(g.ravel(), dE.ravel())[0, 1] > 0.999
    assert np.abs(g - dE).max() < 0.05 * np.abs(dE).max()


@pytest.mark.parametrize("name", NAMES)
def test_direct_lighting_value(gpu, name):
    d, lv, rho, Le = F.direct_lighting(res=32, spp=1024)
    v = scenes.render(gpu, d, seed=1, integrator=_integ(gpu, name))[0, ..., 0]
    ref = F.direct_expected(d, lv, rho, Le)
    assert abs(v.mean() / ref.mean() - 1) < 3e-3


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("cfg", ["ico_surface", "dielectric_medium"])
def test_parity(gpu, oracle, cfg, name):
    if cfg == "ico_surface":
        d, kw = scenes.ico_surface(nder=1, res=24, spp=32), dict(sppe=1024)
    else:
        d = scenes.dielectric_medium(nder=8, res=48, spp=16, freq=10, grid_res=16, crop=(14, 14, 20, 20))
        d.bsdfs = [("roughdielectric", 0.3) + tuple(b[2:]) if b[0] == "roughdielectric" else b for b in d.bsdfs]
        kw = dict(sppe=128, max_bounces=6)
    sg, so = scenes.instantiate(gpu, d), scenes.instantiate(oracle, d)
    k = 16
    a = np.stack([scenes.render(gpu, d, seed=s, scene=sg, integrator=_integ(gpu, name), **kw) for s in range(k)])
    b = np.stack([scenes.render(oracle, d, seed=1000 + s, scene=so, integrator=_integ(oracle, name), **kw)
                  for s in range(k)])
    assert_parity(a.astype(np.float64), b.astype(np.float64), 4, what="%s %s" % (name, cfg), max_mean_z2=1.8)


def test_path_ad_overflow_replay_is_exact(gpu):
    """Edge sub-path records (CHILD_PATH) through the exact overflow replay."""
    d = scenes.ico_surface(nder=1, res=32, spp=64)
    sc = scenes.instantiate(gpu, d)
    ref = scenes.render(gpu, d, seed=4, scene=sc, sppe=64, integrator=gpu.PathTracerAD()).astype(np.float64)
    st_ref = gpu.render_stats(sc)
    assert st_ref["replayed"] == 0 and st_ref["children"] > 0
    gpu.set_queue_budget(sc, 4096 * 128)
    try:
        got = scenes.render(gpu, d, seed=4, scene=sc, sppe=64, integrator=gpu.PathTracerAD()).astype(np.float64)
        st = gpu.render_stats(sc)
    finally:
        gpu.set_queue_budget(sc, 0)
    assert st["replayed"] > 0 and st["children"] == st_ref["children"]
    scale = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
    for lane in range(ref.shape[0]):
        assert np.allclose(got[lane], ref[lane], rtol=1e-3, atol=1e-4 * scale[lane]), lane

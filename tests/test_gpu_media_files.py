"""The drop-in medium constructor Heterogeneous(file, ...) (src/dtrt.cpp:87-93) on the
GPU: a density grid read from a VOL v3 file renders like the same grid handed over in
memory (dtrt_heterogeneous_create_grid), and the albedo-grid branch (albedo given as a
3-channel VOL file, gridvolume.h:88-132) scales scattering like the constant albedo."""
import copy

import numpy as np
import pytest

from dtrt_amd import scenes

pytestmark = pytest.mark.gpu
INTERIOR, LINS = 1 | 2 | 8, 16


def _scene():
    return scenes.cube_medium(nder=3, res=32, spp=8, grid_res=16)


def test_vol_file_renders_like_the_in_memory_grid(gpu, tmp_path):
    d = _scene()
    m = d.media[0]
    path = str(tmp_path / "density.vol")
    scenes.write_vol(path, m.density, m.bmin, m.bmax)
    f = copy.deepcopy(d)
    f.media[0].density_file = path
    a = scenes.render(gpu, d, seed=4)
    b = scenes.render(gpu, f, seed=4)
    assert np.abs(a[0]).max() > 0 and np.abs(a[1:]).max() > 0
    # the same medium record, hence the same paths: equal up to the summation order of atomics
    assert np.allclose(a, b, rtol=1e-6, atol=1e-7 * np.abs(a).max())


def test_albedo_grid_file_matches_constant_albedo_in_the_interior_terms(gpu, tmp_path):
    d = _scene()
    m = d.media[0]
    dens, alb = str(tmp_path / "density.vol"), str(tmp_path / "albedo.vol")
    scenes.write_vol(dens, m.density, m.bmin, m.bmax)
    scenes.write_vol(alb, np.full(m.density.shape + (3,), 0.8, np.float32), m.bmin, m.bmax)
    f = copy.deepcopy(d)
    f.media[0].density_file = dens
    f.media[0].albedo = alb
    for mm in (d.media[0], f.media[0]):
        mm.d_albedo = None
    a = scenes.render(gpu, d, seed=6, mode=INTERIOR, sppe=0, sppse=0)
    b = scenes.render(gpu, f, seed=6, mode=INTERIOR, sppe=0, sppse=0)
    assert np.abs(a[0]).max() > 0
    assert np.allclose(a, b, rtol=1e-5, atol=1e-6 * np.abs(a).max())
    # sigma_s of the Lins term uses the constant albedo, which the albedo-file variant
    # leaves at zero (heterogeneous.h:13-106): that term vanishes there
    la = scenes.render(gpu, d, seed=6, mode=LINS, sppe=0)
    lb = scenes.render(gpu, f, seed=6, mode=LINS, sppe=0)
    assert np.abs(la[1:]).max() > 0 and np.abs(lb[1:]).max() == 0

"""The two estimator precisions. Scenes are created in double (the reference's
precision, include/fwd.h:10-13) by default; the float kernel set is the second leg.
Both must agree with the oracle, and with each other, at 3 sigma."""
import contextlib

import numpy as np
import pytest

from dtrt_amd import scenes
from tests.helpers import assert_3sigma, block_z, renders, assert_parity

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def precision(gpu, bits):
    old = gpu.precision()
    gpu.set_precision(bits)
    try:
        yield
    finally:
        gpu.set_precision(old)


def test_default_precision_is_double(gpu):
    assert gpu.precision() == 64
    sc = scenes.instantiate(gpu, scenes.ico_surface(nder=1, res=8, spp=1))
    assert gpu.scene_precision(sc) == 64
    with precision(gpu, 32):
        assert gpu.scene_precision(scenes.instantiate(gpu, scenes.ico_surface(nder=1, res=8, spp=1))) == 32
    with pytest.raises(ValueError):
        gpu.set_precision(16)


@pytest.mark.parametrize("which", ["ico_surface", "cube_medium"])
def test_fp32_leg_matches_the_oracle(gpu, oracle, which):
    if which == "ico_surface":
        d, sppe = scenes.ico_surface(nder=1, res=24, spp=32), 1024
    else:
        d, sppe = scenes.cube_medium(nder=3, res=64, spp=32, grid_res=32, crop=(20, 20, 24, 24)), 256
    with precision(gpu, 32):
        a = renders(gpu, d, range(16), sppe=sppe)
    b = renders(oracle, d, range(1000, 1016), sppe=sppe)
    assert_parity(a, b, 4, what="fp32 " + which)


def test_fp64_and_fp32_legs_agree(gpu):
    """Same scene, both kernel sets, 24 renders each: block means agree at 3 sigma."""
    d = scenes.cube_medium(nder=3, res=64, spp=32, grid_res=32, crop=(20, 20, 24, 24))
    a = renders(gpu, d, range(24), sppe=256)
    with precision(gpu, 32):
        b = renders(gpu, d, range(100, 124), sppe=256)
    assert_parity(a, b, 4, what="fp64 vs fp32")

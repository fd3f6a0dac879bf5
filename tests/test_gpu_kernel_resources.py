"""Scratch limits of the kernels: the host-side check in scene creation.

Every kernel a scene of each dual width and estimator precision launches reports its
private segment (register spills, private arrays; hipFuncGetAttributes) and occupancy.
All of them stay within the limits the library checks, and a scene whose kernels
exceed a (lowered) limit is refused with an error that names the kernel, instead of
being launched.
"""
import ctypes as C

import numpy as np
import pytest

from dtrt_amd import scenes

pytestmark = pytest.mark.gpu

NAME_LEN = 40


@pytest.fixture(scope="module")
def lib(gpu):
    L = gpu.lib
    L.dtrt_test_kernel_resources.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_int, C.c_void_p, C.c_int]
    L.dtrt_test_scratch_limits.argtypes = [C.c_void_p]
    L.dtrt_test_set_scratch_limits.argtypes = [C.c_uint64, C.c_uint64]
    return L


def _resources(lib, nder, bits):
    names = C.create_string_buffer(NAME_LEN * 32)
    vals = np.zeros((32, 5), np.uint64)
    n = lib.dtrt_test_kernel_resources(nder, bits, names, NAME_LEN, vals.ctypes.data, 32)
    assert 0 < n <= 32
    raw = names.raw
    return [(raw[i * NAME_LEN:(i + 1) * NAME_LEN].split(b"\0")[0].decode(), *map(int, vals[i])) for i in range(n)]


def _limits(lib):
    out = np.zeros(6, np.uint64)
    assert lib.dtrt_test_scratch_limits(out.ctypes.data) == 0
    return [int(x) for x in out]


def test_every_kernel_within_the_scratch_limits(lib):
    lane, device, stack, smin, smax, scur = _limits(lib)
    print("limits: %d B/lane, %.1f MB/device; runtime stack %d B/thread, scratch min %d max %d current %d"
          % (lane, device / 1e6, stack, smin, smax, scur))
    assert (lane, device) == (12 << 10, 768 << 20)
    widest = 0
    for bits in (32, 64):
        for nder in (1, 2, 3, 4, 8, 16):
            for name, priv, vgprs, block, waves, dev_bytes in _resources(lib, nder, bits):
                print("f%d nder %2d %-24s %6d B/lane vgpr %3d block %3d waves/CU %2d %7.1f MB"
                      % (bits, nder, name, priv, vgprs, block, waves, dev_bytes / 1e6))
                assert waves >= 1, name
                assert priv <= lane and dev_bytes <= device, (bits, nder, name, priv, dev_bytes)
                widest = max(widest, priv)
    # the double-precision width-16 refill kernel spills: the check sees real figures
    assert widest > 4096


def test_scene_beyond_a_lowered_limit_is_refused(gpu, lib):
    d = scenes.cube_medium(nder=16, res=16, spp=1, grid_res=8)
    try:
        assert lib.dtrt_test_set_scratch_limits(1024, 0) == 0
        with pytest.raises(RuntimeError, match="bytes of scratch per lane"):
            scenes.instantiate(gpu, d)
        assert lib.dtrt_test_set_scratch_limits(0, 1 << 20) == 0
        with pytest.raises(RuntimeError, match="over the device"):
            scenes.instantiate(gpu, d)
    finally:
        assert lib.dtrt_test_set_scratch_limits(0, 0) == 0
    sc = scenes.instantiate(gpu, d)
    img = scenes.render(gpu, d, seed=1, scene=sc)
    assert np.isfinite(img).all()

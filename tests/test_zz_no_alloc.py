"""dtrt_render_device_noalloc: the forward-mode render with every device buffer in
the library's fixed scratch area. Through the C-ABI, its images equal
dtrt_render_device's for the same scene, options and seed (same random streams, so
the same paths; float summation order aside), on two calls in a row, the second with
a smaller image than the first; an image too large for the area is refused."""
import ctypes as C

import numpy as np
import pytest

from dtrt_amd import scenes

pytestmark = pytest.mark.gpu


def _noalloc(gpu, sc, opts, out):
    lib = gpu.lib
    lib.dtrt_render_device_noalloc.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    st = opts._struct()
    gpu._after_torch(sc)
    return lib.dtrt_render_device_noalloc(0, sc._h, C.byref(st), C.c_void_p(out.data_ptr()))


def _pair(gpu, d, seed):
    import torch
    sc = scenes.instantiate(gpu, d)
    opts = gpu.RenderOptions(seed, d.spp, d.max_bounces, d.sppe, d.sppse, True, -1, d.ddist)
    ref = torch.zeros(d.out_shape(), dtype=torch.float32, device="cuda")
    gpu.render_device(0, sc, opts, ref.data_ptr())
    got = torch.zeros(d.out_shape(), dtype=torch.float32, device="cuda")
    rc = _noalloc(gpu, sc, opts, got)
    assert rc == 0, gpu.lib.dtrt_last_error().decode()
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


@pytest.mark.parametrize("bits", [64, 32])
def test_noalloc_equals_render_device_twice(gpu, bits):
    old = gpu.precision()
    gpu.set_precision(bits)
    try:
        # first a larger medium scene, then a smaller surface scene in the same scratch area
        for d, seed in ((scenes.cube_medium(nder=3, res=96, spp=16, grid_res=32), 5),
                        (scenes.ico_surface(nder=1, res=40, spp=16), 6)):
            ref, got = _pair(gpu, d, seed)
            assert np.abs(ref[0]).max() > 0 and np.abs(ref[1:]).max() > 0
            assert np.allclose(got, ref, rtol=1e-4, atol=1e-5 * np.abs(ref).max())
    finally:
        gpu.set_precision(old)


def test_noalloc_refuses_images_larger_than_the_scratch_area(gpu):
    import torch
    big = scenes.ico_surface(nder=16, res=1400, spp=1)   # 17 x 1400^2 x 3 doubles of accumulators alone: 800 MB
    sc = scenes.instantiate(gpu, big)
    out = torch.zeros(8, dtype=torch.float32, device="cuda")   # never written: refused before any launch
    opts = gpu.RenderOptions(1, 1, big.max_bounces, 0, 0, True, -1, big.ddist)
    assert _noalloc(gpu, sc, opts, out) != 0
    assert b"scratch area" in gpu.lib.dtrt_last_error()

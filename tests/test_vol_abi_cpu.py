"""The drop-in medium constructor ``Heterogeneous(file, ...)`` (src/dtrt.cpp:87-93,
gridvolume.h:88-132) through the product library, without a GPU: a VOL v3 file
written by ``scenes.write_vol`` loads the same voxels (negatives clamped to 0) as the
in-memory grid constructor, for the density and the albedo-grid branch; malformed
files come back as an error, never an abort."""
import ctypes as C
import struct

import numpy as np
import pytest

from dtrt_amd import scenes


@pytest.fixture(scope="module")
def G():
    import dtrt_amd.dtrt as G
    G.lib.dtrt_test_medium_data.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p]
    return G


def _data(G, med, which=0):
    meta = np.zeros(8, np.float64)
    assert G.lib.dtrt_test_medium_data(med._h, which, None, 0, meta.ctypes.data) == 0
    out = np.zeros(int(meta[4]), np.float32)
    assert G.lib.dtrt_test_medium_data(med._h, which, out.ctypes.data, len(out), meta.ctypes.data) == 0
    return out, meta


def _grid(seed=3, shape=(5, 6, 7)):
    rng = np.random.default_rng(seed)
    return (rng.random(shape, dtype=np.float32) * 2.0 - 0.3).astype(np.float32)   # some negatives


def test_vol_file_loads_like_the_in_memory_grid(G, tmp_path):
    g = _grid()
    path = str(tmp_path / "d.vol")
    scenes.write_vol(path, g, (-0.5, -0.4, -0.3), (0.5, 0.6, 0.7))
    mf = G.Heterogeneous(path, np.eye(4, dtype=np.float32), 1.5, G.Spectrum3f(0.8, 0.8, 0.8), 0)
    mg = G.HeterogeneousGrid(g, (7, 6, 5), (-0.5, -0.4, -0.3), (0.5, 0.6, 0.7), np.eye(4, dtype=np.float32), 1.5,
                             G.Spectrum3f(0.8, 0.8, 0.8), 0)
    a, ma = _data(G, mf)
    b, mb = _data(G, mg)
    assert np.array_equal(a, np.maximum(g, 0).ravel()) and np.array_equal(a, b)
    assert ma[:3].tolist() == [7, 6, 5] and ma[5] == np.maximum(g, 0).max() and ma[6] == 1.5
    assert np.array_equal(ma, mb)


def test_vol_albedo_grid_branch(G, tmp_path):
    g = _grid()
    alb = np.random.default_rng(4).random((3, 4, 5, 3), dtype=np.float32)
    p, q = str(tmp_path / "d.vol"), str(tmp_path / "a.vol")
    scenes.write_vol(p, g, (0, 0, 0), (1, 1, 1))
    scenes.write_vol(q, alb, (0, 0, 0), (1, 1, 1))
    m = G.Heterogeneous(p, np.eye(4, dtype=np.float32), 1.0, q, 0)
    a, meta = _data(G, m, 1)
    assert meta[:4].tolist() == [5, 4, 3, 3] and np.array_equal(a, alb.ravel())
    with pytest.raises(Exception, match="three channels"):
        G.Heterogeneous(p, np.eye(4, dtype=np.float32), 1.0, p, 0)   # a 1-channel file as albedo


def _header(magic=b"VOL", ver=3, typ=1, res=(4, 4, 4), ch=1):
    return magic + struct.pack("<B", ver) + struct.pack("<i", typ) + struct.pack("<3i", *res) + \
        struct.pack("<i", ch) + struct.pack("<3f", 0, 0, 0) + struct.pack("<3f", 1, 1, 1)


@pytest.mark.parametrize("case", ["magic", "version", "type", "channels", "truncated", "oversized", "huge",
                                  "negative", "missing"])
def test_malformed_vol_files_are_errors(G, tmp_path, case):
    body = np.ones(64, np.float32).tobytes()
    blobs = {
        "magic": _header(magic=b"VOX") + body,
        "version": _header(ver=2) + body,
        "type": _header(typ=2) + body,
        "channels": _header(ch=2) + body,
        "truncated": _header() + body[:100],
        "oversized": _header(res=(400, 400, 400)) + body,
        "huge": _header(res=(100000, 100000, 100000)) + body,
        "negative": _header(res=(-4, 4, 4)) + body,
    }
    path = str(tmp_path / "bad.vol")
    if case != "missing":
        open(path, "wb").write(blobs[case])
    with pytest.raises(Exception, match="VOL|volume"):
        G.Heterogeneous(path, np.eye(4, dtype=np.float32), 1.0, G.Spectrum3f(0.5, 0.5, 0.5), 0)
    assert b"VOL" in G.fns["last_error"]() or b"volume" in G.fns["last_error"]()

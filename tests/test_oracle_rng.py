"""Pins the oracle's PCG32 (src/sampler.cpp:3-41) to the published demo sequence."""
import json
import os

import numpy as np

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat.json")))


def _seed_idx(initstate, initseq):
    seed = (initstate - 0x853C49E6748FEA9B) % (1 << 64)
    return seed, initseq - 1


def test_pcg32_published_sequence(oracle):
    g = GOLD["pcg32_demo"]
    seed, idx = _seed_idx(g["initstate"], g["initseq"])
    got = oracle.unit("pcg32", [seed >> 32, seed & 0xFFFFFFFF, idx, len(g["outputs"])])
    assert [int(x) for x in got] == [int(h, 16) for h in g["outputs"]]


def test_next1d_is_u32_over_2_pow_32(oracle):
    # sampler.cpp:19-27: the 32-bit output placed at the top of a [1,2) mantissa, minus one
    u = oracle.unit("pcg32", [0, 5, 7, 512], nout=512)
    d = oracle.unit("pcg_next1d", [0, 5, 7, 512], nout=512)
    assert np.array_equal(d, u / 2.0 ** 32)


def test_python_generator_matches_oracle(oracle):
    from dtrt_amd.scenes import Pcg32
    p = Pcg32(123, 9)
    got = [p.next_u32() for _ in range(16)]
    ref = oracle.unit("pcg32", [0, 123, 9, 16])
    assert got == [int(x) for x in ref]

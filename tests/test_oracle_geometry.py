"""Pins the oracle's geometry: edges (shape.cpp:254-300), silhouettes (:302-335),
BVH closest hit (stand-in for Embree), primary-edge projection
(edge_manager.cpp:31-110) and light selection (scene.cpp:45-98)."""
import json
import math
import os

import numpy as np
import pytest

from dtrt_amd import scenes

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat.json")))


def _shape(o, v, f, normals=None):
    o.set_nder(1)
    return o.Shape(np.ascontiguousarray(v, np.float32), np.ascontiguousarray(f, np.int32), None,
                   None if normals is None else np.ascontiguousarray(normals, np.float32), len(v), len(f), -1, 0, -1, -1,
                   None)


@pytest.mark.parametrize("sub", [2, 3])
def test_icosphere_edge_count_is_euler(oracle, sub):
    v, f = scenes.icosphere(sub, 0.5)
    g = GOLD["icosphere_subdiv%d" % sub]
    assert (len(v), len(f)) == (g["vertices"], g["faces"])
    d = scenes.ico_surface(nder=1, res=16, spp=1)
    d.shapes[1] = scenes.ShapeDesc(v, f, bsdf_id=1)
    sc = scenes.instantiate(oracle, d)
    n, length = oracle.scene_query(sc, "num_edges", 1)
    assert int(n) == g["edges"] == 3 * len(f) // 2
    e = np.array([sorted(x) for x in np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])])
    e = np.unique(e, axis=0)
    assert abs(length - np.linalg.norm(v[e[:, 0]].astype(np.float64) - v[e[:, 1]], axis=1).sum()) < 1e-4


def test_coplanar_diagonals_are_not_edges(oracle):
    d = scenes.ico_surface(nder=1, res=16, spp=1)
    sc = scenes.instantiate(oracle, d)
    n, length = oracle.scene_query(sc, "num_edges", 0)     # 4x4 ground quad: 4 boundary edges
    assert int(n) == 4 and abs(length - 16.0) < 1e-6


def test_bvh_matches_brute_force(oracle):
    d = scenes.ico_surface(nder=1, res=16, spp=1)
    sc = scenes.instantiate(oracle, d)
    rng = np.random.default_rng(0)
    n = 5000
    o = rng.uniform(-3, 3, (n, 3))
    t = rng.uniform(-0.6, 0.6, (n, 3)) + [0, 0.5, 0]
    dr = t - o
    dr /= np.linalg.norm(dr, axis=1, keepdims=True)
    rays = np.concatenate([o, dr, np.zeros((n, 1)), np.full((n, 1), np.inf)], 1)
    a = oracle.intersect(sc, rays)
    b = oracle.intersect(sc, rays, brute=True)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()
    hit = a[0] >= 0
    assert hit.mean() > 0.3 and np.allclose(a[2][hit], b[2][hit], rtol=1e-12)


def test_primary_edges_are_projected_vertices(oracle):
    d = scenes.ico_surface(nder=1, res=64, spp=1)
    sc = scenes.instantiate(oracle, d)
    cnt, total = oracle.scene_query(sc, "primary_edges")
    assert int(cnt) > 0
    c2w = d.cam_to_world.astype(np.float64)
    w2c = np.linalg.inv(c2w)
    fov = d.cam_to_ndc[0, 0]

    def project(p):
        q = w2c[:3, :3] @ p + w2c[:3, 3]
        x, y = q[0] / q[2], q[1] / q[2]
        return np.array([(x * fov * 0.5 + 0.5) * d.width, (y * fov * -0.5 + 0.5) * d.height])

    verts = [(s.vertices.astype(np.float64), s.velocities) for s in d.shapes]
    lengths = 0.0
    matched = 0
    for k in range(int(cnt)):
        e = oracle.scene_query(sc, "primary_edge", k)
        p0, p1, g0 = e[0:2], e[2:4], e[4:6]
        lengths += np.linalg.norm(p1 - p0)
        # find a vertex projecting onto p0 (unclipped endpoints)
        for V, vel in verts:
            pr = np.array([project(v) for v in V])
            j = np.argmin(np.linalg.norm(pr - p0, axis=1))
            if np.linalg.norm(pr[j] - p0) < 1e-6:
                h = 1e-6
                dv = vel[0, j] if vel is not None else np.zeros(3)
                fd = (project(V[j] + h * dv) - project(V[j] - h * dv)) / (2 * h)
                assert np.allclose(g0, fd, rtol=1e-4, atol=1e-4)
                matched += 1
                break
    assert matched > 0.9 * cnt
    assert abs(lengths - total) < 1e-6 * total


def test_light_pmf_proportional_to_power(oracle):
    d = scenes.cube_medium(nder=3, res=16, spp=1, grid_res=8)
    sc = scenes.instantiate(oracle, d)
    pmf = oracle.scene_query(sc, "light_pmf")

    def area(s):
        v = s.vertices.astype(np.float64)
        f = s.indices
        return 0.5 * np.linalg.norm(np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]]), axis=1).sum()

    p = np.array([area(d.shapes[1]) * 20.0, area(d.shapes[2]) * 0.5])
    assert np.allclose(pmf, p / p.sum(), rtol=1e-9)

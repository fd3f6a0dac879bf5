"""SceneManager / SceneTransform over scene descriptions (pydtrt/scene_manager.py).

Every tangent input the manager writes for lane i must be the derivative of the
parameter values it writes, with respect to variable i: checked by central finite
differences of set_arguments itself, for each transform type.
"""
import numpy as np
import pytest

from dtrt_amd import scenes
from dtrt_amd.harness import SceneManager, SceneTransform, rotation3
from dtrt_amd.scenes import MediumDesc, SceneDesc, ShapeDesc


def _desc(nder):
    v, f = scenes.icosphere(1, 0.5, (0.2, 0.1, -0.3))
    n = v - np.array([0.2, 0.1, -0.3], np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    q_v, q_f = scenes.quad((0, 2, 0), 1.0, 1, -1)
    w, g = scenes.icosphere(1, 0.3, (-0.6, 0.2, 0.1))
    m = (w - np.array([-0.6, 0.2, 0.1], np.float32)) / 0.3
    return SceneDesc(nder=nder, width=8, height=8, cam_to_world=scenes.look_at((0, 1, 4), (0, 0, 0), (0, 1, 0)),
                     cam_to_ndc=scenes.cam_to_ndc(40), shapes=[ShapeDesc(v, f, 0, normals=n.astype(np.float32)),
                                                              ShapeDesc(q_v, q_f, 2, light_id=0),
                                                              ShapeDesc(w, g, 2, normals=m.astype(np.float32))],
                     bsdfs=[("phong", (0.5, 0.4, 0.3), (0.2, 0.2, 0.2), 20.0), ("roughdielectric", 0.2, 1.5, 1.0),
                            ("diffuse", (0.1, 0.2, 0.3))],
                     lights=[("areaEx", 1, (5.0, 5.0, 5.0), 2.0)], phases=[("hg", 0.3), ("iso",)],
                     media=[MediumDesc("homogeneous", phase_id=1, sigma_t=0.7, albedo=(0.5, 0.6, 0.7)),
                            MediumDesc("heterogeneous", phase_id=0, density=np.ones((4, 4, 4), np.float32),
                                       bmin=(-1, -1, -1), bmax=(1, 2, 1), scale=2.0, albedo=(0.8, 0.8, 0.8))])


TRANSFORMS = {
    "camera": [[SceneTransform("CAMERA_TRANSLATE", [0.3, 0.0, -0.2])], [SceneTransform("CAMERA_ROTATE", [0.1, 0.4, 0.0])]],
    # Velocities are exact derivatives when every later transform of the same shape is a
    # translation; the reference computes them that way too (scene_manager.py:564-598).
    "shape": [[SceneTransform("SHAPE_ROTATE", [0.0, 1.0, 0.5], 0)],
              [SceneTransform("SHAPE_GLOBAL_ROTATE", [0.3, 0.0, 0.2], 2)],
              [SceneTransform("VERTEX_TRANSLATE", [1, 2, 3], 0, 5)],
              [SceneTransform("SHAPE_TRANSLATE", [0, 1, 0], 2)]],
    "material": [[SceneTransform("BSDF_VARY", "phong", 0, [0.1, 0, 0], [0, 0.1, 0], 2.0)],
                 [SceneTransform("BSDF_VARY", "roughdielectric", 1, 0.1, 0.05)],
                 [SceneTransform("BSDF_VARY", "diffuse", 2, [0.1, 0.1, 0.0]), SceneTransform("PHASE_VARY", "hg", 0, 0.2)],
                 [SceneTransform("EMITTER_VARY", "area_lightEx", 0, [1, 0, 0], 0.5)]],
    "medium": [[SceneTransform("MEDIUM_VARY", "homogeneous", 0, 0.3, [0.1, 0, 0])],
               [SceneTransform("MEDIUM_VARY", "heterogeneous", 1, [0.2, 0, 0.1], [0, 0.5, 0.2], [0, 0.1, 0], 0.4)]],
}


def _params(sm):
    d = sm.desc
    out = {"c2w": d.cam_to_world.astype(np.float64), "v0": d.shapes[0].vertices.astype(np.float64),
           "n0": d.shapes[0].normals.astype(np.float64), "v2": d.shapes[2].vertices.astype(np.float64),
           "n2": d.shapes[2].normals.astype(np.float64)}
    out["phong"] = np.r_[d.bsdfs[0][1], d.bsdfs[0][2], d.bsdfs[0][3]]
    out["rd"] = np.r_[d.bsdfs[1][1], d.bsdfs[1][2] / d.bsdfs[1][3]]
    out["diff"] = np.r_[d.bsdfs[2][1]]
    out["g"] = np.r_[d.phases[0][1]]
    out["em"] = np.r_[d.lights[0][2], d.lights[0][3]]
    out["hom"] = np.r_[d.media[0].sigma_t, d.media[0].albedo]
    out["het"] = np.r_[d.media[1].scale, d.media[1].albedo]
    tw = np.asarray(d.media[1].to_world if d.media[1].to_world is not None else np.eye(4), np.float64)
    out["het_pt"] = (tw @ [0.3, -0.2, 0.5, 1.0])[:3]
    return out


def _tangents(sm, i):
    d, nder = sm.desc, sm.desc.nder
    cv = d.cam_vel if d.cam_vel is not None else np.zeros((2, nder, 3))
    c2w = d.cam_to_world.astype(np.float64)
    t = {"c2w_t": cv[0][i], "c2w_R": np.cross(cv[1][i], c2w[:3, :3].T).T}
    for k in (0, 2):
        vel = d.shapes[k].velocities
        t["v%d" % k] = vel[i] if vel is not None else 0.0
        t["n%d" % k] = vel[nder + i] if vel is not None else 0.0
    b = d.bsdfs
    lane = lambda x, k: 0.0 if len(x) <= k or x[k] is None else np.asarray(x[k])[i]
    t["phong"] = np.r_[np.broadcast_to(lane(b[0], 4), 3), np.broadcast_to(lane(b[0], 5), 3), lane(b[0], 6)]
    t["rd"] = np.r_[lane(b[1], 4), lane(b[1], 5)]
    t["diff"] = np.broadcast_to(lane(b[2], 2), 3)
    t["g"] = np.r_[lane(d.phases[0], 2)]
    t["em"] = np.r_[np.broadcast_to(lane(d.lights[0], 4), 3), lane(d.lights[0], 5)]
    m0, m1 = d.media
    z3 = np.zeros(3)
    t["hom"] = np.r_[m0.d_scale[i] if m0.d_scale is not None else 0.0, m0.d_albedo[i] if m0.d_albedo is not None else z3]
    t["het"] = np.r_[m1.d_scale[i] if m1.d_scale is not None else 0.0, m1.d_albedo[i] if m1.d_albedo is not None else z3]
    if m1.trans_vel is not None:
        c = sm._centers_vol_now[1]
        x = _params(sm)["het_pt"]
        t["het_pt"] = m1.trans_vel[i] + np.cross(m1.rot_vel[i], x - c)
    else:
        t["het_pt"] = z3
    return t


@pytest.mark.parametrize("group", sorted(TRANSFORMS))
def test_tangent_inputs_are_derivatives_of_the_parameter_map(group):
    tr = TRANSFORMS[group]
    nder = len(tr)
    sm = SceneManager(_desc(nder), tr)
    rng = np.random.default_rng(len(group))
    theta = rng.normal(scale=0.3, size=nder)
    h = 1e-4
    for i in range(nder):
        sm.set_arguments(theta)
        sm._centers_vol_now = [None, sm.volume_centers_0[1] + sum(
            T.vec_translate * theta[j] for j, ts in enumerate(tr) for T in ts if T.type_name == "MEDIUM_VARY"
            and T.med_type == "heterogeneous")]
        tan = _tangents(sm, i)
        up, dn = theta.copy(), theta.copy()
        up[i] += h
        dn[i] -= h
        sm.set_arguments(up)
        pu = _params(sm)
        sm.set_arguments(dn)
        pd = _params(sm)
        for k in pu:
            fd = (pu[k] - pd[k]) / (2 * h)
            if k == "c2w":
                assert np.allclose(fd[:3, 3], tan["c2w_t"], atol=2e-3), (group, i, k)
                assert np.allclose(fd[:3, :3], tan["c2w_R"], atol=2e-3), (group, i, k)
            else:
                assert np.allclose(fd, tan[k], atol=2e-3), (group, i, k, fd, tan[k])


def test_shape_translate_and_reset():
    d = _desc(2)
    sm = SceneManager(d, [[SceneTransform("SHAPE_TRANSLATE", [1, 0, 0], 0)],
                          [SceneTransform("BSDF_VARY", "roughdielectric", 1, 0.1, 0.0)]])
    sm.set_arguments([-2.0, 0.5])
    assert np.allclose(sm.desc.shapes[0].vertices, d.shapes[0].vertices + [-2.0, 0, 0], atol=1e-6)
    assert sm.desc.bsdfs[1][1] == pytest.approx(0.25) and sm.desc.bsdfs[1][2] == pytest.approx(1.5)
    vel = sm.desc.shapes[0].velocities
    assert vel.shape == (4, len(d.shapes[0].vertices), 3)      # normals present: 2*nder blocks
    assert np.all(vel[0] == [1, 0, 0]) and not vel[1:].any()
    sm.reset()
    assert np.array_equal(sm.desc.shapes[0].vertices, d.shapes[0].vertices)
    assert sm.desc.bsdfs[1][1] == pytest.approx(0.2)


def test_rotation3_matches_rodrigues():
    R = rotation3([0, 0, 2.0], np.pi / 4)          # |axis| * t = pi/2 about z
    assert np.allclose(R @ [1, 0, 0], [0, 1, 0], atol=1e-12)
    assert np.allclose(rotation3([0, 0, 1e-5], 3.0), np.eye(3))


RANGED = [[SceneTransform("BSDF_VARY", "diffuse", 2, [0.1, -0.4, 0.0])],
          [SceneTransform("MEDIUM_VARY", "homogeneous", 0, -1.4, [0.5, 0, 0])],
          [SceneTransform("MEDIUM_VARY", "heterogeneous", 1, [0, 0, 0], [0, 0, 0], [0, 0, -0.2], 0.5)],
          [SceneTransform("BSDF_VARY", "roughdielectric", 1, -0.05, 0.1)]]


def test_check_range_intervals():
    # Hand-derived (scene_manager.py:69-143, 416-437):
    #  var 0: reflectance (0.1, 0.2, 0.3) + t (0.1, -0.4, 0) in [0, 1]  -> t in [-1, 0.5]
    #  var 1: albedo 0.5 + 0.5 t in [0, 1] -> [-1, 1]; sigma_t 0.7 - 1.4 t >= 0 -> t <= 0.5
    #  var 2: albedo 0.8 - 0.2 t in [0, 1] -> [-1, 4]; scale 2 + 0.5 t >= 0 -> t >= -4
    #  var 3: alpha 0.2 - 0.05 t in [0.01, 0.1] -> [2, 3.8]; IOR 1.5 + 0.1 t >= 0 -> t >= -15
    sm = SceneManager(_desc(4), RANGED, check_range=True)
    assert np.allclose(sm.ranges, [[-1, 0.5], [-1, 0.5], [-1, 4], [2, 3.8]])


def test_check_range_reference_quirks_give_empty_ranges():
    # A negative phong exponent rate and a positive alpha rate bound tmax by tmin in
    # the reference, so their ranges are always empty (its assertion fails).
    for T in (SceneTransform("BSDF_VARY", "phong", 0, [0, 0, 0], [0, 0, 0], -2.0),
              SceneTransform("BSDF_VARY", "roughdielectric", 1, 0.05, 0.0)):
        with pytest.raises(ValueError):
            SceneManager(_desc(1), [[T]], check_range=True)
    # the same transforms with the other sign are bounded on one side only
    sm = SceneManager(_desc(1), [[SceneTransform("BSDF_VARY", "phong", 0, [0, 0, 0], [0, 0, 0], 2.0)]], check_range=True)
    assert sm.ranges == [[-10.0, np.inf]]


def test_check_range_clamps_values_drops_tangents_and_returns_the_penalty():
    sm = SceneManager(_desc(4), RANGED, check_range=True)
    g = sm.set_arguments([0.7, -3.0, 0.0, 2.5])
    # clamped: var 0 to 0.5 (width 1.5), var 1 to -1 (width 1.5); 2 (var - bound) width
    assert np.allclose(g, [2 * 0.2 * 1.5, 2 * -2.0 * 1.5, 0.0, 0.0])
    d = sm.desc
    assert np.allclose(d.bsdfs[2][1], [0.15, 0.0, 0.3])
    assert len(d.bsdfs[2]) == 2 or not np.any(d.bsdfs[2][2])          # no reflectance tangent
    m0, m1 = d.media
    assert m0.sigma_t == pytest.approx(0.7 + 1.4) and np.allclose(m0.albedo, [0.0, 0.6, 0.7])
    assert m0.d_scale is None or not np.any(m0.d_scale[1])
    assert np.allclose(m1.d_albedo[2], [0, 0, -0.2]) and m1.d_scale[2] == pytest.approx(0.5)
    assert d.bsdfs[1][1] == pytest.approx(0.2 - 0.05 * 2.5) and d.bsdfs[1][4][3] == pytest.approx(-0.05)
    # inside every range nothing is clamped and every lane has its tangents
    g = sm.set_arguments([0.2, 0.1, 1.0, 3.0])
    assert not np.any(g) and d is not sm.desc
    assert np.allclose(sm.desc.bsdfs[2][2][0], [0.1, -0.4, 0.0])
    assert sm.desc.media[0].d_scale[1] == pytest.approx(-1.4)

"""Small analytic scenes shared by the oracle and GPU tests, with their closed forms."""
import math

import numpy as np

from dtrt_amd import scenes
from dtrt_amd.scenes import MediumDesc, SceneDesc, ShapeDesc, cam_to_ndc, cube, icosphere, look_at, quad


def furnace(nder=1, heterogeneous=False, sigma=0.5, res=16, spp=16, depth=30):
    """Camera inside a uniform two-sided emitting sphere (radiance 1) filled with a
    non-absorbing isotropic medium: L = 1 everywhere and every derivative of the
    medium's density vanishes."""
    v, f = scenes.icosphere(3, 1.0, (0, 0, 0), inward=True)
    shapes = [ShapeDesc(v, f, bsdf_id=0, light_id=0)]
    d_scale = np.zeros(nder, np.float32)
    d_scale[0] = 1.0
    if heterogeneous:
        med = MediumDesc("heterogeneous", phase_id=0, density=np.full((6, 6, 6), sigma, np.float32),
                         bmin=(-1.2, -1.2, -1.2), bmax=(1.2, 1.2, 1.2), scale=1.0, albedo=(1, 1, 1), d_scale=d_scale)
    else:
        med = MediumDesc("homogeneous", phase_id=0, sigma_t=sigma, albedo=(1, 1, 1), d_scale=d_scale)
    return SceneDesc(nder=nder, width=res, height=res, cam_to_world=scenes.look_at((0, 0, 0), (0, 0, -1), (0, 1, 0)),
                     cam_to_ndc=scenes.cam_to_ndc(60.0), shapes=shapes, bsdfs=[("diffuse", (0.0, 0.0, 0.0))],
                     lights=[("area", 0, (1.0, 1.0, 1.0), True)], phases=[("iso",)], media=[med], cam_med=0,
                     spp=spp, max_bounces=depth, sppe=-1, sppse=0, name="furnace")


def direct_lighting(res=24, spp=64, rho=0.5, Le=4.0):
    """Diffuse ground seen from above, lit by a one-sided square light outside the view; depth 1."""
    g_v, g_f = scenes.quad((0, 0, 0), 4.0, 1, +1)
    l_v, l_f = scenes.quad((1.5, 1.0, 0), 1.0, 1, -1)
    shapes = [ShapeDesc(g_v, g_f, bsdf_id=0), ShapeDesc(l_v, l_f, bsdf_id=1, light_id=0)]
    d = SceneDesc(nder=1, width=res, height=res, cam_to_world=scenes.look_at((0, 3, 0), (0, 0, 0), (0, 0, -1)),
                  cam_to_ndc=scenes.cam_to_ndc(40.0), shapes=shapes,
                  bsdfs=[("diffuse", (rho, rho, rho)), ("diffuse", (0.0, 0.0, 0.0))],
                  lights=[("area", 1, (Le, Le, Le), False)], spp=spp, max_bounces=1, sppe=-1, sppse=0, name="direct")
    return d, l_v, rho, Le


def pixel_rays(d):
    """Pixel-centre primary rays (camera.cpp:52-60)."""
    c2w = d.cam_to_world.astype(np.float64)
    ndc2c = np.linalg.inv(d.cam_to_ndc.astype(np.float64))
    ys, xs = np.mgrid[0:d.height, 0:d.width] + 0.5
    x, y = xs / d.width, ys / d.height
    aspect = d.width / d.height
    ndc = np.stack([(x - 0.5) * 2, (y - 0.5) * -2 / aspect, np.ones_like(x)], -1)
    dirs = ndc @ ndc2c.T
    dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
    dirs = dirs @ c2w[:3, :3].T
    return c2w[:3, 3], dirs


def polygon_irradiance(x, n, poly, Le):
    """Irradiance at x (normal n) from a uniform Lambertian polygon of radiance Le (Lambert's formula)."""
    E = 0.0
    k = len(poly)
    for i in range(k):
        a = poly[i] - x
        b = poly[(i + 1) % k] - x
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        c = np.cross(a, b)
        s = np.linalg.norm(c)
        E += math.atan2(s, np.dot(a, b)) * np.dot(n, c / s)
    return 0.5 * Le * abs(E)


def direct_expected(d, light_v, rho, Le):
    o, dirs = pixel_rays(d)
    t = -o[1] / dirs[..., 1]
    hits = o + dirs * t[..., None]
    poly = light_v.astype(np.float64)[[0, 1, 2, 3]]
    out = np.zeros(dirs.shape[:2])
    for i in range(d.height):
        for j in range(d.width):
            out[i, j] = rho / math.pi * polygon_irradiance(hits[i, j], np.array([0, 1.0, 0]), poly, Le)
    return out


def absorbing_slab(heterogeneous=False, sigma=1.3, res=16, spp=4, Le=2.0):
    """Light seen through an absorbing (albedo 0) cube: L = Le exp(-sigma len).

    The lane differentiates sigma_t (homogeneous: dL = -len L) or the density scale
    s of sigma = s * rho (heterogeneous, s = 1: dL = -rho len L)."""
    c_v, c_f = scenes.cube()
    l_v, l_f = scenes.quad((0, 0, -1.0), 6.0, 2, +1)
    d_scale = np.ones(1, np.float32)
    if heterogeneous:
        med = MediumDesc("heterogeneous", phase_id=0, density=np.full((5, 5, 5), sigma, np.float32), scale=1.0,
                         albedo=(0, 0, 0), d_scale=d_scale)
    else:
        med = MediumDesc("homogeneous", phase_id=0, sigma_t=sigma, albedo=(0, 0, 0), d_scale=d_scale)
    shapes = [ShapeDesc(c_v, c_f, bsdf_id=0, med_int=0), ShapeDesc(l_v, l_f, bsdf_id=1, light_id=0)]
    d = SceneDesc(nder=1, width=res, height=res, cam_to_world=scenes.look_at((0.2, 0.1, 3), (0, 0, 0), (0, 1, 0)),
                  cam_to_ndc=scenes.cam_to_ndc(30.0), shapes=shapes, bsdfs=[("null",), ("diffuse", (0, 0, 0))],
                  lights=[("area", 1, (Le, Le, Le), False)], phases=[("iso",)], media=[med], spp=spp, max_bounces=4,
                  sppe=-1, sppse=0, name="slab")
    o, dirs = pixel_rays(d)
    inv = 1.0 / dirs
    t0 = (-0.5 - o) * inv
    t1 = (0.5 - o) * inv
    tn = np.maximum(np.minimum(t0, t1).max(-1), 0)
    tf = np.maximum(t0, t1).min(-1)
    length = np.maximum(tf - tn, 0)
    L = Le * np.exp(-sigma * length)
    return d, L, -(sigma if heterogeneous else 1.0) * length * L


def ico_surface_reflectance_lane(res=24, spp=8, rho=0.8):
    """ico_surface with its single lane on the sphere's diffuse reflectance (nothing moves)."""
    d = scenes.ico_surface(nder=1, res=res, spp=spp)
    d.shapes[1].velocities = None
    d.bsdfs[1] = ("diffuse", (rho, rho, rho), np.ones((1, 3), np.float32))
    return d


def cube_medium_albedo_lane(res=16, spp=8, albedo=0.8, grid_res=16, crop=None):
    """cube_medium with its single lane on the medium's albedo (nothing moves)."""
    d = scenes.cube_medium(nder=1, res=res, spp=spp, grid_res=grid_res, crop=crop)
    d.shapes[0].velocities = None
    m = d.media[0]
    m.trans_vel = None
    m.rot_vel = None
    m.albedo = (albedo, albedo, albedo)
    m.d_albedo = np.ones((1, 3), np.float32)
    return d


def translating_emitter(res=32, spp=4, Le=1.0, sppe=4096):
    """An emissive quad (1x1 at the origin, facing the camera at z = 3) translating in x in
    front of an empty background. Only the primary-edge pass contributes to the tangent;
    its left/right half-image sums are -/+ Le * (edge height in px) * (px per unit)."""
    l_v, l_f = scenes.quad((0, 0, 0), 1.0, 2, +1)
    vel = np.zeros((1, 4, 3), np.float32)
    vel[0, :, 0] = 1.0
    d = SceneDesc(nder=1, width=res, height=res, cam_to_world=scenes.look_at((0, 0, 3), (0, 0, 0), (0, 1, 0)),
                  cam_to_ndc=scenes.cam_to_ndc(40.0), shapes=[ShapeDesc(l_v, l_f, bsdf_id=0, light_id=0, velocities=vel)],
                  bsdfs=[("diffuse", (0.0, 0.0, 0.0))], lights=[("area", 0, (Le, Le, Le), False)], spp=spp,
                  max_bounces=2, sppe=sppe, sppse=0, name="translating_emitter")
    f = d.cam_to_ndc[0, 0]
    ppu = 0.5 * f * res / 3.0                 # pixels per world unit at depth 3
    expected_half = Le * (1.0 * ppu) * ppu     # edge height in px times screen velocity
    return d, expected_half


def translating_light(res=16, spp=256, rho=0.5, Le=4.0, velocity=(0.0, 0.0, 1.0)):
    """Diffuse ground lit by a quad light moving with `velocity`; depth 1, pixel centres.
    The tangent is carried only by the light's boundary edges (secondary edge sampling);
    per pixel it equals d/dtheta of rho/pi * E(x) from Lambert's formula."""
    d, l_v, rho, Le = direct_lighting(res=res, spp=spp, rho=rho, Le=Le)
    vel = np.zeros((1, 4, 3), np.float32)
    vel[0, :, :] = velocity
    d.shapes[1].velocities = vel
    d.sppse = 10
    d.edge_weights = np.array([0.0, 1.0], np.float32)   # sample the light's edges only
    o, dirs = pixel_rays(d)
    t = -o[1] / dirs[..., 1]
    hits = o + dirs * t[..., None]
    poly = l_v.astype(np.float64)
    h = 1e-5
    v = np.asarray(velocity, np.float64)
    dE = np.zeros(dirs.shape[:2])
    for i in range(d.height):
        for j in range(d.width):
            ep = polygon_irradiance(hits[i, j], np.array([0, 1.0, 0]), poly + h * v, Le)
            em = polygon_irradiance(hits[i, j], np.array([0, 1.0, 0]), poly - h * v, Le)
            dE[i, j] = rho / math.pi * (ep - em) / (2 * h)
    return d, dE


def medium_box(move="light", res=12, spp=64, sigma=1.0, albedo=0.8, depth=3, nder=1):
    """A homogeneous isotropic medium in a null cube, lit by a quad light above, on a dark
    background. move = "light": the light translates in x (volume edge term);
    move = "box": the cube stretches along z around its homogeneous medium (Lins boundary term). The
    narrow view keeps every camera ray inside the cube's silhouette: the reference
    keeps null-BSDF shapes out of the primary-edge set (edge_manager.cpp:11), so a
    silhouette crossing the view would be missing from the tangent by design."""
    c_v, c_f = scenes.cube()
    l_v, l_f = scenes.quad((0.3, 1.2, 0), 0.8, 1, -1)
    shapes = [ShapeDesc(c_v, c_f, bsdf_id=0, med_int=0), ShapeDesc(l_v, l_f, bsdf_id=1, light_id=0)]
    def vel(n, axis):
        v = np.zeros((1, n, 3), np.float32)
        v[0, :, axis] = 1.0
        return v
    if move == "light":
        shapes[1].velocities = vel(4, 0)      # light slides along x
    else:                                     # box stretches along z: the exit distance of every camera ray changes
        v = np.zeros((1, 8, 3), np.float32)
        v[0, :, 2] = c_v[:, 2]
        shapes[0].velocities = v
    med = MediumDesc("homogeneous", phase_id=0, sigma_t=sigma, albedo=(albedo, albedo, albedo))
    return SceneDesc(nder=nder, width=res, height=res, cam_to_world=scenes.look_at((0, 0.2, 3), (0, 0, 0), (0, 1, 0)),
                     cam_to_ndc=scenes.cam_to_ndc(10.0), shapes=shapes, bsdfs=[("null",), ("diffuse", (0, 0, 0))],
                     lights=[("area", 1, (10.0, 10.0, 10.0), False)], phases=[("iso",)], media=[med], spp=spp,
                     max_bounces=depth, sppe=-1, sppse=10, edge_weights=np.array([0.0, 1.0], np.float32),
                     name="medium_box_" + move)


def shifted(d, h):
    """Copy of a scene description with every moving shape displaced by h along its lane-0 velocity."""
    import copy
    e = copy.deepcopy(d)
    for s in e.shapes:
        if s.velocities is not None:
            s.vertices = (s.vertices + h * s.velocities[0]).astype(np.float32)
            s.velocities = None
    return e


def feature_sweep(nder=4, res=40, spp=16, crop=(8, 8, 24, 24)):
    """One small scene that exercises the API surface the benchmark scenes leave out:
    Phong (kd, ks, exponent tangents), rough dielectric (alpha, eta tangents) inside a
    null cube holding a homogeneous HG medium (sigma_t, albedo, g tangents), an
    AreaLightEx light (intensity, kappa tangents), a translating and rotating camera,
    Lins with ddistCoeff > 0, and a cropped window."""
    g_v, g_f = quad((0, -0.5, 0), 4.0, 1, +1)
    c_v, c_f = cube(0.0, 1.0)
    lo, hi = np.array([-0.7, -0.45, -0.7], np.float32), np.array([0.7, 0.8, 0.7], np.float32)
    c_v = (lo + c_v * (hi - lo)).astype(np.float32)
    s_v, s_f = icosphere(2, 0.3, (0.0, 0.05, 0.0))
    l_v, l_f = quad((0.3, 1.6, 0.2), 0.8, 1, -1)
    lane = lambda j, v: np.array([v if k == j else (0.0 * v) for k in range(nder)], np.float32)
    phong = ("phong", (0.4, 0.35, 0.3), (0.3, 0.3, 0.3), 20.0, lane(0, np.array([1.0, 0.5, 0.0])),
             lane(1, np.array([0.0, 1.0, 1.0])), lane(2, 5.0))
    rough = ("roughdielectric", 0.25, 1.5, 1.0, lane(2, 1.0), lane(3, 1.0))
    shapes = [ShapeDesc(g_v, g_f, bsdf_id=0), ShapeDesc(c_v, c_f, bsdf_id=1, med_int=0),
              ShapeDesc(s_v, s_f, bsdf_id=2, med_ext=0), ShapeDesc(l_v, l_f, bsdf_id=3, light_id=0)]
    med = MediumDesc("homogeneous", phase_id=0, sigma_t=0.8, albedo=(0.7, 0.7, 0.7),
                     d_scale=lane(1, 1.0), d_albedo=lane(2, np.array([1.0, 0.0, 0.5])))
    cam_vel = np.zeros((2, nder, 3), np.float32)
    cam_vel[0, 1] = (1.0, 0.0, 0.0)      # lane 1: camera translation along x
    cam_vel[1, 3] = (0.0, 1.0, 0.0)      # lane 3: camera rotation about y
    return SceneDesc(nder=nder, width=res, height=res,
                     cam_to_world=look_at((0.4, 1.0, 3.0), (0, 0.1, 0), (0, 1, 0)), cam_to_ndc=cam_to_ndc(40.0),
                     shapes=shapes, bsdfs=[phong, ("null",), rough, ("diffuse", (0.0, 0.0, 0.0))],
                     lights=[("areaEx", 3, (15.0, 15.0, 15.0), 2.0, lane(3, np.array([1.0, 1.0, 1.0])), lane(0, 1.0))],
                     phases=[("hg", 0.4, lane(0, 1.0))], media=[med], cam_vel=cam_vel,
                     spp=spp, max_bounces=6, sppe=256, sppse=4, ddist=2.0, crop=crop, name="features")


def rotating_grid(nder=2, res=48, spp=16, grid_res=16, crop=(12, 12, 24, 24)):
    """The heterogeneous-medium scene with the grid (and its null cube) rotating about y
    in lane 0 and the density scale as lane 1."""
    from dtrt_amd.scenes import cube_medium
    d = cube_medium(nder=nder, res=res, spp=spp, grid_res=grid_res, crop=crop)
    c_v = d.shapes[0].vertices
    vel = np.zeros((nder, len(c_v), 3), np.float32)
    vel[0] = np.cross(np.array([0.0, 1.0, 0.0]), c_v)   # omega x p, centre at the origin
    d.shapes[0].velocities = vel
    m = d.media[0]
    m.trans_vel = np.zeros((nder, 3), np.float32)
    m.rot_vel = np.zeros((nder, 3), np.float32)
    m.rot_vel[0] = (0.0, 1.0, 0.0)
    m.d_scale = np.array([0.0, 1.0] + [0.0] * (nder - 2), np.float32)
    d.sppe, d.sppse, d.max_bounces = 256, 4, 6
    return d


def diffuse_furnace(albedo=0.5, Le=1.0, res=16, spp=16, depth=5):
    """Camera inside a closed inward-facing sphere that both emits Le (two-sided) and
    reflects diffusely with the given albedo; no media. Every path tracer that counts
    the emission seen at the first hit and estimates each further bounce without bias
    gives L = Le * sum_{k=0..depth} albedo^k in every pixel."""
    v, f = scenes.icosphere(3, 1.0, (0, 0, 0), inward=True)
    return SceneDesc(nder=1, width=res, height=res, cam_to_world=scenes.look_at((0, 0, 0), (0, 0, -1), (0, 1, 0)),
                     cam_to_ndc=scenes.cam_to_ndc(60.0), shapes=[ShapeDesc(v, f, bsdf_id=0, light_id=0)],
                     bsdfs=[("diffuse", (albedo, albedo, albedo))], lights=[("area", 0, (Le, Le, Le), True)],
                     spp=spp, max_bounces=depth, sppe=-1, sppse=0, name="diffuse furnace")

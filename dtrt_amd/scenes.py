"""Deterministic synthetic scenes for the five benchmark configurations.

Everything here is generated (no downloads): icospheres, quads, geodesic spheres,
seeded value-noise fBm density grids, VOL v3 files, look-at cameras, and the
per-lane velocity tensors in the layouts the reference expects
(``shape.cpp:37-51``, ``camera.cpp:29-37``, ``heterogeneous.h:187-201``).

A scene is a :class:`SceneDesc`; :func:`instantiate` turns it into renderer
objects through any module returned by :func:`dtrt_amd.abi.bind`.
"""
import math
import struct
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# ----------------------------------------------------------------- PCG32
_MUL = 6364136223846793005
_M64 = (1 << 64) - 1


class Pcg32:
    """PCG32 XSH-RR with the reference's seeding (sampler.cpp:3-9); used for scene generation."""

    def __init__(self, seed, idx):
        self.state = 0
        self.inc = (((idx + 1) << 1) | 1) & _M64
        self.next_u32()
        self.state = (self.state + 0x853C49E6748FEA9B + seed) & _M64
        self.next_u32()

    def next_u32(self):
        old = self.state
        self.state = (old * _MUL + (self.inc | 1)) & _M64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self, n):
        return np.array([self.next_u32() for _ in range(n)], dtype=np.float64) * (1.0 / 4294967296.0)


# -------------------------------------------------------------- geometry
def icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
                  [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(subdiv, radius=1.0, center=(0.0, 0.0, 0.0), inward=False):
    """Loop-free midpoint subdivision of the icosahedron; outward winding unless ``inward``."""
    v, f = icosahedron()
    verts = [tuple(x) for x in v]
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                p = (np.array(verts[a]) + np.array(verts[b])) * 0.5
                p /= np.linalg.norm(p)
                verts.append(tuple(p))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = np.array(nf, dtype=np.int64)
    v = np.array(verts) * radius + np.asarray(center, dtype=np.float64)
    if inward:
        f = f[:, [0, 2, 1]]
    return v.astype(np.float32), f.astype(np.int32)


def geodesic_sphere(freq, radius=1.0, center=(0.0, 0.0, 0.0)):
    """Class-I geodesic sphere: each icosahedron face split into freq^2 triangles (20*freq^2 faces), welded."""
    v0, f0 = icosahedron()
    pts, faces = [], []
    key_to_idx = {}

    def vid(p):
        q = p / np.linalg.norm(p)
        k = tuple(np.round(q * 1e9).astype(np.int64))
        if k not in key_to_idx:
            key_to_idx[k] = len(pts)
            pts.append(q)
        return key_to_idx[k]

    for a, b, c in f0:
        A, B, Cc = v0[a], v0[b], v0[c]
        grid = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                p = A + (B - A) * (i / freq) + (Cc - A) * (j / freq)
                grid[(i, j)] = vid(p)
        for i in range(freq):
            for j in range(freq - i):
                faces.append([grid[(i, j)], grid[(i + 1, j)], grid[(i, j + 1)]])
                if j + i + 1 < freq:
                    faces.append([grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]])
    v = np.array(pts) * radius + np.asarray(center, dtype=np.float64)
    return v.astype(np.float32), np.array(faces, dtype=np.int32)


def quad(center, size, normal_axis, normal_sign):
    """Axis-aligned square (2 triangles) whose face normal is ``normal_sign * e_axis``."""
    c = np.asarray(center, dtype=np.float64)
    a = [(normal_axis + 1) % 3, (normal_axis + 2) % 3]
    h = size / 2.0
    corners = []
    for s, t in [(-1, -1), (1, -1), (1, 1), (-1, 1)]:
        p = c.copy()
        p[a[0]] += s * h
        p[a[1]] += t * h
        corners.append(p)
    v = np.array(corners)
    f = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    n = np.cross(v[1] - v[0], v[2] - v[0])
    if np.sign(n[normal_axis]) != normal_sign:
        f = f[:, [0, 2, 1]]
    return v.astype(np.float32), f


def cube(lo=-0.5, hi=0.5):
    """Closed cube with outward face normals (12 triangles, 8 vertices)."""
    v = np.array([[x, y, z] for z in (lo, hi) for y in (lo, hi) for x in (lo, hi)], dtype=np.float64)
    f = []
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    for a, b, c, d in quads:
        f += [[a, b, c], [a, c, d]]
    f = np.array(f, dtype=np.int32)
    cen = v.mean(axis=0)
    for k in range(len(f)):
        p = v[f[k]]
        n = np.cross(p[1] - p[0], p[2] - p[0])
        if np.dot(n, p.mean(axis=0) - cen) < 0:
            f[k] = f[k][[0, 2, 1]]
    return v.astype(np.float32), f


def vertex_normals(v, f):
    n = np.zeros_like(v, dtype=np.float64)
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    return (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)


# ---------------------------------------------------------------- camera
def look_at(pos, target, up):
    """pydtrt's look-at convention (transform.py gen_look_at_matrix): columns right, up', dir, pos."""
    pos, target, up = [np.asarray(x, dtype=np.float64) for x in (pos, target, up)]
    d = target - pos
    d /= np.linalg.norm(d)
    right = np.cross(d, up / np.linalg.norm(up))
    right /= np.linalg.norm(right)
    nu = np.cross(right, d)
    nu /= np.linalg.norm(nu)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = right, nu, d, pos
    return m.astype(np.float32)


def cam_to_ndc(fov_deg):
    f = 1.0 / math.tan(math.radians(0.5 * fov_deg))
    return np.diag([f, f, 1.0]).astype(np.float32)


# --------------------------------------------------------------- volumes
def fbm_grid(res, seed, base_freq=4, octaves=4, lacunarity=2.0, gain=0.5, lo=0.0, hi=2.0):
    """Value-noise fBm sampled at the grid nodes p = i/(res-1), rescaled linearly to [lo, hi].

    Octave o uses a (F+1)^3 lattice of PCG32(seed, o) uniforms, F = base_freq*lacunarity^o,
    blended with smoothstep weights. Returned as float32[z][y][x].
    """
    x = np.linspace(0.0, 1.0, res)
    total = np.zeros((res, res, res))
    amp = 1.0
    for o in range(octaves):
        F = int(round(base_freq * lacunarity ** o))
        lat = Pcg32(seed, o).uniform((F + 1) ** 3).reshape(F + 1, F + 1, F + 1)
        u = x * F
        i0 = np.minimum(np.floor(u).astype(np.int64), F - 1)
        t = u - i0
        s = t * t * (3.0 - 2.0 * t)
        zi, yi, xi = np.meshgrid(i0, i0, i0, indexing="ij")
        zs, ys, xs = np.meshgrid(s, s, s, indexing="ij")
        acc = 0.0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w = (zs if dz else 1 - zs) * (ys if dy else 1 - ys) * (xs if dx else 1 - xs)
                    acc = acc + w * lat[zi + dz, yi + dy, xi + dx]
        total += amp * acc
        amp *= gain
    mn, mx = total.min(), total.max()
    return (lo + (total - mn) / (mx - mn) * (hi - lo)).astype(np.float32)


def write_vol(path, data, bmin, bmax):
    """VOL v3 writer (the format read by gridvolume.h:88-132): float32, 1 or 3 channels."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    ch = 1 if data.ndim == 3 else data.shape[3]
    rz, ry, rx = data.shape[:3]
    with open(path, "wb") as fh:
        fh.write(b"VOL")
        fh.write(struct.pack("<B", 3))
        fh.write(struct.pack("<i", 1))
        fh.write(struct.pack("<3i", rx, ry, rz))
        fh.write(struct.pack("<i", ch))
        fh.write(struct.pack("<3f", *bmin))
        fh.write(struct.pack("<3f", *bmax))
        fh.write(data.tobytes())


# ----------------------------------------------------------- descriptions
@dataclass
class ShapeDesc:
    vertices: np.ndarray
    indices: np.ndarray
    bsdf_id: int
    light_id: int = -1
    med_int: int = -1
    med_ext: int = -1
    normals: Optional[np.ndarray] = None
    velocities: Optional[np.ndarray] = None   # float32 [nder][nv][3] (+ normal blocks if normals)


@dataclass
class MediumDesc:
    kind: str                                  # 'heterogeneous' | 'homogeneous'
    phase_id: int = 0
    density: Optional[np.ndarray] = None       # float32 [z][y][x]
    bmin: tuple = (-0.5, -0.5, -0.5)
    bmax: tuple = (0.5, 0.5, 0.5)
    to_world: Optional[np.ndarray] = None
    scale: float = 1.0
    albedo: tuple = (0.8, 0.8, 0.8)
    sigma_t: float = 1.0
    trans_vel: Optional[np.ndarray] = None     # [nder][3]
    rot_vel: Optional[np.ndarray] = None       # [nder][3]
    d_scale: Optional[np.ndarray] = None       # [nder]
    d_albedo: Optional[np.ndarray] = None      # [nder][3]
    density_file: Optional[str] = None


@dataclass
class SceneDesc:
    nder: int
    width: int
    height: int
    cam_to_world: np.ndarray
    cam_to_ndc: np.ndarray
    shapes: List[ShapeDesc]
    bsdfs: list                                # ('diffuse', rgb[, d]) | ('null',) | ('phong', kd, ks, e) | ('roughdielectric', a, int, ext[, da, deta])
    lights: list                               # ('area', shape_id, rgb, two_sided) | ('areaEx', shape_id, rgb, kappa)
    phases: list = field(default_factory=list)  # ('hg', g[, dg]) | ('iso',)
    media: List[MediumDesc] = field(default_factory=list)
    clip_near: float = 1e-2
    cam_med: int = -1
    cam_vel: Optional[np.ndarray] = None       # [2][nder][3]
    crop: Optional[tuple] = None               # (ox, oy, w, h)
    edge_weights: Optional[np.ndarray] = None  # per shape
    spp: int = 16
    max_bounces: int = 5
    sppe: int = 4096
    sppse: int = 10
    ddist: float = 0.0
    name: str = ""

    @property
    def npix(self):
        return self.crop[2] * self.crop[3] if self.crop else self.width * self.height

    def out_shape(self):
        h, w = (self.crop[3], self.crop[2]) if self.crop else (self.height, self.width)
        return (self.nder + 1, h, w, 3)


def translate_velocity(nv, nder, lane, axis):
    vel = np.zeros((nder, nv, 3), dtype=np.float32)
    vel[lane, :, axis] = 1.0
    return vel


class _Keep:
    """Holds numpy buffers alive while renderer objects point at them."""

    def __init__(self):
        self.bufs = []

    def __call__(self, a, dtype=np.float32):
        if a is None:
            return None
        b = np.ascontiguousarray(a, dtype=dtype)
        self.bufs.append(b)
        return b


def instantiate(mod, d: SceneDesc):
    """Create Camera/Shapes/.../Scene objects for ``d`` through a bound module ``mod``."""
    keep = _Keep()
    if mod.nder != d.nder:
        mod.set_nder(d.nder)
    cam = mod.Camera(d.width, d.height, keep(d.cam_to_world), keep(d.cam_to_ndc), d.clip_near, d.cam_med,
                     keep(d.cam_vel))
    if d.crop:
        cam.set_rect(*d.crop)
    shapes = []
    for s in d.shapes:
        vel = s.velocities
        if vel is not None and s.normals is not None and vel.shape[0] == d.nder:
            vel = np.concatenate([vel, np.zeros_like(vel)], axis=0)
        shapes.append(mod.Shape(keep(s.vertices), keep(s.indices, np.int32), None, keep(s.normals),
                                len(s.vertices), len(s.indices), s.light_id, s.bsdf_id, s.med_int, s.med_ext,
                                keep(vel)))
    bsdfs = []
    for b in d.bsdfs:
        if b[0] == "diffuse":
            bsdfs.append(mod.BSDF_diffuse(b[1], keep(b[2]) if len(b) > 2 else None))
        elif b[0] == "null":
            bsdfs.append(mod.BSDF_null())
        elif b[0] == "phong":
            bsdfs.append(mod.BSDF_Phong(b[1], b[2], b[3], *[keep(x) for x in b[4:7]]))
        elif b[0] == "roughdielectric":
            bsdfs.append(mod.BSDF_roughdielectric(b[1], b[2], b[3], *[keep(x) for x in b[4:6]]))
        else:
            raise ValueError(b[0])
    lights = []
    for l in d.lights:
        if l[0] == "area":
            lights.append(mod.AreaLight(l[1], keep(np.asarray(l[2], np.float32)), l[3]))
        else:
            lights.append(mod.AreaLightEx(l[1], l[2], l[3], *[keep(x) for x in l[4:6]]))
    phases = []
    for p in d.phases:
        phases.append(mod.HG(p[1], keep(p[2]) if len(p) > 2 else None) if p[0] == "hg" else mod.Isotropic())
    media = []
    for m in d.media:
        if m.kind == "homogeneous":
            media.append(mod.Homogeneous(m.sigma_t, m.albedo, m.phase_id, keep(m.d_scale), keep(m.d_albedo)))
        elif m.density_file:
            media.append(mod.Heterogeneous(m.density_file, keep(m.to_world if m.to_world is not None else np.eye(4)),
                                           m.scale, m.albedo, m.phase_id, keep(m.trans_vel), keep(m.rot_vel),
                                           keep(m.d_scale), keep(m.d_albedo)))
        else:
            rz, ry, rx = m.density.shape
            media.append(mod.HeterogeneousGrid(keep(m.density), (rx, ry, rz), m.bmin, m.bmax,
                                               keep(m.to_world if m.to_world is not None else np.eye(4)), m.scale,
                                               m.albedo, m.phase_id, keep(m.trans_vel), keep(m.rot_vel),
                                               keep(m.d_scale), keep(m.d_albedo)))
    scene = mod.Scene(cam, shapes, bsdfs, lights, phases, media)
    scene.media_handles = media
    if d.edge_weights is not None:
        scene.initEdges(keep(d.edge_weights))
    scene._keep = keep
    return scene


def render(mod, d: SceneDesc, seed=0, spp=None, mode=-1, sppe=None, sppse=None, max_bounces=None, scene=None, out=None,
           integrator=None):
    """Render ``d`` with ``mod`` (product or oracle binding) and a forward-mode
    integrator (default VolPathTracerAD); returns float32[(nder+1), H, W, 3]."""
    scene = scene if scene is not None else instantiate(mod, d)
    opts = mod.RenderOptions(seed, d.spp if spp is None else spp, d.max_bounces if max_bounces is None else max_bounces,
                             d.sppe if sppe is None else sppe, d.sppse if sppse is None else sppse, True, mode, d.ddist)
    if out is None:
        out = np.zeros(d.out_shape(), dtype=np.float32)
    (integrator or mod.VolPathTracerAD()).render(scene, opts, out)
    return out


def render_primal(mod, d: SceneDesc, seed=0, spp=None, max_bounces=None, scene=None, simple=False, kind=None):
    """Primal render of ``d``; returns float32[H, W, 3]. ``kind``: 1 VolPathTracer
    (default), 2 VolPathTracerSimple (also ``simple=True``), 3 PathTracer,
    4 DirectIntegrator, 5 ParticleTracer (``spp`` blocks of 100000 particles)."""
    scene = scene if scene is not None else instantiate(mod, d)
    opts = mod.RenderOptions(seed, d.spp if spp is None else spp, d.max_bounces if max_bounces is None else max_bounces,
                             0, 0, True)
    out = np.zeros(d.out_shape()[1:], dtype=np.float32)
    kind = kind if kind is not None else (2 if simple else 1)
    cls = {1: mod.VolPathTracer, 2: mod.VolPathTracerSimple, 3: mod.PathTracer, 4: mod.DirectIntegrator,
           5: mod.ParticleTracer}[kind]
    cls().render(scene, opts, out)
    return out


# ------------------------------------------------------------- benchmark scenes
def ico_surface(nder=1, res=256, spp=64, crop=None):
    """Ground quad + icosphere (subdiv 3, r 0.5) + area-light quad; no media; sphere x-translation."""
    g_v, g_f = quad((0, 0, 0), 4.0, 1, +1)
    s_v, s_f = icosphere(3, 0.5, (0, 0.5, 0))
    l_v, l_f = quad((0, 2.5, 0), 1.0, 1, -1)
    vel = np.zeros((nder, len(s_v), 3), dtype=np.float32)
    vel[0, :, 0] = 1.0
    shapes = [ShapeDesc(g_v, g_f, bsdf_id=0), ShapeDesc(s_v, s_f, bsdf_id=1, velocities=vel),
              ShapeDesc(l_v, l_f, bsdf_id=2, light_id=0)]
    return SceneDesc(nder=nder, width=res, height=res,
                     cam_to_world=look_at((0, 1.5, 3.5), (0, 0.5, 0), (0, 1, 0)), cam_to_ndc=cam_to_ndc(45.0),
                     shapes=shapes, bsdfs=[("diffuse", (0.5, 0.5, 0.5)), ("diffuse", (0.8, 0.8, 0.8)),
                                           ("diffuse", (0.0, 0.0, 0.0))],
                     lights=[("area", 2, (10.0, 10.0, 10.0), False)], spp=spp, max_bounces=5, sppe=4096, sppse=10,
                     crop=crop, name="ico_surface")


def cube_medium(nder=3, res=512, spp=128, grid_res=64, crop=None, seed=7):
    """Null cube holding a 64^3 fBm heterogeneous medium, HG 0.3; quad light + enclosing emissive sphere."""
    c_v, c_f = cube()
    l_v, l_f = quad((0, 1.5, 0), 1.0, 1, -1)
    e_v, e_f = icosphere(2, 20.0, (0, 0, 0), inward=True)
    cvel = np.zeros((nder, len(c_v), 3), dtype=np.float32)
    tvel = np.zeros((nder, 3), dtype=np.float32)
    for k in range(min(nder, 3)):
        cvel[k, :, k] = 1.0
        tvel[k, k] = 1.0
    dens = fbm_grid(grid_res, seed)
    shapes = [ShapeDesc(c_v, c_f, bsdf_id=0, med_int=0, velocities=cvel),
              ShapeDesc(l_v, l_f, bsdf_id=1, light_id=0),
              ShapeDesc(e_v, e_f, bsdf_id=1, light_id=1)]
    med = MediumDesc("heterogeneous", phase_id=0, density=dens, scale=1.0, albedo=(0.8, 0.8, 0.8), trans_vel=tvel,
                     rot_vel=np.zeros((nder, 3), np.float32))
    return SceneDesc(nder=nder, width=res, height=res,
                     cam_to_world=look_at((0, 0, 3), (0, 0, 0), (0, 1, 0)), cam_to_ndc=cam_to_ndc(40.0),
                     shapes=shapes, bsdfs=[("null",), ("diffuse", (0.0, 0.0, 0.0))],
                     lights=[("area", 1, (20.0, 20.0, 20.0), False), ("area", 2, (0.5, 0.5, 0.5), True)],
                     phases=[("hg", 0.3)], media=[med], edge_weights=np.array([1.0, 1.0, 0.0], np.float32),
                     spp=spp, max_bounces=10, sppe=4096, sppse=10, ddist=0.0, crop=crop, name="cube_medium")


def orbit_views(nviews=32, seed=5, res_choices=tuple(range(128, 769, 64)), spp_choices=(16, 32, 64), radius=3.0):
    """Seeded orbit of cameras around the origin for batched rendering: per view
    (cam_to_world, width, height, spp). Resolutions and spp are drawn with PCG32(seed)."""
    g = Pcg32(seed, 0)
    views = []
    for i in range(nviews):
        u = g.uniform(4)
        phi = 2.0 * np.pi * (i + u[0]) / nviews
        elev = 0.6 * (u[1] - 0.5)
        eye = (radius * np.cos(elev) * np.sin(phi), radius * np.sin(elev), radius * np.cos(elev) * np.cos(phi))
        res = res_choices[min(int(u[2] * len(res_choices)), len(res_choices) - 1)]
        spp = spp_choices[min(int(u[3] * len(spp_choices)), len(spp_choices) - 1)]
        views.append((look_at(eye, (0, 0, 0), (0, 1, 0)), res, res, spp))
    return views


def dielectric_medium(nder=8, res=512, spp=256, freq=71, grid_res=128, crop=None):
    """Displaced geodesic sphere (rough dielectric, IOR 1.5) inside a 128^3 fBm medium; 8 deformation lanes."""
    s_v, s_f = geodesic_sphere(freq, 1.0)
    dirs = s_v.astype(np.float64)
    disp = _smooth_field(dirs, seed=11)
    s_v = (dirs * (0.5 + 0.03 * disp[:, None])).astype(np.float32)
    nrm = vertex_normals(s_v, s_f)
    vel = np.zeros((nder, len(s_v), 3), dtype=np.float32)
    for k in range(nder):
        vel[k] = nrm * _smooth_field(dirs, seed=100 + k)[:, None]
    c_v, c_f = cube(-1.0, 1.0)
    l1_v, l1_f = quad((0, 2.0, 0), 1.0, 1, -1)
    l2_v, l2_f = quad((2.0, 0.5, 0), 1.0, 0, -1)
    dens = fbm_grid(grid_res, 7)
    shapes = [ShapeDesc(s_v, s_f, bsdf_id=2, med_int=-1, med_ext=0, normals=nrm, velocities=vel),
              ShapeDesc(c_v, c_f, bsdf_id=0, med_int=0),
              ShapeDesc(l1_v, l1_f, bsdf_id=1, light_id=0),
              ShapeDesc(l2_v, l2_f, bsdf_id=1, light_id=1)]
    med = MediumDesc("heterogeneous", phase_id=0, density=dens, bmin=(-1, -1, -1), bmax=(1, 1, 1), scale=1.0,
                     albedo=(0.8, 0.8, 0.8))
    return SceneDesc(nder=nder, width=res, height=res,
                     cam_to_world=look_at((0, 0.3, 3.2), (0, 0, 0), (0, 1, 0)), cam_to_ndc=cam_to_ndc(40.0),
                     shapes=shapes, bsdfs=[("null",), ("diffuse", (0.0, 0.0, 0.0)),
                                           ("roughdielectric", 0.01, 1.5, 1.0)],
                     lights=[("area", 2, (15.0, 15.0, 15.0), False), ("area", 3, (8.0, 8.0, 8.0), False)],
                     phases=[("hg", 0.3)], media=[med], spp=spp, max_bounces=12, sppe=4096, sppse=10, crop=crop,
                     name="dielectric_medium")


def displaced_views(nder=8, freq=224, grid_res=256):
    """The batched inverse-rendering scene: dielectric_medium's construction at scale, a displaced
    geodesic sphere of frequency 224 (1,003,520 faces, rough dielectric) inside a 256^3
    fBm medium (64 MiB), 8 deformation lanes. Views come from orbit_views(32, seed 5,
    spp 16..256, radius 3.2)."""
    d = dielectric_medium(nder=nder, res=512, spp=64, freq=freq, grid_res=grid_res)
    d.name = "displaced_views"
    return d


def _smooth_field(dirs, seed, terms=6):
    """Seeded smooth scalar field on the unit sphere in [-1, 1]: a sum of random-direction cosines."""
    u = Pcg32(seed, 0).uniform(terms * 4).reshape(terms, 4)
    out = np.zeros(len(dirs))
    for k in range(terms):
        z = 1 - 2 * u[k, 0]
        r = math.sqrt(max(0.0, 1 - z * z))
        ph = 2 * math.pi * u[k, 1]
        a = np.array([r * math.cos(ph), r * math.sin(ph), z])
        out += np.cos((1.0 + 3.0 * u[k, 2]) * (dirs @ a) * math.pi + 2 * math.pi * u[k, 3])
    return out / terms

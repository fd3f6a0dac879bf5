"""Optimisation harness over scene descriptions: parameter transforms and the loss.

``SceneTransform`` and ``SceneManager`` follow pydtrt/scene_manager.py:11-631: a list
of ``nder`` transform sets, one per optimised variable, maps a variable vector
``var`` to scene parameters (vertex positions, camera pose, material, medium, phase
and emitter values) and to the tangent inputs of lane ``ivar`` (vertex velocities,
camera velocity, ``d_*`` arrays). Here they act on a
:class:`dtrt_amd.scenes.SceneDesc` instead of the flat argument list of
``serialize_scene``; :meth:`SceneManager.scene` instantiates the current scene
through a bound module.

``ADLossFunc`` follows pydtrt/auto_grad.py:21-74: the forward renders the AD image
on the GPU and returns its value channel; the backward returns, per variable, the
multi-scale pyramid dot product of its tangent image with the image gradient (with
clamping and the out-of-range penalty), computed on the device through
:func:`dtrt_amd.adjoint.pyramid_weights`.

``check_range=True`` bounds every variable by the interval in which its material and
medium values stay valid (scene_manager.py:69-143, 416-437); :meth:`set_arguments`
clamps out-of-range variables, leaves their tangent inputs at zero and returns the
quadratic penalty gradient (scene_manager.py:439-454).
"""
import copy
import math
import struct

import numpy as np

from . import scenes

_TYPES = ("CAMERA_TRANSLATE", "SHAPE_TRANSLATE", "VERTEX_TRANSLATE", "CAMERA_ROTATE", "SHAPE_ROTATE",
          "SHAPE_GLOBAL_ROTATE", "BSDF_VARY", "MEDIUM_VARY", "PHASE_VARY", "EMITTER_VARY")


def rotation3(axis, t):
    """Rotation by |axis|*t about axis (identity for |axis| <= 1e-4), as
    transform.gen_rotation_matrix3x3 (transform.py:97-110)."""
    a = np.asarray(axis, np.float64)
    w = math.sqrt(float(a @ a))
    if w <= 1e-4:
        return np.eye(3)
    u, th = a / w, w * float(t)
    c, s = math.cos(th), math.sin(th)
    K = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])
    return c * np.eye(3) + s * K + (1 - c) * np.outer(u, u)


def _vec(x):
    return np.asarray(x, np.float64).reshape(-1)


class SceneTransform:
    """One transform of one variable (scene_manager.py:11-67): same type names and
    positional arguments.

    CAMERA_TRANSLATE(vec) | CAMERA_ROTATE(axis) | SHAPE_TRANSLATE(vec, shape) |
    VERTEX_TRANSLATE(vec, shape, vertex) | SHAPE_ROTATE(axis, shape) |
    SHAPE_GLOBAL_ROTATE(axis, shape) | BSDF_VARY('diffuse', bsdf, d_refl) |
    BSDF_VARY('phong', bsdf, d_kd, d_ks, d_exp) | BSDF_VARY('roughdielectric', bsdf, d_alpha, d_eta) |
    MEDIUM_VARY('homogeneous', med, d_sigmaT, d_albedo) |
    MEDIUM_VARY('heterogeneous', med, translate, rotate, d_albedo, d_scale) |
    PHASE_VARY('hg', phase, d_g) | EMITTER_VARY('area_lightEx', emitter, d_intensity, d_kappa)
    """

    def __init__(self, type_name, *args):
        if type_name not in _TYPES:
            raise ValueError("Transform type [%s] not supported" % type_name)
        self.type_name, self.args = type_name, args
        if type_name in ("CAMERA_TRANSLATE", "CAMERA_ROTATE"):
            self.vec_vel = _vec(args[0])
        elif type_name in ("SHAPE_TRANSLATE", "SHAPE_ROTATE", "SHAPE_GLOBAL_ROTATE"):
            self.vec_vel, self.shape_id = _vec(args[0]), int(args[1])
        elif type_name == "VERTEX_TRANSLATE":
            self.vec_vel, self.shape_id, self.vert_id = _vec(args[0]), int(args[1]), int(args[2])
        elif type_name == "BSDF_VARY":
            self.bsdf_type, self.bsdf_id = args[0], int(args[1])
            if self.bsdf_type == "diffuse":
                self.d_reflectance = _vec(args[2])
            elif self.bsdf_type == "phong":
                self.d_diffuse, self.d_specular, self.d_exponent = _vec(args[2]), _vec(args[3]), float(args[4])
            elif self.bsdf_type == "roughdielectric":
                self.d_alpha, self.d_eta = float(args[2]), float(args[3])
            else:
                raise ValueError(self.bsdf_type)
        elif type_name == "MEDIUM_VARY":
            self.med_type, self.med_id = args[0], int(args[1])
            if self.med_type == "homogeneous":
                self.d_sigmaT, self.d_albedo = float(args[2]), _vec(args[3])
            elif self.med_type == "heterogeneous":
                self.vec_translate, self.vec_rotate = _vec(args[2]), _vec(args[3])
                self.d_albedo, self.d_scalar = _vec(args[4]), float(args[5])
            else:
                raise ValueError(self.med_type)
        elif type_name == "PHASE_VARY":
            if args[0] != "hg":
                raise ValueError(args[0])
            self.phase_type, self.phase_id, self.d_g = args[0], int(args[1]), float(args[2])
        elif type_name == "EMITTER_VARY":
            if args[0] != "area_lightEx":
                raise ValueError(args[0])
            self.emitter_type, self.emitter_id = args[0], int(args[1])
            self.d_intensity, self.d_kappa = _vec(args[2]), float(args[3])

    def determine_range(self, init_val, tmin, tmax):
        """Narrows [tmin, tmax] to the variable values that keep this transform's
        parameters valid (scene_manager.py:69-143): reflectance and albedo channels in
        [0, 1], non-negative phong lobes and exponent, roughdielectric alpha in
        [0.01, 0.1] and a non-negative IOR, non-negative sigma_t and density scale.
        ``init_val`` are the initial parameter values the reference passes. Two of its
        updates bound tmax by tmin (a negative phong exponent rate, a positive alpha
        rate); they are kept as they are."""
        def channels(val, rate, tmin, tmax):   # val + rate * t in [0, 1], per channel
            for ch in range(3):
                if rate[ch] == 0.0:
                    continue
                a = -val[ch] / rate[ch]
                b = (1.0 - val[ch]) / rate[ch]
                tmin = max(tmin, a if rate[ch] > 0 else b)
                tmax = min(tmax, b if rate[ch] > 0 else a)
            return tmin, tmax

        def nonneg(val, rate, tmin, tmax):     # val + rate * t >= 0
            if rate > 0.0:
                tmin = max(tmin, -val / rate)
            elif rate < 0.0:
                tmax = min(tmax, -val / rate)
            return tmin, tmax

        if self.type_name == "BSDF_VARY":
            if self.bsdf_type == "diffuse":
                tmin, tmax = channels(init_val, self.d_reflectance, tmin, tmax)
            elif self.bsdf_type == "phong":
                diffuse, specular, exponent = init_val
                for ch in range(3):
                    tmin, tmax = nonneg(diffuse[ch], self.d_diffuse[ch], tmin, tmax)
                    tmin, tmax = nonneg(specular[ch], self.d_specular[ch], tmin, tmax)
                if self.d_exponent > 0.0:
                    tmin = max(tmin, -exponent / self.d_exponent)
                elif self.d_exponent < 0.0:
                    tmax = min(tmin, -exponent / self.d_exponent)
            elif self.bsdf_type == "roughdielectric":
                a = (0.01 - init_val[0]) / self.d_alpha if self.d_alpha != 0.0 else 0.0
                b = (0.1 - init_val[0]) / self.d_alpha if self.d_alpha != 0.0 else 0.0
                if self.d_alpha > 0.0:
                    tmin = max(tmin, a)
                    tmax = min(tmin, b)
                elif self.d_alpha < 0.0:
                    tmin = max(tmin, b)
                    tmax = min(tmax, a)
                tmin, tmax = nonneg(init_val[1], self.d_eta, tmin, tmax)
        elif self.type_name == "MEDIUM_VARY":
            if self.med_type == "homogeneous":
                sigma_t, albedo = init_val
                tmin, tmax = channels(albedo, self.d_albedo, tmin, tmax)
                tmin, tmax = nonneg(sigma_t, self.d_sigmaT, tmin, tmax)
            else:
                albedo, scale = init_val
                if isinstance(albedo, str):
                    if np.any(self.d_albedo != 0.0):
                        raise ValueError("no range for a varying albedo read from a volume file")
                else:
                    tmin, tmax = channels(albedo, self.d_albedo, tmin, tmax)
                tmin, tmax = nonneg(scale, self.d_scalar, tmin, tmax)
        if not tmax > tmin:
            raise ValueError("empty variable range (%g, %g)" % (tmin, tmax))
        return tmin, tmax


def _volume_box(m):
    """World AABB centre of a heterogeneous medium's grid (scene_manager.py:276-292)."""
    if m.density_file:
        with open(m.density_file, "rb") as fh:
            head = fh.read(48)
        pmin = np.array(struct.unpack("<3f", head[24:36]))
        pmax = np.array(struct.unpack("<3f", head[36:48]))
    else:
        pmin, pmax = np.asarray(m.bmin, np.float64), np.asarray(m.bmax, np.float64)
    tw = np.eye(4) if m.to_world is None else np.asarray(m.to_world, np.float64)
    return (tw @ np.r_[(pmin + pmax) * 0.5, 1.0])[:3]


def _translate4(v):
    m = np.eye(4)
    m[:3, 3] = v
    return m


class SceneManager:
    """Maps a variable vector to scene parameters and tangent inputs
    (scene_manager.py:227-631), over a :class:`SceneDesc` whose ``nder`` equals the
    number of transform sets."""

    def __init__(self, desc, transforms, check_range=False):
        if desc.nder != len(transforms):
            raise ValueError("one transform set per derivative lane (nder = %d)" % desc.nder)
        self.desc0 = copy.deepcopy(desc)
        self.transforms = transforms
        self.shape_centers_0 = [s.vertices.astype(np.float64).sum(0) / len(s.vertices) for s in desc.shapes]
        self.volume_centers_0 = [None if m.kind == "homogeneous" else _volume_box(m) for m in desc.media]
        self.ranges = self.determine_range() if check_range else [[-math.inf, math.inf] for _ in range(desc.nder)]
        self.set_arguments(np.zeros(desc.nder))

    def determine_range(self):
        """Per variable, the intersection of its transforms' valid intervals at the
        initial parameter values (scene_manager.py:416-437)."""
        ranges = []
        for iv, tset in enumerate(self.transforms):
            tmin, tmax = -math.inf, math.inf
            for T in tset:
                if T.type_name == "BSDF_VARY":
                    b = self.desc0.bsdfs[T.bsdf_id]
                    if T.bsdf_type == "diffuse":
                        init = np.asarray(b[1], np.float64)
                    elif T.bsdf_type == "phong":
                        init = (np.asarray(b[1], np.float64), np.asarray(b[2], np.float64), float(b[3]))
                    else:
                        init = (float(b[1]), float(b[2]))
                    tmin, tmax = T.determine_range(init, tmin, tmax)
                elif T.type_name == "MEDIUM_VARY":
                    m = self.desc0.media[T.med_id]
                    alb = m.albedo if isinstance(m.albedo, str) else np.asarray(m.albedo, np.float64)
                    if T.med_type == "homogeneous":
                        tmin, tmax = T.determine_range((float(m.sigma_t), alb), tmin, tmax)
                    else:
                        tmin, tmax = T.determine_range((alb, float(m.scale)), tmin, tmax)
            ranges.append([tmin, tmax])
            print("param #%d range = (%.2f, %.2f)" % (iv, tmin, tmax))
        return ranges

    def reset(self):
        self.set_arguments(np.zeros(self.desc0.nder))

    def set_edge_weight(self, weights):
        self.desc.edge_weights = np.ascontiguousarray(weights, np.float32)

    def set_arguments(self, var_list):
        """Rebuild the current description for ``var_list``; returns the out-of-range
        gradient vector (scene_manager.py:439-454): a variable outside its range is
        clamped to the nearest bound, gets no tangent inputs, and its entry is
        2 (var - bound)(range width)."""
        nder = self.desc0.nder
        var = np.asarray([float(v) for v in var_list], np.float64)
        var_grad = np.zeros(nder, np.float32)
        clamped = [False] * nder
        for iv in range(nder):
            lo, hi = self.ranges[iv]
            if var[iv] < lo or var[iv] > hi:
                clamped[iv] = True
                bound = lo if var[iv] < lo else hi
                var_grad[iv] = 2.0 * (var[iv] - bound) * (hi - lo)
                print("Variable #%d clamped from %.2f to %.2f" % (iv, var[iv], bound))
                var[iv] = bound
        d = copy.deepcopy(self.desc0)
        centers = [c.copy() for c in self.shape_centers_0]
        centers_vol = [None if c is None else c.copy() for c in self.volume_centers_0]
        c2w0 = np.asarray(self.desc0.cam_to_world, np.float64)
        c2w = c2w0.copy()
        cam_vel = np.zeros((2, nder, 3))
        verts = [s.vertices.astype(np.float64) for s in d.shapes]
        nrms = [None if s.normals is None else s.normals.astype(np.float64) for s in d.shapes]
        svel = [np.zeros((2 * nder if s.normals is not None else nder, len(s.vertices), 3)) for s in d.shapes]
        # values only: every tangent input starts at zero (scene_manager.py:355-411)
        bsdfs = [list(b[:{"diffuse": 2, "null": 1, "phong": 4, "roughdielectric": 4}[b[0]]]) for b in d.bsdfs]
        lights = [list(l[:4]) for l in d.lights]
        phases = [list(p[:2]) for p in d.phases]
        for m in d.media:
            m.d_scale = m.d_albedo = m.trans_vel = m.rot_vel = None

        def lane_arr(lst, k, shape):
            while len(lst) <= k:
                lst.append(None)
            if lst[k] is None:
                lst[k] = np.zeros(shape)
            return lst[k]

        for iv, tset in enumerate(self.transforms):
            t = var[iv]
            for T in tset:   # parameter values (scene_manager.py:457-561)
                n = T.type_name
                if n == "CAMERA_TRANSLATE":
                    c2w[:3, 3] += T.vec_vel * t
                elif n == "CAMERA_ROTATE":
                    c2w[:3, :3] = rotation3(c2w0[:3, :3] @ T.vec_vel, t) @ c2w[:3, :3]
                elif n == "SHAPE_TRANSLATE":
                    verts[T.shape_id] += T.vec_vel * t
                    centers[T.shape_id] += T.vec_vel * t
                elif n == "VERTEX_TRANSLATE":
                    verts[T.shape_id][T.vert_id] += T.vec_vel * t
                elif n in ("SHAPE_ROTATE", "SHAPE_GLOBAL_ROTATE"):
                    s = T.shape_id
                    R = rotation3(T.vec_vel, t)
                    c = centers[s] if n == "SHAPE_ROTATE" else np.zeros(3)
                    verts[s] = (verts[s] - c) @ R.T + c
                    if n == "SHAPE_GLOBAL_ROTATE":
                        centers[s] = R @ centers[s]
                    if nrms[s] is not None:
                        nrms[s] = nrms[s] @ R.T
                elif n == "BSDF_VARY":
                    b = bsdfs[T.bsdf_id]
                    if b[0] != T.bsdf_type:
                        raise ValueError("bsdf %d is %s, not %s" % (T.bsdf_id, b[0], T.bsdf_type))
                    if T.bsdf_type == "diffuse":
                        b[1] = tuple(np.asarray(b[1]) + T.d_reflectance * t)
                    elif T.bsdf_type == "phong":
                        b[1] = tuple(np.asarray(b[1]) + T.d_diffuse * t)
                        b[2] = tuple(np.asarray(b[2]) + T.d_specular * t)
                        b[3] = float(b[3]) + T.d_exponent * t
                    else:
                        b0 = self.desc0.bsdfs[T.bsdf_id]
                        b[1] = float(b[1]) + T.d_alpha * t
                        b[2] = (float(b0[2]) / float(b0[3]) + T.d_eta * t) * float(b[3])
                elif n == "MEDIUM_VARY":
                    m = d.media[T.med_id]
                    if m.kind != T.med_type:
                        raise ValueError("medium %d is %s, not %s" % (T.med_id, m.kind, T.med_type))
                    if T.med_type == "homogeneous":
                        m.sigma_t = float(m.sigma_t) + T.d_sigmaT * t
                        m.albedo = tuple(np.asarray(m.albedo) + T.d_albedo * t)
                    else:
                        if not isinstance(m.albedo, str):
                            m.albedo = tuple(np.asarray(m.albedo) + T.d_albedo * t)
                        m.scale = float(m.scale) + T.d_scalar * t
                        cv = centers_vol[T.med_id]
                        R = np.eye(4)
                        R[:3, :3] = rotation3(T.vec_rotate, t)
                        M = np.eye(4) if m.to_world is None else np.asarray(m.to_world, np.float64)
                        M = _translate4(T.vec_translate * t) @ _translate4(cv) @ R @ _translate4(-cv) @ M
                        m.to_world = M.astype(np.float32)
                        centers_vol[T.med_id] = cv + T.vec_translate * t
                elif n == "PHASE_VARY":
                    p = phases[T.phase_id]
                    p[1] = float(p[1]) + T.d_g * t
                elif n == "EMITTER_VARY":
                    l = lights[T.emitter_id]
                    if l[0] != "areaEx":
                        raise ValueError("emitter %d is not an AreaLightEx" % T.emitter_id)
                    l[2] = tuple(np.asarray(l[2]) + T.d_intensity * t)
                    l[3] = float(l[3]) + T.d_kappa * t
            for T in (() if clamped[iv] else tset):   # tangent inputs of lane iv, from the updated values (:556-631)
                n = T.type_name
                if n == "CAMERA_TRANSLATE":
                    cam_vel[0, iv] += T.vec_vel
                elif n == "CAMERA_ROTATE":
                    cam_vel[1, iv] = c2w0[:3, :3] @ T.vec_vel
                elif n == "SHAPE_TRANSLATE":
                    svel[T.shape_id][iv] += T.vec_vel
                elif n == "VERTEX_TRANSLATE":
                    svel[T.shape_id][iv, T.vert_id] += T.vec_vel
                elif n in ("SHAPE_ROTATE", "SHAPE_GLOBAL_ROTATE"):
                    s = T.shape_id
                    c = centers[s] if n == "SHAPE_ROTATE" else np.zeros(3)
                    svel[s][iv] += np.cross(T.vec_vel, verts[s] - c)
                    if nrms[s] is not None:
                        svel[s][nder + iv] += np.cross(T.vec_vel, nrms[s])
                elif n == "BSDF_VARY":
                    b = bsdfs[T.bsdf_id]
                    if T.bsdf_type == "diffuse":
                        lane_arr(b, 2, (nder, 3))[iv] += T.d_reflectance
                    elif T.bsdf_type == "phong":
                        lane_arr(b, 4, (nder, 3))[iv] += T.d_diffuse
                        lane_arr(b, 5, (nder, 3))[iv] += T.d_specular
                        lane_arr(b, 6, (nder,))[iv] += T.d_exponent
                    else:
                        lane_arr(b, 4, (nder,))[iv] += T.d_alpha
                        lane_arr(b, 5, (nder,))[iv] += T.d_eta
                elif n == "MEDIUM_VARY":
                    m = d.media[T.med_id]
                    for k, shp in (("d_albedo", (nder, 3)), ("d_scale", (nder,)), ("trans_vel", (nder, 3)),
                                   ("rot_vel", (nder, 3))):
                        if getattr(m, k) is None:
                            setattr(m, k, np.zeros(shp, np.float32))
                    m.d_albedo[iv] += T.d_albedo
                    if T.med_type == "homogeneous":
                        m.d_scale[iv] += T.d_sigmaT
                    else:
                        m.d_scale[iv] += T.d_scalar
                        m.trans_vel[iv] += T.vec_translate
                        m.rot_vel[iv] = T.vec_rotate
                elif n == "PHASE_VARY":
                    lane_arr(phases[T.phase_id], 2, (nder,))[iv] += T.d_g
                elif n == "EMITTER_VARY":
                    l = lights[T.emitter_id]
                    lane_arr(l, 4, (nder, 3))[iv] += T.d_intensity
                    lane_arr(l, 5, (nder,))[iv] += T.d_kappa
        d.cam_to_world = c2w.astype(np.float32)
        d.cam_vel = cam_vel.astype(np.float32) if np.any(cam_vel) else None
        for s, v, nn, sv in zip(d.shapes, verts, nrms, svel):
            s.vertices = v.astype(np.float32)
            if nn is not None:
                s.normals = nn.astype(np.float32)
            s.velocities = sv.astype(np.float32) if np.any(sv) else None
        d.bsdfs = [tuple(b) for b in bsdfs]
        d.lights = [tuple(l) for l in lights]
        d.phases = [tuple(p) for p in phases]
        self.desc = d
        return var_grad

    def scene(self, mod):
        """Renderer objects of the current description (instantiated through ``mod``)."""
        return scenes.instantiate(mod, self.desc)


def _ad_loss_function():
    import torch

    from . import dtrt as G
    from .adjoint import pyramid_weights

    class ADLossFunc(torch.autograd.Function):
        """auto_grad.py:21-74 over the GPU renderer. ``options`` is a RenderOptions;
        ``input`` the variable tensor (only its gradient is used)."""

        @staticmethod
        def forward(ctx, scene_manager, integrator, options, input, out_of_range=None, penalty_scale=None,
                    pyramid_level=1, pyramid_scale=4.0, index_iter=-1, clamping=0):
            d = scene_manager.desc
            sc = scene_manager.scene(G)
            img = torch.zeros(d.out_shape(), dtype=torch.float32, device="cuda")
            G.render_device(integrator.kind, sc, options, img.data_ptr())
            scene_manager.last_image = img   # the full AD image of this evaluation
            if index_iter > -1:
                torch.save(img.cpu(), "pt_iter%d.pt" % index_iter)
            nder = d.nder
            oor = torch.zeros(nder) if out_of_range is None else torch.as_tensor(out_of_range, dtype=torch.float32)
            pen = torch.ones(nder) if penalty_scale is None else torch.as_tensor(penalty_scale, dtype=torch.float32)
            ctx.save_for_backward(img[1:])
            ctx.pyr = (int(pyramid_level), float(pyramid_scale), int(clamping))
            ctx.oor, ctx.pen = oor, pen
            return img[0].clone()

        @staticmethod
        def backward(ctx, grad_input):
            (derivs,) = ctx.saved_tensors
            W = pyramid_weights(grad_input.to(derivs.device, torch.float32), *ctx.pyr)
            ret = (derivs.double() * W.double()).sum(dim=(1, 2, 3)).float().cpu()
            big = ctx.oor.abs() > 1e-4
            ret[big] = (ctx.oor * ctx.pen)[big]
            return (None, None, None, ret) + (None,) * 6

    return ADLossFunc


class _Lazy:
    """``ADLossFunc`` is built on first use (it needs torch and the GPU library)."""

    def __getattr__(self, name):
        f = _ad_loss_function()
        globals()["ADLossFunc"] = f
        return getattr(f, name)


ADLossFunc = _Lazy()

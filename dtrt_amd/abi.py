"""ctypes binding that presents the reference's ``dtrt`` module surface.

The reference exposes its renderer through pybind11 (``src/dtrt.cpp:24-155``):
object constructors that take raw ``float_ptr``/``int_ptr`` addresses of
torch tensors, a ``Scene``, ``RenderOptions`` and ``VolPathTracerAD.render``.
:func:`bind` builds an object with the same names, overloads, argument meaning
and tensor layouts over any shared library that exports the C-ABI declared in
``include/dtrt.h`` (symbols ``<prefix>camera_create`` ...).
"""
import ctypes as C
import os

VP = C.c_void_p
FP = C.c_void_p   # const float* (raw address)
IP = C.c_void_p   # const int* (raw address)


class RenderOptionsStruct(C.Structure):
    """``dtrt_render_options`` (include/dtrt.h); RenderOptions, integrator.h:10-25."""
    _fields_ = [("seed", C.c_uint64), ("spp", C.c_int32), ("max_bounces", C.c_int32),
                ("sppe", C.c_int32), ("sppse", C.c_int32), ("quiet", C.c_int32),
                ("mode", C.c_int32), ("ddist_coeff", C.c_float)]


# name -> (restype, argtypes); shared by the product library and the oracle.
SIGNATURES = {
    "abi_version": (C.c_int, []),
    "set_nder": (C.c_int, [C.c_int]),
    "nder": (C.c_int, []),
    "angle_eps": (C.c_double, []),
    "last_error": (C.c_char_p, []),
    "destroy": (None, [VP]),
    "camera_create": (VP, [C.c_int, C.c_int, FP, FP, C.c_float, C.c_int, FP]),
    "camera_set_rect": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, C.c_int]),
    "shape_create": (VP, [FP, IP, FP, FP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, FP]),
    "bsdf_diffuse_create": (VP, [FP, FP]),
    "bsdf_null_create": (VP, []),
    "bsdf_phong_create": (VP, [FP, FP, C.c_float, FP, FP, FP]),
    "bsdf_roughdielectric_create": (VP, [C.c_float, C.c_float, C.c_float, FP, FP]),
    "phase_hg_create": (VP, [C.c_float, FP]),
    "phase_isotropic_create": (VP, []),
    "area_light_create": (VP, [C.c_int, FP, C.c_int]),
    "area_light_ex_create": (VP, [C.c_int, FP, C.c_float, FP, FP]),
    "homogeneous_create": (VP, [C.c_float, FP, C.c_int, FP, FP]),
    "heterogeneous_create": (VP, [C.c_char_p, FP, C.c_float, FP, C.c_char_p, C.c_int, FP, FP, FP, FP]),
    "heterogeneous_create_grid": (VP, [FP, C.c_int, C.c_int, C.c_int, FP, FP, FP, C.c_float, FP, C.c_int,
                                       FP, FP, FP, FP]),
    "scene_create": (VP, [VP, C.POINTER(VP), C.c_int, C.POINTER(VP), C.c_int, C.POINTER(VP), C.c_int,
                          C.POINTER(VP), C.c_int, C.POINTER(VP), C.c_int]),
    "scene_init_edges": (C.c_int, [VP, FP]),
    "render": (C.c_int, [C.c_int, VP, C.POINTER(RenderOptionsStruct), FP]),
}


class DtrtError(RuntimeError):
    pass


def _addr(x):
    """Raw address of a float_ptr/int_ptr/None/int/tensor/ndarray argument."""
    if x is None:
        return None
    if isinstance(x, _Ptr):
        return x.addr or None
    if isinstance(x, int):
        return x or None
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    raise TypeError("expected a pointer-like argument, got %r" % type(x))


class _Ptr:
    def __init__(self, addr=0):
        self.addr = int(addr)


class float_ptr(_Ptr):
    """``dtrt.float_ptr(size_t)`` (src/dtrt.cpp:27-28): a raw float address; 0 is null."""


class int_ptr(_Ptr):
    """``dtrt.int_ptr(size_t)`` (src/dtrt.cpp:29-30)."""


class Spectrum3f:
    """``dtrt.Spectrum3f(r, g, b)`` (src/dtrt.cpp:32-33)."""

    def __init__(self, r, g, b):
        self.v = (C.c_float * 3)(float(r), float(g), float(b))


def _spec(x):
    if isinstance(x, Spectrum3f):
        return x.v
    return (C.c_float * 3)(*[float(t) for t in x])


def bind(lib_path, prefix, name="dtrt"):
    """Return a namespace object mirroring the reference ``dtrt`` module."""
    lib = C.CDLL(lib_path)
    fns = {}
    for k, (res, args) in SIGNATURES.items():
        f = getattr(lib, prefix + k)
        f.restype = res
        f.argtypes = args
        fns[k] = f

    def check(h, what):
        if not h:
            raise DtrtError("%s: %s" % (what, fns["last_error"]().decode()))
        return h

    class _Obj:
        _h = None

        def __del__(self):
            h, self._h = self._h, None
            if h:
                try:
                    fns["destroy"](h)
                except Exception:
                    pass

    class Camera(_Obj):
        """camera.cpp:8-38 — velocities: [nder][3] translations then [nder][3] rotation axes."""

        def __init__(self, width, height, cam_to_world, cam_to_ndc, clip_near, med_id, velocities=0):
            self._h = check(fns["camera_create"](int(width), int(height), _addr(cam_to_world), _addr(cam_to_ndc),
                                                 float(clip_near), int(med_id), _addr(velocities)), "Camera")

        def set_rect(self, ox, oy, w, h):
            fns["camera_set_rect"](self._h, int(ox), int(oy), int(w), int(h))

    class Shape(_Obj):
        """shape.cpp:10-52 — argument order of src/dtrt.cpp:39-50 (light_id, bsdf_id, med_int, med_ext)."""

        def __init__(self, vertices, indices, uvs, normals, num_vertices, num_triangles, light_id, bsdf_id,
                     med_int_id, med_ext_id, velocities=0):
            self.num_vertices = int(num_vertices)
            self._has_uvs = bool(_addr(uvs))
            self._has_normals = bool(_addr(normals))
            self._h = check(fns["shape_create"](_addr(vertices), _addr(indices), _addr(uvs), _addr(normals),
                                                int(num_vertices), int(num_triangles), int(light_id), int(bsdf_id),
                                                int(med_int_id), int(med_ext_id), _addr(velocities)), "Shape")

        def has_uvs(self):
            return self._has_uvs

        def has_normals(self):
            return self._has_normals

    class BSDF(_Obj):
        pass

    class BSDF_diffuse(BSDF):
        def __init__(self, reflectance, d_reflectance=None):
            self._h = check(fns["bsdf_diffuse_create"](_spec(reflectance), _addr(d_reflectance)), "BSDF_diffuse")

    class BSDF_null(BSDF):
        def __init__(self):
            self._h = check(fns["bsdf_null_create"](), "BSDF_null")

    class BSDF_Phong(BSDF):
        def __init__(self, kd, ks, exponent, d_kd=None, d_ks=None, d_exponent=None):
            self._h = check(fns["bsdf_phong_create"](_spec(kd), _spec(ks), float(exponent), _addr(d_kd), _addr(d_ks),
                                                     _addr(d_exponent)), "BSDF_Phong")

    class BSDF_roughdielectric(BSDF):
        def __init__(self, alpha, int_ior, ext_ior, d_alpha=None, d_eta=None):
            self._h = check(fns["bsdf_roughdielectric_create"](float(alpha), float(int_ior), float(ext_ior),
                                                               _addr(d_alpha), _addr(d_eta)), "BSDF_roughdielectric")

    class Phase(_Obj):
        pass

    class HG(Phase):
        def __init__(self, g, d_g=None):
            self._h = check(fns["phase_hg_create"](float(g), _addr(d_g)), "HG")

    class Isotropic(Phase):
        def __init__(self):
            self._h = check(fns["phase_isotropic_create"](), "Isotropic")

    class Emitter(_Obj):
        pass

    class AreaLight(Emitter):
        def __init__(self, shape_id, intensity, two_sided=False):
            iv = _spec(intensity) if not isinstance(intensity, (_Ptr, int)) else None
            self._h = check(fns["area_light_create"](int(shape_id), iv if iv is not None else _addr(intensity),
                                                     int(bool(two_sided))), "AreaLight")

    class AreaLightEx(Emitter):
        def __init__(self, shape_id, intensity, kappa, d_intensity=None, d_kappa=None):
            self._h = check(fns["area_light_ex_create"](int(shape_id), _spec(intensity), float(kappa),
                                                        _addr(d_intensity), _addr(d_kappa)), "AreaLightEx")

    class Medium(_Obj):
        pass

    class Homogeneous(Medium):
        def __init__(self, sigma_t, albedo, phase_id, d_sigma_t=None, d_albedo=None):
            self._h = check(fns["homogeneous_create"](float(sigma_t), _spec(albedo), int(phase_id), _addr(d_sigma_t),
                                                      _addr(d_albedo)), "Homogeneous")

    class Heterogeneous(Medium):
        """heterogeneous.h:13-106; overloads of src/dtrt.cpp:87-93.

        Heterogeneous(file, toWorld, scale, albedo:Spectrum3f|str, phase_id
                      [, trans_vel, rot_vel, d_density [, d_albedo]])
        """

        def __init__(self, density_file, to_world, scale, albedo, phase_id, trans_vel=None, rot_vel=None,
                     d_density=None, d_albedo=None):
            afile = albedo.encode() if isinstance(albedo, str) else None
            aval = None if afile else _spec(albedo)
            self._h = check(fns["heterogeneous_create"](density_file.encode(), _addr(to_world), float(scale), aval,
                                                        afile, int(phase_id), _addr(trans_vel), _addr(rot_vel),
                                                        _addr(d_density), _addr(d_albedo)), "Heterogeneous")

    class HeterogeneousGrid(Medium):
        """In-memory density grid (no file): data float32[rz][ry][rx] on box [bmin, bmax]."""

        def __init__(self, density, res, bmin, bmax, to_world, scale, albedo, phase_id, trans_vel=None,
                     rot_vel=None, d_density=None, d_albedo=None):
            rx, ry, rz = [int(t) for t in res]
            self._h = check(fns["heterogeneous_create_grid"](_addr(density), rx, ry, rz, _spec(bmin), _spec(bmax),
                                                             _addr(to_world), float(scale), _spec(albedo),
                                                             int(phase_id), _addr(trans_vel), _addr(rot_vel),
                                                             _addr(d_density), _addr(d_albedo)), "HeterogeneousGrid")

    def _harr(objs):
        arr = (VP * max(1, len(objs)))(*[o._h for o in objs])
        return arr

    class Scene(_Obj):
        """scene.cpp:9-136. Keeps references to its parts (the reference's Scene holds raw pointers)."""

        def __init__(self, camera, shapes, bsdfs, area_lights, phases, media):
            self._parts = (camera, list(shapes), list(bsdfs), list(area_lights), list(phases), list(media))
            self._h = check(fns["scene_create"](camera._h, _harr(shapes), len(shapes), _harr(bsdfs), len(bsdfs),
                                                _harr(area_lights), len(area_lights), _harr(phases), len(phases),
                                                _harr(media), len(media)), "Scene")

        def initEdges(self, weights):
            if fns["scene_init_edges"](self._h, _addr(weights)) != 0:
                raise DtrtError("initEdges: " + fns["last_error"]().decode())

    class RenderOptions:
        """integrator.h:10-25 (seed, spp, max_bounces, sppe, sppse, quiet[, mode[, ddistCoeff]])."""

        def __init__(self, seed, spp, max_bounces, sppe, sppse, quiet, mode=-1, ddistCoeff=0.0):
            self.seed, self.spp, self.max_bounces = int(seed), int(spp), int(max_bounces)
            self.sppe, self.sppse, self.quiet = int(sppe), int(sppse), bool(quiet)
            self.mode, self.ddistCoeff = int(mode), float(ddistCoeff)

        def _struct(self):
            return RenderOptionsStruct(self.seed, self.spp, self.max_bounces, self.sppe, self.sppse,
                                       int(self.quiet), self.mode, self.ddistCoeff)

    class Integrator:
        pass

    class VolPathTracerAD(Integrator):
        """integratorAD.cpp:19-166: fills a caller-zeroed float32[(nder+1)][H][W][3]."""
        kind = 0

        def render(self, scene, options, out):
            st = options._struct()
            rc = fns["render"](self.kind, scene._h, C.byref(st), _addr(out))
            if rc != 0:
                raise DtrtError("render: " + fns["last_error"]().decode())

    class VolPathTracer(Integrator):
        """volpath.cpp:106-141: writes the primal image float32[H][W][3]."""
        kind = 1

        def render(self, scene, options, out):
            st = options._struct()
            rc = fns["render"](self.kind, scene._h, C.byref(st), _addr(out))
            if rc != 0:
                raise DtrtError("render: " + fns["last_error"]().decode())

    class VolPathTracerSimple(VolPathTracer):
        """volpath_simple.cpp:89-124: next-event estimation only; float32[H][W][3]."""
        kind = 2

    class PathTracer(VolPathTracer):
        """path.cpp:64-96: surfaces only, NEE + BSDF sampling with MIS; float32[H][W][3]."""
        kind = 3

    class DirectIntegrator(VolPathTracer):
        """direct.cpp:51-85: direct illumination (one bounce) with MIS; float32[H][W][3]."""
        kind = 4

    class ParticleTracer(VolPathTracer):
        """ptracer.cpp:10-167: 100000 * spp light particles splatted to the camera; float32[H][W][3]."""
        kind = 5

    class PathTracerAD(VolPathTracerAD):
        """pathAD.cpp: surface-only forward-mode path tracer; same layout as VolPathTracerAD."""
        kind = 6

    class DirectAD(VolPathTracerAD):
        """directAD.cpp: surface-only forward-mode direct illumination; same layout as VolPathTracerAD."""
        kind = 7

    class _Module:
        pass

    m = _Module()
    for k, v in dict(float_ptr=float_ptr, int_ptr=int_ptr, Spectrum3f=Spectrum3f, Camera=Camera, Shape=Shape,
                     BSDF=BSDF, BSDF_diffuse=BSDF_diffuse, BSDF_null=BSDF_null, BSDF_Phong=BSDF_Phong,
                     BSDF_roughdielectric=BSDF_roughdielectric, Phase=Phase, HG=HG, Isotropic=Isotropic,
                     Emitter=Emitter, AreaLight=AreaLight, AreaLightEx=AreaLightEx, Medium=Medium,
                     Homogeneous=Homogeneous, Heterogeneous=Heterogeneous, HeterogeneousGrid=HeterogeneousGrid,
                     Scene=Scene, RenderOptions=RenderOptions, Integrator=Integrator,
                     VolPathTracerAD=VolPathTracerAD, VolPathTracer=VolPathTracer,
                     VolPathTracerSimple=VolPathTracerSimple, PathTracer=PathTracer,
                     DirectIntegrator=DirectIntegrator, ParticleTracer=ParticleTracer, PathTracerAD=PathTracerAD,
                     DirectAD=DirectAD).items():
        setattr(m, k, v)
    m.lib = lib
    m.fns = fns
    m.prefix = prefix
    m.__name__ = name

    def set_nder(n):
        if fns["set_nder"](int(n)) != 0:
            raise DtrtError(fns["last_error"]().decode())

    m.set_nder = set_nder
    type(m).nder = property(lambda self: fns["nder"]())
    m.angleEps = fns["angle_eps"]()
    m.abi_version = fns["abi_version"]()
    return m


def env_nder(default=4):
    """Process-wide number of derivative lanes (``DTRT_NDER``), as the reference's compile-time ``nder``."""
    return int(os.environ.get("DTRT_NDER", default))

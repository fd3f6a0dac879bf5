"""Drop-in ``dtrt`` module backed by libdtrt_mi355x.so (gfx950 HIP).

``import dtrt_amd.dtrt as dtrt`` (or ``sys.modules['dtrt'] = dtrt_amd.dtrt``)
gives the names the reference's pybind module exports (src/dtrt.cpp:24-155).
There is no CPU fallback: if the HIP library is missing this import raises.
"""
import ctypes as C
import os

from .abi import bind, env_nder

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libdtrt_mi355x.so")
if not os.path.exists(_LIB):
    raise ImportError("libdtrt_mi355x.so is not built (run `python -m dtrt_amd.build`): " + _LIB)

_m = bind(_LIB, "dtrt_", name="dtrt")
_lib = _m.lib

float_ptr = _m.float_ptr
int_ptr = _m.int_ptr
Spectrum3f = _m.Spectrum3f
Camera = _m.Camera
Shape = _m.Shape
BSDF = _m.BSDF
BSDF_diffuse = _m.BSDF_diffuse
BSDF_null = _m.BSDF_null
BSDF_Phong = _m.BSDF_Phong
BSDF_roughdielectric = _m.BSDF_roughdielectric
Phase = _m.Phase
HG = _m.HG
Isotropic = _m.Isotropic
Emitter = _m.Emitter
AreaLight = _m.AreaLight
AreaLightEx = _m.AreaLightEx
Medium = _m.Medium
Homogeneous = _m.Homogeneous
Heterogeneous = _m.Heterogeneous
HeterogeneousGrid = _m.HeterogeneousGrid
Scene = _m.Scene
RenderOptions = _m.RenderOptions
Integrator = _m.Integrator
VolPathTracerAD = _m.VolPathTracerAD
VolPathTracer = _m.VolPathTracer
VolPathTracerSimple = _m.VolPathTracerSimple
PathTracer = _m.PathTracer
DirectIntegrator = _m.DirectIntegrator
ParticleTracer = _m.ParticleTracer
PathTracerAD = _m.PathTracerAD
DirectAD = _m.DirectAD
angleEps = _m.angleEps
abi_version = _m.abi_version
set_nder = _m.set_nder
fns = _m.fns
lib = _lib
prefix = "dtrt_"


def __getattr__(name):
    if name == "nder":   # process-wide lane count (the reference's compile-time nder)
        return _m.nder
    raise AttributeError(name)


if "DTRT_NDER" in os.environ:
    set_nder(env_nder())

_lib.dtrt_set_precision.argtypes = [C.c_int]
_lib.dtrt_scene_precision.argtypes = [C.c_void_p]


def set_precision(bits):
    """Estimator precision (64 = double, the reference's; 32 = float) of scenes created
    from now on. The default is env DTRT_PRECISION, else 64."""
    if _lib.dtrt_set_precision(int(bits)) != 0:
        raise ValueError(_lib.dtrt_last_error().decode())


def precision():
    return int(_lib.dtrt_precision())


def scene_precision(scene):
    return int(_lib.dtrt_scene_precision(scene._h))

# ------------------------------------------------------------ test / bench hooks
_P = C.c_void_p
_lib.dtrt_test_kernel_times.argtypes = [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_uint64)]
_lib.dtrt_test_scene_info.argtypes = [_P, C.POINTER(C.c_int64)]
_lib.dtrt_test_set_queue_budget.argtypes = [_P, C.c_uint64]
_lib.dtrt_test_set_tree_in_hbm.argtypes = [_P, C.c_int]
_lib.dtrt_test_set_voxel_lanes.argtypes = [_P, C.c_int, C.POINTER(C.c_int), C.c_int]
_lib.dtrt_scene_set_density.argtypes = [_P, C.c_int, _P, C.c_int]
_lib.dtrt_scene_capture_views.argtypes = [_P, C.c_int, C.POINTER(_P), C.POINTER(_m.fns["render"].argtypes[2]._type_),
                                          C.POINTER(_P)]
_lib.dtrt_scene_replay_views.argtypes = [_P, C.POINTER(C.c_uint64)]
_lib.dtrt_render_views.argtypes = [C.c_int, _P, C.c_int, C.POINTER(_P), C.POINTER(_m.fns["render"].argtypes[2]._type_),
                                   C.POINTER(_P), C.c_int]
_lib.dtrt_render_scenes.argtypes = [C.c_int, C.c_int, C.POINTER(_P), C.POINTER(C.c_int), C.POINTER(_P),
                                    C.POINTER(_m.fns["render"].argtypes[2]._type_), C.POINTER(_P), C.c_int]
_lib.dtrt_render_adjoint.argtypes = [C.c_int, _P, C.POINTER(_m.fns["render"].argtypes[2]._type_), C.c_int, _P, _P]
_lib.dtrt_render_views_adjoint.argtypes = [C.c_int, _P, C.c_int, C.POINTER(_P),
                                           C.POINTER(_m.fns["render"].argtypes[2]._type_), C.c_int, _P, _P]
_lib.dtrt_test_shape_edges.argtypes = [_P, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int]
_lib.dtrt_test_trace.argtypes = [_P, _P, C.c_int, _P, _P]
_lib.dtrt_test_li.argtypes = [_P, _P, C.c_int, C.c_int, C.c_int, C.c_uint64, _P]
_lib.dtrt_test_medium.argtypes = [_P, C.c_int, C.c_int, _P, C.c_int, C.c_uint64, _P]
_lib.dtrt_test_grid.argtypes = [_P, C.c_int, _P, C.c_int, _P]
_lib.dtrt_test_philox.argtypes = [_P, _P, _P]
_lib.dtrt_render_stats.argtypes = [_P, C.POINTER(C.c_int64)]
_lib.dtrt_render_device.argtypes = [C.c_int, _P, C.POINTER(_m.fns["render"].argtypes[2]._type_), _P]


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError("%s: %s" % (what, _lib.dtrt_last_error().decode()))


_lib.dtrt_scene_set_stream.argtypes = [_P, _P]


def _after_torch(scene):
    """Orders the scene's next device-memory call after the work torch has queued on its
    current stream (zeroed outputs, loss gradients built just before the call)."""
    import torch
    if torch.cuda.is_initialized():
        _chk(_lib.dtrt_scene_set_stream(scene._h, torch.cuda.current_stream().cuda_stream), "set_stream")


def kernel_times(scene):
    import numpy as np
    ms = (C.c_double * 4)()
    ln = (C.c_int64 * 4)()
    wk = (C.c_uint64 * 12)()
    _chk(_lib.dtrt_test_kernel_times(scene._h, ms, ln, wk), "kernel_times")
    return np.array(ms[:]), np.array(ln[:]), np.array(wk[:], dtype=np.uint64)


_lib.dtrt_scene_autotune.argtypes = [C.c_int, _P, C.POINTER(_m.fns["render"].argtypes[2]._type_),
                                     C.POINTER(C.c_double), C.POINTER(C.c_int64)]


def autotune(scene, options, integrator=0):
    """``dtrt_scene_autotune``: picks and keeps the fastest launch settings for renders
    like ``options``; returns (best ms, launch_settings(scene))."""
    st = options._struct()
    ms = C.c_double()
    s3 = (C.c_int64 * 3)()
    _chk(_lib.dtrt_scene_autotune(integrator, scene._h, C.byref(st), C.byref(ms), s3), "autotune")
    return ms.value, launch_settings(scene)


_lib.dtrt_scene_launch_settings.argtypes = [_P, C.POINTER(C.c_int64)]
_lib.dtrt_scene_set_launch_settings.argtypes = [_P, C.POINTER(C.c_int64)]
_SETTINGS = ("child_blocks_per_cu", "fill_percent", "max_chunk", "queue_budget_bytes", "ad_workgroup", "child_workgroup")


def set_launch_settings(scene, **kw):
    """``dtrt_scene_set_launch_settings``: change some of the scene's launch settings."""
    cur = launch_settings(scene)
    cur.update(kw)
    s6 = (C.c_int64 * 6)(*[int(cur[k]) for k in _SETTINGS])
    _chk(_lib.dtrt_scene_set_launch_settings(scene._h, s6), "set_launch_settings")


def launch_settings(scene):
    """``dtrt_scene_launch_settings``: the scene's current launch settings."""
    s6 = (C.c_int64 * 6)()
    _chk(_lib.dtrt_scene_launch_settings(scene._h, s6), "launch_settings")
    return {"child_blocks_per_cu": int(s6[0]), "fill_percent": int(s6[1]), "max_chunk": int(s6[2]),
            "queue_budget_bytes": int(s6[3]), "ad_workgroup": int(s6[4]), "child_workgroup": int(s6[5])}


def render_views_adjoint(scene, cameras, options, medium, d_images, d_density, integrator=0):
    """``dtrt_render_views_adjoint``: d_images is a list of device tensors [h][w][3] (one
    per view); the density gradient of all views is accumulated into d_density."""
    import torch
    n = len(options)
    cams = (_P * n)(*[None if c is None else c._h for c in (cameras or [None] * n)])
    opts = (_m.fns["render"].argtypes[2]._type_ * n)(*[o._struct() for o in options])
    flat = torch.cat([g.reshape(-1) for g in d_images]).float().contiguous()
    _after_torch(scene)
    _chk(_lib.dtrt_render_views_adjoint(integrator, scene._h, n, cams, opts, int(medium), C.c_void_p(flat.data_ptr()),
                                        C.c_void_p(d_density.data_ptr())), "render_views_adjoint")
    torch.cuda.synchronize()


def render_views(scene, cameras, options, images, integrator=0):
    """``dtrt_render_views``: one pass over several views.

    cameras: list of Camera objects (None = the scene's camera); options: list of
    RenderOptions; images: list of zeroed torch CUDA tensors, or numpy float32 arrays
    [(nder+1), h, w, 3], all of one kind.
    """
    n = len(options)
    cams = (_P * n)(*[None if c is None else c._h for c in (cameras or [None] * n)])
    opts = (_m.fns["render"].argtypes[2]._type_ * n)(*[o._struct() for o in options])
    on_dev = hasattr(images[0], "data_ptr")
    ptrs = (_P * n)(*[im.data_ptr() if on_dev else im.ctypes.data for im in images])
    if on_dev:
        _after_torch(scene)
    _chk(_lib.dtrt_render_views(integrator, scene._h, n, cams, opts, ptrs, int(on_dev)), "render_views")


def render_scenes(scenes, cameras, options, images, integrator=0):
    """``dtrt_render_scenes``: the views of several scenes in one call, each scene on its
    own stream. cameras, options, images: one list per scene, as for render_views."""
    counts = [len(o) for o in options]
    flat_c = [c for i, n in enumerate(counts) for c in (cameras[i] if cameras and cameras[i] else [None] * n)]
    flat_o = [o for ol in options for o in ol]
    flat_i = [im for il in images for im in il]
    on_dev = hasattr(flat_i[0], "data_ptr")
    cams = (_P * len(flat_o))(*[None if c is None else c._h for c in flat_c])
    opts = (_m.fns["render"].argtypes[2]._type_ * len(flat_o))(*[o._struct() for o in flat_o])
    ptrs = (_P * len(flat_o))(*[im.data_ptr() if on_dev else im.ctypes.data for im in flat_i])
    if on_dev:
        for sc in scenes:
            _after_torch(sc)
    hs = (_P * len(scenes))(*[sc._h for sc in scenes])
    nv = (C.c_int * len(scenes))(*counts)
    _chk(_lib.dtrt_render_scenes(integrator, len(scenes), hs, nv, cams, opts, ptrs, int(on_dev)), "render_scenes")


def capture_views(scene, cameras, options, images):
    """``dtrt_scene_capture_views``: render once, then capture the launch sequence as a HIP graph."""
    n = len(options)
    cams = (_P * n)(*[None if c is None else c._h for c in (cameras or [None] * n)])
    opts = (_m.fns["render"].argtypes[2]._type_ * n)(*[o._struct() for o in options])
    ptrs = (_P * n)(*[im.data_ptr() for im in images])
    _after_torch(scene)
    _chk(_lib.dtrt_scene_capture_views(scene._h, n, cams, opts, ptrs), "capture_views")


def replay_views(scene, seeds):
    """``dtrt_scene_replay_views``: 0 = graph replay, 1 = eager fallback (overflow)."""
    s = (C.c_uint64 * len(seeds))(*[int(x) for x in seeds])
    _after_torch(scene)
    rc = _lib.dtrt_scene_replay_views(scene._h, s)
    if rc < 0:
        raise RuntimeError("replay_views: " + _lib.dtrt_last_error().decode())
    return rc


_lib.dtrt_scene_step.argtypes = [_P, C.c_int, C.POINTER(_P), C.POINTER(_m.fns["render"].argtypes[2]._type_),
                                 C.POINTER(_P), C.POINTER(_P), C.c_int,
                                 C.POINTER(_m.fns["render"].argtypes[2]._type_), _P, _P, _P]
_lib.dtrt_scene_capture_step.argtypes = _lib.dtrt_scene_step.argtypes
_lib.dtrt_scene_replay_step.argtypes = [_P, C.POINTER(C.c_uint64)]


def _step_args(scene, cameras, options, images, targets, medium, adj_options, d_density, param_grad, loss):
    n = len(options)
    cams = (_P * n)(*[None if c is None else c._h for c in (cameras or [None] * n)])
    opts = (_m.fns["render"].argtypes[2]._type_ * n)(*[o._struct() for o in options])
    aopts = (_m.fns["render"].argtypes[2]._type_ * n)(*[o._struct() for o in adj_options])
    ims = (_P * n)(*[im.data_ptr() for im in images])
    tgs = (_P * n)(*[t.data_ptr() for t in targets])
    return (scene._h, n, cams, opts, ims, tgs, int(medium), aopts, d_density.data_ptr(),
            None if param_grad is None else param_grad.data_ptr(), None if loss is None else loss.data_ptr())


def step(scene, cameras, options, images, targets, medium, adj_options, d_density, param_grad=None, loss=None):
    """``dtrt_scene_step``: one L2 inverse-rendering step over the views (device tensors:
    images [(nder+1), h, w, 3], targets [h, w, 3], d_density [rz, ry, rx] float32,
    param_grad [nder] and loss [1] float64)."""
    _after_torch(scene)
    _chk(_lib.dtrt_scene_step(*_step_args(scene, cameras, options, images, targets, medium, adj_options, d_density,
                                          param_grad, loss)), "step")


def capture_step(scene, cameras, options, images, targets, medium, adj_options, d_density, param_grad=None, loss=None):
    """``dtrt_scene_capture_step``: run the step once, then capture it as a HIP graph."""
    _after_torch(scene)
    _chk(_lib.dtrt_scene_capture_step(*_step_args(scene, cameras, options, images, targets, medium, adj_options,
                                                  d_density, param_grad, loss)), "capture_step")


def replay_step(scene, seeds):
    """``dtrt_scene_replay_step``: seeds[2v] forward, seeds[2v+1] adjoint; 0 = graph replay,
    1 = eager fallback (queue overflow)."""
    s = (C.c_uint64 * len(seeds))(*[int(x) for x in seeds])
    _after_torch(scene)
    rc = _lib.dtrt_scene_replay_step(scene._h, s)
    if rc < 0:
        raise RuntimeError("replay_step: " + _lib.dtrt_last_error().decode())
    return rc


_lib.dtrt_scene_update_shape.argtypes = [_P, C.c_int, _P, _P, _P]


def update_shape(scene, shape, vertices, normals=None, velocities=None):
    """``dtrt_scene_update_shape``: new vertex positions [nv, 3] (and optionally normals and
    tangents in the Shape constructor's layouts) for shape `shape`, with a GPU BVH refit."""
    import numpy as np

    def host(x):
        return None if x is None else np.ascontiguousarray(x, dtype=np.float32)
    v, n, vel = host(vertices), host(normals), host(velocities)
    _chk(_lib.dtrt_scene_update_shape(scene._h, int(shape), v.ctypes.data, None if n is None else n.ctypes.data,
                                      None if vel is None else vel.ctypes.data), "update_shape")


def set_density(scene, medium, density):
    """``dtrt_scene_set_density``: a torch tensor (host or device) or numpy array [rz][ry][rx]."""
    import numpy as np
    if hasattr(density, "data_ptr"):
        t = density.detach().contiguous().float()
        if t.is_cuda:
            _after_torch(scene)
        _chk(_lib.dtrt_scene_set_density(scene._h, int(medium), t.data_ptr(), int(t.is_cuda)), "set_density")
    else:
        a = np.ascontiguousarray(density, dtype=np.float32)
        _chk(_lib.dtrt_scene_set_density(scene._h, int(medium), a.ctypes.data, 0), "set_density")


def render_adjoint(scene, options, medium, d_image_ptr, d_density_ptr, integrator=0):
    """``dtrt_render_adjoint``: device pointers to d loss/d image [H][W][3] and the accumulated grid gradient."""
    st = options._struct()
    _after_torch(scene)
    _chk(_lib.dtrt_render_adjoint(integrator, scene._h, C.byref(st), int(medium), d_image_ptr, d_density_ptr), "render_adjoint")


def set_voxel_lanes(scene, medium, voxels):
    """Forward-mode lanes w.r.t. single grid nodes (flat [z][y][x] indices; [] clears)."""
    v = (C.c_int * max(1, len(voxels)))(*[int(x) for x in voxels])
    _chk(_lib.dtrt_test_set_voxel_lanes(scene._h, int(medium), v, len(voxels)), "set_voxel_lanes")


def set_queue_budget(scene, nbytes):
    """Child-record queue bytes for this scene's renders (0 restores the default)."""
    _chk(_lib.dtrt_test_set_queue_budget(scene._h, int(nbytes)), "set_queue_budget")


def set_tree_in_hbm(scene, on=True):
    """Traversal reads the BVH from HBM only (True) or from LDS where a launch has room."""
    _chk(_lib.dtrt_test_set_tree_in_hbm(scene._h, 1 if on else 0), "set_tree_in_hbm")


def render_stats(scene):
    a = (C.c_int64 * 10)()
    _chk(_lib.dtrt_render_stats(scene._h, a), "render_stats")
    keys = ["neg_pixels", "nonfinite_der", "dt_cap", "bad_edge_weight", "nonortho", "med_inconsistent",
            "replayed", "ad_samples", "children", "pe_samples"]
    return dict(zip(keys, a[:]))


def scene_info(scene):
    import numpy as np
    a = (C.c_int64 * 8)()
    _chk(_lib.dtrt_test_scene_info(scene._h, a), "scene_info")
    pe_sum = np.array([a[4]], dtype=np.int64).view(np.float64)[0]
    return dict(nodes=a[0], tris=a[1], edges=a[2], primary_edges=a[3], primary_edge_sum=float(pe_sum),
                edges_valid=a[5], vertices=a[6], faces=a[7])


def shape_edges(shape, cap=1 << 22):
    import numpy as np
    e = np.zeros((cap, 4), np.int32)
    l = np.zeros(cap, np.float64)
    n = _lib.dtrt_test_shape_edges(shape._h, e.ctypes.data_as(C.POINTER(C.c_int)), l.ctypes.data_as(C.POINTER(C.c_double)), cap)
    if n < 0:
        raise RuntimeError(_lib.dtrt_last_error().decode())
    return e[:n], l[:n]


def trace(scene, rays):
    import numpy as np
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = len(rays)
    face = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    _chk(_lib.dtrt_test_trace(scene._h, rays.ctypes.data, n, face.ctypes.data, t.ctypes.data), "trace")
    return face, t


def li(scene, rays, max_depth, medium=-1, seed=0):
    import numpy as np
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    out = np.zeros((len(rays), 3), np.float32)
    _chk(_lib.dtrt_test_li(scene._h, rays.ctypes.data, len(rays), int(max_depth), int(medium), int(seed), out.ctypes.data), "li")
    return out


def medium_estimate(scene, medium, kind, segs, seed=0):
    import numpy as np
    segs = np.ascontiguousarray(segs, dtype=np.float32)
    out = np.zeros(len(segs), np.float32)
    _chk(_lib.dtrt_test_medium(scene._h, int(medium), int(kind), segs.ctypes.data, len(segs), int(seed), out.ctypes.data),
         "medium_estimate")
    return out


def grid_lookup(scene, medium, pts, nder):
    import numpy as np
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    out = np.zeros((len(pts), 1 + nder), np.float32)
    _chk(_lib.dtrt_test_grid(scene._h, int(medium), pts.ctypes.data, len(pts), out.ctypes.data), "grid_lookup")
    return out


def philox(ctr, key):
    import numpy as np
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    _chk(_lib.dtrt_test_philox(c.ctypes.data, k.ctypes.data, o.ctypes.data), "philox")
    return o


def render_device(integrator, scene, options, out_ptr):
    """``dtrt_render_device``: ``out_ptr`` is a device address (e.g. ``tensor.data_ptr()`` of a HIP tensor)."""
    st = options._struct()
    _after_torch(scene)
    _chk(_lib.dtrt_render_device(integrator, scene._h, C.byref(st), out_ptr), "render_device")

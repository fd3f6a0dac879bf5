"""Scene ingestion: Wavefront OBJ and the Mitsuba 0.5 XML subset the reference reads.

``load_obj`` follows pydtrt/load_obj.py:13-103 and ``load_mitsuba`` follows
pydtrt/load_mitsuba.py:10-350. Both return plain numpy data: a
:class:`dtrt_amd.scenes.SceneDesc` that :func:`dtrt_amd.scenes.instantiate` turns
into renderer objects through the ``dtrt`` module (or the oracle's binding). The
result is the scene the reference's ``serialize_scene`` + ``render_scene`` would
build from the same file, including its quirks:

* OBJ: corners are deduplicated on (position, uv, normal) index triples; quads are
  split (0,1,2), (0,2,3); a negative position (uv) index is offset by the number of
  vertices (uvs) emitted so far, not by the pool size, and a negative normal index
  counts from the end of the normal pool.
* XML: a ``<bsdf>`` without an ``id`` is dropped; the camera medium is looked up
  among the media declared before the ``<sensor>``; ``nearClip`` and ``fovAxis`` are
  ignored (clip_near 1e-2, the fov applies to both axes); area lights are one-sided;
  shading normals are multiplied by the inverse of the shape's toWorld (not its
  inverse transpose) and not renormalised.

One deliberate difference: file names inside the XML (meshes, VOL grids) are
resolved against the XML file's directory, where the reference changes the working
directory while parsing OBJ files but not when the renderer later opens VOL files.
"""
import math
import os
import re
import xml.etree.ElementTree as ET

import numpy as np

from .scenes import MediumDesc, SceneDesc, ShapeDesc, look_at

_SPACES = re.compile(r" +")


# ------------------------------------------------------------------ OBJ
def load_obj(path):
    """Wavefront OBJ -> (vertices float32 [nv][3], indices int32 [nf][3], uvs float32
    [n][2] or None, normals float32 [n][3] or None); all faces form one mesh."""
    pos_pool, uv_pool, nrm_pool = [], [], []
    verts, uvs, nrms, faces = [], [], [], []
    seen = {}

    def resolve(tok, slot, emitted):
        i = int(tok)
        if i < 0:
            return i + emitted if slot < 2 else i   # the normal slot keeps its negative index
        return i - 1

    def corner(tok):
        parts = tok.split("/")
        pi = resolve(parts[0], 0, len(verts))
        ti = resolve(parts[1], 1, len(uvs)) if len(parts) > 1 and parts[1] != "" else None
        ni = resolve(parts[2], 2, 0) if len(parts) > 2 and parts[2] != "" else None
        key = (pi, ti, ni)
        vid = seen.get(key)
        if vid is None:
            vid = seen[key] = len(verts)
            verts.append(pos_pool[pi])
            if ti is not None:
                uvs.append(uv_pool[ti])
            if ni is not None:
                nrms.append(nrm_pool[ni])
        return vid

    with open(path, "r") as fh:
        for line in fh:
            tok = _SPACES.split(line.strip())
            head = tok[0]
            if head == "v":
                pos_pool.append([float(tok[1]), float(tok[2]), float(tok[3])])
            elif head == "vt":
                uv_pool.append([float(tok[1]), float(tok[2])])
            elif head == "vn":
                nrm_pool.append([float(tok[1]), float(tok[2]), float(tok[3])])
            elif head == "f":
                if len(tok) > 5:
                    raise ValueError("%s: faces with more than 4 corners are not supported" % path)
                c = [corner(t) for t in tok[1:]]
                faces.append([c[0], c[1], c[2]])
                if len(c) == 4:
                    faces.append([c[0], c[2], c[3]])
    return (np.array(verts, np.float32).reshape(-1, 3), np.array(faces, np.int32).reshape(-1, 3),
            np.array(uvs, np.float32) if uvs else None, np.array(nrms, np.float32) if nrms else None)


# ----------------------------------------------------------- transforms
def rotate_matrix(axis, angle_deg):
    """4x4 rotation by ``angle_deg`` about ``axis`` (right-handed; identity for a
    near-zero axis), as the Mitsuba ``<rotate>`` element."""
    a = np.asarray(axis, np.float64)
    n = math.sqrt(float(a @ a))
    m = np.eye(4)
    if n > 1e-4:
        u = a / n
        t = math.radians(angle_deg)
        c, s = math.cos(t), math.sin(t)
        K = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])
        m[:3, :3] = c * np.eye(3) + s * K + (1 - c) * np.outer(u, u)
    return m


def parse_transform(node):
    """Compose the children of a ``<transform>`` element; each one applies after the
    ones before it (M = T_k ... T_1)."""
    m = np.eye(4)
    for ch in node:
        if ch.tag == "matrix":
            t = np.array([float(x) for x in ch.attrib["value"].split()], np.float64).reshape(4, 4)
        elif ch.tag in ("translate", "scale", "rotate"):
            dflt = 1.0 if ch.tag == "scale" else 0.0
            x, y, z = (float(ch.attrib.get(k, dflt)) for k in ("x", "y", "z"))
            if ch.tag == "translate":
                t = np.eye(4)
                t[:3, 3] = (x, y, z)
            elif ch.tag == "scale":
                t = np.diag([x, y, z, 1.0])
            else:
                t = rotate_matrix((x, y, z), float(ch.attrib["angle"]))
        else:
            continue
        m = t @ m
    return m


def parse_vector(s):
    """Comma-separated triple, else whitespace-separated values."""
    parts = [p for p in s.split(",")]
    if len(parts) == 3:
        return np.array([float(p) for p in parts], np.float32)
    return np.array([float(p) for p in s.split()], np.float32)


def _scalar(s):
    return float(parse_vector(s)[0])


def _rgb(s):
    v = parse_vector(s)
    return tuple(float(x) for x in (v if v.size == 3 else np.repeat(v[:1], 3)))


# --------------------------------------------------------------- elements
_RECT_V = np.array([[-1.0, -1.0, 0.0], [-1.0, 1.0, 0.0], [1.0, -1.0, 0.0], [1.0, 1.0, 0.0]], np.float32)
_RECT_F = np.array([[0, 2, 1], [1, 2, 3]], np.int32)

_INTEGRATORS = ("direct", "path", "volpath_simple", "volpath", "directAD", "pathAD", "volpathAD", "ptracer")


def _named(node):
    return {c.attrib["name"]: c for c in node if "name" in c.attrib}


def _bsdf(node):
    kind = node.attrib["type"]
    p = _named(node)
    if kind == "diffuse":
        return ("diffuse", _rgb(p["reflectance"].attrib["value"]) if "reflectance" in p else (0.5, 0.5, 0.5))
    if kind == "null":
        return ("null",)
    if kind == "phong":
        kd = _rgb(p["diffuseReflectance"].attrib["value"]) if "diffuseReflectance" in p else (0.5, 0.5, 0.5)
        ks = _rgb(p["specularReflectance"].attrib["value"]) if "specularReflectance" in p else (0.2, 0.2, 0.2)
        ex = _scalar(p["exponent"].attrib["value"]) if "exponent" in p else 30.0
        return ("phong", kd, ks, ex)
    if kind == "roughdielectric":
        return ("roughdielectric", _scalar(p["alpha"].attrib["value"]), _scalar(p["intIOR"].attrib["value"]),
                _scalar(p["extIOR"].attrib["value"]))
    raise ValueError("unsupported bsdf type: %s" % kind)


def _phase(node):
    if node is None or node.attrib["type"] == "isotropic":
        return ("iso",)
    if node.attrib["type"] == "hg":
        return ("hg", _scalar(_named(node)["g"].attrib["value"]))
    raise ValueError("unsupported phase type: %s" % node.attrib["type"])


def _medium(node, base):
    kind = node.attrib["type"]
    phase = _phase(next((c for c in node if c.tag == "phase"), None))
    p = _named(node)
    if kind == "homogeneous":
        return MediumDesc("homogeneous", sigma_t=_scalar(p["sigmaT"].attrib["value"]),
                          albedo=_rgb(p["albedo"].attrib["value"])), phase
    if kind == "heterogeneous":
        dens = _named(p["density"])
        to_world = parse_transform(dens["toWorld"]) if "toWorld" in dens else np.eye(4)
        alb = _named(p["albedo"])
        if "filename" in alb:
            a_world = parse_transform(alb["toWorld"]) if "toWorld" in alb else np.eye(4)
            if np.abs(to_world - a_world).max() >= 1e-6:
                raise ValueError("albedo and density grids must share their toWorld")
            albedo = os.path.join(base, alb["filename"].attrib["value"])
        else:
            albedo = _rgb(alb["value"].attrib["value"])
        scale = _scalar(p["scale"].attrib["value"]) if "scale" in p else 1.0
        return MediumDesc("heterogeneous", density_file=os.path.join(base, dens["filename"].attrib["value"]),
                          to_world=to_world.astype(np.float32), scale=scale, albedo=albedo), phase
    raise ValueError("unsupported medium type: %s" % kind)


def _shape(node, base, bsdf_ids, med_ids, shape_id):
    kind = node.attrib["type"]
    if kind not in ("obj", "rectangle"):
        raise ValueError("unsupported shape type: %s" % kind)
    to_world, bsdf, m_int, m_ext, light, fname = np.eye(4), -1, -1, -1, None, None
    for ch in node:
        name = ch.attrib.get("name")
        if name == "filename":
            fname = os.path.join(base, ch.attrib["value"])
        elif name == "toWorld":
            to_world = parse_transform(ch)
        if ch.tag == "ref":
            if name == "interior":
                m_int = med_ids[ch.attrib["id"]]
            elif name == "exterior":
                m_ext = med_ids[ch.attrib["id"]]
            elif name is None:   # a named reference of another kind is ignored
                bsdf = bsdf_ids[ch.attrib["id"]]
        elif ch.tag == "emitter":
            q = _named(ch)
            rgb = _rgb(q["radiance"].attrib["value"]) if "radiance" in q else None
            if ch.attrib["type"] == "area":
                light = ("area", shape_id, rgb, False)
            elif ch.attrib["type"] == "areaEx":
                light = ("areaEx", shape_id, rgb, float(q["kappa"].attrib["value"]))
    if fname is None:
        v, f, nrm = _RECT_V.astype(np.float64), _RECT_F, None
    else:
        v, f, _, nrm = load_obj(fname)
    vh = np.c_[np.asarray(v, np.float64), np.ones(len(v))] @ to_world.T
    verts = (vh[:, :3] / vh[:, 3:4]).astype(np.float32)
    if nrm is not None:
        nrm = (np.asarray(nrm, np.float64) @ np.linalg.inv(to_world.T)[:3, :3]).astype(np.float32)
    return ShapeDesc(verts, np.ascontiguousarray(f, np.int32), bsdf_id=bsdf, med_int=m_int, med_ext=m_ext,
                     normals=nrm), light


def _camera(node, med_ids):
    p = _named(node)
    fov = _scalar(p["fov"].attrib["value"]) if "fov" in p else 45.0
    la = next((c for c in p["toWorld"] if c.tag.lower() == "lookat"), None) if "toWorld" in p else None
    if la is None:
        raise ValueError("the sensor needs a <lookat> toWorld")
    c2w = look_at(parse_vector(la.attrib["origin"]), parse_vector(la.attrib["target"]), parse_vector(la.attrib["up"]))
    w, h, crop, spp = 256, 256, [0, 0, -1, -1], None
    med = -1
    for ch in node:
        if ch.tag == "film":
            for g in ch:
                k = g.attrib.get("name")
                if k in ("width", "height", "cropOffsetX", "cropOffsetY", "cropWidth", "cropHeight"):
                    v = int(g.attrib["value"])
                    if k == "width":
                        w = v
                    elif k == "height":
                        h = v
                    else:
                        crop[["cropOffsetX", "cropOffsetY", "cropWidth", "cropHeight"].index(k)] = v
        elif ch.tag == "sampler":
            for g in ch:
                if g.attrib.get("name") == "sampleCount":
                    spp = int(g.attrib["value"])
        elif ch.tag == "ref":
            med = med_ids[ch.attrib["id"]]
    f = 1.0 / math.tan(math.radians(0.5 * fov))
    return dict(cam_to_world=c2w, cam_to_ndc=np.diag([f, f, 1.0]).astype(np.float32), width=w, height=h,
                crop=None if crop[2] == -1 or crop[3] == -1 else tuple(crop), cam_med=med, spp=spp)


def load_mitsuba(path, nder=1):
    """Mitsuba XML scene -> (SceneDesc, integrator type name). Tangent inputs are zero
    (set them on the returned description); ``spp`` and ``max_bounces`` default to
    the file's sampleCount and maxDepth."""
    base = os.path.dirname(os.path.abspath(path))
    root = ET.parse(path).getroot()
    cam, integrator, max_depth = None, None, None
    bsdfs, bsdf_ids, media, med_ids, phases, shapes, lights = [], {}, [], {}, [], [], []
    for ch in root:
        if ch.tag == "sensor":
            cam = _camera(ch, med_ids)
        elif ch.tag == "bsdf":
            b = _bsdf(ch)
            if "id" in ch.attrib:
                bsdf_ids[ch.attrib["id"]] = len(bsdfs)
                bsdfs.append(b)
        elif ch.tag == "medium":
            m, ph = _medium(ch, base)
            if "id" in ch.attrib:
                med_ids[ch.attrib["id"]] = len(media)
                m.phase_id = len(phases)
                media.append(m)
                phases.append(ph)
        elif ch.tag == "shape":
            s, light = _shape(ch, base, bsdf_ids, med_ids, len(shapes))
            shapes.append(s)
            if light is not None:
                lights.append(light)
        elif ch.tag == "integrator":
            integrator = ch.attrib["type"]
            if integrator not in _INTEGRATORS:
                raise ValueError("integrator type [ %s ] not supported" % integrator)
            for g in ch:
                if g.attrib.get("name") == "maxDepth":
                    max_depth = int(g.attrib["value"])
    if cam is None:
        raise ValueError("%s: no sensor" % path)
    for lid, l in enumerate(lights):
        shapes[l[1]].light_id = lid
    spp = cam.pop("spp")
    d = SceneDesc(nder=nder, shapes=shapes, bsdfs=bsdfs, lights=lights, phases=phases, media=media,
                  name=os.path.splitext(os.path.basename(path))[0], **cam)
    if spp:
        d.spp = spp
    if max_depth is not None and max_depth >= 0:
        d.max_bounces = max_depth
    return d, integrator


def integrator_for(mod, name):
    """Renderer object for an integrator type name of the XML: every integrator the
    reference's load_mitsuba.py:318-335 accepts."""
    if name == "volpathAD":
        return mod.VolPathTracerAD()
    if name == "volpath":
        return mod.VolPathTracer()
    if name == "volpath_simple":
        return mod.VolPathTracerSimple()
    if name == "path":
        return mod.PathTracer()
    if name == "direct":
        return mod.DirectIntegrator()
    if name == "ptracer":
        return mod.ParticleTracer()
    if name == "pathAD":
        return mod.PathTracerAD()
    if name == "directAD":
        return mod.DirectAD()
    raise NotImplementedError("integrator %s is not provided" % name)

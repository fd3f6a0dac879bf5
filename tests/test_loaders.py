"""Scene ingestion (pydtrt/load_obj.py, pydtrt/load_mitsuba.py): hand-computed OBJ
cases, transform closed forms, and the reference's example scenes (glass_1, shadow_1;
their XML and OBJ files are committed under tests/golden/scenes)."""
import math
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from dtrt_amd import loaders, scenes

SCENES = os.path.join(os.path.dirname(__file__), "golden", "scenes")


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_obj_dedup_quads_uvs_normals(tmp_path):
    path = _write(tmp_path, "m.obj", "\n".join([
        "v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0",
        "vt 0 0", "vt 1 0", "vt 1 1",
        "vn 0 0 1", "vn 0 0 -1",
        "f 1/1/1 2/2/1 3/3/1 4/1/1",      # quad -> (0,1,2), (0,2,3)
        "f 1/1/1 3/3/1 2/2/2",            # corner 2/2/2 differs in its normal: a new vertex
    ]))
    v, f, uv, n = loaders.load_obj(path)
    assert f.tolist() == [[0, 1, 2], [0, 2, 3], [0, 2, 4]]
    assert v.tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [1, 0, 0]]
    assert uv.tolist() == [[0, 0], [1, 0], [1, 1], [0, 0], [1, 0]]
    assert n.tolist() == [[0, 0, 1]] * 4 + [[0, 0, -1]]
    assert v.dtype == np.float32 and f.dtype == np.int32


def test_obj_negative_indices_follow_the_reference(tmp_path):
    # After the first face 3 vertices have been emitted, so the position index -1
    # becomes -1 + 3 = 2: the third pool vertex (2,0,0), where OBJ's relative
    # indexing would give the last one (3,0,0). Negative normal indices count from
    # the end of the normal pool.
    path = _write(tmp_path, "n.obj", "\n".join([
        "v 0 0 0", "v 1 0 0", "v 2 0 0", "v 3 0 0", "vn 0 1 0", "vn 1 0 0",
        "f 1 2 3", "f -1 3 4", "f 1//-1 2//-2 4//-1",
    ]))
    v, f, uv, n = loaders.load_obj(path)
    assert f.tolist() == [[0, 1, 2], [2, 2, 3], [4, 5, 6]]
    assert v[2].tolist() == [2, 0, 0] and v[3].tolist() == [3, 0, 0]
    assert n.tolist() == [[1, 0, 0], [0, 1, 0], [1, 0, 0]]
    assert uv is None


def test_transform_order_and_rotation():
    node = ET.fromstring('<transform name="toWorld"><translate x="1"/><scale x="2" y="3" z="4"/>'
                         '<rotate z="1" angle="90"/></transform>')
    m = loaders.parse_transform(node)
    p = m @ np.array([1.0, 1.0, 1.0, 1.0])
    # translate (2,1,1) -> scale (4,3,4) -> rotate 90 deg about z: (x,y) -> (-y,x)
    assert np.allclose(p[:3], [-3.0, 4.0, 4.0])
    r = loaders.rotate_matrix((1, 1, 1), 120.0)[:3, :3]          # cyclic permutation x->y->z
    assert np.allclose(r @ [1, 0, 0], [0, 1, 0], atol=1e-12)
    assert np.allclose(loaders.rotate_matrix((0, 0, 0), 30.0), np.eye(4))
    mt = ET.fromstring('<transform><matrix value="1 0 0 5 0 1 0 6 0 0 1 7 0 0 0 1"/><scale x="2"/></transform>')
    assert np.allclose(loaders.parse_transform(mt) @ [0, 0, 0, 1], [10, 6, 7, 1])


def test_glass_1_scene():
    d, integ = loaders.load_mitsuba(os.path.join(SCENES, "glass_1", "scene.xml"), nder=4)
    assert integ == "volpath" and (d.width, d.height, d.spp, d.max_bounces) == (256, 256, 512, 5)
    assert [len(s.indices) for s in d.shapes] == [768, 12, 2, 2]
    assert [s.bsdf_id for s in d.shapes] == [1, 0, 3, 2]
    assert [b[0] for b in d.bsdfs] == ["roughdielectric", "phong", "diffuse", "diffuse"]
    assert d.bsdfs[0][1:] == pytest.approx((0.3, 1.5, 1.0)) and d.bsdfs[1][3] == 30.0
    assert d.lights == [("area", 2, (15.0, 15.0, 15.0), False)] and d.shapes[2].light_id == 0
    assert d.shapes[1].med_int == 0 and d.shapes[1].med_ext == -1 and d.cam_med == -1
    m = d.media[0]
    assert m.kind == "homogeneous" and m.sigma_t == pytest.approx(0.1) and m.albedo == (0.0, 0.0, 0.0)
    assert d.phases == [("iso",)] and m.phase_id == 0
    c2w = d.cam_to_world.astype(np.float64)
    assert np.allclose(c2w[:3, :3].T @ c2w[:3, :3], np.eye(3), atol=1e-6)
    assert np.allclose(c2w[:3, 3], [95.2236, 87.7657, 101.679], atol=1e-4)
    fwd = np.array([94.6555, 87.2046, 101.077]) - c2w[:3, 3]
    assert np.allclose(c2w[:3, 2], fwd / np.linalg.norm(fwd), atol=1e-5)
    assert d.cam_to_ndc[0, 0] == pytest.approx(1 / math.tan(math.radians(5.0)))
    # apple: translate(-0.5,6,-1.4) -> scale 2 -> translate(0.5,-6,1.4): p -> 2p + (-0.5, 6, -1.4)
    v0, _, _, n0 = loaders.load_obj(os.path.join(SCENES, "glass_1", "apple.obj"))
    assert np.allclose(d.shapes[0].vertices, 2 * v0 + [-0.5, 6.0, -1.4], atol=1e-4)
    # normals go through the inverse of toWorld (uniform scale 2 -> halved, not renormalised)
    assert np.allclose(d.shapes[0].normals, 0.5 * n0, atol=1e-6)
    # the light quad: y = 30 plane rotated -45 deg about x, shifted, then -60 deg about y
    lv = d.shapes[2].vertices.astype(np.float64)
    e = np.cross(lv[1] - lv[0], lv[2] - lv[0])
    assert np.linalg.norm(e) == pytest.approx(80 * 80, rel=1e-5)


def test_shadow_1_scene():
    d, integ = loaders.load_mitsuba(os.path.join(SCENES, "shadow_1", "cbox.xml"))
    assert integ == "volpathAD" and d.spp == 1024
    assert d.bsdfs[4] == ("phong", (0.5, 0.5, 0.5), (0.2, 0.2, 0.2), 100.0) and d.bsdfs[5] == ("null",)
    assert [s.bsdf_id for s in d.shapes] == [3, 0, 4, 5]
    assert d.shapes[0].light_id == 0 and d.lights[0][2] == (8.0, 8.0, 8.0)
    assert np.allclose(d.shapes[0].vertices[:, 1], 548.7999 - 0.5)
    assert d.shapes[3].med_int == 0 and d.media[0].albedo == pytest.approx((0.9, 0.5, 0.5))


def test_heterogeneous_medium_rectangle_and_emitters(tmp_path):
    dens = scenes.fbm_grid(8, seed=3)
    scenes.write_vol(str(tmp_path / "density.vol"), dens, (-1, -1, -1), (1, 1, 1))
    scenes.write_vol(str(tmp_path / "albedo.vol"), np.full(dens.shape + (3,), 0.7, np.float32), (-1, -1, -1), (1, 1, 1))
    xml = """<scene version="0.5.0">
      <integrator type="volpathAD"><integer name="maxDepth" value="7"/></integrator>
      <medium type="heterogeneous" id="cube">
        <volume name="density" type="gridvolume"><string name="filename" value="density.vol"/>
          <transform name="toWorld"><scale x="2" y="2" z="2"/></transform></volume>
        <volume name="albedo" type="gridvolume"><string name="filename" value="albedo.vol"/>
          <transform name="toWorld"><scale x="2" y="2" z="2"/></transform></volume>
        <float name="scale" value="3.5"/>
        <phase type="hg"><float name="g" value="0.4"/></phase>
      </medium>
      <sensor type="perspective"><float name="fov" value="30"/>
        <transform name="toWorld"><lookat origin="0, 0, 6" target="0, 0, 0" up="0, 1, 0"/></transform>
        <film type="hdrfilm"><integer name="width" value="40"/><integer name="height" value="30"/>
          <integer name="cropOffsetX" value="4"/><integer name="cropOffsetY" value="2"/>
          <integer name="cropWidth" value="10"/><integer name="cropHeight" value="12"/></film>
        <ref id="cube"/></sensor>
      <bsdf type="null" id="null"/>
      <bsdf type="diffuse"><spectrum name="reflectance" value="0.3"/></bsdf>
      <shape type="rectangle"><ref id="null"/><ref name="interior" id="cube"/>
        <emitter type="areaEx"><spectrum name="radiance" value="2"/><float name="kappa" value="5"/></emitter></shape>
    </scene>"""
    d, integ = loaders.load_mitsuba(_write(tmp_path, "s.xml", xml), nder=2)
    assert integ == "volpathAD" and d.max_bounces == 7 and d.cam_med == 0
    assert (d.width, d.height, d.crop) == (40, 30, (4, 2, 10, 12)) and d.out_shape() == (3, 12, 10, 3)
    assert len(d.bsdfs) == 1                      # the bsdf without an id is dropped
    m = d.media[0]
    assert m.density_file == str(tmp_path / "density.vol") and m.albedo == str(tmp_path / "albedo.vol")
    assert m.scale == 3.5 and np.allclose(m.to_world, np.diag([2, 2, 2, 1])) and d.phases == [("hg", pytest.approx(0.4))]
    assert d.shapes[0].indices.tolist() == [[0, 2, 1], [1, 2, 3]]
    assert d.lights == [("areaEx", 0, (2.0, 2.0, 2.0), 5.0)]


def test_example_scene_renders_through_the_oracle(oracle):
    d, _ = loaders.load_mitsuba(os.path.join(SCENES, "shadow_1", "cbox.xml"), nder=1)
    d.width = d.height = 16
    img = scenes.render(oracle, d, seed=0, spp=4, sppe=0, sppse=0)
    assert np.isfinite(img).all() and img[0].mean() > 0
    assert loaders.integrator_for(oracle, "volpath").kind == 1
    kinds = [loaders.integrator_for(oracle, n).kind for n in
             ("volpathAD", "volpath", "volpath_simple", "path", "direct", "ptracer", "pathAD", "directAD")]
    assert kinds == [0, 1, 2, 3, 4, 5, 6, 7]
    with pytest.raises(NotImplementedError):
        loaders.integrator_for(oracle, "bdpt")

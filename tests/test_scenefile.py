"""Scene files in RTLib's formats: OBJ/MTL parsing and light splitting, the
scene JSON (instances, transforms, camera controller), and the OBJ writer's
round trip of the procedural scenes."""
import json
import math

import numpy as np
import pytest

from rtlib_amd import scenefile, scenes

MTL = """# test materials
newmtl white
Kd 0.5 0.5 0.5
Ks 0 0 0
Ns 10
illum 2
newmtl shiny
Kd 0.2 0.3 0.4
Ks 0.5 0.5 0.5
Ns 30
illum 2
newmtl glass
Kd 0 0 0
Ks 1 1 1
Ni 1.5
illum 7
newmtl lampA
Kd 0.78 0.78 0.78
Ke 17 12 4
newmtl lampB
Ke 1 2 3
"""

OBJ = """mtllib m.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 1 0 1
g floor
usemtl white
f 1/1/1 2/2/1 3/3/1 4/4/1
usemtl shiny
f 1//1 5//1 6//1
g lamp
usemtl lampA
f -3 -2 -1
usemtl white
f 2 3 6
usemtl lampB
f 4 5 6
f 1 2 3 4 5
o glassy
usemtl glass
f 2 4 6
"""


@pytest.fixture
def model_dir(tmp_path):
    (tmp_path / "m.mtl").write_text(MTL)
    (tmp_path / "m.obj").write_text(OBJ)
    return tmp_path


def test_mtl_keys_and_material_types(model_dir):
    mats = {m["name"]: m for m in scenefile.load_mtl(model_dir / "m.mtl")}
    assert mats["shiny"]["diffCol"] == (0.2, 0.3, 0.4) and mats["shiny"]["specCol"] == (0.5, 0.5, 0.5)
    assert mats["shiny"]["shinness"] == 30.0 and mats["glass"]["refrIndx"] == 1.5
    assert mats["lampB"]["diffCol"] == (0.0, 0.0, 0.0)  # tinyobjloader's defaults
    kinds = {n: scenefile.specify_material_type(m) for n, m in mats.items()}
    assert kinds == {"white": "phong", "shiny": "phong", "glass": "glass", "lampA": "light", "lampB": "light"}


def test_obj_groups_triangulation_and_light_split(model_dir):
    m = scenefile.load_obj(model_dir / "m.obj", split_light=False)
    assert list(m.meshes) == ["floor", "lamp", "glassy"]
    fl = m.meshes["floor"]
    # quad -> fan (1,2,3) (1,3,4); v//vn and v/vt/vn forms
    assert fl.tris.tolist() == [[0, 1, 2], [0, 2, 3], [0, 4, 5]]
    assert fl.mat_local.tolist() == [0, 0, 1]
    lamp = m.meshes["lamp"]
    # negative indices count back from the last vertex; the pentagon gives three triangles
    assert lamp.tris.tolist()[:3] == [[3, 4, 5], [1, 2, 5], [3, 4, 5]]
    assert lamp.tris.shape[0] == 6
    m = scenefile.load_obj(model_dir / "m.obj")
    # lamp mixes emissive and white faces: .Surface + .Emission, and the
    # emission part has two materials: one mesh per material
    assert sorted(m.meshes) == ["floor", "glassy", "lamp.Emission.0", "lamp.Emission.1", "lamp.Surface"]
    assert not m.meshes["lamp.Surface"].has_light and m.meshes["lamp.Surface"].tris.shape[0] == 1
    e0, e1 = m.meshes["lamp.Emission.0"], m.meshes["lamp.Emission.1"]
    assert e0.has_light and e1.has_light
    assert e0.tris.shape[0] + e1.tris.shape[0] == 5
    assert {m.materials[e0.materials[0]]["name"], m.materials[e1.materials[0]]["name"]} == {"lampA", "lampB"}


def test_camera_controller_closed_form():
    cc = scenefile.CameraController((0.0, 1.0, 5.0))
    cam = cc.camera(640, 480)
    assert np.allclose(cam["eye"], (0, 1, 5)) and np.allclose(cam["lookat"], (0, 1, 4), atol=1e-12)
    assert cam["fovy"] == 45.0 and cam["up"] == (0.0, 1.0, 0.0)
    # yaw 0 looks along +x; pitch 30 lifts the view by 30 degrees
    cc = scenefile.CameraController((0, 0, 0), yaw=0.0, pitch=30.0)
    assert np.allclose(cc.front(), (math.cos(math.radians(30)), 0.5, 0.0))
    j = scenefile.CameraController.from_json({"Position": [1, 2, 3], "Yaw": -45.0, "Pitch": 0.0, "Zoom": 30.0,
                                              "MouseSensitivity": 0.1, "MovementSpeed": 2.0})
    f = j.front()
    assert np.allclose(f, (math.sqrt(0.5), 0.0, -math.sqrt(0.5))) and j.zoom == 30.0


def _canonical(sc, lm):
    """Per-triangle rows (sorted vertex coordinates, material arrays), sorted."""
    v = np.asarray(sc["verts"], np.float32)[np.asarray(sc["idx"], np.int64)]
    v = np.sort(v.reshape(-1, 3, 3).view([("x", np.float32), ("y", np.float32), ("z", np.float32)]), axis=1)
    v = v.view(np.float32).reshape(-1, 9)
    spec = sc["spec"] if sc.get("spec") is not None else np.zeros((v.shape[0], 4), np.float32)
    rows = np.concatenate([v, sc["mat"][:, None].astype(np.float32), sc["albedo"], spec, sc["emission"]], axis=1)
    order = np.lexsort(rows.T[::-1])
    return rows[order], np.asarray(lm)[order]


def test_round_trip_of_procedural_scene(tmp_path):
    sc, lm = scenes.cornell_materials(32, 32)
    scenefile.save_obj(sc, lm, str(tmp_path / "cornell.obj"))
    sj = scenefile.default_scene_json(str(tmp_path / "cornell.obj"))
    (tmp_path / "scene.json").write_text(json.dumps(sj))
    s2, lm2, cfg = scenefile.load_scene(str(tmp_path / "scene.json"))
    assert cfg["Width"] == 1024 and s2["camera"]["width"] == 1024
    a, la = _canonical(sc, lm)
    b, lb = _canonical(s2, lm2)
    assert np.array_equal(a, b)  # identical triangles and materials, bit for bit
    # the same partition of emissive triangles into light meshes, up to relabelling
    em = a[:, 18:21].sum(axis=1) > 0  # columns: 9 coordinates, type, albedo, spec, Le
    pairs = set(zip(la[em].tolist(), lb[em].tolist()))
    assert len(pairs) == len(set(la[em].tolist())) == len(set(lb[em].tolist())) == 2


def test_instances_transforms_and_nesting(model_dir):
    sj = {
        "ObjModels": {"Assets": {"M": "m.obj"}},
        "World": {
            "Geometries": {"G": {"Type": "ObjModel", "Base": "M"}},
            "GeometryASs": {"GAS": {"Type": "GeometryAS", "Geometries": ["G"]}},
            "Instances": {
                "I0": {"Type": "Instance", "ASType": "Geometry", "Base": "GAS",
                       "Transform": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]},
                "I1": {"Type": "Instance", "ASType": "Geometry", "Base": "GAS",
                       "Transform": [2, 0, 0, 10, 0, 2, 0, 0, 0, 0, 2, 0]},
                "Outer": {"Type": "Instance", "ASType": "Instance", "Base": "Sub",
                          "Transform": [1, 0, 0, 0, 0, 1, 0, 100, 0, 0, 1, 0]},
            },
            "InstanceASs": {"Root": {"Type": "InstanceAS", "Instances": ["I0", "Outer"]},
                            "Sub": {"Type": "InstanceAS", "Instances": ["I1"]}},
        },
        "CameraController": {"Position": [0, 0, 3], "Yaw": -90.0, "Pitch": 0.0, "Zoom": 40.0},
        "Config": {"Width": 64, "Height": 48},
    }
    sc, lm, cfg = scenefile.load_scene(sj, base_dir=str(model_dir))
    nt = sc["idx"].shape[0]
    assert nt == 2 * 10  # floor 3 + lamp 6 + glassy 1 per instance
    p = sc["verts"][sc["idx"]]
    lo, hi = p.reshape(-1, 3).min(0), p.reshape(-1, 3).max(0)
    # I0 spans the unit cube; Outer(I1) = translate y 100 of (scale 2, translate x 10)
    assert np.allclose(lo, 0.0) and np.allclose(hi, [12.0, 102.0, 2.0])
    assert sc["camera"]["fovy"] == 40.0 and sc["camera"]["width"] == 64
    # 2 light meshes per instance
    em = sc["emission"].sum(axis=1) > 0
    assert len(set(lm[em].tolist())) == 4
    with pytest.raises(ValueError):
        scenefile.load_scene(sj, base_dir=str(model_dir), nee_threshold=2)

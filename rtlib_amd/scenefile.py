"""Scene files in RTLib's formats, turned into the arrays rtpg_scene_upload takes.

- Wavefront OBJ/MTL, as RTLib's ObjModel reads them through tinyobjloader
  (Core/Src/BinaryReader.cpp:89-192 mesh groups and light splitting, :385-433
  materials): one mesh per `o`/`g` group, materials under RTLib's keys
  (`diffCol`, `specCol`, `tranCol`, `emitCol`, `shinness`, `refrIndx`,
  `illum`, texture names), emitter meshes split from surfaces.
- The scene JSON (SceneData, Core/Include/RTLib/Core/Scene.h:130-233;
  WorldData, Core/World.h; CameraController, Core/Camera.h:108-252): ObjModels
  assets, Geometries / GeometryASs / Instances / InstanceASs with 3x4 instance
  transforms, the camera controller and the trace config.
- Material types as SpecifyMaterialType (Test/RTLibExtOPX7Test.h:1572-1601):
  emissive -> light, illum 7 -> glass, otherwise Phong.
- Texture coordinates, vertex normals and texture maps (map_Kd / map_Ks /
  map_Ke) as the hitgroup and mesh-light records carry them
  (RTLibExtOPX7Test.h:386-511, 1602-1638; RTLibExtOPX7TestApplication.cpp:
  262-280); a missing texture coordinate is (0.5, 0.5) and a missing normal
  (0, 1, 0), RTLib's defaults (Core/Src/BinaryReader.cpp:323-349).

Meshes are flattened into one triangle list in world space (the tracer has a
single BVH over all instances); every emissive mesh instance becomes one light
mesh for the uniform mesh-then-triangle selection (InitLight,
RTLibExtOPX7TestApplication.cpp:244-290)."""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

from .scenes import MAT_GLASS, MAT_LAMBERT, MAT_PHONG

# tinyobjloader's InitMaterial defaults
_MTL_DEFAULTS = {"diffCol": (0.0, 0.0, 0.0), "specCol": (0.0, 0.0, 0.0), "tranCol": (0.0, 0.0, 0.0),
                 "emitCol": (0.0, 0.0, 0.0), "shinness": 1.0, "refrIndx": 1.0, "illum": 0,
                 "diffTex": "", "specTex": "", "emitTex": "", "shinTex": ""}
_MTL_COLOURS = {"Kd": "diffCol", "Ks": "specCol", "Tf": "tranCol", "Kt": "tranCol", "Ke": "emitCol"}
_MTL_TEXTURES = {"map_Kd": "diffTex", "map_Ks": "specTex", "map_Ke": "emitTex", "map_Ns": "shinTex"}


def default_material(name="default"):
    m = dict(_MTL_DEFAULTS)
    m["name"] = name
    return m


def load_mtl(path):
    """MTL file -> list of materials (dicts with RTLib's keys), in file order."""
    mats = []
    cur = None
    base = os.path.dirname(os.path.abspath(path))
    with open(path, "r", encoding="utf-8", errors="replace") as f:
        for raw in f:
            line = raw.split("#", 1)[0].strip()
            if not line:
                continue
            tok = line.split()
            key = tok[0]
            if key == "newmtl":
                cur = default_material(" ".join(tok[1:]))
                mats.append(cur)
            elif cur is None:
                continue
            elif key in _MTL_COLOURS:
                vals = [float(x) for x in tok[1:4]]
                if len(vals) == 1:
                    vals = vals * 3
                cur[_MTL_COLOURS[key]] = tuple(vals)
            elif key == "Ns":
                cur["shinness"] = float(tok[1])
            elif key == "Ni":
                cur["refrIndx"] = float(tok[1])
            elif key == "illum":
                cur["illum"] = int(tok[1])
            elif key in _MTL_TEXTURES:
                cur[_MTL_TEXTURES[key]] = os.path.join(base, tok[-1])
    return mats


def is_emissive(mat):
    e = mat["emitCol"]
    return e[0] + e[1] + e[2] > 0.0


def specify_material_type(mat):
    """SpecifyMaterialType (RTLibExtOPX7Test.h:1572-1601): 'light', 'glass' or 'phong'."""
    if is_emissive(mat):
        return "light"
    if int(mat.get("illum", 0)) == 7:
        return "glass"
    return "phong"


@dataclass
class Mesh:
    """A MeshUniqueResource: triangles (indices into the model's vertices),
    per-triangle index into `materials` (global material indices), and per
    corner the index of its texture coordinate and normal (-1: none)."""
    name: str
    tris: np.ndarray
    mat_local: np.ndarray
    materials: list
    has_light: bool = False
    tc: np.ndarray = None
    nrm: np.ndarray = None

    def select(self, sel, name, mat_local, materials, has_light):
        return Mesh(name, self.tris[sel], mat_local, materials, has_light,
                    None if self.tc is None else self.tc[sel], None if self.nrm is None else self.nrm[sel])


@dataclass
class ObjModel:
    verts: np.ndarray
    meshes: dict = field(default_factory=dict)    # name -> Mesh, in file order
    materials: list = field(default_factory=list)  # RTLib material dicts
    texcoords: np.ndarray = None                   # vt, N x 2
    normals: np.ndarray = None                     # vn, N x 3

    def mesh_materials(self, mesh):
        return [self.materials[i] for i in mesh.materials]


def _face_index(tok, nv):
    i = int(tok.split("/", 1)[0])
    return i - 1 if i > 0 else nv + i


def _corner_index(tok, k, n):
    """Index k (1: texture coordinate, 2: normal) of a face corner `v/vt/vn`,
    -1 when absent."""
    parts = tok.split("/")
    if len(parts) <= k or parts[k] == "":
        return -1
    i = int(parts[k])
    return i - 1 if i > 0 else n + i


def load_obj(path, split_light=True):
    """OBJ (+ its MTL libraries) -> ObjModel.  Polygons are triangulated as a
    fan from their first vertex.  Faces without `usemtl` take tinyobjloader's
    default material.  Groups of the same name are merged."""
    base = os.path.dirname(os.path.abspath(path))
    verts, vts, vns = [], [], []
    materials, mat_index = [], {}
    groups = {}   # name -> [tris, mats, texcoord corners, normal corners]
    order = []
    cur_name = ""
    cur_mat = -1
    nv = 0
    with open(path, "r", encoding="utf-8", errors="replace") as f:
        for raw in f:
            line = raw.split("#", 1)[0].strip()
            if not line:
                continue
            tok = line.split()
            key = tok[0]
            if key == "v":
                verts.append((float(tok[1]), float(tok[2]), float(tok[3])))
                nv += 1
            elif key == "vt":
                vts.append((float(tok[1]), float(tok[2]) if len(tok) > 2 else 0.0))
            elif key == "vn":
                vns.append((float(tok[1]), float(tok[2]), float(tok[3])))
            elif key == "f":
                idx = [_face_index(t, nv) for t in tok[1:]]
                if len(idx) < 3:
                    continue
                ti = [_corner_index(t, 1, len(vts)) for t in tok[1:]]
                ni = [_corner_index(t, 2, len(vns)) for t in tok[1:]]
                if cur_name not in groups:
                    groups[cur_name] = ([], [], [], [])
                    order.append(cur_name)
                g = groups[cur_name]
                for k in range(1, len(idx) - 1):
                    g[0].append((idx[0], idx[k], idx[k + 1]))
                    g[1].append(cur_mat)
                    g[2].append((ti[0], ti[k], ti[k + 1]))
                    g[3].append((ni[0], ni[k], ni[k + 1]))
            elif key in ("o", "g"):
                cur_name = " ".join(tok[1:]) if len(tok) > 1 else ""
            elif key == "usemtl":
                cur_mat = mat_index.get(" ".join(tok[1:]), -1)
            elif key == "mtllib":
                for lib in tok[1:]:
                    p = os.path.join(base, lib)
                    if os.path.exists(p):
                        for m in load_mtl(p):
                            mat_index.setdefault(m["name"], len(materials))
                            materials.append(m)
    v = np.asarray(verts, np.float32).reshape(-1, 3)
    dflt = -1
    if any(-1 in groups[n][1] for n in order):
        mat_index.setdefault("default", len(materials))
        materials.append(default_material())
        dflt = len(materials) - 1
    model = ObjModel(v, {}, materials, np.asarray(vts, np.float32).reshape(-1, 2),
                     np.asarray(vns, np.float32).reshape(-1, 3))
    for name in order:
        tris = np.asarray(groups[name][0], np.int64).reshape(-1, 3)
        if tris.size and (tris.min() < 0 or tris.max() >= v.shape[0]):
            raise ValueError(f"{path}: face index out of range in group '{name}'")
        tc = np.asarray(groups[name][2], np.int64).reshape(-1, 3)
        nrm = np.asarray(groups[name][3], np.int64).reshape(-1, 3)
        # corners naming a missing vt / vn take RTLib's defaults
        tc[(tc < 0) | (tc >= len(vts))] = -1
        nrm[(nrm < 0) | (nrm >= len(vns))] = -1
        gm = [dflt if m == -1 else m for m in groups[name][1]]
        # per-mesh material list in order of first use (BinaryReader.cpp:356-371)
        local, mats = {}, []
        ml = np.empty(len(gm), np.uint32)
        for j, m in enumerate(gm):
            if m not in local:
                local[m] = len(mats)
                mats.append(m)
            ml[j] = local[m]
        model.meshes[name] = Mesh(name, tris.astype(np.uint32), ml, mats, False, tc, nrm)
    if split_light:
        split_lights(model)
    return model


def split_lights(model):
    """ObjModel::SplitLight (BinaryReader.cpp:89-192): a mesh mixing emissive and
    other materials becomes `name.Surface` and `name.Emission`; an emissive mesh
    with several materials becomes one mesh per material, `name.<i>`."""
    out = {}
    for name, m in model.meshes.items():
        emit = [is_emissive(model.materials[g]) for g in m.materials]
        if any(emit) and not all(emit):
            fe = np.array([emit[i] for i in m.mat_local], bool)
            for suffix, sel, light in ((".Surface", ~fe, False), (".Emission", fe, True)):
                local, mats = {}, []
                ml = np.empty(int(sel.sum()), np.uint32)
                for j, i in enumerate(m.mat_local[sel]):
                    if int(i) not in local:
                        local[int(i)] = len(mats)
                        mats.append(m.materials[int(i)])
                    ml[j] = local[int(i)]
                out[name + suffix] = m.select(sel, name + suffix, ml, mats, light)
        else:
            m.has_light = bool(all(emit)) and len(emit) > 0
            out[name] = m
    final = {}
    for name, m in out.items():
        if m.has_light and len(m.materials) > 1:
            for i, g in enumerate(m.materials):
                sel = m.mat_local == i
                final[f"{name}.{i}"] = m.select(sel, f"{name}.{i}", np.zeros(int(sel.sum()), np.uint32), [g], True)
        else:
            final[name] = m
    model.meshes = final
    return model


# ------------------------------------------------------------------ camera
class CameraController:
    """CameraController (Core/Camera.h:108-252): position, yaw/pitch in degrees,
    world up (0, 1, 0), zoom = vertical field of view in degrees."""

    def __init__(self, position=(0.0, 0.0, 0.0), yaw=-90.0, pitch=0.0, zoom=45.0, sensitivity=0.025, speed=1.0):
        self.position = tuple(float(x) for x in position)
        self.up = (0.0, 1.0, 0.0)
        self.yaw, self.pitch, self.zoom = float(yaw), float(pitch), float(zoom)
        self.sensitivity, self.speed = float(sensitivity), float(speed)

    @classmethod
    def from_json(cls, j):
        return cls(j["Position"], j.get("Yaw", -90.0), j.get("Pitch", 0.0), j.get("Zoom", 45.0),
                   j.get("MouseSensitivity", 0.025), j.get("MovementSpeed", 1.0))

    def to_json(self):
        return {"Position": list(self.position), "Yaw": self.yaw, "Pitch": self.pitch, "Zoom": self.zoom,
                "MouseSensitivity": self.sensitivity, "MovementSpeed": self.speed}

    def front(self):
        y, p = math.radians(self.yaw), math.radians(self.pitch)
        f = np.array([math.cos(y) * math.cos(p), math.sin(p), math.sin(y) * math.cos(p)])
        return f / np.linalg.norm(f)

    def camera(self, width, height):
        """GetCamera(aspect): eye, lookat = eye + front, up, fovy = zoom."""
        eye = np.asarray(self.position, np.float64)
        return {"eye": tuple(float(x) for x in eye), "lookat": tuple(float(x) for x in eye + self.front()),
                "up": self.up, "fovy": self.zoom,
                "width": int(width), "height": int(height)}


# ------------------------------------------------------------------ scene JSON
def default_scene_json(obj_path):
    """GetDefaultSceneJson (RTLibExtOPX7Test.h:780-819): one OBJ asset, identity
    instance under Root, camera at (0, 1, 5) looking down -z, 1024 x 1024."""
    return {
        "ObjModels": {"CacheDir": ".", "Assets": {"CornellBox-Original": obj_path}},
        "World": {
            "Geometries": {"CornellBox-Original": {"Type": "ObjModel", "Base": "CornellBox-Original"}},
            "GeometryASs": {"CornellBox-Original": {"Type": "GeometryAS", "Geometries": ["CornellBox-Original"]}},
            "Instances": {"Instance0": {"Type": "Instance", "ASType": "Geometry", "Base": "CornellBox-Original",
                                        "Transform": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]}},
            "InstanceASs": {"Root": {"Type": "InstanceAS", "Instances": ["Instance0"]}},
        },
        "CameraController": CameraController((0.0, 1.0, 5.0), speed=10.0).to_json(),
        "Config": {"Width": 1024, "Height": 1024, "Samples": 1, "MaxSamples": 10000, "MaxDepth": 8,
                   "SamplesPerSave": 1000, "ImagePath": "."},
    }


def _affine(t):
    m = np.asarray(t, np.float64).reshape(3, 4)
    a = np.eye(4)
    a[:3, :] = m
    return a


def _material_arrays(mats):
    """Per material of a mesh: (type, albedo4, spec4, emission3)."""
    rows = []
    for m in mats:
        kind = specify_material_type(m)
        d = np.asarray(m["diffCol"], np.float32)
        s = np.asarray(m["specCol"], np.float32)
        if kind == "light":
            rows.append((MAT_LAMBERT, (0, 0, 0, 0), (0, 0, 0, 0), tuple(m["emitCol"])))
        elif kind == "glass":
            # reflection is weighted by specCol, refraction by 1 (SimpleTrace.cu:347-407)
            rows.append((MAT_GLASS, (*s, float(m["refrIndx"])), (0, 0, 0, 0), (0, 0, 0)))
        elif float(s.sum()) == 0.0:
            # Phong with a zero specular lobe is Lambert (selection prob. 1 - 1e-10 / diffuse)
            rows.append((MAT_LAMBERT, (*d, 0.0), (0, 0, 0, 0), (0, 0, 0)))
        else:
            rows.append((MAT_PHONG, (*d, 0.0), (*s, float(m["shinness"])), (0, 0, 0)))
    return rows


TEX_NONE = 0xFFFFFFFF


def load_texture(path):
    """Image file -> RGBA8 texels (H x W uint32, R in the low byte), rows
    bottom-up as LoadDefImage stores them (Core/Src/BinaryReader.cpp:768-782,
    stbi_load with 4 channels: grey -> (g, g, g, 255), 16-bit -> high byte)."""
    from PIL import Image

    with Image.open(path) as im:
        if im.mode in ("I;16", "I;16B", "I;16L", "I"):
            g = (np.asarray(im, np.uint32) >> 8).astype(np.uint8)
            a = np.stack([g, g, g, np.full_like(g, 255)], axis=-1)
        else:
            a = np.asarray(im.convert("RGBA"), np.uint8)
    a = a[::-1].astype(np.uint32)
    return np.ascontiguousarray(a[..., 0] | (a[..., 1] << 8) | (a[..., 2] << 16) | (a[..., 3] << 24))


# texture of each slot (diffuse, specular, emission on hit, emission when
# light-sampled): "reference" as the reference binds them -- its hitgroup
# record takes the specular slot from emitTex and the emission slot from
# specTex (RTLibExtOPX7Test.h:1634-1636), its mesh lights emitTex
# (RTLibExtOPX7TestApplication.cpp:272-277); "named" by the MTL keys.
_TEX_SLOTS = {"reference": ("diffTex", "emitTex", "specTex", "emitTex"),
              "named": ("diffTex", "specTex", "emitTex", "emitTex")}


def load_scene(path_or_json, base_dir=None, nee_threshold=None, texture_slots="reference", missing_normal="reference"):
    """Scene JSON (file path or dict) -> (scene dict, light_mesh, config dict).

    scene: the arrays of rtpg_scene_upload plus `camera`, and when the scene
    has texture maps `texcoord`, `tex_id`, `texels`, `tex`, when it has glass
    `vnormal` (rtpg_scene_attributes); light_mesh: light mesh id per triangle
    (for light_mode 1); config: the JSON's Config.  missing_normal: a corner
    without `vn` takes (0, 1, 0) ("reference", RTLib's default) or the face
    normal ("face")."""
    if texture_slots not in _TEX_SLOTS:
        raise ValueError(f"texture_slots must be one of {sorted(_TEX_SLOTS)}")
    if missing_normal not in ("reference", "face"):
        raise ValueError("missing_normal must be 'reference' or 'face'")
    slot_keys = _TEX_SLOTS[texture_slots]
    if isinstance(path_or_json, (str, os.PathLike)):
        base_dir = base_dir or os.path.dirname(os.path.abspath(path_or_json))
        with open(path_or_json, "r", encoding="utf-8") as f:
            sj = json.load(f)
    else:
        sj = path_or_json
        base_dir = base_dir or "."
    cfg = dict(sj.get("Config", {}))
    custom = cfg.get("Custom", {})
    th = int(custom.get("Nee.ThresholdMeshSize", 0) if nee_threshold is None else nee_threshold)
    assets = {}
    for name, p in sj["ObjModels"]["Assets"].items():
        p = p if os.path.isabs(p) else os.path.join(base_dir, p)
        assets[name] = load_obj(p)
    world = sj["World"]
    geoms = world["Geometries"]

    def geometry_meshes(gname):
        g = geoms[gname]
        if g["Type"] != "ObjModel":
            raise ValueError(f"geometry '{gname}': type {g['Type']} is not supported (triangle meshes only)")
        model = assets[g["Base"]]
        if "Meshes" not in g:
            return [(model, mesh, model.mesh_materials(mesh)) for mesh in model.meshes.values()]
        out = []
        for mname, md in g["Meshes"].items():
            mesh = model.meshes[md["Base"]]
            mats = md.get("Materials") or model.mesh_materials(mesh)
            mats = [{**default_material(), **mm} for mm in mats]
            out.append((model, mesh, mats))
        return out

    V, I, ALB, MAT, EM, SP, LM, TC, NR, TID = [], [], [], [], [], [], [], [], [], []
    nv = 0
    n_light = 0
    tex_of, tex_rows, texel_chunks = {}, [], []
    n_texels = 0

    def tex_index(path):
        nonlocal n_texels
        if not path:
            return TEX_NONE
        if path not in tex_of:
            if not os.path.exists(path):
                raise ValueError(f"texture '{path}' not found")
            img = load_texture(path)
            tex_of[path] = len(tex_rows)
            tex_rows.append((n_texels, img.shape[1], img.shape[0], 1))  # LoadFromPath: wrap addressing
            texel_chunks.append(img.ravel())
            n_texels += img.size
        return tex_of[path]

    def emit(gas_name, xf):
        nonlocal nv, n_light
        for gname in world["GeometryASs"][gas_name]["Geometries"]:
            for model, mesh, mats in geometry_meshes(gname):
                if mesh.tris.shape[0] == 0:
                    continue
                used = np.unique(mesh.tris)
                remap = np.full(model.verts.shape[0], -1, np.int64)
                remap[used] = np.arange(used.size)
                p = model.verts[used].astype(np.float64)
                p = p @ xf[:3, :3].T + xf[:3, 3]
                V.append(p.astype(np.float32))
                I.append((remap[mesh.tris] + nv).astype(np.uint32))
                nv += used.size
                # corner attributes: texture coordinates and normals (world space:
                # the inverse transpose of the instance transform)
                nt_m = mesh.tris.shape[0]
                tci = mesh.tc if mesh.tc is not None else np.full((nt_m, 3), -1, np.int64)
                tcs = np.full((nt_m, 3, 2), 0.5, np.float32)
                if model.texcoords is not None and model.texcoords.shape[0]:
                    ok = tci >= 0
                    tcs[ok] = model.texcoords[tci[ok]]
                TC.append(tcs)
                nri = mesh.nrm if mesh.nrm is not None else np.full((nt_m, 3), -1, np.int64)
                nrs = np.zeros((nt_m, 3, 3), np.float64)
                nrs[...] = (0.0, 1.0, 0.0)
                if missing_normal == "face":
                    q = model.verts[mesh.tris.astype(np.int64)].astype(np.float64)
                    nrs[...] = np.cross(q[:, 1] - q[:, 0], q[:, 2] - q[:, 0])[:, None, :]
                if model.normals is not None and model.normals.shape[0]:
                    ok = nri >= 0
                    nrs[ok] = model.normals[nri[ok]]
                nx = np.linalg.inv(xf[:3, :3]).T
                NR.append((nrs @ nx.T).astype(np.float32))
                ids = np.array([[tex_index(m.get(k, "")) for k in slot_keys] for m in mats], np.uint32).reshape(-1, 4)
                rows = _material_arrays(mats)
                ml = mesh.mat_local
                TID.append(ids[ml])
                MAT.append(np.array([rows[i][0] for i in ml], np.uint32))
                ALB.append(np.array([rows[i][1] for i in ml], np.float32).reshape(-1, 4))
                SP.append(np.array([rows[i][2] for i in ml], np.float32).reshape(-1, 4))
                EM.append(np.array([rows[i][3] for i in ml], np.float32).reshape(-1, 3))
                light = mesh.has_light and any(is_emissive(m) for m in mats)
                if light and mesh.tris.shape[0] <= th:
                    raise ValueError(f"mesh '{mesh.name}' ({mesh.tris.shape[0]} triangles) is below "
                                     f"Nee.ThresholdMeshSize {th}: emitters outside light sampling are not supported")
                LM.append(np.full(mesh.tris.shape[0], n_light if light else 0, np.uint32))
                n_light += 1 if light else 0

    def walk(ias_name, xf):
        for iname in world["InstanceASs"][ias_name]["Instances"]:
            inst = world["Instances"][iname]
            t = xf @ _affine(inst["Transform"])
            if inst["ASType"] == "Geometry":
                emit(inst["Base"], t)
            else:
                walk(inst["Base"], t)

    walk("Root", np.eye(4))
    if not I:
        raise ValueError("the scene has no triangles")
    w = int(cfg.get("Width", 1024))
    h = int(cfg.get("Height", 1024))
    cc = CameraController.from_json(sj["CameraController"]) if "CameraController" in sj else CameraController()
    scene = {
        "verts": np.ascontiguousarray(np.concatenate(V), np.float32),
        "idx": np.ascontiguousarray(np.concatenate(I), np.uint32),
        "albedo": np.ascontiguousarray(np.concatenate(ALB), np.float32),
        "mat": np.ascontiguousarray(np.concatenate(MAT), np.uint32),
        "emission": np.ascontiguousarray(np.concatenate(EM), np.float32),
        "spec": np.ascontiguousarray(np.concatenate(SP), np.float32),
        "camera": cc.camera(w, h),
    }
    if tex_rows:
        scene["texcoord"] = np.ascontiguousarray(np.concatenate(TC), np.float32)
        scene["tex_id"] = np.ascontiguousarray(np.concatenate(TID), np.uint32)
        scene["texels"] = np.ascontiguousarray(np.concatenate(texel_chunks), np.uint32)
        scene["tex"] = np.asarray(tex_rows, np.uint32).reshape(-1, 4)
    if np.any(scene["mat"] == MAT_GLASS):
        scene["vnormal"] = np.ascontiguousarray(np.concatenate(NR), np.float32)
    return scene, np.ascontiguousarray(np.concatenate(LM), np.uint32), cfg


# ------------------------------------------------------------------ writing
def save_obj(scene, light_mesh, obj_path, name="scene"):
    """Write a scene dict as OBJ + MTL: one group per (material, light mesh),
    so that load_obj / load_scene give back the same triangles and materials."""
    mtl_path = os.path.splitext(obj_path)[0] + ".mtl"
    mat = np.asarray(scene["mat"])
    alb = np.asarray(scene["albedo"], np.float32)
    em = np.asarray(scene["emission"], np.float32)
    sp = scene.get("spec")
    sp = np.zeros((mat.shape[0], 4), np.float32) if sp is None else np.asarray(sp, np.float32)
    lm = np.zeros(mat.shape[0], np.uint32) if light_mesh is None else np.asarray(light_mesh)
    emissive = em.sum(axis=1) > 0
    keys = np.concatenate([mat[:, None].astype(np.float32), alb, sp, em,
                           np.where(emissive, lm, 0)[:, None].astype(np.float32)], axis=1)
    uniq, inv = np.unique(keys, axis=0, return_inverse=True)
    inv = inv.ravel()

    def c1(x):  # shortest decimal string that reads back as the same float32
        return np.format_float_positional(np.float32(x), unique=True, trim="0")

    def c3(a, b, c):
        return " ".join(c1(x) for x in (a, b, c))

    with open(mtl_path, "w") as f:
        for k, row in enumerate(uniq):
            t = int(row[0])
            f.write(f"newmtl m{k}\n")
            if row[9] + row[10] + row[11] > 0:  # columns: type, albedo 1-4, spec 5-8, Le 9-11, light mesh 12
                f.write(f"Kd 0 0 0\nKs 0 0 0\nKe {c3(*row[9:12])}\nillum 0\n")
            elif t == MAT_GLASS:
                f.write(f"Kd 0 0 0\nKs {c3(*row[1:4])}\nNi {c1(row[4])}\nillum 7\n")
            elif t == MAT_PHONG:
                f.write(f"Kd {c3(*row[1:4])}\nKs {c3(*row[5:8])}\nNs {c1(row[8])}\nillum 2\n")
            elif t == MAT_LAMBERT:
                f.write(f"Kd {c3(*row[1:4])}\nKs 0 0 0\nillum 1\n")
            else:
                raise ValueError(f"material type {t} has no MTL form")
    v = np.asarray(scene["verts"], np.float32)
    idx = np.asarray(scene["idx"], np.int64)
    # one normal per triangle (its face normal) so that glass, which shades
    # with vertex normals, reads back flat
    q = v[idx].astype(np.float64)
    fn = np.cross(q[:, 1] - q[:, 0], q[:, 2] - q[:, 0])
    fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-30)
    with open(obj_path, "w") as f:
        f.write(f"mtllib {os.path.basename(mtl_path)}\n")
        for p in v:
            f.write(f"v {c3(*p)}\n")
        for n in fn:
            f.write(f"vn {c3(*n)}\n")
        for k in range(uniq.shape[0]):
            f.write(f"g {name}_{k}\nusemtl m{k}\n")
            for t in np.nonzero(inv == k)[0]:
                a, b, c = idx[t] + 1
                f.write(f"f {a}//{t + 1} {b}//{t + 1} {c}//{t + 1}\n")
    return mtl_path

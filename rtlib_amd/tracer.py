"""Python mirror of RTLib's tracer/integrator interface over the rtpg C ABI.

Class and method names follow the reference's C++ API so code written
against it reads the same:
  RTSTreeWrapper / RTSTreeController        Utils/OPX7UtilsPathGuiding.h:1333-1920
  RTMortonQuadTreeWrapper / ...Controller   Utils/OPX7UtilsMorton.h:590-993
  DoubleBufferedHashGrid3Buffer             Test/Include/RTLibExtOPX7Test.h:1472-1570
  Context.render (TracerData::Launch)       Test/Include/RTLibExtOPX7Test.h:1345-1348
All work runs through librtpg.so on the GPU; there is no host fallback.
"""
from __future__ import annotations

import ctypes as C
import time
import weakref

import numpy as np

from . import _lib
from ._lib import call, ptr
from .scenes import camera_uvw

FLAG_NEE = 1
FLAG_RIS = 2
FLAG_USE_GRID = 4
FLAG_USE_TREE = 8
FLAG_GUIDE_SAMPLE = 16
FLAG_FINAL = 32
FLAG_LOCATE = 64
FLAG_MIS = 128
FLAG_RR = 256
FLAG_USE_MLP = 512
FLAG_MOMENTS = 1024  # per-pixel sample moments (Context.relvar)
TEX_NONE = 0xFFFFFFFF
TEX_CLAMP = 0
TEX_WRAP = 1

u32 = C.c_uint32
u16 = C.c_uint16


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


class Context:
    """Device, stream, scene, camera and frame (OPX7Context + TracerData)."""

    def __init__(self, device: int = 0):
        _lib.load()
        h = C.c_void_p()
        call("rtpg_context_create", device, C.byref(h))
        self.h = h
        self.width = self.height = 0
        # guides created on this context: the C ABI requires them destroyed first
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)

    def close(self):
        if self.h:
            for c in list(self._children):
                c._release()
            call("rtpg_context_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        call("rtpg_synchronize", self.h)

    # -- scene
    def upload_scene(self, scene: dict, light_mode: int = 0, light_mesh=None):
        v = _f32(scene["verts"])
        i = _u32(scene["idx"])
        alb = _f32(scene["albedo"])
        mat = _u32(scene["mat"])
        em = _f32(scene["emission"])
        sp = None if scene.get("spec") is None else _f32(scene["spec"])
        lm = None if light_mesh is None else _u32(light_mesh)
        call("rtpg_scene_upload", self.h, ptr(v), v.shape[0], ptr(i, u32), i.shape[0], ptr(alb), ptr(mat, u32),
             ptr(em), ptr(sp), light_mode, ptr(lm, u32))
        self.n_tris = i.shape[0]
        if scene.get("tex_id") is not None or scene.get("vnormal") is not None:
            self.set_attributes(scene.get("texcoord"), scene.get("tex_id"), scene.get("vnormal"),
                                scene.get("texels"), scene.get("tex"))
        cam = scene.get("camera")
        if cam is not None:
            self.set_camera_look(cam["eye"], cam["lookat"], cam["up"], cam["fovy"], cam["width"], cam["height"])

    def set_attributes(self, texcoord=None, tex_id=None, vnormal=None, texels=None, tex=None):
        """Texture coordinates (nt x 3 x 2), texture indices per triangle
        (nt x 4: diffuse, specular, emission on hit, emission when
        light-sampled; TEX_NONE = white), vertex normals (nt x 3 x 3), the
        RGBA8 texels (uint32, R in the low byte, rows bottom-up) and the
        textures (n x 4: first texel, width, height, TEX_CLAMP / TEX_WRAP)."""
        tc = None if texcoord is None else _f32(texcoord).reshape(-1)
        ti = None if tex_id is None else _u32(tex_id).reshape(-1)
        vn = None if vnormal is None else _f32(vnormal).reshape(-1)
        tx = None if texels is None else _u32(texels).reshape(-1)
        td = None if tex is None else _u32(tex).reshape(-1, 4)
        call("rtpg_scene_attributes", self.h, ptr(tc), ptr(ti, u32), ptr(vn), ptr(tx, u32),
             0 if tx is None else tx.size, ptr(td, u32), 0 if td is None else td.shape[0])

    def set_albedo(self, albedo_rgb):
        a = _f32(albedo_rgb)
        call("rtpg_scene_set_albedo", self.h, ptr(a))

    def scene_info(self):
        info = np.zeros(6, np.uint32)
        box = np.zeros(6, np.float32)
        call("rtpg_scene_info", self.h, ptr(info, u32), ptr(box))
        keys = ["nodes", "bvh_depth", "stack_bound", "max_leaf", "lights", "tris"]
        d = {k: int(x) for k, x in zip(keys, info)}
        d["aabb_min"] = box[:3].copy()
        d["aabb_max"] = box[3:].copy()
        return d

    # -- camera / frame
    def set_camera(self, eye, u, v, w):
        e, uu, vv, ww = (_f32(x, (3,)) for x in (eye, u, v, w))
        call("rtpg_camera_set", self.h, ptr(e), ptr(uu), ptr(vv), ptr(ww))

    def set_camera_look(self, eye, lookat, up, fovy, width, height):
        e, u, v, w = camera_uvw(eye, lookat, up, fovy, width / height)
        self.set_camera(e, u, v, w)
        if (width, height) != (self.width, self.height):
            self.resize(width, height)

    def resize(self, width, height):
        call("rtpg_frame_resize", self.h, width, height)
        self.width, self.height = width, height

    def clear(self):
        call("rtpg_accum_clear", self.h)

    def accum(self):
        """(sum image H x W x 3 with row 0 at the bottom, samples accumulated)."""
        out = np.zeros((self.height, self.width, 3), np.float32)
        spp = C.c_uint32()
        call("rtpg_accum_download", self.h, ptr(out), C.byref(spp))
        return out, spp.value

    def image(self):
        a, spp = self.accum()
        return a / max(spp, 1)

    def frame(self):
        """Tonemapped RGBA8 display frame (H x W x 4, row 0 at the bottom)."""
        out = np.zeros((self.height, self.width, 4), np.uint8)
        call("rtpg_frame_tonemap", self.h, ptr(out, C.c_uint8))
        return out

    # -- render
    def render(self, spp=1, flags=FLAG_NEE | FLAG_MIS | FLAG_RR, max_depth=8, rr_depth=3, num_candidates=32,
               seed=0, sample_offset=0, bg=(0.0, 0.0, 0.0), fraction=0.3, sdtree=None, hashgrid=None,
               mqtree=None, max_paths=0, mlp=None):
        p = _lib.RenderParams()
        p.flags = flags
        p.spp = spp
        p.max_depth = max_depth
        p.rr_depth = rr_depth
        p.num_candidates = num_candidates
        p.sample_offset = sample_offset
        p.seed = seed
        p.bg[0], p.bg[1], p.bg[2] = bg
        p.fraction = fraction
        p.sdtree = sdtree.h if sdtree is not None else None
        p.hashgrid = hashgrid.h if hashgrid is not None else None
        p.mqtree = mqtree.h if mqtree is not None else None
        p.max_paths = max_paths
        p.mlp = mlp.h if mlp is not None else None
        call("rtpg_render", self.h, C.byref(p))

    # DebugFrameType (SimpleKernel.h:160-171), plus 8 = world position
    DEBUG_NORMAL, DEBUG_DEPTH, DEBUG_DIFFUSE, DEBUG_SPECULAR, DEBUG_EMISSION = 0, 1, 2, 3, 4
    DEBUG_SHINNESS, DEBUG_REFR_INDEX, DEBUG_GRID_VALUE, DEBUG_POSITION = 5, 6, 7, 8

    def render_debug(self, frame_type, seed=0, sample_offset=0, hashgrid=None, mqtree=None):
        """Debug frame: (value H x W x 3 float32, primitive H x W uint32, rgba
        H x W x 4 uint8).  The grid value needs hashgrid and mqtree."""
        p = _lib.RenderParams()
        p.seed = seed
        p.sample_offset = sample_offset
        if hashgrid is not None and mqtree is not None:
            p.flags = FLAG_USE_GRID
            p.hashgrid, p.mqtree = hashgrid.h, mqtree.h
        val = np.zeros((self.height, self.width, 4), np.float32)
        rgba = np.zeros((self.height, self.width, 4), np.uint8)
        call("rtpg_render_debug", self.h, C.byref(p), frame_type, ptr(val), ptr(rgba, C.c_uint8))
        return val[..., :3].copy(), val[..., 3].view(np.uint32).copy(), rgba

    def relvar(self):
        """Per-sample relative variance of the last FLAG_MOMENTS render (>= 2
        spp): mean over pixels of s^2 / (m^2 + 0.01), s^2 the variance of that
        render's samples of y = (r + g + b) / 3, m the pixel's y in the
        accumulated image."""
        v = C.c_double()
        call("rtpg_render_relvar", self.h, C.byref(v))
        return float(v.value)

    def moments(self):
        """(sum, sum of squares) of y = (r + g + b) / 3 per pixel (H x W x 2)
        and the samples per pixel of the last FLAG_MOMENTS render."""
        m = np.zeros((self.height, self.width, 2), np.float32)
        n = C.c_uint32()
        call("rtpg_render_moments", self.h, ptr(m), C.byref(n))
        return m, int(n.value)

    def render_stats(self):
        s = np.zeros(3, np.uint64)
        call("rtpg_render_stats", self.h, ptr(s, C.c_uint64))
        return {"paths": int(s[0]), "rays": int(s[1]), "shadow_rays": int(s[2])}

    def set_emission(self, le_rgb):
        le = _f32(le_rgb)
        call("rtpg_scene_set_emission", self.h, ptr(le))

    def lights(self):
        n = self.scene_info()["lights"]
        out = np.zeros(max(n, 1), np.uint32)
        call("rtpg_scene_lights", self.h, ptr(out, u32))
        return out[:n]

    def adjoint(self, dl_dimage, spp=1, flags=FLAG_NEE | FLAG_MIS, max_depth=6, seed=0, sample_offset=0,
                bg=(0.0, 0.0, 0.0), max_paths=0):
        """Gradient of sum(dl_dimage * image) for the image render(spp, flags,
        seed, ...) would produce: (albedo grad nt x 3, light radiance grad nl x 3)."""
        p = _lib.RenderParams()
        p.flags = flags
        p.spp = spp
        p.max_depth = max_depth
        p.rr_depth = 1000
        p.num_candidates = 1
        p.sample_offset = sample_offset
        p.seed = seed
        p.bg[0], p.bg[1], p.bg[2] = bg
        p.max_paths = max_paths
        g = _f32(dl_dimage, (self.height, self.width, 3))
        nl = self.scene_info()["lights"]
        ga = np.zeros((self.n_tris, 3), np.float32)
        gl = np.zeros((max(nl, 1), 3), np.float32)
        call("rtpg_adjoint", self.h, C.byref(p), ptr(g), ptr(ga), ptr(gl))
        return ga, gl[:nl]

    def set_profile(self, time_kernels=False, count_work=False):
        call("rtpg_render_profile", self.h, (1 if time_kernels else 0) | (2 if count_work else 0))

    def kernel_stats(self):
        ms = np.zeros(5, np.float64)
        work = np.zeros(4, np.uint64)
        launches = np.zeros(5, np.uint64)
        call("rtpg_render_kernel_stats", self.h, ptr(ms, C.c_double), ptr(work, C.c_uint64),
             ptr(launches, C.c_uint64))
        names = ["extend", "shade", "shadow", "accumulate", "commit"]
        return ({n: float(x) for n, x in zip(names, ms)},
                {"extend_nodes": int(work[0]), "extend_tris": int(work[1]), "shadow_nodes": int(work[2]),
                 "shadow_tris": int(work[3])},
                {n: int(x) for n, x in zip(names, launches)})

    # -- ray queries
    def trace_closest(self, rays):
        r = _f32(rays, (-1, 8))
        hits = np.zeros((r.shape[0], 4), np.float32)
        call("rtpg_trace_closest", self.h, ptr(r), r.shape[0], ptr(hits))
        prim = hits[:, 3].view(np.uint32).copy()
        return hits[:, 0].copy(), hits[:, 1].copy(), hits[:, 2].copy(), prim

    def trace_occluded(self, rays):
        r = _f32(rays, (-1, 8))
        occ = np.zeros(r.shape[0], np.uint32)
        call("rtpg_trace_occluded", self.h, ptr(r), r.shape[0], ptr(occ, u32))
        return occ.astype(bool)


class RTSTreeWrapper:
    """RTSTreeWrapperT: spatial binary tree of directional quadtrees on the GPU."""

    TraceStateRecord, TraceStateRecordAndSample, TraceStateSample = 0, 1, 2

    def __init__(self, ctx: Context, aabb_min, aabb_max, max_dtree_depth=20, fraction=0.3):
        h = C.c_void_p()
        lo, hi = _f32(aabb_min, (3,)), _f32(aabb_max, (3,))
        call("rtpg_sdtree_create", ctx.h, ptr(lo), ptr(hi), max_dtree_depth, fraction, C.byref(h))
        self.h = h
        self.ctx = ctx
        self.fraction = fraction
        ctx._adopt(self)

    def Destroy(self):
        if self.h:
            call("rtpg_sdtree_destroy", self.h)
            self.h = None

    _release = Destroy

    def __del__(self):
        try:
            self.Destroy()
        except Exception:
            pass

    def Clear(self):
        call("rtpg_sdtree_clear", self.h)

    def SetSplit(self, mode=0, c=4000.0, max_mb=2000, subdiv_th=0.01):
        call("rtpg_sdtree_set_split", self.h, mode, c, max_mb, subdiv_th)

    def SetOctree(self, enable=True):
        """Refine by 8-way splits (the reference's STree2 variant)."""
        call("rtpg_sdtree_set_octree", self.h, 1 if enable else 0)

    def SetFilter(self, spatial=0, directional=0, seed=0):
        """Recording filters, the reference's SpatialFilter / DirectionalFilter
        (OPX7UtilsPathGuiding.h:31-43): 0 Nearest, 1 Box (stochastic form)."""
        call("rtpg_sdtree_set_filter", self.h, spatial, directional, seed)

    def Reset(self, iteration, samples_per_pass):
        call("rtpg_sdtree_reset", self.h, iteration, samples_per_pass)

    def Build(self):
        call("rtpg_sdtree_build", self.h)

    def Record(self, pos, canon, irradiance, stat_weight):
        p, c, i, w = _f32(pos, (-1, 3)), _f32(canon, (-1, 2)), _f32(irradiance), _f32(stat_weight)
        call("rtpg_sdtree_record", self.h, ptr(p), ptr(c), ptr(i), ptr(w), p.shape[0])

    def Lookup(self, pos):
        p = _f32(pos, (-1, 3))
        out = np.zeros(p.shape[0], np.uint32)
        call("rtpg_sdtree_lookup", self.h, ptr(p), p.shape[0], ptr(out, u32))
        return out

    def Sample(self, dtree_idx, seed):
        d = _u32(dtree_idx)
        out = np.zeros((d.shape[0], 2), np.float32)
        call("rtpg_sdtree_sample", self.h, ptr(d, u32), d.shape[0], seed, ptr(out))
        return out

    def Pdf(self, dtree_idx, canon):
        d, c = _u32(dtree_idx), _f32(canon, (-1, 2))
        out = np.zeros(d.shape[0], np.float32)
        call("rtpg_sdtree_pdf", self.h, ptr(d, u32), ptr(c), d.shape[0], ptr(out))
        return out

    def Sizes(self):
        s = np.zeros(4, np.uint32)
        call("rtpg_sdtree_sizes", self.h, ptr(s, u32))
        return {"stree_nodes": int(s[0]), "dtrees": int(s[1]), "building_nodes": int(s[2]),
                "sampling_nodes": int(s[3])}

    def Download(self):
        """Flattened state (reference Download/Upload layout, see rtpg_sdtree_export)."""
        sz = self.Sizes()
        n, bn, sn = sz["dtrees"], sz["building_nodes"], sz["sampling_nodes"]
        d = {
            "stree": np.zeros(sz["stree_nodes"], np.uint32),
            "b_cnt": np.zeros(n, np.uint32),
            "s_cnt": np.zeros(n, np.uint32),
            "scal": np.zeros((6, n), np.float32),
            "b_sums": np.zeros((bn, 4), np.float32),
            "b_child": np.zeros((bn, 4), np.uint16),
            "s_sums": np.zeros((sn, 4), np.float32),
            "s_child": np.zeros((sn, 4), np.uint16),
        }
        call("rtpg_sdtree_export", self.h, ptr(d["stree"], u32), ptr(d["b_cnt"], u32), ptr(d["s_cnt"], u32),
             ptr(d["scal"]), ptr(d["b_sums"]), ptr(d["b_child"], u16), ptr(d["s_sums"]), ptr(d["s_child"], u16))
        return d

    def Upload(self, d):
        st, bc, sc = _u32(d["stree"]), _u32(d["b_cnt"]), _u32(d["s_cnt"])
        scal = _f32(d["scal"])
        bs, ss = _f32(d["b_sums"]), _f32(d["s_sums"])
        bch = np.ascontiguousarray(d["b_child"], np.uint16)
        sch = np.ascontiguousarray(d["s_child"], np.uint16)
        call("rtpg_sdtree_import", self.h, st.shape[0], ptr(st, u32), bc.shape[0], ptr(bc, u32), ptr(sc, u32),
             ptr(scal), ptr(bs), ptr(bch, u16), ptr(ss), ptr(sch, u16))



class RTSTreeWrapper2(RTSTreeWrapper):
    """RTSTreeWrapper2T (OPX7UtilsPathGuiding.h:2156-2432): the octree spatial
    tree, with its Reset's fixed split rule (sqrt(2^iter spp / 4) * 4000,
    4000 MB guard, DTree threshold 0.01) and its fraction 0.5."""

    def __init__(self, ctx: Context, aabb_min, aabb_max, max_dtree_depth=20):
        super().__init__(ctx, aabb_min, aabb_max, max_dtree_depth, 0.5)
        self.SetSplit(mode=0, c=4000.0, max_mb=4000, subdiv_th=0.01)
        self.SetOctree(True)


class _Ctrl:
    prefix = ""

    def Start(self):
        call(self.prefix + "start", self.h)

    def BegTrace(self):
        call(self.prefix + "beg_trace", self.h)

    def EndTrace(self, samples=0):
        call(self.prefix + "end_trace", self.h, samples)

    def _q(self):
        q = np.zeros(6, np.uint32)
        call(self.prefix + "query", self.h, ptr(q, u32))
        return q

    def GetState(self):
        return int(self._q()[0])

    def GetIteration(self):
        return int(self._q()[1])

    def GetSamplePerTmp(self):
        return int(self._q()[2])

    def GetSampleForPass(self):
        return int(self._q()[3])

    def IsExecuting(self):
        return bool(self._q()[4])

    def __del__(self):
        try:
            if self.h:
                call(self.prefix + "destroy", self.h)
                self.h = None
        except Exception:
            pass


class RTSTreeController(_Ctrl):
    """RTSTreeControllerT: SD-tree pass schedule and state machine."""

    prefix = "rtpg_sdtree_ctrl_"

    def __init__(self, tree: RTSTreeWrapper, sample_for_budget, iteration_for_built=0, ratio_for_budget=0.5,
                 sample_per_launch=1):
        h = C.c_void_p()
        call("rtpg_sdtree_ctrl_create", tree.h, sample_for_budget, iteration_for_built, ratio_for_budget,
             sample_per_launch, C.byref(h))
        self.h = h
        self.tree = tree


class DoubleBufferedHashGrid3Buffer:
    """Spatial hash with current/previous checksum tables."""

    def __init__(self, ctx: Context, aabb_min, aabb_max, bounds=(128, 64, 128), slots=65536, key_mode=0):
        h = C.c_void_p()
        lo, hi, b = _f32(aabb_min, (3,)), _f32(aabb_max, (3,)), _u32(bounds)
        call("rtpg_hashgrid_create", ctx.h, ptr(lo), ptr(hi), ptr(b, u32), slots, key_mode, C.byref(h))
        self.h = h
        self.slots = slots
        self.ctx = ctx
        ctx._adopt(self)

    def _release(self):
        if self.h:
            call("rtpg_hashgrid_destroy", self.h)
            self.h = None

    def Clear(self):
        call("rtpg_hashgrid_clear", self.h)

    def Update(self):
        call("rtpg_hashgrid_update", self.h)

    def Find(self, pos, insert=False):
        p = _f32(pos, (-1, 3))
        out = np.zeros(p.shape[0], np.uint32)
        call("rtpg_hashgrid_find", self.h, ptr(p), p.shape[0], 1 if insert else 0, ptr(out, u32))
        return out

    def Download(self, which=0):
        out = np.zeros(self.slots, np.uint32)
        call("rtpg_hashgrid_download", self.h, which, ptr(out, u32))
        return out

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


class RTMortonQuadTreeWrapper:
    """Per-slot Morton-ordered directional quadtree pyramids (building/sampling)."""

    def __init__(self, ctx: Context, slots, level, fraction=0.3):
        h = C.c_void_p()
        call("rtpg_mqtree_create", ctx.h, slots, level, fraction, C.byref(h))
        self.h = h
        self.slots, self.level, self.fraction = slots, level, fraction
        self.per_cell = ((1 << (2 * (level + 1))) - 1) // 3
        self.ctx = ctx
        ctx._adopt(self)

    def _release(self):
        if self.h:
            call("rtpg_mqtree_destroy", self.h)
            self.h = None

    def Allocate(self):
        call("rtpg_mqtree_allocate", self.h)

    def Update(self):
        call("rtpg_mqtree_update", self.h)

    def LaunchBuildKernel(self):
        call("rtpg_mqtree_build", self.h)

    def Record(self, slot, canon, value):
        s, c, v = _u32(slot), _f32(canon, (-1, 2)), _f32(value)
        call("rtpg_mqtree_record", self.h, ptr(s, u32), ptr(c), ptr(v), s.shape[0])

    def SampleAndPdf(self, slot, seed):
        s = _u32(slot)
        out = np.zeros((s.shape[0], 3), np.float32)
        call("rtpg_mqtree_sample", self.h, ptr(s, u32), s.shape[0], seed, ptr(out))
        return out

    def Pdf(self, slot, canon):
        s, c = _u32(slot), _f32(canon, (-1, 2))
        out = np.zeros(s.shape[0], np.float32)
        call("rtpg_mqtree_pdf", self.h, ptr(s, u32), ptr(c), s.shape[0], ptr(out))
        return out

    def Download(self, which=0):
        out = np.zeros((self.slots, self.per_cell), np.float32)
        call("rtpg_mqtree_download", self.h, which, ptr(out))
        return out

    def Upload(self, weights, which=0):
        w = _f32(weights, (self.slots, self.per_cell))
        call("rtpg_mqtree_upload", self.h, which, ptr(w))

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


class RTMortonQuadTreeController(_Ctrl):
    """RTMortonQuadTreeControllerT: hash-grid guide pass schedule (with Locate)."""

    prefix = "rtpg_mqtree_ctrl_"
    TraceStateRecord, TraceStateRecordAndSample, TraceStateSample, TraceStateLocate = 0, 1, 2, 3

    def __init__(self, tree: RTMortonQuadTreeWrapper, grid: DoubleBufferedHashGrid3Buffer, sample_for_budget,
                 iteration_for_built=3, iteration_for_locate=0, ratio_for_budget=0.5, sample_per_launch=1):
        h = C.c_void_p()
        call("rtpg_mqtree_ctrl_create", tree.h, grid.h if grid is not None else None, sample_for_budget,
             iteration_for_built, iteration_for_locate, ratio_for_budget, sample_per_launch, C.byref(h))
        self.h = h
        self.tree, self.grid = tree, grid


def guide_flags(state: int, base: int, kind: str) -> int:
    """Render flags for a controller state (reference TracePipeline,
    RTLibExtOPX7TestApplication.cpp:2155-2231: Build/Trace/Final/Locate)."""
    f = base | (FLAG_USE_TREE if kind == "tree" else FLAG_USE_GRID)
    if state == 1:
        f |= FLAG_GUIDE_SAMPLE
    elif state == 2:
        f |= FLAG_GUIDE_SAMPLE | FLAG_FINAL
    elif state == 3:
        f |= FLAG_LOCATE
    return f


def render_guided(ctx: Context, ctrl, guide, kind="tree", base_flags=FLAG_NEE | FLAG_MIS | FLAG_RR, max_depth=10,
                  rr_depth=3, seed=0, fraction=0.3, grid=None, max_paths=0, on_pass=None):
    """Runs the controller's whole budget, one render call per pass remainder
    (the reference launches samplesPerLaunch at a time; a pass is rendered here
    in one call).  Returns the number of samples traced."""
    ctrl.Start()
    traced = 0
    while True:
        ctrl.BegTrace()
        if not ctrl.IsExecuting():
            break
        q = ctrl._q()
        state, per_tmp, pass_spp = int(q[0]), int(q[2]), int(q[3])
        if pass_spp == 0:
            # The reference controller stays executing for one extra launch
            # after the budget is spent (EndTrace tests per_all > budget);
            # that launch is skipped here.
            break
        spp = pass_spp - per_tmp
        flags = guide_flags(state, base_flags, kind)
        if kind == "tree":
            ctx.render(spp=spp, flags=flags, max_depth=max_depth, rr_depth=rr_depth, seed=seed,
                       sample_offset=traced, fraction=fraction, sdtree=guide, max_paths=max_paths)
        else:
            ctx.render(spp=spp, flags=flags, max_depth=max_depth, rr_depth=rr_depth, seed=seed,
                       sample_offset=traced, fraction=fraction, hashgrid=grid, mqtree=guide, max_paths=max_paths)
        traced += spp
        ctrl.EndTrace(spp)
        if on_pass is not None:
            on_pass(state, spp, traced)
    return traced


# ---------------------------------------------------------------- learned guide
MLP_WEIGHTS = 14336
MLP_LEVELS = 16


class MLPGuide:
    """Learned directional guide (rtpg_mlp_*): hash-grid encoding + 3 x 64 ReLU
    MLP on the matrix cores -> 8-lobe vMF mixture and the probability c of
    sampling it against the BSDF.  No counterpart in the reference; it sits
    beside RTSTreeWrapper and is passed to Context.render(mlp=...) with
    FLAG_USE_MLP.

    Training samples are dicts of arrays: pos, wo, nrm (n x 3), rough (n),
    omega (n x 3, sampled direction), weight (n, f Li |cos| / p_sample),
    pdf_bsdf (n)."""

    def __init__(self, ctx: Context, aabb_min, aabb_max, log2_table=19, base_res=16.0, max_res=256.0, lr=1e-2,
                 beta1=0.9, beta2=0.99, eps=1e-15, batch=1 << 18, capacity=1 << 20, record_paths=1 << 19, seed=0,
                 guide_depth=0, cache_res=0):
        cfg = _lib.MlpConfig(log2_table, base_res, max_res, lr, beta1, beta2, eps, batch, capacity, record_paths,
                             seed, guide_depth, cache_res)
        h = C.c_void_p()
        call("rtpg_mlp_create", ctx.h, ptr(_f32(aabb_min)), ptr(_f32(aabb_max)), C.byref(cfg), C.byref(h))
        self.h = h.value
        self.ctx = ctx
        self.log2_table = log2_table
        ctx._adopt(self)

    def close(self):
        if getattr(self, "h", None):
            call("rtpg_mlp_destroy", self.h)
            self.h = None

    _release = close

    def cache(self):
        """Spatial cache (cache_res > 0): (mix entries x 48, c entries)."""
        n = C.c_uint32()
        call("rtpg_mlp_cache_download", self.h, None, None, C.byref(n))
        mix = np.zeros((n.value, 48), np.float32)
        c = np.zeros(n.value, np.float32)
        call("rtpg_mlp_cache_download", self.h, ptr(mix), ptr(c), C.byref(n))
        return mix, c

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def reset(self, seed=0):
        call("rtpg_mlp_reset", self.h, seed)

    def info(self):
        i = np.zeros(7, np.uint64)
        res = np.zeros(16, np.float32)
        call("rtpg_mlp_info", self.h, ptr(i, C.c_uint64), ptr(res))
        return {"weights": int(i[0]), "table_floats": int(i[1]), "recorded": int(i[2]), "steps": int(i[3]),
                "levels": int(i[4]), "entries": int(i[5]), "dense_mask": int(i[6]), "res": res}

    @staticmethod
    def _q(pos, wo, nrm, rough):
        pos, wo, nrm = _f32(pos, (-1, 3)), _f32(wo, (-1, 3)), _f32(nrm, (-1, 3))
        rough = _f32(rough, (-1,))
        n = pos.shape[0]
        assert wo.shape[0] == n and nrm.shape[0] == n and rough.shape[0] == n
        return pos, wo, nrm, rough, n

    def forward(self, pos, wo, nrm, rough):
        """(raw outputs n x 41, mixture n x 41: 8 x (pi, mu xyz, kappa), c)."""
        pos, wo, nrm, rough, n = self._q(pos, wo, nrm, rough)
        raw = np.zeros((n, 41), np.float32)
        mix = np.zeros((n, 41), np.float32)
        call("rtpg_mlp_forward", self.h, ptr(pos), ptr(wo), ptr(nrm), ptr(rough), n, ptr(raw), ptr(mix))
        return raw, mix

    def sample(self, mixture, seed=0):
        """One direction per row of a forward() mixture: (n x 4) = w xyz, q(w)."""
        mix = _f32(mixture, (-1, 41))
        out = np.zeros((mix.shape[0], 4), np.float32)
        call("rtpg_mlp_sample", self.h, ptr(mix), mix.shape[0], seed, ptr(out))
        return out

    @staticmethod
    def _s(s):
        pos, wo, nrm, rough, n = MLPGuide._q(s["pos"], s["wo"], s["nrm"], s["rough"])
        om = _f32(s["omega"], (-1, 3))
        w = _f32(s["weight"], (-1,))
        pb = _f32(s["pdf_bsdf"], (-1,))
        assert om.shape[0] == n and w.shape[0] == n and pb.shape[0] == n
        return pos, wo, nrm, rough, om, w, pb, n

    def gradient(self, samples, table=True):
        """(loss, d loss / d weights [14336], d loss / d table or None) of the mean loss."""
        pos, wo, nrm, rough, om, w, pb, n = self._s(samples)
        loss = np.zeros(1, np.float32)
        gw = np.zeros(MLP_WEIGHTS, np.float32)
        gt = np.zeros(MLP_LEVELS * (1 << self.log2_table) * 2, np.float32) if table else None
        call("rtpg_mlp_gradient", self.h, ptr(pos), ptr(wo), ptr(nrm), ptr(rough), ptr(om), ptr(w), ptr(pb), n,
             ptr(loss), ptr(gw), ptr(gt) if table else None)
        return float(loss[0]), gw, gt

    def train(self, samples, steps=1):
        pos, wo, nrm, rough, om, w, pb, n = self._s(samples)
        loss = np.zeros(max(steps, 1), np.float32)
        call("rtpg_mlp_train", self.h, ptr(pos), ptr(wo), ptr(nrm), ptr(rough), ptr(om), ptr(w), ptr(pb), n, steps,
             ptr(loss))
        return loss[:steps]

    def train_recorded(self, steps=1):
        loss = np.zeros(max(steps, 1), np.float32)
        used = np.zeros(1, np.uint32)
        call("rtpg_mlp_train_recorded", self.h, steps, ptr(loss), ptr(used, C.c_uint32))
        return loss[:steps], int(used[0])

    def export(self):
        w = np.zeros(MLP_WEIGHTS, np.float32)
        t = np.zeros(MLP_LEVELS * (1 << self.log2_table) * 2, np.float32)
        call("rtpg_mlp_export", self.h, ptr(w), ptr(t))
        return w, t

    def import_(self, weights=None, table=None):
        w = _f32(weights) if weights is not None else None
        t = _f32(table) if table is not None else None
        call("rtpg_mlp_import", self.h, ptr(w) if w is not None else None, ptr(t) if t is not None else None)


def render_mlp_guided(ctx: Context, mlp: MLPGuide, budget, base_flags=FLAG_NEE | FLAG_MIS | FLAG_RR, max_depth=8,
                      rr_depth=3, seed=0, train_passes=8, pass_spp=1, steps_per_pass=4, max_paths=0, on_pass=None,
                      fallback=None, fraction=0.3, judge=None):
    """Online schedule: train_passes passes of pass_spp samples record training
    samples (guided by the network once it has been trained at least once),
    each followed by steps_per_pass Adam steps; the rest of the budget is one
    final guided pass.  Every pass accumulates into the image.  Returns the
    number of samples traced.

    fallback: an RTSTreeWrapper that the training passes record into as well
    (reset with the reference's refinement rule before and built after every
    pass), so that an SD-tree trained on the same paths is at hand.

    judge: None, or a dict that enables the fallback decision and receives
    its report.  Options: probe_spp (4), max_train_share (0.25), margin
    (0.25).  Plain path tracing renders a first probe pass, which times it;
    training stops (and plain path tracing renders the rest) when the
    training passes would take more than max_train_share of the budget's time
    at the plain rate: before the first pass by a lower bound of their time,
    after it by the measured one.  After training, probe passes of plain path
    tracing, the learned guide and the fallback SD-tree measure their
    per-sample relative variance (Context.relvar) and cost; the rest of the
    budget goes to the estimator with the lowest variance x time when that is
    below the learned guide's by more than the margin, else to the learned
    guide.  Probe passes count towards the budget and the image; every pass
    is an unbiased estimate, so the mixture is too."""
    traced = 0
    trained = False
    probe = int(judge.get("probe_spp", 4)) if judge is not None else 0
    cands = {}
    if fallback is not None:
        fallback.Clear()

    def probe_render(flags, **kw):
        """A FLAG_MOMENTS pass of probe spp: (per-sample relative variance
        against the image accumulated so far, cost per spp in seconds)."""
        ctx.synchronize()
        t = time.perf_counter()
        ctx.render(spp=probe, flags=flags | FLAG_MOMENTS, max_depth=max_depth, rr_depth=rr_depth, seed=seed,
                   sample_offset=traced, max_paths=max_paths, **kw)
        ctx.synchronize()
        cost = (time.perf_counter() - t) / probe
        return ctx.relvar(), cost

    if judge is not None and budget >= probe + 1:
        cands["plain"] = probe_render(base_flags)
        traced += probe
    stop_training = False
    if "plain" in cands:
        # lower bound of the training passes' time: the recording passes at
        # the plain rate plus, per Adam step, one pass over the fp32 table and
        # weights (parameter, two moments and gradient read and written, the
        # bf16 copy written) at the 8 TB/s HBM peak
        info = mlp.info()
        adam_bytes = 34.0 * (info["table_floats"] + info["weights"])
        t_min = train_passes * (pass_spp * cands["plain"][1] + steps_per_pass * adam_bytes / 8e12)
        share = float(judge.get("max_train_share", 0.25))
        stop_training = t_min > share * cands["plain"][1] * budget
    for k in range(train_passes):
        if stop_training or traced + pass_spp + probe > budget:
            break
        t0 = time.perf_counter()
        if fallback is not None and k > 0:
            fallback.Reset(k, pass_spp)
        flags = base_flags | FLAG_USE_MLP | (FLAG_GUIDE_SAMPLE if trained else 0)
        ctx.render(spp=pass_spp, flags=flags, max_depth=max_depth, rr_depth=rr_depth, seed=seed,
                   sample_offset=traced, mlp=mlp, max_paths=max_paths, sdtree=fallback)
        traced += pass_spp
        loss, used = mlp.train_recorded(steps_per_pass)
        if fallback is not None:
            fallback.Build()
        trained = trained or used > 0
        if on_pass is not None:
            on_pass(k, pass_spp, traced, loss, used)
        if judge is not None and k == 0 and "plain" in cands:
            ctx.synchronize()
            t_pass = time.perf_counter() - t0
            share = float(judge.get("max_train_share", 0.25))
            if t_pass * train_passes > share * cands["plain"][1] * budget:
                stop_training = True
    rest = budget - traced
    choice = "learned"
    if judge is not None:
        # the probes that rank the estimators run back to back, each scored
        # against the image accumulated by then (all of them past the
        # training passes, so the normalisers are close and steadier than a
        # probe's own few samples, whose mean would cap the bright outliers
        # that dominate the relative error)
        n_probes = 3 if fallback is not None else 2
        if not (stop_training or not trained) and rest >= (n_probes + 1) * probe:
            cands["plain"] = probe_render(base_flags)
            traced += probe
            cands["learned"] = probe_render(base_flags | FLAG_USE_MLP | FLAG_FINAL | FLAG_GUIDE_SAMPLE, mlp=mlp)
            traced += probe
            if fallback is not None:
                cands["sdtree"] = probe_render(base_flags | FLAG_USE_TREE | FLAG_FINAL | FLAG_GUIDE_SAMPLE,
                                               sdtree=fallback, fraction=fraction)
                traced += probe
        report = {c: {"relvar": v, "ms_per_spp": 1e3 * t} for c, (v, t) in cands.items()}
        eff = {c: r["relvar"] * r["ms_per_spp"] for c, r in report.items()}
        if stop_training or not trained:
            choice = "plain"
        elif "learned" in eff:
            # the learned guide stays unless a fallback is measured clearly
            # more efficient (a few probe samples per pixel estimate the
            # variance loosely)
            best = min(eff, key=eff.get)
            margin = float(judge.get("margin", 0.25))
            choice = best if eff[best] * (1.0 + margin) < eff["learned"] else "learned"
        judge["candidates"] = report
        judge["choice"] = choice
        judge["training_stopped"] = stop_training
        rest = budget - traced
    if rest > 0:
        if choice == "plain":
            flags, kw = base_flags, {}
        elif choice == "sdtree":
            flags, kw = base_flags | FLAG_USE_TREE | FLAG_FINAL | FLAG_GUIDE_SAMPLE, {"sdtree": fallback,
                                                                                      "fraction": fraction}
        else:
            flags = base_flags | FLAG_USE_MLP | FLAG_FINAL | (FLAG_GUIDE_SAMPLE if trained else 0)
            kw = {"mlp": mlp}
        ctx.render(spp=rest, flags=flags, max_depth=max_depth, rr_depth=rr_depth, seed=seed, sample_offset=traced,
                   max_paths=max_paths, **kw)
        traced += rest
        if on_pass is not None:
            on_pass(-1, rest, traced, None, 0)
    return traced



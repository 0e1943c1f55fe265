"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
u16p = C.POINTER(C.c_uint16)


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


class Params(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("flags", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_depth", C.c_uint32), ("num_candidates", C.c_uint32), ("spp", C.c_uint32),
                ("sample_offset", C.c_uint32), ("seed", C.c_uint64), ("bg", C.c_float * 3),
                ("fraction", C.c_float), ("eye", C.c_float * 3), ("u", C.c_float * 3), ("v", C.c_float * 3),
                ("w", C.c_float * 3), ("px_begin", C.c_uint32), ("px_end", C.c_uint32), ("threads", C.c_uint32)]


class SdTree(C.Structure):
    _fields_ = [("stree", u32p), ("aabb_min", C.c_float * 3), ("size", C.c_float), ("s_base", u32p),
                ("s_sum", f32p), ("s_sw", f32p), ("s_sums", f32p), ("s_child", u16p), ("b_base", u32p),
                ("b_sw", f32p), ("b_sums", f32p), ("b_child", u16p), ("filter", C.c_uint32)]


class Grid(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("size", C.c_float * 3), ("bounds", C.c_uint32 * 3),
                ("slots", C.c_uint32), ("key_mode", C.c_uint32)]


class HashGuideS(C.Structure):
    _fields_ = [("grid", Grid), ("cur", u32p), ("prv", u32p), ("level", C.c_uint32), ("per_cell", C.c_uint32),
                ("s_w", f32p), ("b_w", f32p)]


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.orc_to_uniform_float.restype = C.c_float
        L.orc_to_uniform_float.argtypes = [C.c_uint32]
        L.orc_xorshift32.restype = C.c_uint32
        L.orc_xorshift32.argtypes = [u32p]
        L.orc_pcg1d.restype = C.c_uint32
        L.orc_pcg1d.argtypes = [C.c_uint32]
        L.orc_xxhash32.restype = C.c_uint32
        L.orc_xxhash32.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_morton2.restype = C.c_uint32
        L.orc_morton2.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_morton3.restype = C.c_uint32
        L.orc_morton3.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_dtree_reset.restype = C.c_uint32
        L.orc_dtree_reset.argtypes = [f32p, u16p, C.c_float, C.c_uint32, C.c_float, u16p, u32p]
        L.orc_stree_refine.restype = C.c_uint32
        L.orc_stree_refine.argtypes = [u32p, C.c_uint32, f32p, u32p, u32p, C.c_uint64, C.c_int, u32p, C.c_uint32,
                                       u32p, u32p, C.c_int]
        L.orc_split_threshold.restype = C.c_uint64
        L.orc_split_threshold.argtypes = [C.c_int, C.c_float, C.c_int, C.c_int]
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [f32p, C.c_uint32, u32p, C.c_uint32, f32p, u32p, f32p, f32p]
        L.orc_scene_set_light_pdf.argtypes = [C.c_void_p, f32p]
        L.orc_scene_set_attributes.argtypes = [C.c_void_p, f32p, u32p, f32p, u32p, C.c_uint64, u32p, C.c_uint32]
        L.orc_tex_sample.argtypes = [u32p, u32p, C.c_float, C.c_float, f32p]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_debug_frame.argtypes = [C.c_void_p, C.POINTER(Params), C.c_uint32, f32p, C.POINTER(C.c_int32)]
        L.orc_render.restype = C.c_uint64
        L.orc_render.argtypes = [C.c_void_p, C.POINTER(Params), C.POINTER(SdTree), f32p]
        L.orc_render_grid.restype = C.c_uint64
        L.orc_render_grid.argtypes = [C.c_void_p, C.POINTER(Params), C.POINTER(HashGuideS), f32p]
        L.orc_ctrl_schedule.restype = C.c_uint32
        L.orc_ctrl_schedule.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32,
                                        C.c_uint32, u32p, u32p, u32p, u32p]
        L.orc_adjoint.argtypes = [C.c_void_p, C.POINTER(Params), f32p, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _p(a, t=f32p):
    return None if a is None else a.ctypes.data_as(t)


def _f(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    return a if shape is None else a.reshape(shape)


def _u(a):
    return np.ascontiguousarray(a, np.uint32)


# ---- random numbers
def rng_stream(seed, pixel, sample, n):
    out = np.zeros(n, np.uint32)
    lib().orc_rng_stream(C.c_uint64(seed), pixel, sample, n, _p(out, u32p))
    return out


def pcg32_raw(initstate, seq, n):
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32_raw(C.c_uint64(initstate), C.c_uint64(seq), n, _p(out, u32p))
    return out


def rng_floats(seed, pixel, sample, n):
    out = np.zeros(n, np.float32)
    lib().orc_rng_floats(C.c_uint64(seed), pixel, sample, n, _p(out))
    return out


def to_uniform_float(v):
    return lib().orc_to_uniform_float(v)


def xorshift32(seed, n):
    s = C.c_uint32(seed)
    return np.array([lib().orc_xorshift32(C.byref(s)) for _ in range(n)], np.uint32)


def pcg1d(v):
    return lib().orc_pcg1d(v)


def morton2(x, y):
    return lib().orc_morton2(x, y)


def morton3(x, y, z):
    return lib().orc_morton3(x, y, z)


def dir_to_canonical(d):
    d = _f(d, (-1, 3))
    out = np.zeros((d.shape[0], 2), np.float32)
    for i in range(d.shape[0]):
        lib().orc_dir_to_canonical(_p(d[i]), _p(out[i]))
    return out


def canonical_to_dir(uv):
    uv = _f(uv, (-1, 2))
    out = np.zeros((uv.shape[0], 3), np.float32)
    for i in range(uv.shape[0]):
        lib().orc_canonical_to_dir(_p(uv[i]), _p(out[i]))
    return out


# ---- hash grid / Morton quadtree
def grid(origin, size, bounds, slots, key_mode=0):
    g = Grid()
    for k in range(3):
        g.origin[k], g.size[k], g.bounds[k] = origin[k], size[k], bounds[k]
    g.slots, g.key_mode = slots, key_mode
    return g


def grid_find(g, cur, prv, pos, insert):
    pos = _f(pos, (-1, 3))
    out = np.zeros(pos.shape[0], np.uint32)
    lib().orc_grid_find(C.byref(g), _p(cur, u32p), _p(prv, u32p), _p(pos), pos.shape[0], int(insert),
                        _p(out, u32p))
    return out


def mq_record(w, level, slot, canon, val):
    lib().orc_mq_record(_p(w), w.shape[1], level, _p(_u(slot), u32p), _p(_f(canon)), _p(_f(val)), len(slot))


def mq_build(w, level):
    lib().orc_mq_build(_p(w), level, w.shape[1], w.shape[0])


def mq_sample(w, level, slot, seed):
    slot = _u(slot)
    out = np.zeros((slot.shape[0], 3), np.float32)
    lib().orc_mq_sample(_p(w), w.shape[1], level, _p(slot, u32p), slot.shape[0], C.c_uint64(seed), _p(out))
    return out


def mq_pdf(w, level, slot, canon):
    slot = _u(slot)
    out = np.zeros(slot.shape[0], np.float32)
    lib().orc_mq_pdf(_p(w), w.shape[1], level, _p(slot, u32p), _p(_f(canon)), slot.shape[0], _p(out))
    return out


# ---- SD-tree on the exported layout (see rtpg_sdtree_export)
def _bases(cnt):
    return np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.uint32)


def sdtree_build(st):
    """RTSTreeWrapperT::Build on a state dict; returns a new dict."""
    out = {k: np.array(v, copy=True) for k, v in st.items()}
    n = out["b_cnt"].shape[0]
    bs = np.ascontiguousarray(out["b_sums"], np.float32)
    bc = np.ascontiguousarray(out["b_child"], np.uint16)
    s_sum = np.zeros(n, np.float32)
    s_area = np.zeros(n, np.float32)
    lib().orc_dtree_build(n, _p(_u(out["b_cnt"]), u32p), _p(bs), _p(bc, u16p), _p(s_sum), _p(s_area))
    out["b_sums"] = bs
    scal = out["scal"]
    scal[1] = s_sum
    scal[2] = s_area
    scal[3] = scal[0]
    scal[4] = s_sum
    scal[5] = s_area
    out["s_sums"] = bs.copy()
    out["s_child"] = bc.copy()
    out["s_cnt"] = out["b_cnt"].copy()
    return out


def sdtree_reset(st, iteration, spp, mode=0, c=4000.0, max_mb=2000, th=0.01, max_depth=20, octree=False):
    """RTSTreeWrapperT::Reset(iter, spp) on a state dict (octree: the
    RTSTree2T refinement); returns a new dict."""
    if iteration <= 0:
        return {k: np.array(v, copy=True) for k, v in st.items()}
    L = lib()
    thr = L.orc_split_threshold(mode, c, iteration, spp)
    stree = _u(st["stree"])
    n_dt = st["b_cnt"].shape[0]
    b_sw = _f(st["scal"][0])
    cap = max(4 * stree.shape[0] + 1024, 1 << 16)
    while True:
        out = np.zeros(cap, np.uint32)
        mp = np.zeros(cap, np.uint32)
        nd = C.c_uint32()
        nn = L.orc_stree_refine(_p(stree, u32p), stree.shape[0], _p(b_sw), _p(_u(st["b_cnt"]), u32p),
                                _p(_u(st["s_cnt"]), u32p), C.c_uint64(thr), max_mb, _p(out, u32p), cap,
                                _p(mp, u32p), C.byref(nd), 1 if octree else 0)
        if nn:
            break
        cap *= 4
    out, mp = out[:nn], mp[:nd.value]
    s_base = _bases(st["s_cnt"])
    s_sums = _f(st["s_sums"], (-1, 4))
    s_child = np.ascontiguousarray(st["s_child"], np.uint16).reshape(-1, 4)
    scal = st["scal"]
    b_list, s_list_sums, s_list_child, b_cnt, s_cnt = [], [], [], [], []
    new_scal = np.zeros((6, nd.value), np.float32)
    tmp = np.zeros((65536, 4), np.uint16)
    for j, src in enumerate(mp):
        sb, scn = int(s_base[src]), int(st["s_cnt"][src])
        ss = np.ascontiguousarray(s_sums[sb:sb + scn])
        sc = np.ascontiguousarray(s_child[sb:sb + scn])
        depth = C.c_uint32()
        cnt = L.orc_dtree_reset(_p(ss), _p(sc, u16p), C.c_float(scal[4][src]), max_depth, C.c_float(th),
                                _p(tmp, u16p), C.byref(depth))
        b_list.append(tmp[:cnt].copy())
        b_cnt.append(cnt)
        s_list_sums.append(ss.copy())
        s_list_child.append(sc.copy())
        s_cnt.append(scn)
        new_scal[3, j], new_scal[4, j], new_scal[5, j] = scal[3][src], scal[4][src], scal[5][src]
    bch = np.concatenate(b_list)
    return {
        "stree": out,
        "b_cnt": np.array(b_cnt, np.uint32),
        "s_cnt": np.array(s_cnt, np.uint32),
        "scal": new_scal,
        "b_sums": np.zeros((bch.shape[0], 4), np.float32),
        "b_child": bch,
        "s_sums": np.concatenate(s_list_sums),
        "s_child": np.concatenate(s_list_child),
    }


def stree_lookup(stree, aabb_min, size, pos):
    pos = _f(pos, (-1, 3))
    out = np.zeros(pos.shape[0], np.uint32)
    lo = _f(aabb_min, (3,))
    lib().orc_stree_lookup(_p(_u(stree), u32p), _p(lo), C.c_float(size), _p(pos), pos.shape[0], _p(out, u32p))
    return out


def sdtree_record(st, aabb_min, size, pos, canon, irr, sw, filter=0, seed=0, batch=0):
    """Sequential record into a copy of the state's building trees.  filter:
    bit 0 spatial box, bit 1 directional box, in the stochastic form (draws of
    record i from the filter stream (seed, i, batch, 0))."""
    out = {k: np.array(v, copy=True) for k, v in st.items()}
    pos = _f(pos, (-1, 3))
    b_sw = np.ascontiguousarray(out["scal"][0], np.float32).copy()
    bs = np.ascontiguousarray(out["b_sums"], np.float32)
    lo = _f(aabb_min, (3,))
    lib().orc_sdtree_record(_p(_u(out["stree"]), u32p), _p(lo), C.c_float(size), _p(_bases(out["b_cnt"]), u32p),
                            _p(b_sw), _p(bs), _p(np.ascontiguousarray(out["b_child"], np.uint16), u16p), _p(pos),
                            _p(_f(canon)), _p(_f(irr)), _p(_f(sw)), C.c_uint32(pos.shape[0]), C.c_uint32(filter),
                            C.c_uint64(seed), C.c_uint32(batch))
    out["scal"][0] = b_sw
    out["b_sums"] = bs
    return out


def sdtree_record_box(st, aabb_min, size, pos, canon, irr, sw, spatial, directional):
    """The reference's deterministic filtered record (spatial, directional:
    0 Nearest, 1 Box) into a copy of the state's building trees."""
    out = {k: np.array(v, copy=True) for k, v in st.items()}
    pos = _f(pos, (-1, 3))
    b_sw = np.ascontiguousarray(out["scal"][0], np.float32).copy()
    bs = np.ascontiguousarray(out["b_sums"], np.float32)
    lo = _f(aabb_min, (3,))
    lib().orc_sdtree_record_box(_p(_u(out["stree"]), u32p), _p(lo), C.c_float(size),
                                _p(_bases(out["b_cnt"]), u32p), _p(b_sw), _p(bs),
                                _p(np.ascontiguousarray(out["b_child"], np.uint16), u16p), _p(pos), _p(_f(canon)),
                                _p(_f(irr)), _p(_f(sw)), C.c_uint32(pos.shape[0]), C.c_int(spatial),
                                C.c_int(directional))
    out["scal"][0] = b_sw
    out["b_sums"] = bs
    return out


def sdtree_sample(st, cells, seed):
    cells = _u(cells)
    out = np.zeros((cells.shape[0], 2), np.float32)
    lib().orc_sdtree_sample(_p(_bases(st["s_cnt"]), u32p), _p(_f(st["scal"][4])), _p(_f(st["scal"][3])),
                            _p(_f(st["s_sums"])), _p(np.ascontiguousarray(st["s_child"], np.uint16), u16p),
                            _p(cells, u32p), cells.shape[0], C.c_uint64(seed), _p(out))
    return out


def sdtree_pdf(st, cells, canon):
    cells = _u(cells)
    out = np.zeros(cells.shape[0], np.float32)
    lib().orc_sdtree_pdf(_p(_bases(st["s_cnt"]), u32p), _p(_f(st["scal"][4])), _p(_f(st["scal"][3])),
                         _p(_f(st["s_sums"])), _p(np.ascontiguousarray(st["s_child"], np.uint16), u16p),
                         _p(cells, u32p), _p(_f(canon)), cells.shape[0], _p(out))
    return out


# ---- geometry
def brute_closest(verts, idx, rays):
    v, i, r = _f(verts), _u(idx), _f(rays, (-1, 8))
    hits = np.zeros((r.shape[0], 4), np.float32)
    lib().orc_brute_closest(_p(v), _p(i, u32p), i.shape[0], _p(r), r.shape[0], _p(hits))
    return hits[:, 0].copy(), hits[:, 1].copy(), hits[:, 2].copy(), hits[:, 3].view(np.uint32).copy()


# ---- integrator
class Scene:
    def __init__(self, sc, light_mesh=None):
        L = lib()
        self._keep = [_f(sc["verts"]), _u(sc["idx"]), _f(sc["albedo"]), _u(sc["mat"]), _f(sc["emission"]),
                      None if sc.get("spec") is None else _f(sc["spec"])]
        v, i, a, m, e, s = self._keep
        self.h = L.orc_scene_create(_p(v), v.shape[0], _p(i, u32p), i.shape[0], _p(a), _p(m, u32p), _p(e), _p(s))
        if light_mesh is not None:
            # uniform light mesh, then uniform triangle within it
            verts, idx = sc["verts"].astype(np.float32), sc["idx"]
            em = np.asarray(sc["emission"])
            lt = []
            for t in range(idx.shape[0]):
                if em[t].sum() <= 0:
                    continue
                p0, p1, p2 = verts[idx[t]]
                if 0.5 * np.linalg.norm(np.cross(p1 - p0, p2 - p0)) > 0:
                    lt.append(t)
            meshes = np.asarray(light_mesh)[lt]
            uniq, counts = np.unique(meshes, return_counts=True)
            cnt = dict(zip(uniq.tolist(), counts.tolist()))
            pdf = np.array([1.0 / (len(uniq) * cnt[m]) for m in meshes.tolist()], np.float32)
            L.orc_scene_set_light_pdf(C.c_void_p(self.h), _p(pdf))
        if sc.get("tex_id") is not None or sc.get("vnormal") is not None:
            tid = None if sc.get("tex_id") is None else _u(np.asarray(sc["tex_id"]).reshape(-1))
            tc = None if tid is None else _f(np.asarray(sc["texcoord"]).reshape(-1))
            vn = None if sc.get("vnormal") is None else _f(np.asarray(sc["vnormal"]).reshape(-1))
            tx = None if tid is None else _u(np.asarray(sc["texels"]).reshape(-1))
            td = None if tid is None else _u(np.asarray(sc["tex"]).reshape(-1))
            self._keep += [tc, tid, vn, tx, td]
            L.orc_scene_set_attributes(C.c_void_p(self.h), _p(tc), _p(tid, u32p), _p(vn), _p(tx, u32p),
                                       0 if tx is None else tx.size, _p(td, u32p), 0 if td is None else td.size // 4)

    def __del__(self):
        try:
            lib().orc_scene_destroy(self.h)
        except Exception:
            pass


def tex_sample(texels, tex, u, v):
    """Filtered RGB of one RGBA8 texture (first texel, width, height, wrap) at
    normalised (u, v)."""
    tx = _u(np.asarray(texels).reshape(-1))
    td = _u(np.asarray(tex).reshape(4))
    out = np.zeros(3, np.float32)
    lib().orc_tex_sample(_p(tx, u32p), _p(td, u32p), float(u), float(v), _p(out))
    return out


def render(scene: Scene, camera_uvw, width, height, spp, flags, max_depth=8, rr_depth=3, num_candidates=32, seed=0,
           sample_offset=0, bg=(0, 0, 0), fraction=0.3, px_range=None, threads=None, guide=None, accum=None):
    """Renders with the CPU integrator; returns (accum H x W x 3 sums, rays).
    guide: a Guide (SD-tree) or a HashGuide (hash grid + Morton pyramids)."""
    p = Params()
    p.width, p.height, p.flags, p.max_depth, p.rr_depth = width, height, flags, max_depth, rr_depth
    p.num_candidates, p.spp, p.sample_offset, p.seed = num_candidates, spp, sample_offset, seed
    for k in range(3):
        p.bg[k] = bg[k]
        p.eye[k], p.u[k], p.v[k], p.w[k] = (float(x[k]) for x in camera_uvw)
    p.fraction = fraction
    p.px_begin, p.px_end = (0, width * height) if px_range is None else px_range
    p.threads = threads or os.cpu_count() or 1
    if accum is None:
        accum = np.zeros((height, width, 3), np.float32)
    if isinstance(guide, HashGuide):
        rays = lib().orc_render_grid(scene.h, C.byref(p), C.byref(guide.struct), _p(accum))
        return accum, rays
    gp = None
    if guide is not None:
        gp = C.byref(guide.struct)
    rays = lib().orc_render(scene.h, C.byref(p), gp, _p(accum))
    return accum, rays


def debug_frame(scene: Scene, camera_uvw, width, height, frame_type, seed=0, sample_offset=0):
    """Debug frame (orc_debug_frame): (value H x W x 3, triangle H x W, -1 on a miss)."""
    p = Params()
    p.width, p.height, p.seed, p.sample_offset = width, height, seed, sample_offset
    for k in range(3):
        p.eye[k], p.u[k], p.v[k], p.w[k] = (float(x[k]) for x in camera_uvw)
    val = np.zeros((height, width, 3), np.float32)
    tri = np.zeros((height, width), np.int32)
    lib().orc_debug_frame(scene.h, C.byref(p), frame_type, _p(val), tri.ctypes.data_as(C.POINTER(C.c_int32)))
    return val, tri


def adjoint(scene: Scene, camera_uvw, width, height, spp, flags, dldi, n_lights, max_depth=6, seed=0,
            sample_offset=0, bg=(0, 0, 0)):
    """FP64 reference gradient (see orc_adjoint): (grad_albedo nt x 3, grad_le nl x 3)."""
    p = Params()
    p.width, p.height, p.flags, p.max_depth, p.rr_depth = width, height, flags, max_depth, 1000
    p.num_candidates, p.spp, p.sample_offset, p.seed = 1, spp, sample_offset, seed
    for k in range(3):
        p.bg[k] = bg[k]
        p.eye[k], p.u[k], p.v[k], p.w[k] = (float(x[k]) for x in camera_uvw)
    p.fraction = 0.0
    p.px_begin, p.px_end = 0, width * height
    p.threads = 1
    nt = scene._keep[1].shape[0]
    ga = np.zeros((nt, 3), np.float64)
    gl = np.zeros((max(n_lights, 1), 3), np.float64)
    d = _f(dldi, (-1,))
    lib().orc_adjoint(scene.h, C.byref(p), _p(d), ga.ctypes.data_as(C.POINTER(C.c_double)),
                      gl.ctypes.data_as(C.POINTER(C.c_double)))
    return ga, gl[:n_lights]


class Guide:
    """Binds a state dict (exported layout) to the integrator: sampling trees
    read, building trees recorded into (in place, in `state`)."""

    def __init__(self, state, aabb_min, size, filter=0):
        self.state = state
        st = state
        self.s_base = _bases(st["s_cnt"])
        self.b_base = _bases(st["b_cnt"])
        self.stree = _u(st["stree"])
        self.scal = np.ascontiguousarray(st["scal"], np.float32)
        st["scal"] = self.scal
        self.s_sums = _f(st["s_sums"])
        self.s_child = np.ascontiguousarray(st["s_child"], np.uint16)
        self.b_sums = np.ascontiguousarray(st["b_sums"], np.float32)
        st["b_sums"] = self.b_sums
        self.b_child = np.ascontiguousarray(st["b_child"], np.uint16)
        g = SdTree()
        g.stree = _p(self.stree, u32p)
        for k in range(3):
            g.aabb_min[k] = aabb_min[k]
        g.size = size
        g.s_base = _p(self.s_base, u32p)
        g.s_sum = self.scal[4].ctypes.data_as(f32p)
        g.s_sw = self.scal[3].ctypes.data_as(f32p)
        g.s_sums = _p(self.s_sums)
        g.s_child = _p(self.s_child, u16p)
        g.b_base = _p(self.b_base, u32p)
        g.b_sw = self.scal[0].ctypes.data_as(f32p)
        g.b_sums = _p(self.b_sums)
        g.b_child = _p(self.b_child, u16p)
        g.filter = filter
        self.struct = g


class HashGuide:
    """Binds hash-grid tables and Morton pyramids to the integrator: cur is
    inserted into in a Locate pass, prv probed, the sampling pyramids s_w read
    and the building pyramids b_w recorded into (all in place)."""

    def __init__(self, origin, size, bounds, slots, key_mode, cur, prv, level, s_w, b_w):
        self.cur = np.ascontiguousarray(cur, np.uint32)
        self.prv = np.ascontiguousarray(prv, np.uint32)
        self.s_w = np.ascontiguousarray(s_w, np.float32)
        self.b_w = np.ascontiguousarray(b_w, np.float32)
        assert self.cur.shape == (slots,) and self.prv.shape == (slots,)
        per_cell = ((1 << (2 * (level + 1))) - 1) // 3
        assert self.s_w.shape == (slots, per_cell) and self.b_w.shape == (slots, per_cell)
        g = HashGuideS()
        g.grid = grid(origin, size, bounds, slots, key_mode)
        g.cur, g.prv = _p(self.cur, u32p), _p(self.prv, u32p)
        g.level, g.per_cell = level, per_cell
        g.s_w, g.b_w = _p(self.s_w), _p(self.b_w)
        self.struct = g


def ctrl_schedule(morton, budget, iter_built=0, iter_locate=0, ratio=0.5, spl=1, cap=1 << 16):
    """Launch-by-launch schedule of the reference's pass controllers (see
    orc_ctrl_schedule): list of (state, iteration, pass samples, events)."""
    st, it, ps, ev = (np.zeros(cap, np.uint32) for _ in range(4))
    n = lib().orc_ctrl_schedule(int(morton), budget, iter_built, iter_locate, float(ratio), spl, cap, _p(st, u32p),
                                _p(it, u32p), _p(ps, u32p), _p(ev, u32p))
    return [(int(st[k]), int(it[k]), int(ps[k]), int(ev[k])) for k in range(n)]


def ctrl_passes(morton, budget, iter_built=0, iter_locate=0, ratio=0.5, spl=1):
    """The schedule grouped into passes: list of (state, samples) of every pass
    that traces samples, in order (the controllers' extra launch after the
    budget, whose pass holds no samples, is left out)."""
    out = []
    for state, _, pass_spp, ev in ctrl_schedule(morton, budget, iter_built, iter_locate, ratio, spl):
        if pass_spp == 0:
            continue
        if not out or out[-1][2]:
            out.append([state, 0, False])
        out[-1][1] += spl
        out[-1][2] = bool(ev & 2)
    return [(s, n) for s, n, _ in out]
"""ctypes binding of librtpg.so (the C ABI declared in include/rtpg.h).

The library is loaded from the package directory.  There is no fallback: if
the HIP library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "librtpg.so"

c_u32p = C.POINTER(C.c_uint32)
c_u16p = C.POINTER(C.c_uint16)
c_u64p = C.POINTER(C.c_uint64)
c_fp = C.POINTER(C.c_float)
c_vp = C.c_void_p


class RenderParams(C.Structure):
    _fields_ = [
        ("flags", C.c_uint32),
        ("spp", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("rr_depth", C.c_uint32),
        ("num_candidates", C.c_uint32),
        ("sample_offset", C.c_uint32),
        ("seed", C.c_uint64),
        ("bg", C.c_float * 3),
        ("fraction", C.c_float),
        ("sdtree", c_vp),
        ("hashgrid", c_vp),
        ("mqtree", c_vp),
        ("max_paths", C.c_uint32),
        ("mlp", c_vp),
    ]


class MlpConfig(C.Structure):
    _fields_ = [
        ("log2_table", C.c_uint32),
        ("base_res", C.c_float),
        ("max_res", C.c_float),
        ("lr", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("eps", C.c_float),
        ("batch", C.c_uint32),
        ("capacity", C.c_uint32),
        ("record_paths", C.c_uint32),
        ("seed", C.c_uint64),
        ("guide_depth", C.c_uint32),
        ("cache_res", C.c_uint32),
    ]


# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "rtpg_last_error": [],
    "rtpg_version": [],
    "rtpg_context_create": [C.c_int, C.POINTER(c_vp)],
    "rtpg_context_destroy": [c_vp],
    "rtpg_context_set_stream": [c_vp, c_vp],
    "rtpg_synchronize": [c_vp],
    "rtpg_scene_upload": [c_vp, c_fp, C.c_uint32, c_u32p, C.c_uint32, c_fp, c_u32p, c_fp, c_fp, C.c_uint32, c_u32p],
    "rtpg_scene_set_albedo": [c_vp, c_fp],
    "rtpg_scene_attributes": [c_vp, c_fp, c_u32p, c_fp, c_u32p, C.c_uint64, c_u32p, C.c_uint32],
    "rtpg_scene_info": [c_vp, c_u32p, c_fp],
    "rtpg_camera_set": [c_vp, c_fp, c_fp, c_fp, c_fp],
    "rtpg_frame_resize": [c_vp, C.c_uint32, C.c_uint32],
    "rtpg_accum_clear": [c_vp],
    "rtpg_accum_download": [c_vp, c_fp, c_u32p],
    "rtpg_frame_tonemap": [c_vp, C.POINTER(C.c_uint8)],
    "rtpg_render": [c_vp, C.POINTER(RenderParams)],
    "rtpg_render_debug": [c_vp, C.POINTER(RenderParams), C.c_uint32, c_fp, C.POINTER(C.c_uint8)],
    "rtpg_render_stats": [c_vp, c_u64p],
    "rtpg_render_relvar": [c_vp, C.POINTER(C.c_double)],
    "rtpg_render_moments": [c_vp, c_vp, C.POINTER(C.c_uint32)],
    "rtpg_render_profile": [c_vp, C.c_uint32],
    "rtpg_adjoint": [c_vp, C.POINTER(RenderParams), c_fp, c_fp, c_fp],
    "rtpg_scene_set_emission": [c_vp, c_fp],
    "rtpg_scene_lights": [c_vp, c_u32p],
    "rtpg_render_kernel_stats": [c_vp, C.POINTER(C.c_double), c_u64p, c_u64p],
    "rtpg_mlp_create": [c_vp, c_fp, c_fp, C.POINTER(MlpConfig), C.POINTER(c_vp)],
    "rtpg_mlp_destroy": [c_vp],
    "rtpg_mlp_cache_download": [c_vp, c_fp, c_fp, C.POINTER(C.c_uint32)],
    "rtpg_mlp_reset": [c_vp, C.c_uint64],
    "rtpg_mlp_forward": [c_vp, c_fp, c_fp, c_fp, c_fp, C.c_uint32, c_fp, c_fp],
    "rtpg_mlp_sample": [c_vp, c_fp, C.c_uint32, C.c_uint64, c_fp],
    "rtpg_mlp_gradient": [c_vp, c_fp, c_fp, c_fp, c_fp, c_fp, c_fp, c_fp, C.c_uint32, c_fp, c_fp, c_fp],
    "rtpg_mlp_train": [c_vp, c_fp, c_fp, c_fp, c_fp, c_fp, c_fp, c_fp, C.c_uint32, C.c_uint32, c_fp],
    "rtpg_mlp_train_recorded": [c_vp, C.c_uint32, c_fp, c_u32p],
    "rtpg_mlp_export": [c_vp, c_fp, c_fp],
    "rtpg_mlp_import": [c_vp, c_fp, c_fp],
    "rtpg_mlp_info": [c_vp, c_u64p, c_fp],
    "rtpg_trace_closest": [c_vp, c_fp, C.c_uint32, c_fp],
    "rtpg_trace_occluded": [c_vp, c_fp, C.c_uint32, c_u32p],
    "rtpg_trace_closest_noalloc": [c_vp, c_vp, C.c_uint32, c_vp],
    "rtpg_sdtree_create": [c_vp, c_fp, c_fp, C.c_uint32, C.c_float, C.POINTER(c_vp)],
    "rtpg_sdtree_destroy": [c_vp],
    "rtpg_sdtree_clear": [c_vp],
    "rtpg_sdtree_set_split": [c_vp, C.c_int, C.c_float, C.c_int, C.c_float],
    "rtpg_sdtree_reset": [c_vp, C.c_int, C.c_int],
    "rtpg_sdtree_build": [c_vp],
    "rtpg_sdtree_record": [c_vp, c_fp, c_fp, c_fp, c_fp, C.c_uint32],
    "rtpg_sdtree_set_filter": [c_vp, C.c_int, C.c_int, C.c_uint64],
    "rtpg_sdtree_set_octree": [c_vp, C.c_int],
    "rtpg_sdtree_lookup": [c_vp, c_fp, C.c_uint32, c_u32p],
    "rtpg_sdtree_sample": [c_vp, c_u32p, C.c_uint32, C.c_uint64, c_fp],
    "rtpg_sdtree_pdf": [c_vp, c_u32p, c_fp, C.c_uint32, c_fp],
    "rtpg_sdtree_sizes": [c_vp, c_u32p],
    "rtpg_sdtree_export": [c_vp, c_u32p, c_u32p, c_u32p, c_fp, c_fp, c_u16p, c_fp, c_u16p],
    "rtpg_sdtree_import": [c_vp, C.c_uint32, c_u32p, C.c_uint32, c_u32p, c_u32p, c_fp, c_fp, c_u16p, c_fp, c_u16p],
    "rtpg_sdtree_ctrl_create": [c_vp, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, C.POINTER(c_vp)],
    "rtpg_sdtree_ctrl_destroy": [c_vp],
    "rtpg_sdtree_ctrl_start": [c_vp],
    "rtpg_sdtree_ctrl_beg_trace": [c_vp],
    "rtpg_sdtree_ctrl_end_trace": [c_vp, C.c_uint32],
    "rtpg_sdtree_ctrl_query": [c_vp, c_u32p],
    "rtpg_hashgrid_create": [c_vp, c_fp, c_fp, c_u32p, C.c_uint32, C.c_uint32, C.POINTER(c_vp)],
    "rtpg_hashgrid_destroy": [c_vp],
    "rtpg_hashgrid_clear": [c_vp],
    "rtpg_hashgrid_update": [c_vp],
    "rtpg_hashgrid_find": [c_vp, c_fp, C.c_uint32, C.c_int, c_u32p],
    "rtpg_hashgrid_download": [c_vp, C.c_int, c_u32p],
    "rtpg_mqtree_create": [c_vp, C.c_uint32, C.c_uint32, C.c_float, C.POINTER(c_vp)],
    "rtpg_mqtree_destroy": [c_vp],
    "rtpg_mqtree_allocate": [c_vp],
    "rtpg_mqtree_update": [c_vp],
    "rtpg_mqtree_build": [c_vp],
    "rtpg_mqtree_record": [c_vp, c_u32p, c_fp, c_fp, C.c_uint32],
    "rtpg_mqtree_sample": [c_vp, c_u32p, C.c_uint32, C.c_uint64, c_fp],
    "rtpg_mqtree_pdf": [c_vp, c_u32p, c_fp, C.c_uint32, c_fp],
    "rtpg_mqtree_download": [c_vp, C.c_int, c_fp],
    "rtpg_mqtree_upload": [c_vp, C.c_int, c_fp],
    "rtpg_mqtree_info": [c_vp, c_u32p],
    "rtpg_mqtree_ctrl_create": [c_vp, c_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32,
                                C.POINTER(c_vp)],
    "rtpg_mqtree_ctrl_destroy": [c_vp],
    "rtpg_mqtree_ctrl_start": [c_vp],
    "rtpg_mqtree_ctrl_beg_trace": [c_vp],
    "rtpg_mqtree_ctrl_end_trace": [c_vp, C.c_uint32],
    "rtpg_mqtree_ctrl_query": [c_vp, c_u32p],
}
_RESTYPE = {"rtpg_last_error": C.c_char_p, "rtpg_version": C.c_char_p}

_lib = None


class RtpgError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, name: str, code: int, msg: str):
        super().__init__(f"{name} failed ({code}): {msg}")
        self.code = code


def load() -> C.CDLL:
    """Loads librtpg.so (building nothing): raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m rtlib_amd.build`")
    lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise RtpgError(name, rc, lib.rtpg_last_error().decode(errors="replace"))


def ptr(a, ctype=C.c_float):
    """Pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))

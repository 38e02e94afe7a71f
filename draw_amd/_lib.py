"""ctypes binding of libdrawhip.so (include/drawhip.h).

The library is built in-tree by ``python -m draw_amd.build``; importing this
module without it raises ImportError -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdrawhip.so")

DH_OK = 0
ERRORS = {
    0: "DH_OK", -1: "DH_E_ARG", -2: "DH_E_INDEX", -3: "DH_E_TEXCOORD", -4: "DH_E_CAMERA", -5: "DH_E_PLANE",
    -6: "DH_E_HIP", -7: "DH_E_NOMEM", -8: "DH_E_HANDLE", -9: "DH_E_UNSUPPORTED", -10: "DH_E_BUDGET",
    -11: "DH_E_CAPACITY",
}


class DhFrameParams(C.Structure):
    _fields_ = [("M", C.c_float * 16), ("cam", C.c_float * 3), ("light", C.c_float * 3),
                ("planes", (C.c_float * 4) * 6), ("W", C.c_int32), ("H", C.c_int32), ("off_x", C.c_float),
                ("off_y", C.c_float), ("clear_depth", C.c_float), ("clear_bgra", C.c_uint8 * 4)]


class DhCamera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("dir", C.c_float * 3), ("fov_x_deg", C.c_float), ("near_", C.c_float),
                ("far_", C.c_float)]


class DhMaterialDesc(C.Structure):
    _fields_ = [("ka", C.c_float * 3), ("kd", C.c_float * 3), ("ks", C.c_float * 3), ("alpha", C.c_float),
                ("power", C.c_int32), ("map_kd", C.c_uint64), ("map_ka", C.c_uint64)]


class DhView(C.Structure):
    _fields_ = [("fr", DhFrameParams), ("canvas", C.c_uint64), ("meshes", C.POINTER(C.c_uint64)),
                ("n_meshes", C.c_uint32), ("pad", C.c_uint32)]


DH_BATCH_GRAPH = 1
DH_CANVAS_NO_WINNER = 1


class DhServerConfig(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("workers", C.c_uint32), ("max_views", C.c_uint32),
                ("max_tris", C.c_uint64)]


class DhJob(C.Structure):
    _fields_ = [("fr", DhFrameParams), ("meshes", C.POINTER(C.c_uint64)), ("n_meshes", C.c_uint32),
                ("flags", C.c_uint32)]


class DhJobInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("batch_views", C.c_uint32), ("canvas", C.c_uint64),
                ("t_submit_ms", C.c_double), ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double)]


class DhServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("launches", C.c_uint64), ("bytes_in_use", C.c_uint64), ("bytes_peak", C.c_uint64),
                ("queued", C.c_uint64)]


class DhObjSummary(C.Structure):
    _fields_ = [("n_pos", C.c_uint32), ("n_nrm", C.c_uint32), ("n_uv", C.c_uint32), ("n_meshes", C.c_uint32),
                ("n_textures", C.c_uint32), ("pad", C.c_uint32), ("pos", C.POINTER(C.c_float)),
                ("nrm", C.POINTER(C.c_float)), ("uv", C.POINTER(C.c_float))]


class DhObjMeshDesc(C.Structure):
    _fields_ = [("tri", C.POINTER(C.c_uint32)), ("n_tri", C.c_uint32), ("texture_idx", C.c_uint32),
                ("name", C.c_char_p)]


class DhObjTextureDesc(C.Structure):
    _fields_ = [("ka", C.c_float * 3), ("kd", C.c_float * 3), ("ks", C.c_float * 3), ("alpha", C.c_float),
                ("name", C.c_char_p), ("map_kd", C.POINTER(C.c_uint8)), ("kd_w", C.c_int32), ("kd_h", C.c_int32),
                ("kd_comp", C.c_int32), ("map_ka", C.POINTER(C.c_uint8)), ("ka_w", C.c_int32), ("ka_h", C.c_int32),
                ("ka_comp", C.c_int32)]


class DhGuiVertex(C.Structure):
    _fields_ = [("pos", C.c_float * 2), ("uv", C.c_float * 2), ("col", C.c_uint8 * 4)]


class DhGuiCmd(C.Structure):
    _fields_ = [("clip_rect", C.c_float * 4), ("idx_offset", C.c_uint32), ("count", C.c_uint32),
                ("vtx_offset", C.c_uint32), ("pad", C.c_uint32)]


DH_JOB_AUTORELEASE = 1
DH_JOB_DONE, DH_JOB_QUEUED, DH_JOB_RUNNING = 0, 1, 2


class DhStats(C.Structure):
    _fields_ = [("tris_in", C.c_uint64), ("tris_out", C.c_uint64), ("bin_entries", C.c_uint64),
                ("frags", C.c_uint64), ("retries", C.c_uint64), ("ms_geometry", C.c_float),
                ("ms_binning", C.c_float), ("ms_raster", C.c_float), ("ms_shade", C.c_float),
                ("ms_frame", C.c_float)]


# Every symbol declared in include/drawhip.h, with its ctypes signature.
_P = C.c_void_p
_H = C.c_uint64
SIGNATURES = {
    "dh_frame_uniforms": [C.POINTER(DhCamera), C.c_int32, C.c_int32, C.POINTER(C.c_float), C.POINTER(DhFrameParams)],
    "dh_frame_uniforms_canvas": [C.POINTER(DhCamera)] + [C.c_int32] * 6 + [C.POINTER(C.c_float),
                                                                          C.POINTER(DhFrameParams)],
    "dh_ctx_create": [C.c_int32, C.POINTER(_P)],
    "dh_ctx_create_shared": [_P, C.POINTER(_P)],
    "dh_ctx_destroy": [_P],
    "dh_last_error": [_P],
    "dh_set_tiling": [_P, C.c_int32, C.c_int32],
    "dh_set_profiling": [_P, C.c_int32],
    "dh_set_bin_order": [_P, C.c_int32],
    "dh_set_graph": [_P, C.c_int32],
    "dh_render_shard": [_P, C.POINTER(DhFrameParams), C.POINTER(C.c_uint64), C.c_uint32, C.c_uint32, C.c_uint64,
                        C.c_void_p],
    "dh_vis_merge_min": [_P, C.c_void_p, C.c_void_p, C.c_uint64],
    "dh_shade_shard": [_P, C.c_void_p, C.c_uint32],
    "dh_device_alloc": [_P, C.c_uint64, C.POINTER(C.c_void_p)],
    "dh_device_free": [_P, C.c_void_p],
    "dh_device_read": [_P, C.c_void_p, C.c_void_p, C.c_uint64],
    "dh_set_raster_kernel": [_P, C.c_int32, C.c_int32, C.c_int32],
    "dh_texture_create": [_P, _P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(_H)],
    "dh_material_create": [_P, C.POINTER(DhMaterialDesc), C.POINTER(_H)],
    "dh_object_create": [_P, _P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, C.POINTER(_H)],
    "dh_mesh_create": [_P, _H, _P, C.c_uint32, _H, C.POINTER(_H)],
    "dh_canvas_create": [_P, C.c_int32, C.c_int32, C.POINTER(_H)],
    "dh_canvas_create_ex": [_P, C.c_int32, C.c_int32, C.c_uint32, C.POINTER(_H)],
    "dh_mesh_read": [_P, _H, _P],
    "dh_release": [_P, _H],
    "dh_render": [_P, C.POINTER(DhFrameParams), C.POINTER(_H), C.c_uint32, _H],
    "dh_render_batch": [_P, C.POINTER(DhView), C.c_uint32, C.c_uint32],
    "dh_canvas_read": [_P, _H, _P, _P, _P],
    "dh_canvas_device_ptrs": [_P, _H, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)],
    "dh_canvas_read_rgba": [_P, _H, _P],
    "dh_canvas_read_rgba_noalloc": [_P, _H, _P],
    "dh_canvas_export_png": [_P, _H, C.c_char_p],
    "dh_write_png": [C.c_char_p, _P, C.c_int32, C.c_int32],
    "dh_canvas_export_jpg": [_P, _H, C.c_char_p, C.c_int32],
    "dh_write_jpg": [C.c_char_p, _P, C.c_int32, C.c_int32, C.c_int32],
    "dh_get_stats": [_P, C.POINTER(DhStats)],
    "dh_draw_gui": [_P, _H, _P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, _H],
    "dh_image_info": [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_char_p,
                      C.c_uint32],
    "dh_image_load": [C.c_char_p, _P, C.c_uint64, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                      C.c_char_p, C.c_uint32],
    "dh_obj_load": [C.c_char_p, C.POINTER(_P), C.c_char_p, C.c_uint32],
    "dh_obj_info": [_P, C.POINTER(DhObjSummary)],
    "dh_obj_mesh": [_P, C.c_uint32, C.POINTER(DhObjMeshDesc)],
    "dh_obj_texture": [_P, C.c_uint32, C.POINTER(DhObjTextureDesc)],
    "dh_obj_upload": [_P, _P, C.POINTER(_H), C.POINTER(_H), C.c_uint32],
    "dh_obj_free": [_P],
    "dh_srv_create": [_P, C.POINTER(DhServerConfig), C.POINTER(_P)],
    "dh_srv_submit": [_P, C.POINTER(DhJob), C.POINTER(C.c_uint64)],
    "dh_srv_poll": [_P, C.c_uint64, C.c_int32, C.POINTER(DhJobInfo)],
    "dh_srv_release": [_P, C.c_uint64],
    "dh_srv_stats": [_P, C.POINTER(DhServerStats)],
    "dh_srv_destroy": [_P],
}

_lib = None


def lib() -> C.CDLL:
    """Load libdrawhip.so (raises ImportError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m draw_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_char_p if name == "dh_last_error" else C.c_int
        _lib = L
    return _lib


class DrawHipError(RuntimeError):
    def __init__(self, code: int, what: str, detail: str = ""):
        self.code = code
        super().__init__(f"{what} failed: {ERRORS.get(code, code)}" + (f" ({detail})" if detail else ""))

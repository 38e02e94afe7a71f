"""ctypes wrapper of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by draw_amd.  It renders a draw_amd.scene
``Scene`` exactly as the reference's Scene::render would (scene/mod.rs:901-1249)
and returns the frame (BGRA, top-down), depth (bottom-up), per-pixel winner id
(bottom-up; the emission index of the opaque clipped triangle that last wrote
the pixel) and counters.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ERRORS = {0: "OK", -1: "E_ARG", -2: "E_INDEX", -3: "E_TEXCOORD", -4: "E_CAMERA", -5: "E_PLANE"}


class OrFrame(C.Structure):
    _fields_ = [("M", C.c_float * 16), ("cam", C.c_float * 3), ("light", C.c_float * 3),
                ("planes", (C.c_float * 4) * 6), ("W", C.c_int32), ("H", C.c_int32),
                ("off_x", C.c_float), ("off_y", C.c_float), ("clear_depth", C.c_float),
                ("clear_bgra", C.c_uint8 * 4)]


class OrCameraDesc(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("dir", C.c_float * 3), ("fov_x_deg", C.c_float),
                ("near_", C.c_float), ("far_", C.c_float)]


class OrMaterial(C.Structure):
    _fields_ = [("ka", C.c_float * 3), ("kd", C.c_float * 3), ("ks", C.c_float * 3), ("alpha", C.c_float),
                ("power", C.c_int32),
                ("map_kd", C.c_void_p), ("kd_w", C.c_int32), ("kd_h", C.c_int32), ("kd_comp", C.c_int32),
                ("map_ka", C.c_void_p), ("ka_w", C.c_int32), ("ka_h", C.c_int32), ("ka_comp", C.c_int32)]


class OrMesh(C.Structure):
    _fields_ = [("tri", C.c_void_p), ("n_tri", C.c_uint32), ("material", C.c_int32)]


class OrObject(C.Structure):
    _fields_ = [("pos", C.c_void_p), ("n_pos", C.c_uint32), ("nrm", C.c_void_p), ("n_nrm", C.c_uint32),
                ("uv", C.c_void_p), ("n_uv", C.c_uint32), ("meshes", C.POINTER(OrMesh)),
                ("n_meshes", C.c_uint32)]


class OrStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("tris_in", "tris_culled", "tris_emitted", "tris_emitted_tr",
                                          "px_tested", "frags_covered", "frags_written")]


class OrAttr(C.Structure):
    _fields_ = [("sx", C.c_float), ("sy", C.c_float), ("depth", C.c_float), ("normal", C.c_float * 3),
                ("light", C.c_float * 3), ("halfway", C.c_float * 3), ("uv", C.c_float * 2)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.or_frame_uniforms.argtypes = [C.POINTER(OrCameraDesc), C.c_int32, C.c_int32, C.POINTER(C.c_float),
                                        C.POINTER(OrFrame)]
        L.or_frame_uniforms_canvas.argtypes = [C.POINTER(OrCameraDesc)] + [C.c_int32] * 6 + [
            C.POINTER(C.c_float), C.POINTER(OrFrame)]
        L.or_render.argtypes = [C.POINTER(OrFrame), C.POINTER(OrObject), C.c_uint32, C.POINTER(OrMaterial),
                                C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(OrStats)]
        L.or_render_range.argtypes = [C.POINTER(OrFrame), C.POINTER(OrObject), C.c_uint32, C.POINTER(OrMaterial),
                                      C.c_uint32, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.POINTER(OrStats)]
        L.or_render_grid.argtypes = [C.POINTER(OrFrame), C.POINTER(OrObject), C.c_uint32, C.POINTER(OrMaterial),
                                     C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.POINTER(OrStats)]
        L.or_vertex_visual.argtypes = [C.POINTER(C.c_float)] * 4
        L.or_vertex_visual.restype = None
        L.or_clip.argtypes = [C.POINTER(OrFrame), C.POINTER(C.c_float), C.POINTER(OrAttr), C.POINTER(C.c_float),
                              C.POINTER(OrAttr)]
        L.or_draw_triangle.argtypes = [C.POINTER(OrFrame), C.POINTER(OrAttr), C.POINTER(OrAttr), C.POINTER(OrAttr),
                                       C.POINTER(OrMaterial), C.c_int, C.c_uint32, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.POINTER(OrStats)]
        L.or_clear.argtypes = [C.POINTER(OrFrame), C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_clear.restype = None
        _LIB = L
    return _LIB


def _f3(v):
    return (C.c_float * 3)(*[float(np.float32(x)) for x in v])


def frame_uniforms(scene) -> OrFrame:
    """a1: the oracle's own Camera::new + gen_transformation_matrix +
    gen_view_planes for ``scene``."""
    cam = scene.camera
    cd = OrCameraDesc(_f3(cam.position), _f3(cam.direction), float(cam.fov_x), float(cam.near), float(cam.far))
    fr = OrFrame()
    if getattr(scene, "canvas", None) is None:
        rc = lib().or_frame_uniforms(C.byref(cd), int(scene.width), int(scene.height), _f3(scene.light_source),
                                     C.byref(fr))
    else:
        (cw, ch), (ox, oy) = scene.canvas, scene.canvas_offset
        rc = lib().or_frame_uniforms_canvas(C.byref(cd), int(scene.width), int(scene.height), int(cw), int(ch),
                                            int(ox), int(oy), _f3(scene.light_source), C.byref(fr))
    if rc != 0:
        raise RuntimeError(f"or_frame_uniforms failed: {ERRORS.get(rc, rc)}")
    return fr


@dataclass
class OracleResult:
    rc: int
    frame: np.ndarray     # (H, W, 4) uint8, row 0 = top
    depth: np.ndarray     # (H, W) float32, row 0 = bottom
    winner: np.ndarray    # (H, W) uint32, row 0 = bottom
    stats: dict


def _material(tex, keep):
    kd = np.ascontiguousarray(tex.map_kd.img)
    ka = np.ascontiguousarray(tex.map_ka.img)
    keep += [kd, ka]
    return OrMaterial(_f3(tex.ka), _f3(tex.kd), _f3(tex.ks), float(np.float32(tex.alpha)), int(tex.power),
                      kd.ctypes.data, tex.map_kd.width, tex.map_kd.height, tex.map_kd.components,
                      ka.ctypes.data, tex.map_ka.width, tex.map_ka.height, tex.map_ka.components)


def _marshal(scene, keep):
    mats = []
    objs = []
    for o in scene.objects:
        base = len(mats)
        for t in o.textures:
            mats.append(_material(t, keep))
        meshes = (OrMesh * max(1, len(o.meshes)))()
        for i, m in enumerate(o.meshes):
            tri = np.ascontiguousarray(m.triangles, np.uint32)
            keep.append(tri)
            mi = base + m.texture_idx if 0 <= m.texture_idx < len(o.textures) else -1
            meshes[i] = OrMesh(tri.ctypes.data, tri.shape[0], mi)
        keep.append(meshes)
        pos = np.ascontiguousarray(o.vertices, np.float32)
        nrm = np.ascontiguousarray(o.normals_vertices, np.float32)
        uv = np.ascontiguousarray(o.texture_vertices, np.float32)
        keep += [pos, nrm, uv]
        objs.append(OrObject(pos.ctypes.data, pos.shape[0], nrm.ctypes.data, nrm.shape[0], uv.ctypes.data,
                             uv.shape[0], meshes, len(o.meshes)))
    mat_arr = (OrMaterial * max(1, len(mats)))(*mats)
    obj_arr = (OrObject * max(1, len(objs)))(*objs)
    return obj_arr, len(objs), mat_arr, len(mats)


def render(scene, frame: OrFrame | None = None, tri_range=None, grid=None) -> OracleResult:
    """Scene::render on the CPU.  ``frame`` defaults to the oracle's own a1
    uniforms; ``tri_range`` = (begin, end) restricts the opaque submission
    stream (for bounded timing samples); ``grid`` = (sx, sy, ox, oy) evaluates
    only the pixels with x % sx == ox and y % sy == oy (exact there)."""
    fr = frame if frame is not None else frame_uniforms(scene)
    W, H = fr.W, fr.H
    bgra = np.empty((H, W, 4), np.uint8)
    depth = np.empty((H, W), np.float32)
    winner = np.empty((H, W), np.uint32)
    keep = []
    objs, n_obj, mats, n_mat = _marshal(scene, keep)
    st = OrStats()
    if grid is not None:
        rc = lib().or_render_grid(C.byref(fr), objs, n_obj, mats, n_mat, *[int(g) for g in grid], bgra.ctypes.data,
                                  depth.ctypes.data, winner.ctypes.data, C.byref(st))
    elif tri_range is None:
        rc = lib().or_render(C.byref(fr), objs, n_obj, mats, n_mat, bgra.ctypes.data, depth.ctypes.data,
                             winner.ctypes.data, C.byref(st))
    else:
        rc = lib().or_render_range(C.byref(fr), objs, n_obj, mats, n_mat, int(tri_range[0]), int(tri_range[1]),
                                   bgra.ctypes.data, depth.ctypes.data, winner.ctypes.data, C.byref(st))
    stats = {n: int(getattr(st, n)) for n, _ in OrStats._fields_}
    return OracleResult(rc, bgra, depth, winner, stats)


def make_attr(sx, sy, depth, normal, light, halfway, uv) -> OrAttr:
    return OrAttr(float(sx), float(sy), float(depth), _f3(normal), _f3(light), _f3(halfway),
                  (C.c_float * 2)(*[float(np.float32(x)) for x in uv]))


def draw_triangle(frame: OrFrame, a: OrAttr, b: OrAttr, c: OrAttr, tex, bgra, depth, winner=None,
                  depth_update=True, tid=0):
    """Canvas::draw_triangle_with_attributes onto caller arrays."""
    keep = []
    mat = _material(tex, keep)
    st = OrStats()
    rc = lib().or_draw_triangle(C.byref(frame), C.byref(a), C.byref(b), C.byref(c), C.byref(mat),
                                int(bool(depth_update)), int(tid), bgra.ctypes.data, depth.ctypes.data,
                                None if winner is None else winner.ctypes.data, C.byref(st))
    return rc, {n: int(getattr(st, n)) for n, _ in OrStats._fields_}


def blank_frame(W: int, H: int, clear_depth: float = 100000.0) -> OrFrame:
    fr = OrFrame()
    fr.W, fr.H = W, H
    fr.clear_depth = clear_depth
    for i, v in enumerate((255, 186, 155, 255)):
        fr.clear_bgra[i] = v
    return fr


def vertex_visual(v, cam, light):
    out = (C.c_float * 10)()
    lib().or_vertex_visual(_f3(v), _f3(cam), _f3(light), out)
    return np.frombuffer(out, np.float32).copy()


def clip(frame: OrFrame, pos9, attrs):
    pos = (C.c_float * 9)(*[float(np.float32(x)) for x in pos9])
    at = (OrAttr * 3)(*attrs)
    out_pos = (C.c_float * (12 * 9))()
    out_attr = (OrAttr * 36)()
    n = lib().or_clip(C.byref(frame), pos, at, out_pos, out_attr)
    return n, np.frombuffer(out_pos, np.float32).reshape(12, 3, 3)[:n].copy(), [out_attr[i] for i in range(3 * n)]


# ------------------------------------------------------------ GUI overlay
GUI_VERTEX = np.dtype([("pos", "<f4", (2,)), ("uv", "<f4", (2,)), ("col", "u1", (4,))])          # 20 B
GUI_CMD = np.dtype([("clip_rect", "<f4", (4,)), ("idx_offset", "<u4"), ("count", "<u4"), ("vtx_offset", "<u4"),
                    ("pad", "<u4")])                                                               # 32 B


def gui_draw(bgra, depth, vtx, idx, cmds, font) -> int:
    """or_gui_draw on (H, W, 4) BGRA (rows top-down, modified in place) and
    (H, W) depth (rows bottom-up); font: a TextureMap-like with .img."""
    L = lib()
    if not hasattr(L.or_gui_draw, "_set"):
        L.or_gui_draw.argtypes = [C.c_int32, C.c_int32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                  C.c_uint32, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
        L.or_gui_draw._set = True
    H, W = depth.shape
    vtx = np.ascontiguousarray(vtx, GUI_VERTEX)
    idx = np.ascontiguousarray(idx, np.uint32)
    cmds = np.ascontiguousarray(cmds, GUI_CMD)
    img = np.ascontiguousarray(font.img, np.uint8)
    depth = np.ascontiguousarray(depth, np.float32)
    assert bgra.flags.c_contiguous and bgra.dtype == np.uint8
    return L.or_gui_draw(W, H, vtx.ctypes.data, len(vtx), idx.ctypes.data, len(idx), cmds.ctypes.data, len(cmds),
                         img.ctypes.data, img.shape[1], img.shape[0], img.shape[2], bgra.ctypes.data,
                         depth.ctypes.data)

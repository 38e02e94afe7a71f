"""GPU renderer: Scene::render (scene/mod.rs:901) on libdrawhip.

``Renderer`` owns one libdrawhip context (one HIP stream, device-resident
resources).  It uploads a ``Scene``'s textures, materials, objects and meshes
once (cached per Python object) and renders frames into device canvases;
``render(scene, canvas)`` copies the result into a host ``Canvas`` the way the
reference fills its canvas.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Tuple

import numpy as np

from . import _lib
from ._lib import DhCamera, DhFrameParams, DhMaterialDesc, DhStats, DrawHipError
from .scene import Canvas, Scene


def _f3(v):
    return (C.c_float * 3)(*[float(np.float32(x)) for x in v])


def frame_uniforms(scene: Scene) -> DhFrameParams:
    """a1 on the host: Camera::new + gen_transformation_matrix + gen_view_planes
    (dh_frame_uniforms), or with the scene's own canvas size and offset
    (dh_frame_uniforms_canvas) when it has one."""
    cam = scene.camera
    cd = DhCamera(_f3(cam.position), _f3(cam.direction), float(cam.fov_x), float(cam.near), float(cam.far))
    fr = DhFrameParams()
    if scene.canvas is None:
        rc = _lib.lib().dh_frame_uniforms(C.byref(cd), int(scene.width), int(scene.height),
                                          _f3(scene.light_source), C.byref(fr))
    else:
        cw, ch = scene.canvas_size
        ox, oy = scene.canvas_offset
        rc = _lib.lib().dh_frame_uniforms_canvas(C.byref(cd), int(scene.width), int(scene.height), cw, ch,
                                                 int(ox), int(oy), _f3(scene.light_source), C.byref(fr))
    if rc != 0:
        raise DrawHipError(rc, "dh_frame_uniforms")
    return fr


class Renderer:
    def __init__(self, device: int = 0, tile: int = 16, waves: int = 4, profiling: bool = False):
        self.L = _lib.lib()
        ctx = C.c_void_p()
        rc = self.L.dh_ctx_create(int(device), C.byref(ctx))
        if rc != 0:
            raise DrawHipError(rc, "dh_ctx_create", "no HIP device" if rc == -9 else "")
        self.ctx = ctx
        self._tex: Dict[int, int] = {}
        self._mat: Dict[Tuple[int, int], int] = {}
        self._obj: Dict[int, Tuple[int, List[int]]] = {}
        self._canvas: Dict[tuple, int] = {}
        # canvases created from here on keep a winner id per pixel (the
        # reference's Canvas has none: False drops that buffer and its writes)
        self.winners = True
        self._nowin = set()  # canvases created without winners
        self._keep = []
        self.set_tiling(tile, waves)
        self.set_profiling(profiling)

    # ------------------------------------------------------------ plumbing
    def _check(self, rc: int, what: str):
        if rc != 0:
            detail = self.L.dh_last_error(self.ctx)
            raise DrawHipError(rc, what, detail.decode() if detail else "")

    def close(self):
        if getattr(self, "ctx", None):
            self.L.dh_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tiling(self, tile: int, waves: int):
        self._check(self.L.dh_set_tiling(self.ctx, int(tile), int(waves)), "dh_set_tiling")

    def set_raster(self, kind: int, tile: int, waves: int):
        """Raster kernel kind (0: quadrant-owning waves, 1: triangle-parallel)
        with its tile size and waves per tile."""
        self._check(self.L.dh_set_raster_kernel(self.ctx, int(kind), int(tile), int(waves)), "dh_set_raster_kernel")

    def set_bin_order(self, mode):
        """dh_set_bin_order: 0 unordered, 1 submission order, 2 near first
        (default); False/True mean 0/1."""
        mode = int(mode) if not isinstance(mode, bool) else int(mode)
        self._check(self.L.dh_set_bin_order(self.ctx, mode), "dh_set_bin_order")

    def shard_context(self) -> "Renderer":
        """Another context (own stream and scratch) on this one's resources
        (dh_ctx_create_shared): one per shard of a sort-last frame."""
        out = C.c_void_p()
        self._check(self.L.dh_ctx_create_shared(self.ctx, C.byref(out)), "dh_ctx_create_shared")
        child = Renderer.__new__(Renderer)
        child.L = self.L
        child.ctx = out
        child._tex, child._mat, child._obj, child._canvas, child._keep = (
            self._tex, self._mat, self._obj, self._canvas, self._keep)
        child.winners, child._nowin = self.winners, self._nowin
        return child

    # ------------------------------------------------- sort-last shards
    def device_alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self._check(self.L.dh_device_alloc(self.ctx, int(nbytes), C.byref(p)), "dh_device_alloc")
        return p.value

    def device_free(self, dev: int):
        self._check(self.L.dh_device_free(self.ctx, C.c_void_p(dev)), "dh_device_free")

    def device_read(self, dev: int, nbytes: int, dtype=np.uint8) -> np.ndarray:
        out = np.empty(int(nbytes) // np.dtype(dtype).itemsize, dtype)
        self._check(self.L.dh_device_read(self.ctx, out.ctypes.data, C.c_void_p(dev), int(nbytes)), "dh_device_read")
        return out

    def render_shard(self, frame: DhFrameParams, meshes: List[int], shard: int, canvas: int, vis_dev: int):
        """dh_render_shard: this shard's submission range into the packed u64
        visibility buffer vis_dev (W*H entries)."""
        arr = (C.c_uint64 * max(1, len(meshes)))(*meshes)
        self._check(self.L.dh_render_shard(self.ctx, C.byref(frame), arr, len(meshes), int(shard), canvas,
                                           C.c_void_p(vis_dev)), "dh_render_shard")

    def vis_merge_min(self, acc_dev: int, src_dev: int, n: int):
        self._check(self.L.dh_vis_merge_min(self.ctx, C.c_void_p(acc_dev), C.c_void_p(src_dev), int(n)),
                    "dh_vis_merge_min")

    def shade_shard(self, merged_dev: int, shard: int):
        self._check(self.L.dh_shade_shard(self.ctx, C.c_void_p(merged_dev), int(shard)), "dh_shade_shard")

    def set_graph(self, on: bool):
        """dh_set_graph: replay repeated frame structures from a hipGraph."""
        self._check(self.L.dh_set_graph(self.ctx, int(bool(on))), "dh_set_graph")

    def set_profiling(self, on: bool):
        self._check(self.L.dh_set_profiling(self.ctx, int(bool(on))), "dh_set_profiling")

    # ----------------------------------------------------------- resources
    def texture(self, tm) -> int:
        key = id(tm)
        if key not in self._tex:
            img = np.ascontiguousarray(tm.img, np.uint8)
            h = C.c_uint64()
            self._check(self.L.dh_texture_create(self.ctx, img.ctypes.data, tm.width, tm.height, tm.components,
                                                 C.byref(h)), "dh_texture_create")
            self._tex[key] = h.value
            self._keep.append(tm)
        return self._tex[key]

    def material(self, tex) -> int:
        key = (id(tex), 0)
        if key not in self._mat:
            d = DhMaterialDesc(_f3(tex.ka), _f3(tex.kd), _f3(tex.ks), float(np.float32(tex.alpha)), int(tex.power),
                               self.texture(tex.map_kd), self.texture(tex.map_ka))
            h = C.c_uint64()
            self._check(self.L.dh_material_create(self.ctx, C.byref(d), C.byref(h)), "dh_material_create")
            self._mat[key] = h.value
            self._keep.append(tex)
        return self._mat[key]

    def upload(self, obj) -> List[int]:
        """Object + its meshes -> mesh handles in submission order."""
        key = id(obj)
        if key not in self._obj:
            pos = np.ascontiguousarray(obj.vertices, np.float32)
            nrm = np.ascontiguousarray(obj.normals_vertices, np.float32)
            uv = np.ascontiguousarray(obj.texture_vertices, np.float32)
            oh = C.c_uint64()
            self._check(self.L.dh_object_create(self.ctx, pos.ctypes.data, pos.shape[0], nrm.ctypes.data,
                                                nrm.shape[0], uv.ctypes.data, uv.shape[0], C.byref(oh)),
                        "dh_object_create")
            meshes = []
            for m in obj.meshes:
                mat = self.material(obj.textures[m.texture_idx]) if 0 <= m.texture_idx < len(obj.textures) else 0
                tri = np.ascontiguousarray(m.triangles, np.uint32)
                mh = C.c_uint64()
                self._check(self.L.dh_mesh_create(self.ctx, oh.value, tri.ctypes.data, tri.shape[0], mat,
                                                  C.byref(mh)), "dh_mesh_create")
                meshes.append(mh.value)
            self._obj[key] = (oh.value, meshes)
            self._keep.append(obj)
        return self._obj[key][1]

    def _new_canvas(self, W: int, H: int) -> int:
        h = C.c_uint64()
        flags = 0 if self.winners else _lib.DH_CANVAS_NO_WINNER
        self._check(self.L.dh_canvas_create_ex(self.ctx, int(W), int(H), flags, C.byref(h)), "dh_canvas_create_ex")
        if not self.winners:
            self._nowin.add(h.value)
        return h.value

    def canvas(self, W: int, H: int) -> int:
        key = (W, H, self.winners)
        if key not in self._canvas:
            self._canvas[key] = self._new_canvas(W, H)
        return self._canvas[key]

    def mesh_triangles(self, handle: int, n_tri: int) -> np.ndarray:
        """A mesh's current triangle triplets, (n_tri, 9) uint32 (dh_mesh_read)."""
        tri = np.empty((n_tri, 9), np.uint32)
        self._check(self.L.dh_mesh_read(self.ctx, int(handle), tri.ctypes.data), "dh_mesh_read")
        return tri

    def sync_transparent_order(self, scene: Scene):
        """Copy the device order of every transparent mesh back into the host
        meshes: the reference sorts a transparent mesh's triangle vector in
        place each frame (scene/mod.rs:1100), and so does the device."""
        for o in scene.objects:
            handles = self.upload(o)
            for m, h in zip(o.meshes, handles):
                if o.material(m).alpha < 1.0:
                    m.triangles = self.mesh_triangles(h, m.triangles.shape[0])

    def mesh_handles(self, scene: Scene) -> List[int]:
        hs: List[int] = []
        for o in scene.objects:
            hs += self.upload(o)
        return hs

    # ------------------------------------------------------------ rendering
    def render_device(self, scene: Scene, frame: DhFrameParams | None = None, meshes: List[int] | None = None) -> int:
        """Render into the device canvas of the scene's size; returns its handle."""
        fr = frame if frame is not None else frame_uniforms(scene)
        hs = meshes if meshes is not None else self.mesh_handles(scene)
        arr = (C.c_uint64 * max(1, len(hs)))(*hs)
        cv = self.canvas(fr.W, fr.H)
        self._check(self.L.dh_render(self.ctx, C.byref(fr), arr, len(hs), cv), "dh_render")
        return cv

    def canvas_slot(self, slot, W: int, H: int) -> int:
        """A device canvas owned by one batch slot (views of a batch need
        distinct canvases even when their sizes agree)."""
        key = ("slot", slot, W, H, self.winners)
        if key not in self._canvas:
            self._canvas[key] = self._new_canvas(W, H)
        return self._canvas[key]

    def batch_views(self, items, slot_base: int = 0):
        """items: [(scene, frame or None)] -> (dh_view array, canvas handles)."""
        views = (_lib.DhView * max(1, len(items)))()
        canvases, keep = [], []
        for k, (scene, frame) in enumerate(items):
            fr = frame if frame is not None else frame_uniforms(scene)
            hs = self.mesh_handles(scene)
            arr = (C.c_uint64 * max(1, len(hs)))(*hs)
            keep.append(arr)
            cv = self.canvas_slot(slot_base + k, fr.W, fr.H)
            canvases.append(cv)
            views[k] = _lib.DhView(fr, cv, C.cast(arr, C.POINTER(C.c_uint64)), len(hs), 0)
        return views, canvases, keep

    def render_batch(self, items, graph: bool = False, prepared=None):
        """Batched Scene::render of [(scene, frame or None)] in one pipeline
        pass (dh_render_batch); returns the canvas handles, one per view.
        ``prepared`` = the result of batch_views(items) to skip re-marshalling."""
        views, canvases, keep = prepared if prepared is not None else self.batch_views(items)
        flags = _lib.DH_BATCH_GRAPH if graph else 0
        self._check(self.L.dh_render_batch(self.ctx, views, len(canvases), flags), "dh_render_batch")
        del keep
        return canvases

    def read(self, canvas_handle: int, W: int, H: int):
        """(bgra, depth, winner) of a canvas; winner is None for a canvas
        created without winners."""
        bgra = np.empty((H, W, 4), np.uint8)
        depth = np.empty((H, W), np.float32)
        winner = None if canvas_handle in self._nowin else np.empty((H, W), np.uint32)
        self._check(self.L.dh_canvas_read(self.ctx, canvas_handle, bgra.ctypes.data, depth.ctypes.data,
                                          None if winner is None else winner.ctypes.data), "dh_canvas_read")
        return bgra, depth, winner

    def read_into(self, canvas_handle: int, bgra: np.ndarray, depth: np.ndarray | None = None,
                  winner: np.ndarray | None = None) -> None:
        """dh_canvas_read into caller-owned contiguous arrays (any may be None),
        as Scene::render's caller reads the frame and depth into its Canvas."""
        def ptr(a, dtype):
            if a is None:
                return None
            assert a.dtype == dtype and a.flags["C_CONTIGUOUS"]
            return a.ctypes.data
        self._check(self.L.dh_canvas_read(self.ctx, canvas_handle, ptr(bgra, np.uint8), ptr(depth, np.float32),
                                          ptr(winner, np.uint32)), "dh_canvas_read")

    def draw_gui(self, canvas_handle: int, vtx: np.ndarray, idx: np.ndarray, cmds: np.ndarray, font) -> None:
        """Gui::render onto a device canvas (dh_draw_gui): vtx/cmds in the
        draw_amd.gui record layouts, idx u32, font a TextureMap (RGBA)."""
        from .gui import GUI_CMD, GUI_VERTEX
        vtx = np.ascontiguousarray(vtx, GUI_VERTEX)
        idx = np.ascontiguousarray(idx, np.uint32)
        cmds = np.ascontiguousarray(cmds, GUI_CMD)
        self._check(self.L.dh_draw_gui(self.ctx, int(canvas_handle), vtx.ctypes.data, len(vtx), idx.ctypes.data,
                                       len(idx), cmds.ctypes.data, len(cmds), self.texture(font)), "dh_draw_gui")

    def read_rgba(self, canvas_handle: int, W: int, H: int) -> np.ndarray:
        """The frame as App::export_frame_as prepares it: RGBA8, rows top-down."""
        rgba = np.empty((H, W, 4), np.uint8)
        self._check(self.L.dh_canvas_read_rgba(self.ctx, canvas_handle, rgba.ctypes.data), "dh_canvas_read_rgba")
        return rgba

    def read_rgba_noalloc(self, canvas_handle: int, W: int, H: int) -> np.ndarray:
        """read_rgba through dh_canvas_read_rgba_noalloc (fixed device staging,
        no allocation; canvases up to 3840x2160)."""
        rgba = np.empty((H, W, 4), np.uint8)
        self._check(self.L.dh_canvas_read_rgba_noalloc(self.ctx, canvas_handle, rgba.ctypes.data),
                    "dh_canvas_read_rgba_noalloc")
        return rgba

    def export_png(self, canvas_handle: int, path: str):
        self._check(self.L.dh_canvas_export_png(self.ctx, canvas_handle, str(path).encode()), "dh_canvas_export_png")

    def export_jpg(self, canvas_handle: int, path: str, quality: int = 100):
        """dh_canvas_export_jpg: the RGBA frame as a baseline JPEG (quality 100
        is what the reference's export call amounts to)."""
        self._check(self.L.dh_canvas_export_jpg(self.ctx, canvas_handle, str(path).encode(), int(quality)),
                    "dh_canvas_export_jpg")

    def render(self, scene: Scene, canvas: Canvas | None = None, frame: DhFrameParams | None = None):
        """Scene::render(&mut canvas): returns (canvas, winner ids)."""
        fr = frame if frame is not None else frame_uniforms(scene)
        cv = self.render_device(scene, fr)
        if canvas is None:
            canvas = Canvas(fr.W, fr.H)
        bgra, depth, winner = self.read(cv, fr.W, fr.H)
        canvas.frame[...] = bgra
        canvas.depth[...] = depth
        return canvas, winner

    def stats(self) -> dict:
        st = DhStats()
        self._check(self.L.dh_get_stats(self.ctx, C.byref(st)), "dh_get_stats")
        return {n: getattr(st, n) for n, _ in DhStats._fields_}

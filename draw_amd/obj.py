"""OBJ/MTL ingest: Object::load_from_file (object.rs:106-419) through the
native loader of libdrawhip (dh_obj_*), returned as the host scene model so
the same object can be uploaded (Renderer.upload) and checked."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DhObjMeshDesc, DhObjSummary, DhObjTextureDesc, DrawHipError
from .scene import IndexedMesh, Object, Texture, TextureMap


def _arr(ptr, n, dtype, shape):
    if n == 0:
        return np.zeros(shape, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True).reshape(shape)


def load_obj(path: str) -> Object:
    """Parse ``path`` (and its MTL libraries and PNG textures) the way the
    reference's Object::load_from_file does; raises DrawHipError on failure."""
    L = _lib.lib()
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    rc = L.dh_obj_load(str(path).encode(), C.byref(h), err, 512)
    if rc != 0:
        raise DrawHipError(rc, "dh_obj_load", err.value.decode(errors="replace"))
    try:
        info = DhObjSummary()
        L.dh_obj_info(h, C.byref(info))
        pos = _arr(info.pos, 3 * info.n_pos, np.float32, (info.n_pos, 3))
        nrm = _arr(info.nrm, 3 * info.n_nrm, np.float32, (info.n_nrm, 3))
        uv = _arr(info.uv, 2 * info.n_uv, np.float32, (info.n_uv, 2))
        textures = []
        for i in range(info.n_textures):
            t = DhObjTextureDesc()
            L.dh_obj_texture(h, i, C.byref(t))
            maps = []
            for ptr, w, hh, c in ((t.map_kd, t.kd_w, t.kd_h, t.kd_comp), (t.map_ka, t.ka_w, t.ka_h, t.ka_comp)):
                if ptr:
                    maps.append(TextureMap(_arr(ptr, w * hh * c, np.uint8, (hh, w, c))))
                else:
                    maps.append(TextureMap.default())
            textures.append(Texture(ka=tuple(t.ka), kd=tuple(t.kd), ks=tuple(t.ks), alpha=float(t.alpha),
                                    map_kd=maps[0], map_ka=maps[1], name=t.name.decode()))
        meshes = []
        for i in range(info.n_meshes):
            m = DhObjMeshDesc()
            L.dh_obj_mesh(h, i, C.byref(m))
            meshes.append(IndexedMesh(_arr(m.tri, 9 * m.n_tri, np.uint32, (m.n_tri, 9)), int(m.texture_idx),
                                      m.name.decode()))
        name = str(path).rsplit("/", 1)[-1]
        return Object(pos, nrm, uv, meshes, textures, name=name)
    finally:
        L.dh_obj_free(h)


def upload_obj(ctx, path: str):
    """dh_obj_upload: parse and upload in one call; returns (object handle,
    mesh handles)."""
    L = _lib.lib()
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    rc = L.dh_obj_load(str(path).encode(), C.byref(h), err, 512)
    if rc != 0:
        raise DrawHipError(rc, "dh_obj_load", err.value.decode(errors="replace"))
    try:
        info = DhObjSummary()
        L.dh_obj_info(h, C.byref(info))
        obj = C.c_uint64()
        meshes = (C.c_uint64 * max(1, info.n_meshes))()
        rc = L.dh_obj_upload(ctx, h, C.byref(obj), meshes, max(1, info.n_meshes))
        if rc != 0:
            raise DrawHipError(rc, "dh_obj_upload")
        return obj.value, [meshes[i] for i in range(info.n_meshes)]
    finally:
        L.dh_obj_free(h)

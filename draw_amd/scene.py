"""Host-side scene model mirroring the reference renderer's data types.

The names and meanings follow mororo18/draw:

* ``TextureMap``  -- an RGB(A)8 image (scene/mod.rs:102-126), row 0 = top row.
* ``Texture``     -- a material: ka, kd, ks, alpha, map_ka, map_kd
                     (scene/mod.rs:206-253).  ``power`` is the specular exponent
                     that canvas.rs:724 hard-codes to 2.
* ``IndexedMesh`` -- triangles as (pos, uv, normal) index triplets plus a
                     texture index (scene/mesh.rs:30-35).
* ``Object``      -- vertex, normal and texture-coordinate arrays shared by its
                     meshes (scene/object.rs:18-31).
* ``Camera``      -- position, direction, fov_x, near, far (scene/mod.rs:282-357);
                     the reference hard-codes near=-10, far=-510, fov_x=135.
* ``Scene``       -- width, height, camera, light and objects (scene/mod.rs:749-786).
* ``Canvas``      -- BGRA8 frame (top-down rows) + f32 depth (bottom-up rows)
                     (canvas.rs:353-433).

Arrays are numpy; this module does no rendering itself (``draw_amd.render`` does,
on the GPU).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

CLEAR_BGRA = (255, 186, 155, 255)   # Pixel::azul_bb, canvas.rs:131-133 (B,G,R,A)
CLEAR_DEPTH = 100000.0              # Canvas::init_depth(100000.0), app/mod.rs:80


@dataclass
class TextureMap:
    """scene/mod.rs:102-126: ``img`` is (height, width, components) uint8."""
    img: np.ndarray

    def __post_init__(self):
        img = np.ascontiguousarray(self.img, dtype=np.uint8)
        if img.ndim != 3 or img.shape[2] not in (3, 4) or img.shape[0] < 1 or img.shape[1] < 1:
            raise ValueError("TextureMap expects (h, w, 3|4) uint8 with h, w >= 1")
        self.img = img

    @property
    def width(self) -> int:
        return int(self.img.shape[1])

    @property
    def height(self) -> int:
        return int(self.img.shape[0])

    @property
    def components(self) -> int:
        return int(self.img.shape[2])

    @staticmethod
    def default() -> "TextureMap":
        """TextureMap::default: a 1x1 white RGB map (scene/mod.rs:128-135)."""
        return TextureMap(np.full((1, 1, 3), 255, np.uint8))

    @staticmethod
    def load_from_file(path) -> "TextureMap":
        """TextureMap::load_from_file (scene/mod.rs:174-202): a PNG or JPEG
        file decoded by libdrawhip (dh_image_info, then dh_image_load)."""
        import ctypes as C

        from . import _lib
        L = _lib.lib()
        w, h, c = C.c_int32(), C.c_int32(), C.c_int32()
        err = C.create_string_buffer(512)
        rc = L.dh_image_info(str(path).encode(), C.byref(w), C.byref(h), C.byref(c), err, 512)
        if rc != 0:
            raise _lib.DrawHipError(rc, "dh_image_info", err.value.decode(errors="replace"))
        img = np.empty((h.value, w.value, c.value), np.uint8)
        rc = L.dh_image_load(str(path).encode(), img.ctypes.data, img.nbytes, C.byref(w), C.byref(h), C.byref(c),
                             err, 512)
        if rc != 0:
            raise _lib.DrawHipError(rc, "dh_image_load", err.value.decode(errors="replace"))
        return TextureMap(img)


@dataclass
class Texture:
    """Material, scene/mod.rs:206-216."""
    ka: Sequence[float] = (0.9, 0.9, 0.9)
    kd: Sequence[float] = (0.4, 0.4, 0.4)
    ks: Sequence[float] = (0.5, 0.5, 0.5)
    alpha: float = 1.0
    map_ka: TextureMap = field(default_factory=TextureMap.default)
    map_kd: TextureMap = field(default_factory=TextureMap.default)
    power: int = 2
    name: str = "default"

    @staticmethod
    def default() -> "Texture":
        """Texture::default, scene/mod.rs:237-252."""
        return Texture()

    @staticmethod
    def with_diffuse_map(diff_map: TextureMap) -> "Texture":
        """Texture::with_diffuse_map, scene/mod.rs:219-234."""
        return Texture(ka=(1.0, 1.0, 1.0), kd=(1.0, 1.0, 1.0), ks=(1.0, 1.0, 1.0), alpha=1.0,
                       map_ka=TextureMap.default(), map_kd=diff_map)


@dataclass
class IndexedMesh:
    """scene/mesh.rs:30-35.  ``triangles``: (T, 9) uint32 =
    (pos a, b, c, uv a, b, c, normal a, b, c); ``texture_idx`` indexes the
    owning object's ``textures`` (a missing one renders with Texture::default,
    scene/mod.rs:1072-1075)."""
    triangles: np.ndarray
    texture_idx: int = 0
    name: str = "mesh"

    def __post_init__(self):
        t = np.ascontiguousarray(self.triangles, dtype=np.uint32)
        if t.ndim != 2 or t.shape[1] != 9:
            raise ValueError("IndexedMesh.triangles must be (T, 9)")
        self.triangles = t


@dataclass
class Object:
    """scene/object.rs:18-31.  Vertex arrays are float32: vertices (N, 3),
    normals (Nn, 3), texture_vertices (Nt, 2)."""
    vertices: np.ndarray
    normals_vertices: np.ndarray
    texture_vertices: np.ndarray
    meshes: List[IndexedMesh]
    textures: List[Texture]
    name: str = "object"

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float32).reshape(-1, 3)
        self.normals_vertices = np.ascontiguousarray(self.normals_vertices, dtype=np.float32).reshape(-1, 3)
        tv = np.ascontiguousarray(self.texture_vertices, dtype=np.float32)
        if tv.ndim == 2 and tv.shape[1] == 3:   # the reference stores UVs as Vec3; z is never read
            tv = np.ascontiguousarray(tv[:, :2])
        self.texture_vertices = tv.reshape(-1, 2)

    @property
    def triangle_count(self) -> int:
        return int(sum(m.triangles.shape[0] for m in self.meshes))

    def material(self, mesh: IndexedMesh) -> Texture:
        if 0 <= mesh.texture_idx < len(self.textures):
            return self.textures[mesh.texture_idx]
        return Texture.default()

    def opaque_meshes(self) -> List[IndexedMesh]:
        """Object::new's split: alpha < 1 is transparent (object.rs:45-53)."""
        return [m for m in self.meshes if not (self.material(m).alpha < 1.0)]

    def transparent_meshes(self) -> List[IndexedMesh]:
        return [m for m in self.meshes if self.material(m).alpha < 1.0]


@dataclass
class Camera:
    """scene/mod.rs:282-357 with near, far and fov_x as parameters (the
    reference fixes them at -10, -510 and 135 degrees).  ``near`` and ``far``
    use the reference's sign convention: far < near < 0."""
    position: Sequence[float] = (0.0, 0.0, 150.0)
    direction: Optional[Sequence[float]] = None
    fov_x: float = 135.0
    near: float = -10.0
    far: float = -510.0

    def __post_init__(self):
        if self.direction is None:
            # Scene::new: camera_dir = camera_pos * -1.0 (scene/mod.rs:764)
            self.direction = tuple(float(np.float32(p) * np.float32(-1.0)) for p in self.position)


@dataclass
class Scene:
    """scene/mod.rs:749-786: the reference's defaults are camera (0,0,150) and
    light (0,300,300).  ``width`` x ``height`` is the scene (screen) size.
    ``canvas`` = (W, H) draws into a canvas of that size moved by
    ``canvas_offset`` (Canvas::new + Canvas::apply_offset, as the reference
    app does: an 800x600 canvas centred on the screen, app/mod.rs:59-63,
    81-84); None draws into a canvas of the scene's size at offset 0."""
    width: int
    height: int
    camera: Camera = field(default_factory=Camera)
    light_source: Sequence[float] = (0.0, 300.0, 300.0)
    objects: List[Object] = field(default_factory=list)
    canvas: Optional[Tuple[int, int]] = None
    canvas_offset: Tuple[int, int] = (0, 0)

    @property
    def canvas_size(self) -> Tuple[int, int]:
        return (int(self.width), int(self.height)) if self.canvas is None else (int(self.canvas[0]),
                                                                                int(self.canvas[1]))

    def add_obj(self, obj: Object) -> int:
        """Scene::add_obj (scene/mod.rs:788-801); returns the object id."""
        self.objects.append(obj)
        return len(self.objects) - 1

    @property
    def triangle_count(self) -> int:
        return int(sum(o.triangle_count for o in self.objects))


class Canvas:
    """canvas.rs:353-433: ``frame`` is (H, W, 4) BGRA with row 0 the TOP row
    (draw_pixel_coord flips y, canvas.rs:955); ``depth`` is (H, W) float32 with
    row 0 the BOTTOM row (canvas.rs:415)."""

    def __init__(self, width: int, height: int, depth_max: float = CLEAR_DEPTH):
        self.width = int(width)
        self.height = int(height)
        self.depth_max = float(depth_max)
        self.frame = np.empty((self.height, self.width, 4), np.uint8)
        self.depth = np.empty((self.height, self.width), np.float32)
        self.clear()

    def clear(self):
        """Canvas::clear, canvas.rs:425-433."""
        self.frame[...] = np.array(CLEAR_BGRA, np.uint8)
        self.depth[...] = np.float32(self.depth_max)

    def as_bytes(self) -> bytes:
        """Canvas::as_bytes_slice (canvas.rs:974-978)."""
        return self.frame.tobytes()

"""Seeded synthetic workloads for the per-frame raster path.

Every generator is deterministic in its seed (numpy PCG64) and returns a
``Scene`` of the reference data model (draw_amd.scene).  Geometry is built in
float64 and stored as float32; the renderer only ever sees the float32 arrays,
so the GPU path and the CPU oracle render identical inputs.

Workloads (sizes as named):

* ``torus_scene``      640x480, torus R=1.0 r=0.4 with 48x24 segments (2304
                       triangles), per-vertex normals, UV seam vertices,
                       256x256 RGBA8 checker with a seeded noise tint.
* ``soup_scene``       1920x1080, 1,000,000-triangle soup: centroids uniform in
                       the view frustum, screen areas log-normal (median 40 px,
                       sigma 1.5), 20% of triangles straddling the near plane,
                       random UVs, 2048x2048 RGBA8 noise texture.
* ``shells_scene``     3840x2160, ~20M micro-triangles on 8 nested jittered
                       sphere shells, 5% flipped winding, 5% outside the
                       frustum, 8192x8192 texture.
* ``mesh_pool``        tori, UV spheres and soups with log-uniform triangle
                       counts, used by the batched and serving modes.

Camera convention (the reference's): camera at (0,0,d) looking at the origin;
``near``/``far`` negative with far < near (scene/mod.rs:308-310, 328-329).
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from .scene import Camera, IndexedMesh, Object, Scene, Texture, TextureMap

F32 = np.float32


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(int(seed)))


# ----------------------------------------------------------------- textures
def checker_texture(size: int, square: int, rng: np.random.Generator) -> TextureMap:
    """RGBA8 checker (``square``-pixel squares) times a uniform noise tint."""
    y, x = np.mgrid[0:size, 0:size]
    base = np.where(((x // square) + (y // square)) % 2 == 0, 255.0, 64.0)
    tint = rng.random((size, size, 3))
    rgb = np.floor(base[..., None] * (0.5 + 0.5 * tint)).astype(np.uint8)
    img = np.empty((size, size, 4), np.uint8)
    img[..., :3] = rgb
    img[..., 3] = 255
    return TextureMap(img)


def noise_texture(size: int, rng: np.random.Generator) -> TextureMap:
    img = rng.integers(0, 256, size=(size, size, 4), dtype=np.uint8)
    img[..., 3] = 255
    return TextureMap(img)


def _phong(tex: TextureMap, power: int = 2) -> Texture:
    """ka=0.1, kd=0.7, ks=0.2 with a diffuse map and the default ambient map."""
    return Texture(ka=(0.1, 0.1, 0.1), kd=(0.7, 0.7, 0.7), ks=(0.2, 0.2, 0.2), alpha=1.0,
                   map_ka=TextureMap.default(), map_kd=tex, power=int(power))


# ------------------------------------------------------------------- meshes
def _rotation(rng: np.random.Generator) -> np.ndarray:
    a, b, c = rng.uniform(0.0, 2.0 * math.pi, 3)
    rx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    ry = np.array([[math.cos(b), 0, math.sin(b)], [0, 1, 0], [-math.sin(b), 0, math.cos(b)]])
    rz = np.array([[math.cos(c), -math.sin(c), 0], [math.sin(c), math.cos(c), 0], [0, 0, 1]])
    return rz @ ry @ rx


def torus_object(R: float, r: float, nu: int, nv: int, texture: Texture,
                 rot: np.ndarray | None = None, offset=(0.0, 0.0, 0.0)) -> Object:
    """Torus around the z axis: ``nu`` segments on the major circle, ``nv`` on
    the minor one, 2*nu*nv triangles.  Positions/normals are periodic
    (nu*nv each); UVs carry seam duplicates ((nu+1)*(nv+1)), so the three index
    triplets of a triangle differ as in OBJ files."""
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    th = 2.0 * math.pi * i / nu
    ph = 2.0 * math.pi * j / nv
    px = (R + r * np.cos(ph)) * np.cos(th)
    py = (R + r * np.cos(ph)) * np.sin(th)
    pz = r * np.sin(ph) * np.ones_like(th)
    pos = np.stack([px, py, pz], -1).reshape(-1, 3)
    nrm = np.stack([np.cos(ph) * np.cos(th), np.cos(ph) * np.sin(th), np.sin(ph) * np.ones_like(th)], -1).reshape(-1, 3)
    iu, iv = np.meshgrid(np.arange(nu + 1), np.arange(nv + 1), indexing="ij")
    uv = np.stack([iu / nu, iv / nv], -1).reshape(-1, 2)
    if rot is not None:
        pos = pos @ rot.T
        nrm = nrm @ rot.T
    pos = pos + np.asarray(offset, np.float64)

    def pidx(a, b):
        return (a % nu) * nv + (b % nv)

    def tidx(a, b):
        return a * (nv + 1) + b

    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    p00, p10, p11, p01 = pidx(ii, jj), pidx(ii + 1, jj), pidx(ii + 1, jj + 1), pidx(ii, jj + 1)
    t00, t10, t11, t01 = tidx(ii, jj), tidx(ii + 1, jj), tidx(ii + 1, jj + 1), tidx(ii, jj + 1)
    # Orientation: (b-a) x (c-b) must point along the outward normal (the
    # reference culls when (cam - center) . n <= 0, scene/mod.rs:1020).
    a, b, c = pos[p00[0]], pos[p10[0]], pos[p11[0]]
    outward = np.cross(b - a, c - b) @ nrm[p00[0]] > 0
    if outward:
        t1 = (p00, p10, p11, t00, t10, t11)
        t2 = (p00, p11, p01, t00, t11, t01)
    else:
        t1 = (p00, p11, p10, t00, t11, t10)
        t2 = (p00, p01, p11, t00, t01, t11)
    tris = []
    for (pa, pb, pc, ta, tb, tc) in (t1, t2):
        tris.append(np.stack([pa, pb, pc, ta, tb, tc, pa, pb, pc], -1))
    tri = np.stack(tris, 1).reshape(-1, 9).astype(np.uint32)
    return Object(pos.astype(F32), nrm.astype(F32), uv.astype(F32), [IndexedMesh(tri, 0, "torus")],
                  [texture], name="torus")


def uv_sphere_object(radius: float, n_tri: int, texture: Texture, rng: np.random.Generator,
                     jitter: float = 0.0, center=(0.0, 0.0, 0.0)) -> Object:
    """UV sphere with about ``n_tri`` triangles (nu*nv*2), optional per-vertex
    radial jitter; outward winding."""
    nv = max(3, int(round(math.sqrt(n_tri / 4.0))))
    nu = max(3, int(round(n_tri / (2.0 * nv))))
    i = np.arange(nu)[:, None]
    j = np.arange(nv + 1)[None, :]
    th = 2.0 * math.pi * i / nu
    ph = math.pi * j / nv - math.pi / 2
    nx = np.cos(ph) * np.cos(th)
    ny = np.sin(ph) * np.ones_like(th)
    nz = np.cos(ph) * np.sin(th)
    nrm = np.stack([nx, ny, nz], -1).reshape(-1, 3)
    rr = radius * (1.0 + (jitter * rng.standard_normal(nrm.shape[0]) if jitter > 0 else 0.0))
    pos = nrm * np.asarray(rr).reshape(-1, 1) + np.asarray(center, np.float64)
    iu, iv = np.meshgrid(np.arange(nu + 1), np.arange(nv + 1), indexing="ij")
    uv = np.stack([iu / nu, iv / nv], -1).reshape(-1, 2)

    def pidx(a, b):
        return (a % nu) * (nv + 1) + b

    def tidx(a, b):
        return a * (nv + 1) + b

    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    p00, p10, p11, p01 = pidx(ii, jj), pidx(ii + 1, jj), pidx(ii + 1, jj + 1), pidx(ii, jj + 1)
    t00, t10, t11, t01 = tidx(ii, jj), tidx(ii + 1, jj), tidx(ii + 1, jj + 1), tidx(ii, jj + 1)
    # outward orientation test on a mid-latitude quad
    k = (nv // 2)
    a, b, c = pos[p00[k]], pos[p10[k]], pos[p11[k]]
    outward = np.cross(b - a, c - b) @ nrm[p00[k]] > 0
    if outward:
        t1 = (p00, p10, p11, t00, t10, t11)
        t2 = (p00, p11, p01, t00, t11, t01)
    else:
        t1 = (p00, p11, p10, t00, t11, t10)
        t2 = (p00, p01, p11, t00, t01, t11)
    tri = np.stack([np.stack([pa, pb, pc, ta, tb, tc, pa, pb, pc], -1) for (pa, pb, pc, ta, tb, tc) in (t1, t2)],
                   1).reshape(-1, 9).astype(np.uint32)
    return Object(pos.astype(F32), nrm.astype(F32), uv.astype(F32), [IndexedMesh(tri, 0, "sphere")],
                  [texture], name="uv_sphere")


def _frame_basis(n: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Orthonormal e1, e2 perpendicular to the unit vectors n (K, 3)."""
    helper = np.where(np.abs(n[:, :1]) < 0.9, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    e1 = np.cross(n, helper)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(n, e1)
    return e1, e2


def soup_object(n: int, W: int, H: int, fov_x_deg: float, near: float, far: float, cam_z: float,
                texture: Texture, rng: np.random.Generator, cross_frac: float = 0.2,
                median_px: float = 40.0, sigma: float = 1.5) -> Object:
    """Unindexed triangle soup in front of a camera at (0,0,cam_z) looking down
    -z.  ``near``/``far`` are positive distances here.  Screen areas follow
    log-normal(ln median_px, sigma); a ``cross_frac`` share straddles the near
    plane with a clipped screen area drawn from the same law."""
    tanx = math.tan(math.radians(fov_x_deg) / 2.0)
    tany = tanx * H / W
    f_px = (W / 2.0) / tanx
    n_cross = int(round(cross_frac * n))
    area = np.exp(math.log(median_px) + sigma * rng.standard_normal(n))
    area = np.minimum(area, 0.5 * W * H)
    # distances: volume-uniform in [near, far] (pdf ~ d^2), kept off the near plane
    d_lo = max(2.0 * near, 0.2)
    u = rng.random(n)
    d = np.cbrt(u * (far ** 3 - d_lo ** 3) + d_lo ** 3)
    d[:n_cross] = near
    xc = (2.0 * rng.random(n) - 1.0) * d * tanx * 0.98
    yc = (2.0 * rng.random(n) - 1.0) * d * tany * 0.98
    # random triangle shapes in 2D, normalised to unit area
    q = rng.standard_normal((n, 3, 2))
    q -= q.mean(axis=1, keepdims=True)
    a2 = np.abs((q[:, 1, 0] - q[:, 0, 0]) * (q[:, 2, 1] - q[:, 0, 1]) -
                (q[:, 2, 0] - q[:, 0, 0]) * (q[:, 1, 1] - q[:, 0, 1])) * 0.5
    q /= np.sqrt(np.maximum(a2, 1e-6))[:, None, None]
    # orientation: plane normal within 60 degrees of the view axis
    tilt = np.arccos(1.0 - rng.random(n) * 0.5)
    azim = rng.uniform(0.0, 2.0 * math.pi, n)
    nrm_t = np.stack([np.sin(tilt) * np.cos(azim), np.sin(tilt) * np.sin(azim), np.cos(tilt)], -1)
    e1, e2 = _frame_basis(nrm_t)
    side = np.sqrt(area / np.cos(tilt)) * d / f_px
    off = (q[..., 0:1] * e1[:, None, :] + q[..., 1:2] * e2[:, None, :]) * side[:, None, None]
    center = np.stack([xc, yc, -d], -1)
    pc = center[:, None, :] + off
    # near-plane straddlers: vertices at depths near*(1 -+ s) around the plane
    if n_cross:
        s = side[:n_cross] / near
        zsh = np.stack([-(1.0 - 0.5 * s), -(1.0 + 0.5 * s), -(1.0 + (rng.random(n_cross) - 0.5) * s)], -1) * near
        pc[:n_cross, :, 2] = zsh
    # random winding: about half the soup faces away and is culled
    flip = rng.random(n) < 0.5
    pc[flip] = pc[flip][:, ::-1, :]
    # per-vertex normals: the face normal towards the camera plus noise
    fn = np.cross(pc[:, 1] - pc[:, 0], pc[:, 2] - pc[:, 1])
    fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-30)
    vn = fn[:, None, :] + 0.3 * rng.standard_normal((n, 3, 3))
    vn /= np.maximum(np.linalg.norm(vn, axis=2, keepdims=True), 1e-30)
    pos = (pc + np.array([0.0, 0.0, cam_z])).reshape(-1, 3)
    uv = rng.random((n * 3, 2))
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    tri = np.concatenate([idx, idx, idx], 1)
    # shuffle so straddlers are spread through the submission order
    perm = rng.permutation(n)
    tri = tri[perm]
    return Object(pos.astype(F32), vn.reshape(-1, 3).astype(F32), uv.astype(F32),
                  [IndexedMesh(tri, 0, "soup")], [texture], name="soup")


# ---------------------------------------------------------------- workloads
def torus_scene(seed: int = 1, width: int = 640, height: int = 480, power: int = 2) -> Scene:
    """Torus R=1.0, r=0.4, 48x24 segments (2304 triangles), 256x256 checker,
    camera at distance 3 (fov_x 60, near 0.1, far 100), light (0, 6, 6)."""
    rng = _rng(seed)
    rot = _rotation(rng)
    tex = _phong(checker_texture(256, 8, rng), power)
    obj = torus_object(1.0, 0.4, 48, 24, tex, rot)
    cam = Camera(position=(0.0, 0.0, 3.0), fov_x=60.0, near=-0.1, far=-100.0)
    return Scene(width, height, cam, (0.0, 6.0, 6.0), [obj])


def soup_scene(seed: int = 2, n_tri: int = 1_000_000, width: int = 1920, height: int = 1080,
               tex_size: int = 2048, power: int = 2, cross_frac: float = 0.2) -> Scene:
    """Seeded triangle soup: 1M triangles at 1920x1080 by default."""
    rng = _rng(seed)
    tex = _phong(noise_texture(tex_size, rng), power)
    obj = soup_object(n_tri, width, height, 60.0, 0.1, 100.0, 3.0, tex, rng, cross_frac=cross_frac)
    cam = Camera(position=(0.0, 0.0, 3.0), fov_x=60.0, near=-0.1, far=-100.0)
    return Scene(width, height, cam, (0.0, 6.0, 6.0), [obj])


def shells_scene(seed: int = 3, n_tri: int = 20_000_000, width: int = 3840, height: int = 2160,
                 tex_size: int = 8192, n_shells: int = 8, power: int = 2) -> Scene:
    """Micro-triangle stress: ``n_shells`` nested jittered UV-sphere shells
    (sigma 1e-3 radial jitter) totalling ``n_tri`` triangles, sized so the mean
    projected area of a front-facing triangle is about 0.5 px; 5% of triangles
    have flipped winding and 5% are moved outside the frustum."""
    rng = _rng(seed)
    tex = _phong(noise_texture(tex_size, rng), power)
    cam_d = 3.0
    fov = 60.0
    f_px = (width / 2.0) / math.tan(math.radians(fov) / 2.0)
    per = n_tri // n_shells
    # projected area of the front hemisphere ~ pi*R_px^2 shared by per/2 triangles
    r_px = math.sqrt(0.5 * (per / 2.0) / math.pi)
    r_out = min(r_px * cam_d / f_px, 1.2)
    objs: List[Object] = []
    for s in range(n_shells):
        rad = r_out * (1.0 - 0.06 * s)
        o = uv_sphere_object(rad, per, tex, rng, jitter=1e-3)
        tri = o.meshes[0].triangles
        t = tri.shape[0]
        flip = rng.random(t) < 0.05
        tri[flip] = tri[flip][:, [0, 2, 1, 3, 5, 4, 6, 8, 7]]
        out = rng.random(t) < 0.05
        if out.any():
            # move the chosen triangles' own copies of their vertices behind the camera
            k = int(out.sum())
            base = o.vertices.shape[0]
            verts = o.vertices[tri[out][:, :3].ravel()].astype(np.float64)
            verts[:, 2] += 2.0 * cam_d + 1.0
            o.vertices = np.concatenate([o.vertices, verts.astype(F32)])
            tri[out, 0:3] = (base + np.arange(3 * k)).reshape(k, 3)
        o.meshes[0].triangles = tri
        objs.append(o)
    cam = Camera(position=(0.0, 0.0, cam_d), fov_x=fov, near=-0.1, far=-100.0)
    return Scene(width, height, cam, (0.0, 2.0 * cam_d, 2.0 * cam_d), objs)


def app_scene(seed: int = 6, screen=(1920, 1080), canvas=(800, 600), n_tri: int = 60_000,
              cross_frac: float = 0.1, tex_size: int = 512, median_px: float = 40.0, sigma: float = 1.5) -> Scene:
    """The reference application's own configuration (app/mod.rs:57-84,
    scene/mod.rs:760-786, 308-310): scene at screen size, camera at
    (0, 0, 150) looking at the origin with near -10, far -510 and fov_x 135,
    light (0, 300, 300), drawn into an 800x600 canvas centred on the screen
    (offset ((screen - canvas) / 2)).  Objects: a torus rescaled to a largest
    vertex norm of 100 (as Object::load_from_file rescales, object.rs:159-170)
    and a seeded soup filling the screen's view frustum, so many triangles
    cross the canvas border or lie wholly outside it."""
    rng = _rng(seed)
    tex = _phong(noise_texture(tex_size, rng))
    torus = torus_object(1.0, 0.4, 48, 24, _phong(checker_texture(256, 8, rng)), _rotation(rng))
    v = torus.vertices.astype(np.float64)
    torus.vertices = (v * (100.0 / np.linalg.norm(v, axis=1).max())).astype(F32)
    soup = soup_object(n_tri, screen[0], screen[1], 135.0, 10.0, 510.0, 150.0, tex, rng, cross_frac=cross_frac,
                       median_px=median_px, sigma=sigma)
    cam = Camera(position=(0.0, 0.0, 150.0))
    ox, oy = (screen[0] - canvas[0]) // 2, (screen[1] - canvas[1]) // 2
    return Scene(screen[0], screen[1], cam, (0.0, 300.0, 300.0), [torus, soup], canvas=tuple(canvas),
                 canvas_offset=(ox, oy))


def sliver_scene(seed: int = 23, n_tri: int = 16_000, width: int = 160, height: int = 120) -> Scene:
    """Slivers that cross both clip planes: one vertex between the camera and
    the near plane, two beyond the far plane, so each front-facing one clips
    into about three triangles (half are back-facing).  16000 of them emit
    about 24000 clipped triangles, more than a fresh context's first capacity
    (n + n/4 + 1024)."""
    rng = _rng(seed)
    tex = _phong(noise_texture(64, rng))
    n = n_tri
    a = np.stack([rng.uniform(-0.01, 0.01, n), rng.uniform(-0.01, 0.01, n), np.full(n, 3.0 - 0.05)], -1)
    b = np.stack([rng.uniform(-60.0, 60.0, n), rng.uniform(-45.0, 45.0, n), np.full(n, 3.0 - 150.0)], -1)
    c = b + np.stack([rng.uniform(-8.0, 8.0, n), rng.uniform(-8.0, 8.0, n), rng.uniform(-20.0, 20.0, n)], -1)
    pc = np.stack([a, b, c], 1)
    flip = rng.random(n) < 0.5
    pc[flip] = pc[flip][:, ::-1, :]
    nrm = np.tile(np.array([0.0, 0.0, 1.0], F32), (3 * n, 1))
    uv = rng.random((3 * n, 2))
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    obj = Object(pc.reshape(-1, 3).astype(F32), nrm, uv.astype(F32),
                 [IndexedMesh(np.concatenate([idx, idx, idx], 1), 0, "slivers")], [tex], name="slivers")
    cam = Camera(position=(0.0, 0.0, 3.0), fov_x=60.0, near=-0.1, far=-100.0)
    return Scene(width, height, cam, (0.0, 6.0, 6.0), [obj])


def small_random_scene(seed: int, n_tri: int = 300, width: int = 96, height: int = 64,
                       cross_frac: float = 0.2, tex_size: int = 16, power: int = 2) -> Scene:
    """Tiny soup for exhaustive parity checks."""
    rng = _rng(seed)
    tex = _phong(noise_texture(tex_size, rng), power)
    obj = soup_object(n_tri, width, height, 60.0, 0.1, 100.0, 3.0, tex, rng, cross_frac=cross_frac,
                      median_px=30.0)
    cam = Camera(position=(0.0, 0.0, 3.0), fov_x=60.0, near=-0.1, far=-100.0)
    return Scene(width, height, cam, (0.0, 6.0, 6.0), [obj])


RESOLUTIONS = [(256, 256), (640, 480), (1280, 720), (1920, 1080)]


def random_camera(rng: np.random.Generator, target=(0.0, 0.0, 0.0), dist=(2.6, 4.0)) -> Camera:
    """A camera on a random direction around ``target`` looking at it (kept off
    the +-y poles, where the reference's up-vector basis degenerates)."""
    while True:
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        if abs(v[1]) < 0.9:
            break
    dd = rng.uniform(*dist)
    pos = np.asarray(target, np.float64) + v * dd
    direction = np.asarray(target, np.float64) - pos
    return Camera(position=tuple(float(x) for x in pos.astype(F32)),
                  direction=tuple(float(x) for x in direction.astype(F32)),
                  fov_x=60.0, near=-0.1, far=-100.0)


def mesh_pool(seed: int, n_meshes: int, tri_lo: int, tri_hi: int, tex_size: int = 1024) -> List[Object]:
    """Meshes with log-uniform triangle counts in [tri_lo, tri_hi], cycling
    torus / UV sphere / soup, each with its own ``tex_size``^2 RGBA8 texture.
    Every mesh fits a radius-1.4 ball around the origin."""
    rng = _rng(seed)
    out: List[Object] = []
    for m in range(n_meshes):
        nt = int(round(math.exp(rng.uniform(math.log(tri_lo), math.log(tri_hi)))))
        tex = _phong(noise_texture(tex_size, rng) if m % 2 else checker_texture(tex_size, 16, rng))
        kind = m % 3
        if kind == 0:
            nv = max(3, int(round(math.sqrt(nt / 4.0))))
            nu = max(3, int(round(nt / (2.0 * nv))))
            out.append(torus_object(1.0, 0.4, nu, nv, tex, _rotation(rng)))
        elif kind == 1:
            out.append(uv_sphere_object(1.2, nt, tex, rng))
        else:
            out.append(_ball_soup(nt, 1.2, tex, rng))
    return out


def batched_scenes(seed: int = 4, n_meshes: int = 8, cams_per_mesh: int = 32, tri_lo: int = 1000,
                   tri_hi: int = 500_000, tex_size: int = 1024, resolutions=RESOLUTIONS) -> List[Scene]:
    """Batched workload: ``n_meshes`` seeded meshes (tori, UV spheres and
    soups with log-uniform triangle counts in [tri_lo, tri_hi], each with its
    own tex_size^2 RGBA8 texture), each seen by ``cams_per_mesh`` random
    cameras at a resolution drawn uniformly from ``resolutions``.  Returns one
    Scene per view, views of a mesh sharing the same Object."""
    rng = _rng(seed)
    meshes = mesh_pool(seed + 1000, n_meshes, tri_lo, tri_hi, tex_size)
    views: List[Scene] = []
    for obj in meshes:
        for _ in range(cams_per_mesh):
            W, H = resolutions[int(rng.integers(len(resolutions)))]
            cam = random_camera(rng)
            views.append(Scene(W, H, cam, (0.0, 6.0, 6.0), [obj]))
    return views


def window_object(n: int, half: float, z: float, texture: Texture, x0: float = 0.0, y0: float = 0.0) -> Object:
    """An n x n grid of quads (2n^2 triangles) in the plane z = ``z``, centred
    on (x0, y0), facing +z, UVs spanning [0, 1]^2.  Symmetric about x = x0, so
    mirrored triangles have equal centroid distances from a camera on that
    plane of symmetry (ties for the transparent sort)."""
    g = np.linspace(-half, half, n + 1, dtype=np.float64)
    xs, ys = np.meshgrid(g + x0, g + y0, indexing="xy")
    verts = np.stack([xs.ravel(), ys.ravel(), np.full(xs.size, z)], 1).astype(F32)
    uv = np.stack([(xs.ravel() - x0 + half) / (2 * half), (ys.ravel() - y0 + half) / (2 * half)], 1)
    uv = np.clip(uv, 0.0, 1.0).astype(F32)
    tris = []
    for j in range(n):
        for i in range(n):
            a, b = j * (n + 1) + i, j * (n + 1) + i + 1
            c, d = (j + 1) * (n + 1) + i + 1, (j + 1) * (n + 1) + i
            for t in ((a, b, c), (a, c, d)):   # counter-clockwise seen from +z
                tris.append([t[0], t[1], t[2], t[0], t[1], t[2], 0, 0, 0])
    normals = np.array([[0.0, 0.0, 1.0]], F32)
    return Object(verts, normals, uv, [IndexedMesh(np.array(tris, np.uint32), 0)], [texture])


SERVING_RESOLUTIONS = [((854, 480), 0.4), ((1280, 720), 0.3), ((1920, 1080), 0.25), ((3840, 2160), 0.05)]


def serving_requests(seed: int = 5, n_requests: int = 4096, n_meshes: int = 64, tri_lo: int = 1000,
                     tri_hi: int = 2_000_000, tex_size: int = 1024, resolutions=SERVING_RESOLUTIONS,
                     pool: List[Object] | None = None) -> Tuple[List[Object], List[Scene]]:
    """Serving workload: a pool of ``n_meshes`` seeded meshes (log-uniform
    triangle counts in [tri_lo, tri_hi]) and ``n_requests`` requests, each a
    mesh drawn uniformly from the pool, a random camera and a resolution drawn
    with the given probabilities.  Returns (pool, one Scene per request)."""
    rng = _rng(seed)
    if pool is None:
        pool = mesh_pool(seed + 2000, n_meshes, tri_lo, tri_hi, tex_size)
    sizes = [r for r, _ in resolutions]
    p = np.array([q for _, q in resolutions], np.float64)
    p /= p.sum()
    reqs: List[Scene] = []
    for _ in range(n_requests):
        obj = pool[int(rng.integers(len(pool)))]
        W, H = sizes[int(rng.choice(len(sizes), p=p))]
        reqs.append(Scene(W, H, random_camera(rng), (0.0, 6.0, 6.0), [obj]))
    return pool, reqs


def _ball_soup(n: int, radius: float, tex: Texture, rng: np.random.Generator) -> Object:
    """Soup of small triangles with centroids uniform in a ball."""
    c = rng.standard_normal((n, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    c *= radius * np.cbrt(rng.random(n))[:, None]
    size = 2.5 * radius / math.sqrt(max(n, 1))
    q = rng.standard_normal((n, 3, 3)) * size
    pc = c[:, None, :] + q
    fn = np.cross(pc[:, 1] - pc[:, 0], pc[:, 2] - pc[:, 1])
    fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-30)
    vn = np.repeat(fn[:, None, :], 3, 1) + 0.2 * rng.standard_normal((n, 3, 3))
    vn /= np.maximum(np.linalg.norm(vn, axis=2, keepdims=True), 1e-30)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    tri = np.concatenate([idx, idx, idx], 1)
    return Object(pc.reshape(-1, 3).astype(F32), vn.reshape(-1, 3).astype(F32),
                  rng.random((3 * n, 2)).astype(F32), [IndexedMesh(tri, 0, "ball_soup")], [tex],
                  name="ball_soup")


def split_submission(scene: Scene, n_shards: int) -> Scene:
    """The scene's first object with its first mesh cut into ``n_shards``
    consecutive submission ranges (meshes over the same vertex arrays): the
    shards of a sort-last frame, in submission order."""
    import copy
    o = scene.objects[0]
    tri = o.meshes[0].triangles
    cuts = np.linspace(0, tri.shape[0], n_shards + 1).astype(int)
    o2 = copy.copy(o)
    o2.meshes = []
    for k in range(n_shards):
        m = copy.copy(o.meshes[0])
        m.triangles = np.ascontiguousarray(tri[cuts[k]:cuts[k + 1]])
        o2.meshes.append(m)
    s2 = copy.copy(scene)
    s2.objects = [o2]
    return s2

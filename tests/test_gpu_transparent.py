"""The transparent pass on the GPU (scene/mod.rs:1088-1247): per-object order
(an object's transparent meshes after its opaque ones, before the next
object's), back-to-front stable sort kept across frames, no culling, no depth
writes, u8 blend.  Frames are compared with the oracle bit for bit."""
import numpy as np
import pytest

from draw_amd import workloads
from draw_amd.scene import Camera, Scene, Texture
from parity import assert_same, oracle

pytestmark = pytest.mark.gpu


def _glass(rng, alpha, size=64):
    tex = workloads.checker_texture(size, 8, rng)
    return Texture(ka=(0.2, 0.2, 0.2), kd=(0.8, 0.6, 0.4), ks=(0.3, 0.3, 0.3), alpha=alpha, map_kd=tex)


def _scene(seed=7, width=320, height=240):
    rng = np.random.default_rng(seed)
    base = workloads.torus_scene(seed, width=width, height=height)
    torus = base.objects[0]
    # object 0: the torus + a transparent window in front of it (one object, two meshes)
    win = workloads.window_object(6, 0.9, 0.8, _glass(rng, 0.5))
    tri = win.meshes[0].triangles.copy()
    tri[:, :3] += torus.vertices.shape[0]
    tri[:, 3:6] += torus.texture_vertices.shape[0]
    tri[:, 6:] += torus.normals_vertices.shape[0]
    from draw_amd.scene import IndexedMesh, Object
    obj0 = Object(np.concatenate([torus.vertices, win.vertices]),
                  np.concatenate([torus.normals_vertices, win.normals_vertices]),
                  np.concatenate([torus.texture_vertices, win.texture_vertices]),
                  [IndexedMesh(torus.meshes[0].triangles, 0), IndexedMesh(tri, 1)],
                  [torus.textures[0], win.textures[0]])
    # object 1: an opaque plate in front of the window over part of the view (drawn after
    # object 0's transparent mesh) and its own transparent window further in front
    plate = workloads.window_object(2, 0.3, 1.1, Texture(), x0=0.4, y0=0.2)
    glass2 = workloads.window_object(4, 0.5, 1.4, _glass(rng, 0.3), x0=-0.3, y0=-0.1)
    tri2 = glass2.meshes[0].triangles.copy()
    tri2[:, :3] += plate.vertices.shape[0]
    tri2[:, 3:6] += plate.texture_vertices.shape[0]
    tri2[:, 6:] += plate.normals_vertices.shape[0]
    obj1 = Object(np.concatenate([plate.vertices, glass2.vertices]),
                  np.concatenate([plate.normals_vertices, glass2.normals_vertices]),
                  np.concatenate([plate.texture_vertices, glass2.texture_vertices]),
                  [IndexedMesh(tri2, 1), IndexedMesh(plate.meshes[0].triangles, 0)],   # transparent listed first
                  [plate.textures[0], glass2.textures[0]])
    base.objects = [obj0, obj1]
    return base


def test_transparent_objects_bit_exact(renderer):
    s = _scene()
    c, w = renderer.render(s)
    assert_same(c.frame, c.depth, w, oracle(s), label="transparent frame")
    assert np.any(c.frame[:, :, 3] == 0), "no pixel was blended"


def test_sort_order_persists_across_frames(renderer):
    s = _scene(seed=8)
    w_mesh = s.objects[0].meshes[1]
    original = w_mesh.triangles.copy()
    c, w = renderer.render(s)
    assert_same(c.frame, c.depth, w, oracle(s), label="frame 1")
    renderer.sync_transparent_order(s)
    # the host model of the device's in-place sort: stable, descending distance
    v = s.objects[0].vertices.astype(np.float32)
    cam = np.array(s.camera.position, np.float32)
    ctr = ((v[original[:, 0]] + v[original[:, 1]]) + v[original[:, 2]]) / np.float32(3.0)
    d = ctr - cam
    dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    order = np.argsort(-dist, kind="stable")
    assert np.array_equal(w_mesh.triangles, original[order])
    assert len(np.unique(dist)) < len(dist), "the window should have equal distances (ties)"
    # frame 2 from another camera starts from the permuted order
    s.camera = Camera(position=(0.35, 0.25, 2.9), direction=None, fov_x=60.0, near=-0.1, far=-100.0)
    c2, w2 = renderer.render(s)
    assert_same(c2.frame, c2.depth, w2, oracle(s), label="frame 2")


def test_batch_rejects_transparent_meshes(renderer):
    from draw_amd._lib import DrawHipError
    with pytest.raises(DrawHipError, match="DH_E_UNSUPPORTED"):
        renderer.render_batch([(_scene(seed=9), None)])

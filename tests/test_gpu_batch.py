"""Batched rendering: many views with uneven triangle counts and resolutions
through one pipeline pass (dh_render_batch), optionally replayed from a
captured hipGraph.  Each view must equal its standalone render and the
oracle, bit for bit (winner ids included)."""
import numpy as np
import pytest

from draw_amd import workloads
from draw_amd.render import frame_uniforms
from parity import assert_same, oracle

pytestmark = pytest.mark.gpu


def _read(renderer, cv, scene):
    return renderer.read(cv, scene.width, scene.height)


def _mixed_views():
    views = [workloads.torus_scene(1), workloads.small_random_scene(21, n_tri=300, width=97, height=61),
             workloads.torus_scene(2, width=320, height=200, power=32),
             workloads.small_random_scene(22, n_tri=800, width=160, height=120, cross_frac=0.4)]
    # the same object seen from several cameras (shared mesh, different views)
    rng = np.random.default_rng(3)
    shared = workloads.torus_scene(5, width=128, height=96)
    for _ in range(4):
        s = workloads.torus_scene(5, width=int(rng.integers(40, 200)), height=int(rng.integers(30, 150)))
        s.objects = shared.objects
        s.camera = workloads.random_camera(rng)
        views.append(s)
    return views


def test_batch_views_equal_standalone_and_oracle(renderer):
    views = _mixed_views()
    cvs = renderer.render_batch([(s, None) for s in views])
    outs = [_read(renderer, cv, s) for cv, s in zip(cvs, views)]
    for k, (s, (f, d, w)) in enumerate(zip(views, outs)):
        ref = oracle(s)
        assert_same(f, d, w, ref, label=f"batch view {k} ({s.width}x{s.height})")
        c, w1 = renderer.render(s)
        assert np.array_equal(c.frame, f) and np.array_equal(c.depth.view(np.uint32), d.view(np.uint32))
        assert np.array_equal(w1, w)


def test_batch_graph_replays_match_eager(renderer):
    """hipGraph replays, with eager work between them, each equal to the eager
    batch; new uniforms between replays give the new eager result."""
    views = workloads.batched_scenes(seed=8, n_meshes=4, cams_per_mesh=4, tri_lo=500, tri_hi=20_000, tex_size=128,
                                     resolutions=[(96, 64), (160, 120), (256, 256)])
    items = [(s, None) for s in views]
    eager = [_read(renderer, cv, s) for cv, s in zip(renderer.render_batch(items), views)]
    for rep in range(4):
        cvs = renderer.render_batch(items, graph=True)
        got = [_read(renderer, cv, s) for cv, s in zip(cvs, views)]
        for (a, b, c), (x, y, z) in zip(got, eager):
            assert np.array_equal(a, x) and np.array_equal(b.view(np.uint32), y.view(np.uint32))
            assert np.array_equal(c, z)
        renderer.render(workloads.small_random_scene(30 + rep, n_tri=100))   # eager work in between
    # same structure, new cameras: the replay must follow the new uniforms
    rng = np.random.default_rng(9)
    for s in views:
        s.camera = workloads.random_camera(rng)
    items = [(s, frame_uniforms(s)) for s in views]
    cvs = renderer.render_batch(items, graph=True)
    got = [_read(renderer, cv, s) for cv, s in zip(cvs, views)]
    for k, (s, (f, d, w)) in enumerate(zip(views, got)):
        ref = oracle(s)
        assert_same(f, d, w, ref, label=f"replayed view {k}")


def test_batch_256_views_configured_size_full_frames(renderer):
    """The batched workload at its configured size: 8 seeded meshes (tori, UV
    spheres, soups; 1k-500k triangles log-uniform, 1024x1024 textures) x 32
    cameras, resolutions uniform over 256x256..1920x1080, one graph-replayed
    dh_render_batch.  Every one of the 256 views equals the oracle on every
    pixel, and a handful equal their standalone dh_render."""
    views = workloads.batched_scenes(seed=4)
    assert len(views) == 256 and max(s.triangle_count for s in views) > 50_000
    items = [(s, None) for s in views]
    renderer.render_batch(items, graph=True)           # captures
    cvs = renderer.render_batch(items, graph=True)     # replays
    for k, s in enumerate(views):
        f, d, w = _read(renderer, cvs[k], s)
        assert_same(f, d, w, oracle(s), label=f"view {k} ({s.width}x{s.height}, {s.triangle_count} tris)")
    for k in range(0, 256, 51):
        s = views[k]
        f, d, w = _read(renderer, cvs[k], s)
        c, w1 = renderer.render(s)
        assert np.array_equal(c.frame, f) and np.array_equal(c.depth.view(np.uint32), d.view(np.uint32))
        assert np.array_equal(w1, w)


@pytest.mark.parametrize("kind,tile,waves", [(0, 16, 4), (1, 16, 4)])
def test_batch_of_tiny_views_matches_oracle(renderer, kind, tile, waves):
    """150 views of 3-24 pixels a side, most of them smaller than a wave's 64
    pixels: a shading wave's pixels then span several views, and the raster
    workgroups search 150 view bases.  Both raster kinds (kind 1 shades in a
    separate pass over the visibility buffer)."""
    rng = np.random.default_rng(12)
    views = [workloads.small_random_scene(100 + k, n_tri=40, width=int(rng.integers(3, 25)),
                                          height=int(rng.integers(3, 25)), tex_size=8) for k in range(150)]
    renderer.set_raster(kind, tile, waves)
    try:
        cvs = renderer.render_batch([(s, None) for s in views])
        outs = [_read(renderer, cv, s) for cv, s in zip(cvs, views)]
    finally:
        renderer.set_raster(0, 16, 4)
    assert sum(s.width * s.height < 64 for s in views) > 10
    for k, (s, (f, d, w)) in enumerate(zip(views, outs)):
        assert_same(f, d, w, oracle(s), label=f"tiny view {k} ({s.width}x{s.height}, kind {kind})")

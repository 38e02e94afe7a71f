"""Serving: the continuous-batching server (dh_srv_*).  Every job's result in
a mixed, memory-limited stream of jobs must equal the same job rendered alone
(through the server and through dh_render) bit for bit, and the oracle."""
import numpy as np
import pytest

from draw_amd import workloads
from draw_amd._lib import DrawHipError
from draw_amd.render import frame_uniforms
from draw_amd.serve import Server, poisson_load
from parity import assert_same, oracle

pytestmark = pytest.mark.gpu

SMALL_RES = [((96, 64), 0.3), ((160, 120), 0.3), ((320, 240), 0.3), ((200, 150), 0.1)]


def _requests(renderer, seed, n):
    pool, scenes = workloads.serving_requests(seed=seed, n_requests=n, n_meshes=6, tri_lo=300, tri_hi=30_000,
                                              tex_size=128, resolutions=SMALL_RES)
    return scenes, [(frame_uniforms(s), renderer.mesh_handles(s)) for s in scenes]


def _alone_need(scene):
    # canvas + a generous bound on one job's scratch
    return 16 * scene.width * scene.height + 450 * scene.triangle_count + (12 << 20)


def test_mix_equals_alone_under_memory_pressure(renderer):
    scenes, reqs = _requests(renderer, seed=31, n=40)
    budget = 3 * max(_alone_need(s) for s in scenes)
    srv = Server(renderer, budget_bytes=budget, workers=2, max_views=16)
    try:
        tickets = [srv.submit(fr, hs) for fr, hs in reqs]           # a burst: all queued at once
        mixed = []
        for t, s in zip(tickets, scenes):
            res = srv.poll(t, wait=True)
            assert res.status == 0
            mixed.append(srv.read(res, s.width, s.height))
            srv.release(t)
        st = srv.stats()
        assert st["completed"] == len(scenes) and st["failed"] == 0
        assert st["launches"] < len(scenes), "jobs were not batched"
        assert st["bytes_peak"] <= budget
        # the same jobs one at a time through the server, and through dh_render
        for k, ((fr, hs), s) in enumerate(zip(reqs, scenes)):
            t = srv.submit(fr, hs)
            res = srv.poll(t, wait=True)
            assert res.status == 0 and res.batch_views == 1
            alone = srv.read(res, s.width, s.height)
            srv.release(t)
            c, w = renderer.render(s, frame=fr)
            f, d, wm = mixed[k]
            for a, b in zip(alone, mixed[k]):
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f"job {k}: mix differs from alone"
            assert np.array_equal(c.frame, f) and np.array_equal(c.depth.view(np.uint32), d.view(np.uint32))
            assert np.array_equal(w, wm)
            if k % 8 == 0:
                assert_same(f, d, wm, oracle(s), label=f"served job {k}")
    finally:
        srv.close()


def test_job_over_budget_is_rejected(renderer):
    scenes, reqs = _requests(renderer, seed=32, n=2)
    srv = Server(renderer, budget_bytes=1 << 20, workers=1)
    try:
        with pytest.raises(DrawHipError, match="DH_E_BUDGET"):
            srv.submit(*reqs[0])
    finally:
        srv.close()


def test_poisson_stream_completes(renderer):
    scenes, reqs = _requests(renderer, seed=33, n=64)
    srv = Server(renderer, budget_bytes=512 << 20, workers=2)
    try:
        out = poisson_load(srv, reqs, rate_hz=400.0, seconds=0.5, seed=1)
        assert out["completed"] == out["requests"] > 50 and out["failed"] == 0
        assert 0 < out["latency_ms"]["p50"] <= out["latency_ms"]["p99"] < 5000
        st = srv.stats()
        assert st["queued"] == 0 and st["completed"] == out["requests"]
    finally:
        srv.close()


def _digest(bufs) -> bytes:
    import hashlib
    h = hashlib.blake2b(digest_size=16)
    for b in bufs:
        h.update(np.ascontiguousarray(b).view(np.uint8).tobytes())
    return h.digest()


def test_serving_configured_pool_mix_equals_alone_and_oracle(renderer):
    """The serving workload at its configured size: the 64-mesh resident pool
    (1k-2M triangles log-uniform, 14.3M in total, 1024x1024 textures), random
    cameras, resolutions 854x480 / 1280x720 / 1920x1080 / 3840x2160 with
    p = .4/.3/.25/.05 and the 200 GB budget.  192 requests (every 4K one among
    the first 1024 included) go in as one burst over two workers; each
    result equals the same job rendered alone by dh_render byte for byte, and
    every 12th job and every 4K job equal the oracle on every pixel."""
    pool, scenes = workloads.serving_requests(seed=5)
    assert len(pool) == 64 and sum(o.triangle_count for o in pool) > 10_000_000
    pick = list(range(160)) + [k for k in range(160, 1024) if scenes[k].width == 3840][:32]
    jobs = [scenes[k] for k in pick]
    reqs = [(frame_uniforms(sc), renderer.mesh_handles(sc)) for sc in jobs]
    n4k = sum(sc.width == 3840 for sc in jobs)
    assert n4k >= 8 and max(sc.triangle_count for sc in jobs) > 1_000_000
    srv = Server(renderer, budget_bytes=200 << 30, workers=2, max_views=64)
    try:
        tickets = [srv.submit(fr, hs) for fr, hs in reqs]
        mixed, batch_views = [], []
        for t, sc in zip(tickets, jobs):
            res = srv.poll(t, wait=True)
            assert res.status == 0
            batch_views.append(res.batch_views)
            mixed.append(_digest(srv.read(res, sc.width, sc.height)))
            srv.release(t)
        st = srv.stats()
        assert st["completed"] == len(jobs) and st["failed"] == 0
        assert st["launches"] < len(jobs) and max(batch_views) > 1, "jobs were not batched"
    finally:
        srv.close()
    for k, ((fr, hs), sc) in enumerate(zip(reqs, jobs)):
        cv = renderer.render_device(sc, fr, hs)
        f, d, w = renderer.read(cv, sc.width, sc.height)
        assert _digest((f, d, w)) == mixed[k], f"job {k} ({sc.width}x{sc.height}): mix differs from alone"
        if k % 12 == 0 or sc.width == 3840:
            assert_same(f, d, w, oracle(sc), label=f"served job {k} ({sc.width}x{sc.height}, "
                                                   f"{sc.triangle_count} tris)")


def _stacked_quads(n=40, width=1280, height=720):
    """n full-screen quads at slightly different depths: every triangle lands
    in every tile, far beyond the 4 bin entries per triangle a launch is
    first sized for."""
    from draw_amd.scene import Camera, IndexedMesh, Object, Scene, Texture
    quads = [workloads.window_object(1, 3.0, -0.01 * k, Texture()) for k in range(n)]
    verts = np.concatenate([q.vertices for q in quads])
    uv = np.concatenate([q.texture_vertices for q in quads])
    tri = np.concatenate([q.meshes[0].triangles + np.array([4 * k] * 6 + [0] * 3, np.uint32)
                          for k, q in enumerate(quads)])
    obj = Object(verts, quads[0].normals_vertices, uv, [IndexedMesh(tri, 0)], [Texture()])
    cam = Camera(position=(0.0, 0.0, 3.0), direction=None, fov_x=60.0, near=-0.1, far=-100.0)
    return Scene(width, height, cam, (0.0, 6.0, 6.0), [obj])


def test_launch_outgrowing_its_estimate_stays_within_budget(renderer):
    s = _stacked_quads()
    fr, hs = frame_uniforms(s), renderer.mesh_handles(s)
    budget = 256 << 20
    srv = Server(renderer, budget_bytes=budget, workers=1)
    try:
        t = srv.submit(fr, hs)
        res = srv.poll(t, wait=True)
        assert res.status == 0
        got = srv.read(res, s.width, s.height)
        srv.release(t)
        assert srv.stats()["bytes_peak"] <= budget
    finally:
        srv.close()
    c, w = renderer.render(s, frame=fr)
    assert np.array_equal(got[0], c.frame) and np.array_equal(got[2], w)
    assert_same(c.frame, c.depth, w, oracle(s, grid=(9, 7, 2, 3)),
                mask=_grid(s, (9, 7, 2, 3)), label="stacked quads")


def _grid(s, g):
    m = np.zeros((s.height, s.width), bool)
    m[g[3]::g[1], g[2]::g[0]] = True
    return m

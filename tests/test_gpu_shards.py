"""Sort-last sharding on one GPU: a frame split by submission range into
shards, each rendered on its own context into a packed u64 visibility buffer,
merged with an element-wise minimum (what an RCCL ncclMin reduction does
across GPUs) and shaded per shard, equals the single-context frame byte for
byte; the merged winners map back to the single frame's winner ids."""
import numpy as np
import pytest

from draw_amd import workloads
from draw_amd.render import frame_uniforms

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_shards,seed", [(4, 8), (8, 9), (2, 10)])
def test_sort_last_shards_match_single_render(renderer, n_shards, seed):
    s = workloads.split_submission(
        workloads.soup_scene(seed, n_tri=120_000, width=640, height=480, tex_size=256, cross_frac=0.1), n_shards)
    W, H = s.width, s.height
    fr = frame_uniforms(s)
    meshes = renderer.mesh_handles(s)
    assert len(meshes) == n_shards
    cv = renderer.render_device(s, fr, meshes)
    ref_bgra, ref_depth, ref_win = renderer.read(cv, W, H)

    ctxs = [renderer.shard_context() for _ in range(n_shards)]
    canvas = renderer.canvas_slot("shards", W, H)
    bufs = [renderer.device_alloc(W * H * 8) for _ in range(n_shards)]
    try:
        base, b = [], 0
        for k, cx in enumerate(ctxs):
            cx.render_shard(fr, [meshes[k]], k, canvas, bufs[k])
            base.append(b)
            b += cx.stats()["tris_out"]
        for k in range(1, n_shards):
            renderer.vis_merge_min(bufs[0], bufs[k], W * H)
        for k, cx in enumerate(ctxs):
            cx.shade_shard(bufs[0], k)
        merged = renderer.device_read(bufs[0], W * H * 8, np.uint64).reshape(H, W)
        bgra, depth, win = renderer.read(canvas, W, H)
    finally:
        for p in bufs:
            renderer.device_free(p)
        for cx in ctxs:
            cx.close()
    assert np.array_equal(bgra, ref_bgra)
    assert np.array_equal(depth.view(np.uint32), ref_depth.view(np.uint32))
    # winner ids: the shard's local id + the clipped triangles of the shards before it
    lo = (merged & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    none = merged == np.uint64(0xFFFFFFFFFFFFFFFF)
    shard = lo >> 29
    glob = (lo & np.uint32((1 << 29) - 1)) + np.asarray(base, np.uint32)[np.minimum(shard, n_shards - 1)]
    assert np.array_equal(none, ref_win == 0xFFFFFFFF)
    assert np.array_equal(glob[~none], ref_win[~none])
    assert np.array_equal(win[~none], (lo & np.uint32((1 << 29) - 1))[~none])
    # the merged depth key is the winner's depth
    assert (shard[~none] < n_shards).all()


def test_shard_errors(renderer):
    s = workloads.torus_scene(1)
    fr = frame_uniforms(s)
    meshes = renderer.mesh_handles(s)
    cx = renderer.shard_context()
    buf = renderer.device_alloc(s.width * s.height * 8)
    try:
        from draw_amd._lib import DrawHipError
        with pytest.raises(DrawHipError):
            cx.render_shard(fr, meshes, 8, renderer.canvas(s.width, s.height), buf)  # shard index too large
        with pytest.raises(DrawHipError):
            cx.shade_shard(buf, 0)  # no shard frame on this context
        cx.render_shard(fr, meshes, 0, renderer.canvas(s.width, s.height), buf)
        with pytest.raises(DrawHipError):
            cx.shade_shard(buf, 1)  # another shard's index
        cx.shade_shard(buf, 0)
    finally:
        renderer.device_free(buf)
        cx.close()


def test_shard_graph_replay_identical(renderer):
    """A repeated shard frame replays its captured launch sequence; the
    visibility it writes is the eager frame's, bit for bit."""
    s = workloads.split_submission(workloads.soup_scene(11, n_tri=40_000, width=320, height=240, tex_size=256), 3)
    W, H = s.width, s.height
    fr = frame_uniforms(s)
    meshes = renderer.mesh_handles(s)
    cx = renderer.shard_context()
    bufs = [renderer.device_alloc(W * H * 8) for _ in range(2)]
    try:
        out = []
        for rep in range(3):
            cx.render_shard(fr, [meshes[1]], 1, renderer.canvas(W, H), bufs[rep % 2])
            out.append(renderer.device_read(bufs[rep % 2], W * H * 8, np.uint64))
    finally:
        for p in bufs:
            renderer.device_free(p)
        cx.close()
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
    assert (out[0] != np.uint64(0xFFFFFFFFFFFFFFFF)).any()

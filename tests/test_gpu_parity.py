"""GPU parity: libdrawhip (through its C ABI) against the CPU oracle on the
same seeded scenes.  Bar: winner ids, depth bits and BGRA bytes identical."""
import numpy as np
import pytest

from draw_amd import workloads
from draw_amd.render import frame_uniforms
from oracle import oracle as O
from parity import assert_same, max_lsb, oracle

pytestmark = pytest.mark.gpu

TILINGS = [(8, 1), (16, 1), (16, 2), (16, 4), (32, 4), (32, 16)]


def _gpu(renderer, scene, tile=16, waves=4, kind=0):
    renderer.set_raster(kind, tile, waves)
    try:
        canvas, winner = renderer.render(scene)
    finally:
        renderer.set_tiling(16, 4)
    return canvas.frame, canvas.depth, winner


def test_torus_640x480_bit_exact(renderer):
    s = workloads.torus_scene(1)
    ref = oracle(s)
    f, d, w = _gpu(renderer, s)
    assert_same(f, d, w, ref, label="torus")
    st = renderer.stats()
    assert st["tris_out"] == ref.stats["tris_emitted"]


def test_torus_power32_bit_exact(renderer):
    s = workloads.torus_scene(4, power=32)
    ref = oracle(s)
    f, d, w = _gpu(renderer, s)
    assert_same(f, d, w, ref, label="torus p32")


@pytest.mark.parametrize("kind,tile,waves", [(0, t, w) for t, w in TILINGS] + [(1, 16, 4), (1, 16, 8), (1, 32, 4),
                                                                             (1, 32, 8), (1, 16, 2)])
def test_small_soups_all_tilings(renderer, kind, tile, waves):
    for seed in range(3):
        s = workloads.small_random_scene(100 + seed, n_tri=400, width=97, height=61, cross_frac=0.3)
        ref = oracle(s)
        f, d, w = _gpu(renderer, s, tile, waves, kind)
        assert_same(f, d, w, ref, label=f"soup seed {seed} kind {kind} tile {tile}x{waves}")


def test_soup_20k_320x240_bit_exact(renderer):
    s = workloads.soup_scene(7, n_tri=20_000, width=320, height=240, tex_size=256, cross_frac=0.05)
    ref = oracle(s)
    f, d, w = _gpu(renderer, s)
    assert_same(f, d, w, ref, label="soup20k")
    assert renderer.stats()["tris_out"] == ref.stats["tris_emitted"]


def test_soup_1m_1080p_full_frame_bit_exact(renderer):
    """The 1920x1080, 1M-triangle soup (20% near-plane straddlers) against the
    oracle on every one of its 2,073,600 pixels: winner ids, depth bits and
    BGRA bytes identical, and the emitted-triangle count exactly."""
    s = workloads.soup_scene(2)
    f, d, w = _gpu(renderer, s, 32, 8)
    emitted = renderer.stats()["tris_out"]
    ref = oracle(s)
    assert_same(f, d, w, ref, label="soup1M 1080p full frame")
    assert emitted == ref.stats["tris_emitted"]
    assert max_lsb(f, ref.frame) == 0


def test_shells_20m_4k_full_frame_bit_exact(renderer):
    """The 3840x2160 stress workload: ~20M micro-triangles on 8 nested shells
    (5% flipped, 5% outside the frustum, 8192x8192 texture) against the oracle
    on every one of its 8,294,400 pixels, and the emitted-triangle count."""
    s = workloads.shells_scene(3)
    f, d, w = _gpu(renderer, s, 32, 8)
    emitted = renderer.stats()["tris_out"]
    ref = oracle(s)
    assert_same(f, d, w, ref, label="shells20M 4K full frame")
    assert emitted == ref.stats["tris_emitted"]


@pytest.mark.parametrize("name", ["shells200k_320", "soup120k_160"])
@pytest.mark.parametrize("kind,tile,waves", [(0, 16, 4), (0, 32, 8), (1, 16, 4)])
@pytest.mark.parametrize("mode", [2, 1])
def test_dense_frames_bit_exact(renderer, name, kind, tile, waves, mode):
    """Frames with at least one input triangle per pixel (past the small-frame
    limit), where many unclipped triangles of at most 3x3 pixels cover none and
    the emits drop them, in both ordered bin modes (near first, submission
    order): the dropped triangles keep their clipped-triangle ids, so winners,
    depth and colour match the oracle on every pixel, as does the emitted
    count."""
    s = {"shells200k_320": lambda: workloads.shells_scene(31, n_tri=200_000, width=320, height=240, tex_size=256,
                                                           n_shells=4),
         "soup120k_160": lambda: workloads.soup_scene(32, n_tri=120_000, width=160, height=120, tex_size=256,
                                                      cross_frac=0.1)}[name]()
    assert sum(o.triangle_count for o in s.objects) >= s.width * s.height
    try:
        renderer.set_bin_order(mode)
        f, d, w = _gpu(renderer, s, tile, waves, kind)
        emitted = renderer.stats()["tris_out"]
    finally:
        renderer.set_bin_order(2)
    ref = oracle(s)
    assert_same(f, d, w, ref, label=f"{name} kind {kind} tile {tile}x{waves}")
    assert emitted == ref.stats["tris_emitted"]


def test_soup_1m_tiling_invariance(renderer):
    s = workloads.soup_scene(2)
    base = _gpu(renderer, s, 16, 4)
    for kind, tile, waves in [(0, 8, 1), (0, 32, 4), (1, 32, 8), (1, 16, 4)]:
        other = _gpu(renderer, s, tile, waves, kind)
        for a, b in zip(base, other):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f"kind {kind} tile {tile}x{waves} differs"


def test_bin_order_modes_identical(renderer):
    """Unordered (atomic), submission-ordered and near-first bin lists give the
    same frame: the per-pixel visibility key does not depend on list order."""
    s = workloads.soup_scene(5, n_tri=200_000, width=1280, height=720, tex_size=512)
    frames = []
    try:
        for mode in (0, 1, 2):
            renderer.set_bin_order(mode)
            frames.append(_gpu(renderer, s))
            frames.append(_gpu(renderer, s, 32, 4))
            frames.append(_gpu(renderer, s, 16, 4, kind=1))
    finally:
        renderer.set_bin_order(2)
    for other in frames[1:]:
        for x, y in zip(frames[0], other):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    assert_same(*frames[0], oracle(s), label="soup200k full frame")


@pytest.mark.parametrize("kind,tile,waves", [(0, 32, 8), (0, 16, 4), (0, 16, 2), (1, 32, 4), (0, 8, 1)])
def test_small_frame_binning_matches_general_path(renderer, kind, tile, waves):
    """A frame of at most 16384 input triangles and 4096 tiles is binned by one
    workgroup (k_bin_small, unordered lists) in the default bin order; the
    general path (submission-ordered lists, mode 1) gives the same frame bytes
    and the same number of bin entries, and both give the oracle's frame.  The
    soup's near-plane straddlers clip into triangles many tiles wide; at 8x8
    tiles the frame has 14400 tiles and takes the general path in both modes."""
    s = workloads.soup_scene(7, n_tri=16_000, width=1280, height=720, tex_size=256)
    ref = oracle(s)
    got = {}
    try:
        for mode in (2, 1):
            renderer.set_bin_order(mode)
            got[mode] = (_gpu(renderer, s, tile, waves, kind), renderer.stats()["bin_entries"])
    finally:
        renderer.set_bin_order(2)
    assert got[2][1] == got[1][1] > 0
    for x, y in zip(got[2][0], got[1][0]):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    assert_same(*got[2][0], ref, label=f"16k soup, kind {kind} tile {tile}x{waves}")


def test_near_first_back_to_front_submission(renderer):
    """A soup submitted back to front (the near-first order's worst case for
    submission-ordered lists) gives the oracle's frame in every bin mode."""
    import copy
    s = workloads.soup_scene(6, n_tri=60_000, width=640, height=480, tex_size=256)
    o = s.objects[0]
    tri = o.meshes[0].triangles
    zc = o.vertices.astype(np.float64)[tri[:, :3]].mean(axis=1)[:, 2]
    o2 = copy.copy(o)
    m2 = copy.copy(o.meshes[0])
    m2.triangles = np.ascontiguousarray(tri[np.argsort(zc, kind="stable")])
    o2.meshes = [m2]
    s.objects = [o2]
    ref = oracle(s)
    try:
        for mode in (1, 2):
            renderer.set_bin_order(mode)
            f, d, w = _gpu(renderer, s)
            assert_same(f, d, w, ref, label=f"back to front, bin mode {mode}")
    finally:
        renderer.set_bin_order(2)


def test_multi_object_and_default_material(renderer):
    rng = np.random.default_rng(5)
    t1 = workloads.torus_scene(9)
    s2 = workloads.small_random_scene(11, n_tri=300, width=640, height=480, cross_frac=0.1)
    obj_a = t1.objects[0]
    obj_b = s2.objects[0]
    obj_b.meshes[0].texture_idx = 7      # missing texture -> Texture::default (scene/mod.rs:1072-1075)
    t1.objects.append(obj_b)
    ref = oracle(t1)
    f, d, w = _gpu(renderer, t1)
    assert_same(f, d, w, ref, label="two objects")
    del rng, obj_a


def test_texcoord_out_of_range_is_an_error(renderer):
    s = workloads.torus_scene(1)
    obj = s.objects[0]
    obj.texture_vertices = obj.texture_vertices * np.float32(3.0)   # v up to 3 -> index wraps out of range
    ref = O.render(s)
    assert ref.rc == -3
    from draw_amd._lib import DrawHipError
    with pytest.raises(DrawHipError) as e:
        renderer.render(s)
    assert e.value.code == -3


def test_repeat_frames_identical(renderer):
    s = workloads.soup_scene(3, n_tri=50_000, width=640, height=480, tex_size=512)
    a = _gpu(renderer, s)
    b = _gpu(renderer, s)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_frame_uniforms_match_oracle():
    for s in [workloads.torus_scene(1), workloads.soup_scene(2, n_tri=10)]:
        assert bytes(frame_uniforms(s)) == bytes(O.frame_uniforms(s))


def test_graph_replay_matches_eager_and_oracle(renderer):
    """dh_render replays its captured launch sequence while the frame
    structure repeats; new uniforms (camera, light) each frame must give the
    eager frame's bytes, and the oracle's."""
    import copy
    from draw_amd.scene import Camera
    s = workloads.soup_scene(4, n_tri=30_000, width=640, height=480, tex_size=256)
    meshes = renderer.mesh_handles(s)
    scenes = []
    for k, (pos, light) in enumerate([((0.0, 0.0, 150.0), (0.0, 300.0, 300.0)),
                                      ((20.0, -10.0, 140.0), (100.0, 200.0, 300.0)),
                                      ((-30.0, 25.0, 120.0), (-50.0, 300.0, 100.0))]):
        v = copy.copy(s)
        v.camera = Camera(position=pos)
        v.light_source = light
        scenes.append(v)
    renderer.set_graph(False)
    eager = []
    for v in scenes:
        cv = renderer.render_device(v, frame_uniforms(v), meshes)
        eager.append(renderer.read(cv, 640, 480))
    renderer.set_graph(True)
    for rep in range(2):
        for k in [0, 1, 2, 1, 0]:
            cv = renderer.render_device(scenes[k], frame_uniforms(scenes[k]), meshes)
            got = renderer.read(cv, 640, 480)
            for a, b in zip(got, eager[k]):
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f"replay {rep} view {k}"
    ref = oracle(scenes[1])
    assert_same(*eager[1], ref, label="moved camera")


def test_bin_capacity_regrowth_large_triangles():
    """A fresh context whose first frame has far more bin entries than its
    initial capacity (40 overlapping screen-filling quads: 80 wide
    triangles): the frame reports the overflow, regrows and is retried, and
    the result is still the reference's."""
    import math
    from draw_amd.render import Renderer
    from draw_amd.scene import Camera, Scene
    rng = np.random.default_rng(17)
    tex = workloads._phong(workloads.noise_texture(64, rng))
    objs = []
    for k in range(40):
        z = -0.5 - 0.75 * k                      # distance 3.5 .. 33 from the camera
        half = (3.0 - z) * math.tan(math.radians(30.0)) * float(rng.uniform(0.9, 1.4))
        objs.append(workloads.window_object(1, half, z, tex, x0=float(rng.uniform(-0.3, 0.3)),
                                            y0=float(rng.uniform(-0.3, 0.3))))
    s = Scene(320, 240, Camera(position=(0.0, 0.0, 3.0), fov_x=60.0, near=-0.1, far=-100.0), (0.0, 6.0, 6.0),
              objs)
    ref = oracle(s)
    r = Renderer(0)
    try:
        for mode in (2, 1):
            r.set_bin_order(mode)
            r.set_raster(0, 16, 4)
            canvas, winner = r.render(s)
            assert_same(canvas.frame, canvas.depth, winner, ref, label=f"large quads, bin order {mode}")
            if mode == 2:
                st = r.stats()
                assert st["retries"] >= 1, st
                assert st["bin_entries"] > 4 * (st["tris_out"] + 1024), st
    finally:
        r.close()


def test_canvas_without_winners(renderer):
    """A canvas created with DH_CANVAS_NO_WINNER (the reference's Canvas has
    no winner buffer) gets the same frame and depth bytes, in a plain frame
    and in a batch; asking it for winners is an argument error."""
    import ctypes as C
    s = workloads.soup_scene(7, n_tri=20_000, width=320, height=240, tex_size=256, cross_frac=0.05)
    items = [(workloads.torus_scene(1, width=160, height=120), None), (s, None)]
    fr = frame_uniforms(s)
    f1, d1, w1 = renderer.read(renderer.render_device(s, fr), fr.W, fr.H)
    b1 = [renderer.read(cv, it[0].width, it[0].height) for cv, it in zip(renderer.render_batch(items), items)]
    renderer.winners = False
    try:
        cv = renderer.render_device(s, fr)
        f2, d2, w2 = renderer.read(cv, fr.W, fr.H)
        b2 = [renderer.read(c2, it[0].width, it[0].height)
              for c2, it in zip(renderer.render_batch(items, prepared=renderer.batch_views(items, slot_base=100)), items)]
        win = np.empty((fr.H, fr.W), np.uint32)
        assert renderer.L.dh_canvas_read(renderer.ctx, cv, None, None, win.ctypes.data) == -1
    finally:
        renderer.winners = True
    assert w2 is None and w1 is not None
    assert np.array_equal(f1, f2) and np.array_equal(d1.view(np.uint32), d2.view(np.uint32))
    for (fa, da, _), (fb, db, wb) in zip(b1, b2):
        assert wb is None
        assert np.array_equal(fa, fb) and np.array_equal(da.view(np.uint32), db.view(np.uint32))


@pytest.mark.parametrize("n_tri", [16384, 16385])
def test_small_frame_limit(renderer, n_tri):
    """16384 input triangles is the largest frame the single-workgroup
    geometry and binning kernels take; one more takes the general path.  Both
    match the oracle."""
    s = workloads.soup_scene(11, n_tri=n_tri, width=320, height=240, tex_size=256)
    ref = oracle(s)
    f, d, w = _gpu(renderer, s, 32, 8)
    assert_same(f, d, w, ref, label=f"soup of {n_tri}")
    assert renderer.stats()["tris_out"] == ref.stats["tris_emitted"]


@pytest.mark.parametrize("kind,tile,waves", [(0, 16, 4), (1, 32, 4)])
def test_small_frame_capacity_regrowth(kind, tile, waves):
    """A fresh context's first small frame (16000 input triangles: the
    single-workgroup geometry and binning kernels) emits more clipped
    triangles than its initial capacity, and their bin entries overflow the
    pair buffers: the frame reports the overflow, regrows, is retried and
    matches the oracle, in both bin orders."""
    from draw_amd.render import Renderer
    s = workloads.sliver_scene(23)
    ref = oracle(s)
    assert ref.stats["tris_emitted"] > 16000 + 16000 // 4 + 1024
    for mode in (2, 1):
        r = Renderer(0)
        try:
            r.set_bin_order(mode)
            r.set_raster(kind, tile, waves)
            canvas, winner = r.render(s)
            assert_same(canvas.frame, canvas.depth, winner, ref, label=f"slivers, bin order {mode}")
            st = r.stats()
            assert st["retries"] >= 1, st
            assert st["tris_out"] == ref.stats["tris_emitted"], st
        finally:
            r.close()


@pytest.mark.parametrize("screen,kind,tile,waves", [((1920, 1080), 0, 32, 8), ((2560, 1440), 0, 16, 4),
                                                    ((1920, 1080), 1, 32, 4)])
def test_reference_app_configuration_offset_canvas(renderer, screen, kind, tile, waves):
    """The reference application's own set-up (app/mod.rs:57-84): the scene at
    screen size (camera (0,0,150), near -10, far -510, fov_x 135, light
    (0,300,300)) drawn into an 800x600 canvas centred on the screen.  Most of
    the soup lies outside the canvas or crosses its border; the frame equals
    the oracle's on every canvas pixel, and batched views with offset
    canvases equal their standalone renders."""
    s = workloads.app_scene(6, screen=screen)
    fr = frame_uniforms(s)
    assert (fr.W, fr.H) == (800, 600)
    assert (fr.off_x, fr.off_y) == ((screen[0] - 800) // 2, (screen[1] - 600) // 2)
    ref = oracle(s)
    f, d, w = _gpu(renderer, s, tile, waves, kind)
    assert_same(f, d, w, ref, label=f"app config {screen} kind {kind}")
    assert renderer.stats()["tris_out"] == ref.stats["tris_emitted"]
    views = [s, workloads.app_scene(7, screen=(1280, 1024), n_tri=5_000)]
    cvs = renderer.render_batch([(v, None) for v in views])
    for k, (v, cv) in enumerate(zip(views, cvs)):
        got = renderer.read(cv, 800, 600)
        c, w1 = renderer.render(v)
        assert np.array_equal(got[0], c.frame) and np.array_equal(got[2], w1), f"batched view {k}"

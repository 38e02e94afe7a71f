"""GUI overlay (f4).  The oracle's or_gui_draw is pinned against an
independent numpy float32 restatement on random draw lists, closed forms
(each pixel of a quad exactly once, clip rectangles, the depth-0 test), and
the error paths of the reference (index panics, get_rgba_slice panics)."""
import numpy as np
import pytest

import numpy_ref as NR
from draw_amd import gui
from oracle import oracle as O


def _canvas(W, H, seed=0):
    rng = np.random.default_rng(seed)
    bgra = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    depth = rng.uniform(0.5, 10, (H, W)).astype(np.float32)
    return bgra, depth


@pytest.mark.parametrize("seed", range(4))
def test_oracle_equals_numpy_restatement(seed):
    W, H = 96, 64
    v, i, c = gui.gui_drawlist(seed, W, H, n_quads=24, n_cmds=3)
    font = gui.font_atlas(seed)
    b1, d = _canvas(W, H, seed)
    d[5:9, 10:30] = 0.0                                  # depth 0: `0 < depth` fails there
    b2 = b1.copy()
    assert O.gui_draw(b1, d, v, i, c, font) == 0
    assert NR.gui_draw(b2, d, v, i, c, font) == 0
    assert np.array_equal(b1, b2)
    assert not np.array_equal(b1, _canvas(W, H, seed)[0]), "nothing was drawn"


def _quad(x0, y0, x1, y1, col=(10, 200, 30, 255), uv=(0.01, 0.01)):
    v = np.zeros(4, gui.GUI_VERTEX)
    for k, (x, y) in enumerate([(x0, y0), (x1, y0), (x1, y1), (x0, y1)]):
        v[k] = ((x, y), uv, col)
    return v, np.array([0, 1, 2, 0, 2, 3], np.uint32)


def test_quad_writes_each_pixel_of_its_rectangle_once():
    W, H = 40, 30
    v, i = _quad(5.2, 3.7, 17.4, 11.1)
    font = gui.font_atlas(1)
    cmd = np.zeros(1, gui.GUI_CMD)
    cmd[0] = ((0, 0, W, H), 0, 6, 0, 0)
    b = np.zeros((H, W, 4), np.uint8)
    b[:, :, 3] = 255
    d = np.ones((H, W), np.float32)
    assert O.gui_draw(b, d, v, i, cmd, font) == 0
    written = b[:, :, 3] == 0
    b[:, :, 3] = 0
    # snapped corners: x 5..17, imgui y 4..11 -> rows; the two triangles share the diagonal
    xs = np.where(written.any(0))[0]
    ys = np.where(written.any(1))[0]
    assert written.sum() == (xs.max() - xs.min() + 1) * (ys.max() - ys.min() + 1), "holes or overlaps"
    # texel alpha 255, vertex alpha 1: three truncated barycentric terms (up to 2 below the
    # colour), and an interpolated alpha that may round below 1 adds one truncating blend
    # step over the black background (1 more)
    got = set(map(tuple, b[written][:, :3].tolist()))
    assert got <= {(30 - k, 200 - j, 10 - m) for k in range(4) for j in range(4) for m in range(4)}
    assert (30, 200, 10) in got or (29, 199, 9) in got


def test_clip_rectangle_limits_writes():
    W, H = 40, 30
    v, i = _quad(0, 0, W, H)
    font = gui.font_atlas(1)
    cmd = np.zeros(1, gui.GUI_CMD)
    cmd[0] = ((10, 5, 20, 12), 0, 6, 0, 0)
    b = np.full((H, W, 4), 255, np.uint8)
    d = np.ones((H, W), np.float32)
    assert O.gui_draw(b, d, v, i, cmd, font) == 0
    written = b[:, :, 3] == 0
    cols = np.where(written.any(0))[0]
    rows = np.where(written.any(1))[0]                   # canvas rows are top-down, like imgui
    assert (cols.min(), cols.max()) == (10, 20)
    assert (rows.min(), rows.max()) == (5 - 1, 12 - 1)   # clip y flipped with H - y, rows with H - 1 - y
    cmd[0]["clip_rect"] = (20, 5, 10, 12)                # inverted: the command is skipped
    b2 = np.full((H, W, 4), 255, np.uint8)
    assert O.gui_draw(b2, d, v, i, cmd, font) == 0 and np.all(b2 == 255)


def test_error_paths():
    W, H = 20, 20
    font = gui.font_atlas(2)
    d = np.ones((H, W), np.float32)
    v, i = _quad(2, 2, 10, 10)
    cmd = np.zeros(1, gui.GUI_CMD)
    cmd[0] = ((0, 0, W, H), 0, 6, 2, 0)                  # vertex offset pushes indices past the buffer
    assert O.gui_draw(np.zeros((H, W, 4), np.uint8), d, v, i, cmd, font) == -2
    cmd[0] = ((0, 0, W, H), 3, 6, 0, 0)                  # index slice past the buffer
    assert O.gui_draw(np.zeros((H, W, 4), np.uint8), d, v, i, cmd, font) == -2
    # u = 1 on the last stored texel row (imgui v = 1) reads past the image: get_rgba_slice panics
    v2, i2 = _quad(2, 2, 10, 10, uv=(1.0, 1.0))
    cmd[0] = ((0, 0, W, H), 0, 6, 0, 0)
    assert O.gui_draw(np.zeros((H, W, 4), np.uint8), d, v2, i2, cmd, font) == -3
    assert NR.gui_draw(np.zeros((H, W, 4), np.uint8), d, v2, i2, cmd, font) == -3
    # u = 1 on another row reads the first texel of the next row (release indexing), no panic
    v3, i3 = _quad(2, 2, 10, 10, uv=(1.0, 0.5))
    b3, b4 = np.zeros((H, W, 4), np.uint8), np.zeros((H, W, 4), np.uint8)
    assert O.gui_draw(b3, d, v3, i3, cmd, font) == 0 and NR.gui_draw(b4, d, v3, i3, cmd, font) == 0
    assert np.array_equal(b3, b4)


def test_record_layouts_match_the_abi():
    import ctypes as C
    from draw_amd import _lib
    assert gui.GUI_VERTEX.itemsize == C.sizeof(_lib.DhGuiVertex) == 20
    assert gui.GUI_CMD.itemsize == C.sizeof(_lib.DhGuiCmd) == 32
    assert O.GUI_VERTEX == gui.GUI_VERTEX and O.GUI_CMD == gui.GUI_CMD


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1])
def test_gpu_overlay_on_a_rendered_frame_bit_exact(renderer, seed):
    from draw_amd import workloads
    s = workloads.torus_scene(seed + 3, width=320, height=240)
    cv = renderer.render_device(s)
    v, i, c = gui.gui_drawlist(seed, s.width, s.height, n_quads=120, n_cmds=5)
    font = gui.font_atlas(seed)
    renderer.draw_gui(cv, v, i, c, font)
    f, d, w = renderer.read(cv, s.width, s.height)
    ref = O.render(s)
    assert O.gui_draw(ref.frame, ref.depth, v, i, c, font) == 0
    assert np.array_equal(f, ref.frame)
    assert np.array_equal(d.view(np.uint32), ref.depth.view(np.uint32))   # depth untouched by the GUI


@pytest.mark.gpu
def test_gpu_overlay_errors(renderer):
    from draw_amd._lib import DrawHipError
    from draw_amd import workloads
    s = workloads.torus_scene(1, width=64, height=48)
    cv = renderer.render_device(s)
    font = gui.font_atlas(2)
    v, i = _quad(2, 2, 10, 10, uv=(1.0, 1.0))
    cmd = np.zeros(1, gui.GUI_CMD)
    cmd[0] = ((0, 0, 64, 48), 0, 6, 0, 0)
    with pytest.raises(DrawHipError, match="DH_E_TEXCOORD"):
        renderer.draw_gui(cv, v, i, cmd, font)
    cmd[0] = ((0, 0, 64, 48), 0, 6, 1, 0)
    with pytest.raises(DrawHipError, match="DH_E_INDEX"):
        renderer.draw_gui(cv, v, i, cmd, font)

"""GUI overlay (Gui::render, app/gui.rs:382-478, over Canvas::draw_triangle,
canvas.rs:435-575): imgui-style draw data -- vertices (pos, uv, RGBA colour),
u32 indices, commands (clip rect, index offset/count, vertex offset) and an
RGBA font texture -- drawn onto a canvas by dh_draw_gui, and a seeded
synthetic draw-list generator."""
from __future__ import annotations

import numpy as np

from .scene import TextureMap

GUI_VERTEX = np.dtype([("pos", "<f4", (2,)), ("uv", "<f4", (2,)), ("col", "u1", (4,))])          # ImDrawVert
GUI_CMD = np.dtype([("clip_rect", "<f4", (4,)), ("idx_offset", "<u4"), ("count", "<u4"), ("vtx_offset", "<u4"),
                    ("pad", "<u4")])


def font_atlas(seed: int, w: int = 128, h: int = 64) -> TextureMap:
    """RGBA 'font atlas': white glyph cells with seeded coverage in alpha, one
    opaque white texel block (where imgui keeps its solid-fill texel)."""
    rng = np.random.default_rng(seed)
    img = np.full((h, w, 4), 255, np.uint8)
    img[..., 3] = (rng.random((h, w)) < 0.45) * rng.integers(0, 256, (h, w))
    img[:4, :4, 3] = 255
    return TextureMap(img)


def gui_drawlist(seed: int, W: int, H: int, n_quads: int = 200, n_cmds: int = 4):
    """Panels (solid quads sampling the opaque texel) and glyph quads (sampling
    atlas cells), colours with alpha, split over n_cmds commands with their own
    clip rectangles and vertex offsets, like an imgui draw list."""
    rng = np.random.default_rng(seed)
    vtx, idx, cmds = [], [], []
    per = max(1, n_quads // n_cmds)
    for c in range(n_cmds):
        v0 = len(vtx)
        i0 = len(idx)
        for q in range(per):
            x0, y0 = rng.uniform(-10, W), rng.uniform(-10, H)
            if q % 5 == 0:   # panel
                w, h = rng.uniform(20, W / 3), rng.uniform(10, H / 4)
                uv = [(0.5 / 128, 0.5 / 64)] * 4
            else:            # glyph
                w, h = rng.uniform(4, 14), rng.uniform(6, 18)
                gx, gy = rng.integers(0, 16), rng.integers(0, 8)
                u0, v0_ = gx / 16.0, gy / 8.0 + 1 / 512.0   # imgui v = 0 flips to 1: floor(v h) = h wraps
                u1, v1 = u0 + 1 / 16.0 - 1 / 512.0, v0_ + 1 / 8.0 - 1 / 512.0   # stay inside the atlas
                uv = [(u0, v0_), (u1, v0_), (u1, v1), (u0, v1)]
            col = rng.integers(0, 256, 4)
            col[3] = rng.choice([255, rng.integers(0, 256)])
            base = len(vtx) - v0
            for (px, py), t in zip([(x0, y0), (x0 + w, y0), (x0 + w, y0 + h), (x0, y0 + h)], uv):
                vtx.append(((px, py), t, tuple(int(x) for x in col)))
            idx += [base, base + 1, base + 2, base, base + 2, base + 3]
        cx0, cy0 = rng.uniform(-5, W / 2), rng.uniform(-5, H / 2)
        clip = (cx0, cy0, rng.uniform(cx0, W + 5), rng.uniform(cy0, H + 5))
        cmds.append((clip, i0, len(idx) - i0, v0, 0))
    v = np.zeros(len(vtx), GUI_VERTEX)
    for k, (p, t, c) in enumerate(vtx):
        v[k] = (p, t, c)
    cm = np.zeros(len(cmds), GUI_CMD)
    for k, c in enumerate(cmds):
        cm[k] = c
    return v, np.asarray(idx, np.uint32), cm

"""dh_canvas_read_rgba_noalloc: the export readback through the fixed device
staging array (no allocation) gives exactly what dh_canvas_read_rgba gives,
on two calls in a row (the second on a smaller canvas, so stale staging from
the first call would show), and refuses canvases beyond the array."""
import numpy as np
import pytest

from draw_amd import workloads
from draw_amd._lib import DrawHipError

pytestmark = pytest.mark.gpu


def _swap(bgra):
    return bgra[..., [2, 1, 0, 3]]


def test_read_rgba_noalloc_matches_read_rgba(renderer):
    scenes = [workloads.soup_scene(9, n_tri=40_000, width=1920, height=1080, tex_size=256),
              workloads.torus_scene(1)]
    for s in scenes:
        cv = renderer.render_device(s)
        want = renderer.read_rgba(cv, s.width, s.height)
        got = renderer.read_rgba_noalloc(cv, s.width, s.height)
        assert np.array_equal(got, want), f"{s.width}x{s.height}"
        bgra, _, _ = renderer.read(cv, s.width, s.height)
        assert np.array_equal(got, _swap(bgra))
    # the largest canvas the staging holds, then one pixel row more
    full = renderer.canvas(3840, 2160)
    a = renderer.read_rgba_noalloc(full, 3840, 2160)
    assert np.array_equal(a, renderer.read_rgba(full, 3840, 2160))
    over = renderer.canvas(3840, 2161)
    with pytest.raises(DrawHipError, match="DH_E_CAPACITY"):
        renderer.read_rgba_noalloc(over, 3840, 2161)

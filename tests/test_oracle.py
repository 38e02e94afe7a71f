"""Pins the CPU oracle (oracle/oracle.c) -- the reference has no tests or
golden vectors of its own, so each part of the oracle is pinned here by an
independent numpy-float32 restatement (numpy_ref.py), by closed forms of the
edge-function rasteriser and by invariants, chosen so that a dropped term, a
wrong sign, a swapped operand or a wrong tie rule fails one of them."""
import math

import numpy as np
import pytest

import numpy_ref as N
from draw_amd import workloads
from draw_amd.scene import Camera, IndexedMesh, Object, Scene, Texture, TextureMap
from oracle import oracle as O

F = np.float32
CLEAR = np.array([255, 186, 155, 255], np.uint8)


# ------------------------------------------------------------ full pipeline
@pytest.mark.parametrize("make", [
    lambda: workloads.torus_scene(1),
    lambda: workloads.torus_scene(2, power=32),
    lambda: workloads.small_random_scene(5, n_tri=300),
    lambda: workloads.small_random_scene(6, n_tri=400, width=61, height=97, cross_frac=0.5),
])
def test_oracle_equals_numpy_restatement(make):
    s = make()
    ref = O.render(s)
    assert ref.rc == 0
    f, d, w, nid = N.render(s)
    assert nid == ref.stats["tris_emitted"]
    assert np.array_equal(f, ref.frame)
    assert np.array_equal(d.view(np.uint32), ref.depth.view(np.uint32))
    assert np.array_equal(w, ref.winner)


def test_frame_uniforms_equal_numpy_restatement():
    rng = np.random.default_rng(0)
    for i in range(20):
        s = workloads.torus_scene(1, width=int(rng.integers(16, 4000)), height=int(rng.integers(16, 2200)))
        s.camera = workloads.random_camera(rng)
        s.camera.fov_x = float(rng.uniform(20, 150))
        fr = O.frame_uniforms(s)
        M, planes, cam, light = N.frame_uniforms(s)
        assert np.array_equal(np.array(fr.M[:], F).view(np.uint32), M.ravel().view(np.uint32))
        got = np.array([[fr.planes[p][k] for k in range(4)] for p in range(6)], F)
        assert np.array_equal(got.view(np.uint32), planes.view(np.uint32))


# ------------------------------------------------ a1 against the mathematics
def test_transform_maps_frustum_to_screen():
    """M = viewport*orth*persp*view: the look-at point lands on the screen
    centre ((W-1)/2, (H-1)/2), a point on the right frustum plane at the
    near-plane half-width lands on x = W - 0.5 (closed form in float64)."""
    W, H = 640, 480
    s = workloads.torus_scene(1, width=W, height=H)
    fr = O.frame_uniforms(s)
    M = np.array(fr.M[:], np.float64).reshape(4, 4)

    def proj(p):
        o = M @ np.array([*p, 1.0])
        return o[0] / o[3], o[1] / o[3]

    x, y = proj((0.0, 0.0, 0.0))
    assert abs(x - (W - 1) / 2) < 1e-3 and abs(y - (H - 1) / 2) < 1e-3
    t = math.tan(math.radians(30.0))
    d = 3.0 - 1.0                        # point at z=1: distance 2 from the camera at z=3
    x, y = proj((d * t, 0.0, 1.0))       # right edge of the view at that depth
    assert abs(x - ((W - 1) / 2 + W / 2)) < 2e-3
    x, y = proj((0.0, d * t * H / W, 1.0))  # top edge
    assert abs(y - ((H - 1) / 2 + H / 2)) < 2e-3
    # every plane keeps the interior point positive and the camera eye ray inside
    P = np.array([[fr.planes[p][k] for k in range(4)] for p in range(6)], np.float64)
    assert np.all(P[:, :3] @ np.array([0.0, 0.0, 1.0]) + P[:, 3] > 0)


def test_view_planes_closed_form():
    """Near plane at distance |near| along the view direction, far at |far|."""
    s = workloads.torus_scene(1)
    fr = O.frame_uniforms(s)
    P = np.array([[fr.planes[p][k] for k in range(4)] for p in range(6)], np.float64)
    n = P[:, :3] / np.linalg.norm(P[:, :3], axis=1, keepdims=True)
    k = P[:, 3] / np.linalg.norm(P[:, :3], axis=1)
    # camera at (0,0,3) looking down -z: near plane z = 3 - 0.1, far plane z = 3 - 100
    assert np.allclose(n[0], [0, 0, -1], atol=1e-6) and abs(k[0] - 2.9) < 1e-5
    assert np.allclose(n[1], [0, 0, 1], atol=1e-6) and abs(k[1] + (3 - 100)) < 1e-3


def test_camera_asserts_are_errors():
    s = workloads.torus_scene(1)
    s.camera.near = 0.1          # Camera::new asserts near < 0 (scene/mod.rs:328)
    with pytest.raises(RuntimeError):
        O.frame_uniforms(s)
    s = workloads.torus_scene(1)
    s.camera.fov_x = 180.0       # asserts fov_x < 180
    with pytest.raises(RuntimeError):
        O.frame_uniforms(s)


# --------------------------------------------------------------- a2 per-vertex
def test_vertex_visual_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(200):
        v, cam, light = (rng.standard_normal(3) * 10).astype(F), (rng.standard_normal(3) * 10).astype(F), \
            (rng.standard_normal(3) * 10).astype(F)
        out = O.vertex_visual(v, cam, light)
        depth = np.linalg.norm(v.astype(np.float64) - cam.astype(np.float64))
        assert abs(out[9] - depth) <= 2 * np.spacing(F(depth))
        lv = (v.astype(np.float64) - light) / np.linalg.norm(v.astype(np.float64) - light)
        assert np.allclose(out[0:3], lv, atol=1e-6)          # light points from the light to the vertex
        ev = (v.astype(np.float64) - cam) / depth
        assert np.allclose(out[3:6], ev, atol=1e-6)
        h = lv + ev
        assert np.allclose(out[6:9], h / np.linalg.norm(h), atol=1e-5)


# ------------------------------------------------------------- raster KATs
def _attr(x, y, depth=1.0, uv=(0.5, 0.5), n=(0, 0, 1), l=(0, 0, -1), h=(0, 0, 1)):
    return O.make_attr(x, y, depth, n, l, h, uv)


def _canvas(W, H):
    fr = O.blank_frame(W, H)
    bgra = np.empty((H, W, 4), np.uint8)
    depth = np.empty((H, W), F)
    winner = np.empty((H, W), np.uint32)
    O.lib().or_clear(fr, bgra.ctypes.data, depth.ctypes.data, winner.ctypes.data)
    return fr, bgra, depth, winner


@pytest.mark.parametrize("n", [1, 2, 3, 10, 37])
@pytest.mark.parametrize("flip", [False, True])
def test_right_triangle_pixel_count(n, flip):
    """Legs on x=0 and y=0: the point (-1,-1) lies outside both legs, so their
    pixels are excluded and the hypotenuse is included: n(n-1)/2 pixels."""
    fr, bgra, depth, winner = _canvas(64, 64)
    a, b, c = _attr(0, 0), _attr(n, 0), _attr(0, n)
    if flip:
        b, c = c, b
    rc, st = O.draw_triangle(fr, a, b, c, Texture(), bgra, depth, winner)
    assert rc == 0
    assert st["frags_covered"] == n * (n - 1) // 2
    ys, xs = np.nonzero(winner != 0xFFFFFFFF)
    assert np.all(xs >= 1) and np.all(ys >= 1) and np.all(xs + ys <= n)


def test_quad_covers_each_pixel_once():
    """A quad split along a diagonal covers (x1-x0)*(y1-y0) pixels, each once:
    the right and top edges are included, the left and bottom ones not."""
    rng = np.random.default_rng(3)
    for _ in range(30):
        x0, y0 = (int(v) for v in rng.integers(-5, 20, 2))
        x1, y1 = x0 + int(rng.integers(1, 30)), y0 + int(rng.integers(1, 30))
        fr, bgra, depth, winner = _canvas(64, 64)
        p0, p1, p2, p3 = _attr(x0, y0), _attr(x1, y0), _attr(x1, y1), _attr(x0, y1)
        tex = Texture()
        rc1, s1 = O.draw_triangle(fr, p0, p1, p2, tex, bgra, depth, winner, tid=1)
        rc2, s2 = O.draw_triangle(fr, p0, p2, p3, tex, bgra, depth, winner, tid=2)
        assert rc1 == 0 and rc2 == 0
        xs, ys = np.meshgrid(np.arange(64), np.arange(64))
        want = (xs > x0) & (xs <= x1) & (ys > y0) & (ys <= y1)
        assert s1["frags_covered"] + s2["frags_covered"] == int(want.sum())
        assert np.array_equal(winner != 0xFFFFFFFF, want)


def test_depth_at_vertex_pixel_is_vertex_depth():
    """At the vertex pixel alpha = f/f = 1 and beta = gama = 0 exactly."""
    rng = np.random.default_rng(4)
    for _ in range(50):
        da, db, dc = (F(v) for v in rng.uniform(0.1, 500, 3))
        fr, bgra, depth, winner = _canvas(32, 32)
        a, b, c = _attr(20, 20, da), _attr(2, 20, db), _attr(20, 3, dc)   # both edges at A include it
        rc, _ = O.draw_triangle(fr, a, b, c, Texture(), bgra, depth, winner)
        assert rc == 0
        assert depth[20, 20] == da


def test_strict_depth_test_first_submission_wins():
    fr, bgra, depth, winner = _canvas(16, 16)
    for tid in (7, 8):
        O.draw_triangle(fr, _attr(0, 0, 5.0), _attr(15, 0, 5.0), _attr(0, 15, 5.0), Texture(), bgra, depth, winner,
                        tid=tid)
    assert set(np.unique(winner)) == {7, 0xFFFFFFFF}
    O.draw_triangle(fr, _attr(0, 0, 4.0), _attr(15, 0, 4.0), _attr(0, 15, 4.0), Texture(), bgra, depth, winner, tid=9)
    assert set(np.unique(winner)) == {9, 0xFFFFFFFF}


def test_texel_rule_and_byte_conversion():
    """ks=0, ka=kd=1, white ambient: colour = u8(f32(f32(t/255)*255)) of the
    texel at (floor(u*(w-1)), h-1-floor(v*(h-1))), for every byte value."""
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (7, 11, 4), dtype=np.uint8)
    img[0, 0, :3] = [0, 1, 254]
    tex = Texture(ka=(1, 1, 1), kd=(1, 1, 1), ks=(0, 0, 0), map_kd=TextureMap(img))
    for (u, v) in [(0.0, 0.0), (1.0, 1.0), (0.37, 0.81), (0.999, 0.001), (0.5, 0.5)]:
        fr, bgra, depth, winner = _canvas(16, 16)
        at = [_attr(0, 0, uv=(u, v)), _attr(15, 0, uv=(u, v)), _attr(0, 15, uv=(u, v))]
        rc, st = O.draw_triangle(fr, *at, tex, bgra, depth, winner)
        assert rc == 0 and st["frags_covered"] > 0
        ui = int(np.floor(F(F(u) * F(10))))
        vi = 6 - int(np.floor(F(F(v) * F(6))))
        rgb = img[vi, ui, :3]
        want = [int(np.trunc(F(F(F(int(t)) / F(255)) * F(255)))) for t in rgb]
        ys, xs = np.nonzero(winner != 0xFFFFFFFF)
        px = bgra[15 - ys[0], xs[0]]
        assert list(px) == [want[2], want[1], want[0], 255]


def test_byte_round_trip_is_exact_in_binary32():
    """RN(RN(t/255)*255) == t for every byte t, so with kd=ka=1, ks=0 a texel
    comes back unchanged through normalized_as_vec3 / from_normalized_vec3."""
    got = [int(np.trunc(F(F(F(t) / F(255)) * F(255)))) for t in range(256)]
    assert got == list(range(256))


def test_backface_light_gives_full_ambient_term():
    """light . normal <= 0 -> s = 1; halfway perpendicular to the normal -> no
    specular: colour = kd*t/255 * (ka + ks) per channel."""
    img = np.full((1, 1, 3), 200, np.uint8)
    tex = Texture(ka=(0.1, 0.2, 0.3), kd=(0.7, 0.6, 0.5), ks=(0.2, 0.4, 0.05), map_kd=TextureMap(img))
    fr, bgra, depth, winner = _canvas(16, 16)
    at = [_attr(0, 0, n=(0, 0, 1), l=(0, 0, -1), h=(1, 0, 0)), _attr(15, 0, n=(0, 0, 1), l=(0, 0, -1), h=(1, 0, 0)),
          _attr(0, 15, n=(0, 0, 1), l=(0, 0, -1), h=(1, 0, 0))]
    rc, _ = O.draw_triangle(fr, *at, tex, bgra, depth, winner)
    assert rc == 0
    diff = F(F(200) / F(255))
    want = []
    for i in range(3):
        c_r = F(diff * F(tex.kd[i]))
        col = F(c_r * F(F(tex.ka[i]) + F(F(tex.ks[i]) * F(1))))
        want.append(int(np.trunc(F(col * F(255)))))
    assert list(bgra[15 - 1, 1]) == [want[2], want[1], want[0], 255]


def test_zero_area_triangle_writes_nothing():
    fr, bgra, depth, winner = _canvas(32, 32)
    rc, st = O.draw_triangle(fr, _attr(1, 1), _attr(10, 10), _attr(20, 20), Texture(), bgra, depth, winner)
    assert rc == 0 and st["frags_covered"] == 0 and st["px_tested"] > 0


def test_clear_frame():
    s = Scene(40, 30, Camera(position=(0, 0, 3), fov_x=60, near=-0.1, far=-100))
    r = O.render(s)
    assert r.rc == 0
    assert np.all(r.frame == CLEAR) and np.all(r.depth == F(100000.0)) and np.all(r.winner == 0xFFFFFFFF)


def test_offscreen_bbox_collapses_to_column_zero():
    """Rectangle::clip turns an empty axis into [0, 0] (canvas.rs:340-347): a
    triangle right of the screen still scans column 0 (and covers nothing)."""
    fr, bgra, depth, winner = _canvas(16, 16)
    rc, st = O.draw_triangle(fr, _attr(20, 2), _attr(30, 2), _attr(20, 12), Texture(), bgra, depth, winner)
    assert rc == 0 and st["frags_covered"] == 0 and st["px_tested"] == 11


# ---------------------------------------------------------------- a3/a4 KATs
def test_backface_cull():
    s = workloads.small_random_scene(9, n_tri=1)
    obj = s.objects[0]
    cam = np.array(s.camera.position, np.float64)
    p = np.array([[-0.2, -0.2, 0.0], [0.2, -0.2, 0.0], [0.0, 0.2, 0.0]], np.float32)   # CCW seen from +z
    obj.vertices = p
    obj.normals_vertices = np.tile([0, 0, 1], (3, 1)).astype(np.float32)
    obj.texture_vertices = np.full((3, 2), 0.5, np.float32)
    obj.meshes = [IndexedMesh(np.array([[0, 1, 2, 0, 1, 2, 0, 1, 2]], np.uint32), 0)]
    r = O.render(s)
    assert r.stats["tris_culled"] == 0 and r.stats["tris_emitted"] == 1
    obj.meshes = [IndexedMesh(np.array([[0, 2, 1, 0, 2, 1, 0, 2, 1]], np.uint32), 0)]
    r = O.render(s)
    assert r.stats["tris_culled"] == 1 and r.stats["tris_emitted"] == 0
    del cam


def _clip_setup():
    s = workloads.torus_scene(1)
    return O.frame_uniforms(s)


def _at0():
    return [O.make_attr(0, 0, 1.0 + i, (0, 0, 1), (0, 0, -1), (0, 0, 1), (0.1 * i, 0.2)) for i in range(3)]


def test_clip_inside_is_identity():
    fr = _clip_setup()
    p = [-0.1, 0.0, 0.0, 0.1, 0.0, 0.0, 0.0, 0.1, 0.0]
    n, pos, _ = O.clip(fr, p, _at0())
    assert n == 1 and np.array_equal(pos[0].ravel(), np.array(p, F))


@pytest.mark.parametrize("behind,count", [(1, 2), (2, 1), (3, 0)])
def test_clip_near_plane_counts(behind, count):
    """Near plane at z = 2.9 (camera z=3, near 0.1).  One vertex behind -> two
    triangles, two behind -> one, all behind -> none; new vertices sit on the
    plane, pulled by EPS along the edge (scene/mod.rs:715-720)."""
    fr = _clip_setup()
    zs = [2.95 if i < behind else 2.0 for i in range(3)]
    p = [-0.01, -0.01, zs[0], 0.01, -0.01, zs[1], 0.0, 0.01, zs[2]]
    n, pos, _ = O.clip(fr, p, _at0())
    assert n == count
    if n:
        z = pos[:, :, 2].ravel()
        assert np.all(z <= 2.9 + 1e-6)                    # nothing left behind the near plane
        assert np.any(np.abs(z - 2.9) < 1e-5)             # the cut lies on the plane


def test_clip_outputs_at_most_four():
    s = workloads.torus_scene(1)
    s.camera.far = -1.0           # near 0.1, far 1: a long triangle crosses both
    fr = O.frame_uniforms(s)
    rng = np.random.default_rng(6)
    seen = set()
    for _ in range(2000):
        z = rng.uniform(1.0, 3.5, 3)
        p = np.stack([rng.uniform(-0.2, 0.2, 3), rng.uniform(-0.2, 0.2, 3), z], 1).ravel()
        n, _, _ = O.clip(fr, p, _at0())
        assert 0 <= n <= 4
        seen.add(n)
    assert {0, 1, 2, 3, 4} <= seen or {1, 2, 4} <= seen


# ------------------------------------------------------------------- errors
def test_index_out_of_range_is_an_error():
    s = workloads.small_random_scene(1, n_tri=5)
    s.objects[0].meshes[0].triangles[2, 6] = 10 ** 6     # normal index
    assert O.render(s).rc == -2


def test_texcoord_out_of_range_is_an_error():
    s = workloads.torus_scene(1)
    s.objects[0].texture_vertices = s.objects[0].texture_vertices * np.float32(3.0)
    assert O.render(s).rc == -3


# ------------------------------------------------------- transparent pass (f1)
def test_transparent_blend_formula():
    """bg*(1-a) + c*a with u8 truncation per term and wrapping u8 add
    (canvas.rs:136-169, 916-921); depth tested, not written."""
    img = np.full((1, 1, 3), 255, np.uint8)
    tex = Texture(ka=(1, 1, 1), kd=(1, 1, 1), ks=(0, 0, 0), alpha=0.3, map_kd=TextureMap(img))
    fr, bgra, depth, winner = _canvas(8, 8)
    at = [_attr(0, 0, 2.0), _attr(7, 0, 2.0), _attr(0, 7, 2.0)]
    rc, st = O.draw_triangle(fr, *at, tex, bgra, depth, winner, depth_update=False)
    assert rc == 0 and st["frags_written"] > 0
    a = F(0.3)
    want = [(int(np.trunc(F(F(bg) * F(F(1) - a)))) + int(np.trunc(F(F(255) * a)))) & 255 for bg in CLEAR[:3]]
    assert list(bgra[7 - 1, 1]) == want + [0]
    assert depth[1, 1] == F(100000.0)


# ------------------------------------------ hand-computed pins of sub-terms
def _dot(a, b):
    """Vec3::dot, linalg.rs:182-184: (x0*y0 + x1*y1) + x2*y2 in binary32."""
    a = [F(v) for v in a]
    b = [F(v) for v in b]
    return F(F(F(a[0] * b[0]) + F(a[1] * b[1])) + F(a[2] * b[2]))


def _unit(v):
    v = np.asarray(v, np.float64)
    return tuple(F(x) for x in v / np.linalg.norm(v))


@pytest.mark.parametrize("power", [2, 32])
def test_specular_term_hand_computed(power):
    """At vertex pixel A (alpha = 1, beta = gama = 0 exactly) the lerped N, L, H
    are A's own, so the colour follows canvas.rs:713-741 step by step:
    col = c_r*(c_a + c_l*(1 - max(0, L.N))) + c_l*(H.N)^power, with H.N far
    from 0 so that the specular term moves every channel.  (H.N)^32 is
    repeated squaring in binary32, the form integer powers take."""
    t = 180
    img = np.full((1, 1, 3), t, np.uint8)
    tex = Texture(ka=(0.1, 0.15, 0.05), kd=(0.7, 0.5, 0.6), ks=(0.2, 0.3, 0.25), power=power, map_kd=TextureMap(img))
    n = _unit((0.1, 0.2, 1.0))
    l = _unit((0.3, -0.4, 0.8))
    h = _unit((0.05, 0.1, 1.0))
    other = dict(n=_unit((1, 0, 0)), l=_unit((0, 1, 0)), h=_unit((0, 0, -1)))
    fr, bgra, depth, winner = _canvas(32, 32)
    a = _attr(20, 20, 1.0, n=n, l=l, h=h)
    b, c = _attr(2, 20, 1.0, **other), _attr(20, 3, 1.0, **other)
    rc, _ = O.draw_triangle(fr, a, b, c, tex, bgra, depth, winner)
    assert rc == 0 and winner[20, 20] == 0
    diff = F(F(t) / F(255))
    dln = _dot(l, n)
    s = F(F(1) - max(dln, F(0)))
    hn = _dot(h, n)
    if power == 2:
        h2 = F(hn * hn)
    else:
        h2 = hn
        for _ in range(5):
            h2 = F(h2 * h2)
    assert F(0.5) < hn and F(0) < dln < F(1)
    want, no_spec = [], []
    for i in range(3):
        c_r = F(diff * F(tex.kd[i]))
        t1 = F(c_r * F(F(tex.ka[i]) + F(F(tex.ks[i]) * s)))
        want.append(int(np.trunc(F(F(t1 + F(F(tex.ks[i]) * h2)) * F(255)))))
        no_spec.append(int(np.trunc(F(t1 * F(255)))))
    got = list(bgra[31 - 20, 20])
    assert got == [want[2], want[1], want[0], 255]
    assert all(w != x for w, x in zip(want, no_spec)), "the specular term must change every channel"


def _near_plane():
    fr = _clip_setup()
    return fr, tuple(F(fr.planes[0][k]) for k in range(3)), F(fr.planes[0][3])


def _clip_case():
    """A triangle with vertex O behind the near plane (z = 2.9) and P, Q in
    front; attributes deliberately unlike any recomputed value (depths 1, 2,
    3; light and halfway unit vectors far apart; uv distinct)."""
    O_ = (F(0.0137), F(-0.0211), F(2.9531))
    P = (F(-0.0419), F(0.0087), F(2.1173))
    Q = (F(0.0311), F(0.0243), F(1.7351))
    attrs = [O.make_attr(0, 0, d, _unit(nv), _unit(lv), _unit(hv), uv) for d, nv, lv, hv, uv in [
        (1.0, (0, 0, 1), (1, 0, 0), (0, 1, 0), (0.125, 0.25)),
        (2.0, (0, 1, 0), (0, 0, 1), (1, 0, 0), (0.75, 0.5)),
        (3.0, (1, 0, 0), (0, 1, 0), (0, 0, 1), (0.375, 0.875))]]
    return O_, P, Q, attrs


def _lerp_edge(p, o, fp, nrm, eps):
    """scene/mod.rs:715-717: t = f(p) / n.(p - o) - EPS; p + (o - p)*t."""
    d = tuple(F(p[i] - o[i]) for i in range(3))
    t = F(F(fp / _dot(nrm, d)) - eps)
    return t, tuple(F(p[i] + F(F(o[i] - p[i]) * t)) for i in range(3))


def test_clip_eps_nudge_hand_computed():
    """Near-plane split of a triangle whose first vertex is behind the plane:
    the odd vertex becomes c (rotation, scene/mod.rs:691-711), and the new
    vertices are P + (O - P)*t_P and Q + (O - Q)*t_Q with t = f/n.(v - O) - EPS
    (EPS = 1e-7, linalg.rs:6).  The case is chosen so that t and t - EPS give
    different binary32 vertices: dropping EPS fails."""
    fr, nrm, k = _near_plane()
    O_, P, Q, attrs = _clip_case()
    f = {name: F(_dot(nrm, v) + k) for name, v in (("O", O_), ("P", P), ("Q", Q))}
    assert f["O"] <= 0 < f["P"] and f["Q"] > 0
    eps = F(1e-7)
    tP, Ap = _lerp_edge(P, O_, f["P"], nrm, eps)
    tQ, Bp = _lerp_edge(Q, O_, f["Q"], nrm, eps)
    _, Ap0 = _lerp_edge(P, O_, f["P"], nrm, F(0))
    _, Bp0 = _lerp_edge(Q, O_, f["Q"], nrm, F(0))
    assert Ap != Ap0 and Bp != Bp0, "case must distinguish t from t - EPS"
    n, pos, _ = O.clip(fr, [*O_, *P, *Q], attrs)
    assert n == 2
    want = np.array([[P, Ap, Bp], [P, Q, Bp]], F)
    assert np.array_equal(pos.view(np.uint32), want.view(np.uint32))


def test_clipped_attributes_are_lerped_not_recomputed():
    """The new vertices' attributes are a_attr + (c_attr - a_attr)*t
    component-wise (canvas.rs:242-291, scene/mod.rs:716-720): depth, light and
    halfway are lerped, never recomputed from the new position nor
    renormalised."""
    fr, nrm, k = _near_plane()
    O_, P, Q, attrs = _clip_case()
    fP = F(_dot(nrm, P) + k)
    tP, Ap = _lerp_edge(P, O_, fP, nrm, F(1e-7))
    aO, aP = attrs[0], attrs[1]
    n, _, out = O.clip(fr, [*O_, *P, *Q], attrs)
    new_a = out[1]                      # second vertex of [P, A', B']

    def lerp(x, y):
        return F(F(x) + F(F(F(y) - F(x)) * tP))
    assert new_a.depth == lerp(aP.depth, aO.depth)
    for field in ("normal", "light", "halfway"):
        got = [F(v) for v in getattr(new_a, field)]
        want = [lerp(getattr(aP, field)[i], getattr(aO, field)[i]) for i in range(3)]
        assert got == want, field
    for i in range(2):
        assert F(new_a.uv[i]) == lerp(aP.uv[i], aO.uv[i])
    light = np.array([F(v) for v in new_a.light], np.float64)
    assert abs(np.linalg.norm(light) - 1.0) > 1e-3, "lerped light is not renormalised"
    cam = np.array(fr.cam[:], np.float64)
    assert abs(new_a.depth - np.linalg.norm(np.array(Ap, np.float64) - cam)) > 0.5, "depth is not recomputed"


# -------------------------------------- canvas offset (the reference app)
def test_frame_uniforms_canvas_closed_form():
    """Scene::new(screen) + Canvas::new(canvas) + apply_offset: M and the
    planes are the screen-size uniforms bit for bit; W, H are the canvas size
    and the offset is the integer offset as f32.  Canvas = screen with offset
    0 is or_frame_uniforms itself."""
    import copy
    rng = np.random.default_rng(21)
    for _ in range(20):
        sw, sh = int(rng.integers(640, 4000)), int(rng.integers(480, 2200))
        cw, ch = int(rng.integers(64, sw + 1)), int(rng.integers(64, sh + 1))
        ox, oy = int(rng.integers(-50, sw)), int(rng.integers(-50, sh))
        s = workloads.torus_scene(1, width=sw, height=sh)
        s.camera = workloads.random_camera(rng)
        full = O.frame_uniforms(s)
        s2 = copy.copy(s)
        s2.canvas, s2.canvas_offset = (cw, ch), (ox, oy)
        fr = O.frame_uniforms(s2)
        assert bytes(fr)[:64 + 24 + 96] == bytes(full)[:64 + 24 + 96]
        assert (fr.W, fr.H, fr.off_x, fr.off_y) == (cw, ch, float(ox), float(oy))
        assert fr.clear_depth == full.clear_depth and bytes(fr.clear_bgra) == bytes(full.clear_bgra)
        s2.canvas, s2.canvas_offset = (sw, sh), (0, 0)
        assert bytes(O.frame_uniforms(s2)) == bytes(full)


def test_canvas_offset_moves_the_triangle_and_its_tie_point():
    """pos_map_center(screen_coord - offset) (canvas.rs:585-587): a quad drawn
    at screen (p + offset + 0.25) into a canvas with that offset covers the
    canvas pixels the quad at (p + 0.25) covers with offset 0 -- including the
    tie rule, evaluated at canvas (-1, -1): right and top edges in, left and
    bottom out.  A flipped sign would move the quad off the canvas."""
    rng = np.random.default_rng(8)
    for _ in range(20):
        x0, y0 = (int(v) for v in rng.integers(-5, 20, 2))
        x1, y1 = x0 + int(rng.integers(1, 30)), y0 + int(rng.integers(1, 30))
        ox, oy = int(rng.integers(100, 3000)), int(rng.integers(100, 2000))
        fr, bgra, depth, winner = _canvas(64, 64)
        fr.off_x, fr.off_y = float(ox), float(oy)
        q = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
        p = [_attr(x + ox + 0.25, y + oy + 0.25) for x, y in q]
        O.draw_triangle(fr, p[0], p[1], p[2], Texture(), bgra, depth, winner, tid=1)
        O.draw_triangle(fr, p[0], p[2], p[3], Texture(), bgra, depth, winner, tid=2)
        xs, ys = np.meshgrid(np.arange(64), np.arange(64))
        want = (xs > x0) & (xs <= x1) & (ys > y0) & (ys <= y1)
        assert np.array_equal(winner != 0xFFFFFFFF, want)

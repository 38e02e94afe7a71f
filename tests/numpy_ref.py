"""Independent numpy-float32 restatement of the reference raster path.

Used only by tests to pin the C oracle.  It is written from the reference
sources (not from oracle.c), in numpy's vectorised style: each triangle's
bounding box is evaluated as whole float32 arrays.  numpy's float32 ufuncs
(add, subtract, multiply, divide, sqrt, floor) are IEEE binary32 operations
with no fused multiply-add, so the results are bit-comparable with rustc's.
``tan`` is taken from the platform libm (``tanf``), as the Rust std does.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math

import numpy as np

F = np.float32
_libm = ctypes.CDLL(ctypes.util.find_library("m"))
_libm.tanf.restype = ctypes.c_float
_libm.tanf.argtypes = [ctypes.c_float]


def tanf(x):
    return F(_libm.tanf(float(F(x))))


def v(*xs):
    return np.array(xs, dtype=F)


def norm(a):  # linalg.rs:167-171
    s = F(a[0] * a[0]) + F(a[1] * a[1])
    s = F(s + F(a[2] * a[2]))
    return F(np.sqrt(s))


def dot(a, b):  # linalg.rs:182-184
    return F(F(F(a[0] * b[0]) + F(a[1] * b[1])) + F(a[2] * b[2]))


def cross(a, b):  # linalg.rs:186-200
    return v(F(a[1] * b[2]) - F(a[2] * b[1]), F(a[2] * b[0]) - F(a[0] * b[2]), F(a[0] * b[1]) - F(a[1] * b[0]))


def matmul(a, b):  # linalg.rs:328-344
    c = np.zeros((4, 4), F)
    for i in range(4):
        for j in range(4):
            acc = F(0.0)
            for k in range(4):
                acc = F(acc + F(a[i, k] * b[k, j]))
            c[i, j] = acc
    return c


def matvec(a, x):  # linalg.rs:346-360
    out = np.zeros(4, F)
    for i in range(4):
        acc = F(0.0)
        for j in range(4):
            acc = F(acc + F(a[i, j] * x[j]))
        out[i] = acc
    return out


def frame_uniforms(scene):
    """Camera::new, gen_transformation_matrix, gen_view_planes."""
    W, H = scene.width, scene.height
    cam = scene.camera
    near, far, fov = F(cam.near), F(cam.far), F(cam.fov_x)
    ratio = F(F(W) / F(H))
    rad = F(fov * F(F(math.pi) / F(180.0)))
    right = F(abs(near) * tanf(F(rad / F(2.0))))
    left = -right
    top = F(F(F(1.0) / ratio) * right)
    bottom = -top
    pos = v(*cam.position)
    d = v(*cam.direction)
    d = (d / norm(d)).astype(F)
    up = v(0, 1, 0)

    def basis():
        w = ((d / norm(d)).astype(F) * F(-1.0)).astype(F)
        txw = cross(up, w)
        u = (txw / norm(txw)).astype(F)
        vv = cross(w, u)
        return (u / norm(u)).astype(F), (vv / norm(vv)).astype(F), (w / norm(w)).astype(F)

    u, vv, w = basis()
    B = np.array([[u[0], vv[0], w[0], 0], [u[1], vv[1], w[1], 0], [u[2], vv[2], w[2], 0], [0, 0, 0, 1]], F)
    T = np.array([[1, 0, 0, -pos[0]], [0, 1, 0, -pos[1]], [0, 0, 1, -pos[2]], [0, 0, 0, 1]], F)
    view = matmul(B.T.copy(), T)
    n, f, r, l, t, b = near, far, right, left, top, bottom
    persp = np.array([[n, 0, 0, 0], [0, n, 0, 0], [0, 0, F(n + f), -F(n * f)], [0, 0, 1, 0]], F)
    orth = np.array([[F(F(2) / F(r - l)), 0, 0, F(-F(r + l) / F(r - l))],
                     [0, F(F(2) / F(t - b)), 0, F(-F(t + b) / F(t - b))],
                     [0, 0, F(F(2) / F(n - f)), F(-F(n + f) / F(n - f))],
                     [0, 0, 0, 1]], F)
    nx, ny = F(W), F(H)
    vp = np.array([[F(nx / F(2)), 0, 0, F(F(nx - F(1)) / F(2))], [0, F(ny / F(2)), 0, F(F(ny - F(1)) / F(2))],
                   [0, 0, 1, 0], [0, 0, 0, 1]], F)
    M = matmul(matmul(matmul(vp, orth), persp), view)

    def world(p):
        o = matvec(B, np.array([p[0], p[1], p[2], 1], F))
        q = v(F(o[0] / o[3]), F(o[1] / o[3]), F(o[2] / o[3]))
        return (q + pos).astype(F)

    urn, uln, lrn, lln = world((r, t, n)), world((l, t, n)), world((r, b, n)), world((l, b, n))
    xc, yc = F(F(l + r) / F(2)), F(F(b + t) / F(2))
    ucf = world((xc, F(F(f * t) / n), f))
    lcf = world((xc, F(F(f * b) / n), f))
    rcf = world((F(F(f * r) / n), yc, f))
    lfc = world((F(F(f * l) / n), yc, f))
    vis = ((urn + lcf).astype(F) / F(2)).astype(F)

    def plane(a, bb, c):
        p = (bb - a).astype(F)
        q = (c - bb).astype(F)
        nn = cross(p, q)
        k = -dot(nn, a)
        if F(dot(nn, vis) + k) < 0:
            nn = cross(q, p)
            k = -dot(nn, a)
            assert F(dot(nn, vis) + k) > 0
        return np.array([nn[0], nn[1], nn[2], k], F)

    planes = np.stack([plane(urn, lrn, lln), plane(lfc, rcf, ucf), plane(rcf, urn, lrn), plane(lfc, lln, uln),
                       plane(ucf, uln, urn), plane(lcf, lln, lrn)])
    return M, planes, pos, v(*scene.light_source)


# ------------------------------------------------------------------ raster
def _usize(x):
    """Rust `f32 as usize`."""
    x = float(x)
    if not (x > 0):
        return 0
    if x >= 2.0 ** 64:
        return 2 ** 64 - 1
    return int(x)


def _u8(a):
    a = np.asarray(a, F)
    out = np.zeros(a.shape, np.uint8)
    ok = a > 0
    out[ok] = np.minimum(np.trunc(a[ok]), 255).astype(np.uint8)
    return out


M64 = (1 << 64) - 1


def sample_rgb(tex, u, vv):
    """TextureMap::get_rgb_slice (scene/mod.rs:154-168), release semantics, for
    arrays of u, v.  Returns (rgb uint8 (n,3), ok bool)."""
    w, h, comp = tex.width, tex.height, tex.components
    img = tex.img.reshape(-1)
    fu = np.floor((u * F(F(w) - F(1))).astype(F))
    fv = np.floor((vv * F(F(h) - F(1))).astype(F))
    out = np.zeros((len(u), 3), np.uint8)
    ok = True
    for i in range(len(u)):
        ui = _usize(fu[i])
        vi = ((h - 1) - _usize(fv[i])) & M64
        off = ((vi * w + ui) * comp) & M64
        end = (off + 3) & M64
        if off > end or end > img.size:
            ok = False
            break
        out[i] = img[off:off + 3]
    return out, ok


def draw_triangle(W, H, attrs, tex, frame, depthbuf, winner, tid, depth_update=True, off=(0.0, 0.0)):
    """Canvas::draw_triangle_with_attributes (canvas.rs:577-750).  attrs: three
    dicts with s (2), depth, n, l, h (3), uv (2).  Mutates frame/depth/winner;
    returns (ok, n_covered)."""
    with np.errstate(all="ignore"):
        cen = []
        for a in attrs:
            x = np.floor(F(F(F(a["s"][0]) + F(-F(off[0]))) + F(0.5)))
            y = np.floor(F(F(F(a["s"][1]) + F(-F(off[1]))) + F(0.5)))
            cen.append((F(x), F(y)))
        (ax, ay), (bx, by), (cx, cy) = cen

        def edge(px, py, qx, qy, x, y):
            return (((F(py - qy) * x).astype(F) + (F(qx - px) * y).astype(F)).astype(F) + F(px * qy)).astype(F) \
                - F(qx * py)

        def lo(*vals):
            r = F(np.inf)
            for q in vals:
                if q < r:
                    r = q
            return r

        def hi(*vals):
            r = F(-np.inf)
            for q in vals:
                if q > r:
                    r = q
            return r

        def axis(a0, a1, smax):
            mn, mx = min(a0, a1), max(a0, a1)
            mx = min(mx, smax)
            if mn > mx:
                mn = mx = 0
            return mn, mx

        x0, x1 = axis(_usize(lo(ax, bx, cx)), _usize(hi(ax, bx, cx)), W - 1)
        y0, y1 = axis(_usize(lo(ay, by, cy)), _usize(hi(ay, by, cy)), H - 1)
        fa = F(edge(bx, by, cx, cy, ax, ay))
        fb = F(edge(cx, cy, ax, ay, bx, by))
        fg = F(edge(ax, ay, bx, by, cx, cy))
        ta = F(fa * F(edge(bx, by, cx, cy, F(-1), F(-1)))) > 0
        tb = F(fb * F(edge(cx, cy, ax, ay, F(-1), F(-1)))) > 0
        tg = F(fg * F(edge(ax, ay, bx, by, F(-1), F(-1)))) > 0
        ys, xs = np.mgrid[y0:y1 + 1, x0:x1 + 1]
        xf = xs.astype(F).ravel()
        yf = ys.astype(F).ravel()
        al = (edge(bx, by, cx, cy, xf, yf) / fa).astype(F)
        be = (edge(cx, cy, ax, ay, xf, yf) / fb).astype(F)
        ga = (edge(ax, ay, bx, by, xf, yf) / fg).astype(F)
        cov = (al >= 0) & (be >= 0) & (ga >= 0) & ((al > 0) | ta) & ((be > 0) | tb) & ((ga > 0) | tg)
        if not cov.any():
            return True, 0
        al, be, ga = al[cov], be[cov], ga[cov]
        px, py = xs.ravel()[cov], ys.ravel()[cov]
        A, B, Cc = attrs
        d = ((al * F(A["depth"])).astype(F) + (be * F(B["depth"])).astype(F)).astype(F) + (ga * F(Cc["depth"])).astype(F)
        d = d.astype(F)

        def lerp(key, i):
            return (((al * F(A[key][i])).astype(F) + (be * F(B[key][i])).astype(F)).astype(F)
                    + (ga * F(Cc[key][i])).astype(F)).astype(F)

        u, vv = lerp("uv", 0), lerp("uv", 1)
        dk, ok1 = sample_rgb(tex.map_kd, u, vv)
        ak, ok2 = sample_rgb(tex.map_ka, u, vv)
        if not (ok1 and ok2):
            return False, 0
        diff = (dk.astype(F) / F(255)).astype(F)
        amb = (ak.astype(F) / F(255)).astype(F)
        N = [lerp("n", i) for i in range(3)]
        L = [lerp("l", i) for i in range(3)]
        Hh = [lerp("h", i) for i in range(3)]

        def dot3(p, q):
            return (((p[0] * q[0]).astype(F) + (p[1] * q[1]).astype(F)).astype(F) + (p[2] * q[2]).astype(F)).astype(F)

        dln = dot3(L, N)
        mx = np.where(dln > 0, dln, F(0)).astype(F)
        s = (F(1) - mx).astype(F)
        hn = dot3(Hh, N)
        h2 = F(1) if tex.power == 0 else None
        # powi by binary exponentiation (low bit first)
        p, base, res = int(tex.power), hn.copy(), None
        while p:
            if p & 1:
                res = base.copy() if res is None else (res * base).astype(F)
            p >>= 1
            if p:
                base = (base * base).astype(F)
        h2 = np.ones_like(hn) if res is None else res
        ks, kd, ka = [np.asarray(x, F) for x in (tex.ks, tex.kd, tex.ka)]
        col = []
        for i in range(3):
            c_r = (diff[:, i] * kd[i]).astype(F)
            c_a = (amb[:, i] * ka[i]).astype(F)
            inner = (c_a + (F(ks[i]) * s).astype(F)).astype(F)
            col.append(((c_r * inner).astype(F) + (F(ks[i]) * h2).astype(F)).astype(F))
        r8, g8, b8 = [_u8((c * F(255)).astype(F)) for c in col]
        # draw_pixel_coord_with_depth (canvas.rs:906-930), sequential for ties
        alpha = F(tex.alpha)
        for k in range(len(px)):
            x, y = int(px[k]), int(py[k])
            if alpha < 1:
                bg = frame[H - 1 - y, x]
                ia = F(F(1) - alpha)
                nc = [(int(_u8(F(F(bg[i]) * ia))) + int(_u8(F(F(c) * alpha)))) & 255
                      for i, c in enumerate((b8[k], g8[k], r8[k]))] + [0]
            else:
                nc = [b8[k], g8[k], r8[k], 255]
            if d[k] < depthbuf[y, x]:
                frame[H - 1 - y, x] = nc
                if depth_update:
                    depthbuf[y, x] = d[k]
                    if winner is not None:
                        winner[y, x] = tid
        return True, int(len(px))


# --------------------------------------------------------------- the scene
def render(scene):
    """Scene::render's opaque pass for a scene without transparent meshes.
    Returns frame (H,W,4) top-down, depth (H,W) bottom-up, winner, n_emitted."""
    M, planes, cam, light = frame_uniforms(scene)
    W, H = scene.width, scene.height
    frame = np.empty((H, W, 4), np.uint8)
    frame[...] = np.array([255, 186, 155, 255], np.uint8)
    depth = np.full((H, W), F(100000.0), F)
    winner = np.full((H, W), 0xFFFFFFFF, np.uint32)
    nid = 0
    with np.errstate(all="ignore"):
        for obj in scene.objects:
            P = obj.vertices.astype(F)
            ed = (P - cam).astype(F)
            ld = (P - light).astype(F)

            def vnorm(a):
                s = ((a[:, 0] * a[:, 0]).astype(F) + (a[:, 1] * a[:, 1]).astype(F)).astype(F)
                return np.sqrt((s + (a[:, 2] * a[:, 2]).astype(F)).astype(F)).astype(F)

            lv = (ld / vnorm(ld)[:, None]).astype(F)
            ev = (ed / vnorm(ed)[:, None]).astype(F)
            dep = vnorm(ed)
            hsum = (lv + ev).astype(F)
            hv = (hsum / vnorm(hsum)[:, None]).astype(F)
            for mesh in obj.opaque_meshes():
                tex = obj.material(mesh)
                for t9 in mesh.triangles:
                    pi, ti, ni = t9[0:3], t9[3:6], t9[6:9]
                    pts = [P[i] for i in pi]
                    at = [dict(s=(F(0), F(0)), depth=dep[pi[k]], n=obj.normals_vertices[ni[k]].astype(F),
                               l=lv[pi[k]], h=hv[pi[k]], uv=obj.texture_vertices[ti[k]].astype(F)) for k in range(3)]
                    nn = cross((pts[1] - pts[0]).astype(F), (pts[2] - pts[1]).astype(F))
                    ctr = ((((v(0, 0, 0) + pts[0]).astype(F) + pts[1]).astype(F) + pts[2]).astype(F) / F(3)).astype(F)
                    if dot((cam - ctr).astype(F), nn) <= 0:
                        continue
                    for cp, ca in clip(planes, pts, at):
                        for k in range(3):
                            o = matvec(M, np.array([cp[k][0], cp[k][1], cp[k][2], 1], F))
                            ca[k] = dict(ca[k], s=(F(o[0] / o[3]), F(o[1] / o[3])))
                        ok, _ = draw_triangle(W, H, ca, tex, frame, depth, winner, nid)
                        nid += 1
                        if not ok:
                            raise RuntimeError("texture index out of range")
    return frame, depth, winner, nid


def clip(planes, pts, at):
    """Triangle::clip_against_planes (scene/mod.rs:43-90, 641-746)."""
    EPS = F(0.0000001)

    def func(pl, p):
        return F(dot(pl[:3], p) + pl[3])

    pool = []
    if all(not (func(planes[i], pts[0]) <= 0 and func(planes[i], pts[1]) <= 0 and func(planes[i], pts[2]) <= 0)
           for i in (2, 3, 4, 5)):
        pool = [(list(pts), list(at))]
    for pl in (planes[0], planes[1]):
        nxt = []
        for (p, a) in pool:
            f = [func(pl, q) for q in p]
            if f[0] > 0 and f[1] > 0 and f[2] > 0:
                nxt.append((p, a))
                continue
            if f[0] <= 0 and f[1] <= 0 and f[2] <= 0:
                continue
            order = [0, 1, 2]
            if F(f[0] * f[2]) >= 0:
                order = [2, 0, 1]
            elif F(f[1] * f[2]) >= 0:
                order = [1, 2, 0]
            A, B, Cc = [p[i] for i in order]
            aA, aB, aC = [a[i] for i in order]
            fc = f[order[2]]

            def lerp_pt(x, y, t):
                return (x + ((y - x).astype(F) * t).astype(F)).astype(F)

            def lerp_at(x, y, t):
                r = {}
                r["s"] = tuple(F(F(x["s"][i]) + F(F(F(y["s"][i]) + F(-F(x["s"][i]))) * t)) for i in range(2))
                r["depth"] = F(F(x["depth"]) + F(F(F(y["depth"]) - F(x["depth"])) * t))
                for k in ("n", "l", "h", "uv"):
                    r[k] = (x[k] + ((y[k] - x[k]).astype(F) * t).astype(F)).astype(F)
                return r

            nrm = pl[:3]
            ta = F(F(func(pl, A) / dot(nrm, (A - Cc).astype(F))) - EPS)
            nA, naA = lerp_pt(A, Cc, ta), lerp_at(aA, aC, ta)
            tb = F(F(func(pl, B) / dot(nrm, (B - Cc).astype(F))) - EPS)
            nB, naB = lerp_pt(B, Cc, tb), lerp_at(aB, aC, tb)
            if fc <= 0:
                nxt.append(([A, nA, nB], [aA, naA, naB]))
                nxt.append(([A, B, nB], [aA, aB, naB]))
            else:
                nxt.append(([Cc, nA, nB], [aC, naA, naB]))
        pool = nxt
    return pool


# ---------------------------------------------------------------- GUI overlay
def _rect_from(x0, y0, x1, y1):  # Rectangle::from_coords canvas.rs:315-330 -> (x, y, w, h)
    return (min(x0, x1), min(y0, y1), max(x0, x1) - min(x0, x1), max(y0, y1) - min(y0, y1))


def _rect_clip(a, b):  # Rectangle::clip canvas.rs:332-350
    x0, y0 = max(a[0], b[0]), max(a[1], b[1])
    x1, y1 = min(a[0] + a[2], b[0] + b[2]), min(a[1] + a[3], b[1] + b[3])
    if x0 > x1:
        x0 = x1 = 0
    if y0 > y1:
        y0 = y1 = 0
    return _rect_from(x0, y0, x1, y1)


def _pmul(p, s):  # Pixel * f32: per channel (c as f32 * s) as u8
    return [int(_u8(F(c) * F(s))) for c in p]


def _padd(a, b):  # Pixel + Pixel: u8 wrapping
    return [(x + y) & 255 for x, y in zip(a, b)]


def gui_draw(bgra, depth, vtx, idx, cmds, font):
    """Gui::render (gui.rs:382-478) + Canvas::draw_triangle (canvas.rs:435-575),
    scalar float32; returns 0, -2 (index) or -3 (texture slice)."""
    H, W = depth.shape
    img = font.img.reshape(-1)
    fh, fw, fc = font.img.shape

    def ef(p, q, x, y):
        return ((F(p[1] - q[1]) * F(x) + F(q[0] - p[0]) * F(y)) + F(p[0] * q[1])) - F(q[0] * p[1])

    for cm in cmds:
        cr = [F(c) for c in cm["clip_rect"]]
        if cr[0] > cr[2] or cr[1] > cr[3]:
            continue
        cr[1] = F(H) - cr[1]
        cr[3] = F(H) - cr[3]
        clip = _rect_from(_usize(cr[0]), _usize(cr[1]), _usize(cr[2]), _usize(cr[3]))
        io, cnt, vo = int(cm["idx_offset"]), int(cm["count"]), int(cm["vtx_offset"])
        if io + cnt > len(idx):
            return -2
        for t in range(0, cnt - cnt % 3, 3):
            vs = []
            for k in range(3):
                vi = int(idx[io + t + k]) + vo
                if vi >= len(vtx):
                    return -2
                q = vtx[vi]
                s = (F(q["pos"][0]), (F(H) - F(1.0)) - F(q["pos"][1]))
                vs.append(dict(c=(np.floor(s[0] + F(0.5)), np.floor(s[1] + F(0.5))),
                               uv=(F(q["uv"][0]), F(1.0) - F(q["uv"][1])),
                               rgb=[int(x) for x in q["col"][:3]], a=F(q["col"][3]) / F(255.0)))
            a, b, c = vs[0]["c"], vs[1]["c"], vs[2]["c"]
            xs, ys = [a[0], b[0], c[0]], [a[1], b[1], c[1]]
            def mn(l):  # canvas.rs:491-500 (NaN never replaces the running value)
                r = F(np.inf)
                for z in l:
                    if z < r:
                        r = z
                return r

            def mx(l):
                r = F(-np.inf)
                for z in l:
                    if z > r:
                        r = z
                return r
            drawable = _rect_from(_usize(mn(xs)), _usize(mn(ys)), _usize(mx(xs)), _usize(mx(ys)))
            drawable = _rect_clip(drawable, _rect_from(0, 0, W - 1, H - 1))
            valid = _rect_clip(clip, drawable)
            fa, fb, fg = ef(b, c, a[0], a[1]), ef(c, a, b[0], b[1]), ef(a, b, c[0], c[1])
            oa, ob, og = ef(b, c, -1, -1), ef(c, a, -1, -1), ef(a, b, -1, -1)
            for y in range(valid[1], valid[1] + valid[3] + 1):
                for x in range(valid[0], valid[0] + valid[2] + 1):
                    al, be, ga = ef(b, c, x, y) / fa, ef(c, a, x, y) / fb, ef(a, b, x, y) / fg
                    if not (al >= 0 and be >= 0 and ga >= 0):
                        continue
                    if not ((al > 0 or fa * oa > 0) and (be > 0 or fb * ob > 0) and (ga > 0 or fg * og > 0)):
                        continue
                    col = _padd(_padd(_pmul(vs[0]["rgb"], al), _pmul(vs[1]["rgb"], be)), _pmul(vs[2]["rgb"], ga))
                    calpha = (al * vs[0]["a"] + be * vs[1]["a"]) + ga * vs[2]["a"]
                    u = (vs[0]["uv"][0] * al + vs[1]["uv"][0] * be) + vs[2]["uv"][0] * ga
                    vv = (vs[0]["uv"][1] * al + vs[1]["uv"][1] * be) + vs[2]["uv"][1] * ga
                    ui = _usize(np.floor(u * F(fw)))
                    vi = ((fh - 1) - _usize(np.floor(vv * F(fh)))) & M64
                    off = ((vi * fw + ui) * fc) & M64
                    if off + 4 > fw * fh * fc:
                        return -3
                    tex = [int(z) for z in img[off:off + 4]]
                    ta = F(tex[3]) / F(255.0)
                    col = _padd(_pmul(col, ta), _pmul(tex[:3], F(1.0) - ta))
                    fin = calpha * ta
                    px = bgra[H - 1 - y, x]
                    out = _padd(_pmul([int(px[2]), int(px[1]), int(px[0])], F(1.0) - fin), _pmul(col, fin)) \
                        if fin < F(1.0) else col
                    if F(0.0) < depth[y, x]:
                        px[0], px[1], px[2], px[3] = out[2], out[1], out[0], 0
    return 0

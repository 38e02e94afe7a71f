"""OBJ/MTL ingest (Object::load_from_file, object.rs:106-419).  Expected
values are computed here independently: the OBJ/MTL/PNG inputs are written by
the test, the post-processing (rescale to a largest norm of 100, normalised
normals, quad split, dummy texture coordinates, generated normals) is restated
in numpy float32 in the reference's operation order."""
import os
import struct
import zlib

import numpy as np
import pytest

from draw_amd._lib import DrawHipError
from draw_amd.obj import load_obj

F = np.float32
REF_MODELS = "/root/reference/models"


def _png(img: np.ndarray, filters=(0, 1, 2, 3, 4), palette=None, depth=8) -> bytes:
    """A PNG written from the specification: rows cycle through the given
    filter types.  img: (h, w, c) uint8 (or uint16 for depth 16); with a
    palette, img is (h, w) of indices."""
    h, w = img.shape[:2]
    if palette is not None:
        ctype, raw_rows = 3, [img[y].astype(np.uint8).tobytes() for y in range(h)]
        bpp = 1
    else:
        c = img.shape[2]
        ctype = {3: 2, 4: 6}[c]
        if depth == 16:
            raw_rows = [img[y].astype(">u2").tobytes() for y in range(h)]
            bpp = 2 * c
        else:
            raw_rows = [img[y].astype(np.uint8).tobytes() for y in range(h)]
            bpp = c
    out, prev = b"", bytes(len(raw_rows[0]))
    for y, row in enumerate(raw_rows):
        ft = filters[y % len(filters)]
        enc = bytearray()
        for i, x in enumerate(row):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            cc = prev[i - bpp] if i >= bpp else 0
            if ft == 0:
                p = 0
            elif ft == 1:
                p = a
            elif ft == 2:
                p = b
            elif ft == 3:
                p = (a + b) // 2
            else:
                pp = a + b - cc
                pa, pb, pc = abs(pp - a), abs(pp - b), abs(pp - cc)
                p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else cc)
            enc.append((x - p) & 255)
        out += bytes([ft]) + bytes(enc)
        prev = row

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if palette is not None:
        png += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).ravel()))
    return png + chunk(b"IDAT", zlib.compress(out)) + chunk(b"IEND", b"")


def _norm(v):
    return np.sqrt((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2])


def _write_model(tmp_path, fmt="png"):
    """The test model; its diffuse map as a PNG, or as a progressive JPEG
    (fmt "jpg", written by Pillow)."""
    rng = np.random.default_rng(3)
    tex = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)
    if fmt == "png":
        (tmp_path / "wood.png").write_bytes(_png(tex))
    else:
        from PIL import Image
        tex = np.kron(tex, np.ones((8, 8, 1), np.uint8))
        Image.fromarray(tex).save(tmp_path / "wood.jpg", quality=90, progressive=True)
    (tmp_path / "m.mtl").write_text(
        "# materials\nnewmtl Wood\nNs 10\nKa 0.1 0.2 0.3\nKd 0.7 0.6 0.5\nKs 0.25 0.25 0.25\nd 1.0\n"
        f"map_Kd wood.{fmt}\n"
        "\nnewmtl Glass\nKa 0 0 0\nKd 0.1 0.2 0.9\nKs 1 1 1\nd 0.35\n")
    (tmp_path / "m.obj").write_text(
        "mtllib m.mtl\no Thing\n"
        "v 1 0 0\nv 0 2 0\nv -1 0 0\nv 0 -2 0\nv 0 0 3.5\nv 0.5 0.5 -1\n"
        "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
        "vn 0 0 2\nvn 0 1 1\n"
        "g top\nusemtl Wood\n"
        "f 1/1/1 2/2/1 3/3/2 4/4/2\n"          # quad, all attributes
        "f -2/-1/-1 1/1/1 2/2/2\n"              # negative indices
        "g side\nusemtl Glass\n"
        "f 1 2 5\nf 2//2 3//2 5//2\n"           # missing uv (dummy), partly missing normals
        "usemtl Nothing\nf 3 4 6\n")            # unknown material -> default, split group
    return tex


def test_synthetic_model_matches_reference_rules(tmp_path):
    tex = _write_model(tmp_path)
    o = load_obj(str(tmp_path / "m.obj"))
    raw = np.array([[1, 0, 0], [0, 2, 0], [-1, 0, 0], [0, -2, 0], [0, 0, 3.5], [0.5, 0.5, -1]], F)
    factor = F(100.0) / _norm(raw).max()
    assert np.array_equal(o.vertices, raw * factor)
    # textures: default first, then the library in order
    assert [t.name for t in o.textures] == ["default", "Wood", "Glass"]
    assert o.textures[1].alpha == 1.0 and o.textures[2].alpha == F(0.35)
    assert np.array_equal(o.textures[1].map_kd.img, tex)
    assert o.textures[2].map_kd.width == 1          # no map: TextureMap::default
    assert tuple(o.textures[1].ka) == (F(0.1), F(0.2), F(0.3))
    # meshes: "top" (Wood), "side" (Glass), "side" again (default: unknown material)
    assert [(m.name, m.texture_idx) for m in o.meshes] == [("top", 1), ("side", 2), ("side", 0)]
    top = o.meshes[0].triangles
    assert top.tolist() == [[0, 1, 2, 0, 1, 2, 0, 0, 1], [2, 3, 0, 2, 3, 0, 1, 1, 0], [4, 0, 1, 3, 0, 1, 1, 0, 1]]
    # normals: the two file normals normalised, then one generated block per group lacking normals
    n_file = np.array([[0, 0, 2], [0, 1, 1]], F)
    assert np.array_equal(o.normals_vertices[:2], n_file / _norm(n_file)[:, None])
    side = o.meshes[1].triangles
    assert o.texture_vertices.shape[0] == 4 + 2                   # one dummy per group lacking uv
    assert np.all(o.texture_vertices[4:] == 0)
    assert side[:, 3:6].tolist() == [[4, 4, 4], [4, 4, 4]]
    # generated normals for "side": sums of (b-a)x(c-b) over the group's triangles, normalised
    v = o.vertices
    gen = np.zeros_like(v)
    for t in side:
        a, b, c = v[t[0]], v[t[1]], v[t[2]]
        p, q = b - a, c - b
        nrm = np.array([p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]], F)
        for k in range(3):
            gen[t[k]] = gen[t[k]] + nrm
    with np.errstate(invalid="ignore", divide="ignore"):
        gen_n = gen / _norm(gen)[:, None]
    block = o.normals_vertices[2:2 + v.shape[0]]
    np.testing.assert_array_equal(block, gen_n)          # NaN where a vertex is unused, as 0/0 in f32
    assert side[0, 6:].tolist() == [2 + 0, 2 + 1, 2 + 4]  # first face lacked normals
    assert side[1, 6:].tolist() == [1, 1, 1]              # second face had its own


@pytest.mark.parametrize("kind", ["rgb_filters", "rgba16", "palette4", "palette_trns"])
def test_png_decoding(tmp_path, kind):
    rng = np.random.default_rng(5)
    if kind == "rgb_filters":
        img = rng.integers(0, 256, (9, 11, 3), dtype=np.uint8)
        data, want = _png(img), img
    elif kind == "rgba16":
        img = rng.integers(0, 65536, (4, 6, 4), dtype=np.uint16)
        data, want = _png(img, depth=16), (img >> 8).astype(np.uint8)
    else:
        pal = rng.integers(0, 256, (16, 3), dtype=np.uint8)
        idx = rng.integers(0, 16, (6, 8), dtype=np.uint8)
        packed = np.array([[(r[2 * i] << 4) | r[2 * i + 1] for i in range(4)] for r in idx], np.uint8)
        data = _png(packed[..., None][..., 0], palette=pal, filters=(0, 1))
        # patch bit depth 4 into IHDR (the rows above are already nibble-packed)
        ihdr = struct.pack(">IIBBBBB", 8, 6, 4, 3, 0, 0, 0)
        start = 8 + 8
        data = data[:start] + ihdr + struct.pack(">I", zlib.crc32(b"IHDR" + ihdr) & 0xFFFFFFFF) + data[start + 17:]
        want = pal[idx]
        if kind == "palette_trns":
            trns = rng.integers(0, 256, 16, dtype=np.uint8)
            t = b"tRNS" + bytes(trns)
            chunk = struct.pack(">I", 16) + t + struct.pack(">I", zlib.crc32(t) & 0xFFFFFFFF)
            k = data.index(b"IDAT") - 4
            data = data[:k] + chunk + data[k:]
            want = np.concatenate([want, trns[idx][..., None]], 2)
    (tmp_path / "t.png").write_bytes(data)
    (tmp_path / "t.mtl").write_text("newmtl M\nKa 1 1 1\nKd 1 1 1\nKs 0 0 0\nmap_Kd t.png\n")
    (tmp_path / "t.obj").write_text("mtllib t.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl M\nf 1 2 3\n")
    o = load_obj(str(tmp_path / "t.obj"))
    assert np.array_equal(o.textures[1].map_kd.img, want)


def test_loader_errors(tmp_path):
    (tmp_path / "a.obj").write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")
    with pytest.raises(DrawHipError, match="DH_E_INDEX"):
        load_obj(str(tmp_path / "a.obj"))
    (tmp_path / "b.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nv 2 2 0\nf 1 2 3 4 5\n")
    with pytest.raises(DrawHipError, match="3 or 4"):
        load_obj(str(tmp_path / "b.obj"))
    (tmp_path / "c.mtl").write_text("newmtl X\nKd 1 1 1\n")
    (tmp_path / "c.obj").write_text("mtllib c.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    with pytest.raises(DrawHipError, match="lacks"):
        load_obj(str(tmp_path / "c.obj"))
    # a damaged JPEG header, then an image format that is not decoded
    (tmp_path / "d.jpg").write_bytes(b"\xff\xd8\xff\xe0" + bytes(16))
    (tmp_path / "d.mtl").write_text("newmtl X\nKa 1 1 1\nKd 1 1 1\nKs 1 1 1\nmap_Kd d.jpg\n")
    (tmp_path / "d.obj").write_text("mtllib d.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    with pytest.raises(DrawHipError, match="DH_E_ARG"):
        load_obj(str(tmp_path / "d.obj"))
    (tmp_path / "d.jpg").write_bytes(b"GIF89a" + bytes(16))
    with pytest.raises(DrawHipError, match="DH_E_UNSUPPORTED"):
        load_obj(str(tmp_path / "d.obj"))


def test_png_header_larger_than_its_data_is_an_error(tmp_path):
    """An IHDR claiming 65535 x 65535 16-bit RGBA (34 GB decoded) over a few
    bytes of IDAT is refused before anything is allocated for it."""
    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    png = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", 65535, 65535, 16, 6, 0, 0, 0)) +
           chunk(b"IDAT", zlib.compress(bytes(64))) + chunk(b"IEND", b""))
    (tmp_path / "big.png").write_bytes(png)
    (tmp_path / "big.mtl").write_text("newmtl M\nKa 1 1 1\nKd 1 1 1\nKs 0 0 0\nmap_Kd big.png\n")
    (tmp_path / "big.obj").write_text("mtllib big.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl M\nf 1 2 3\n")
    with pytest.raises(DrawHipError, match="DH_E_ARG"):
        load_obj(str(tmp_path / "big.obj"))


@pytest.mark.skipif(not os.path.isdir(REF_MODELS), reason="the reference's model files are not on this machine")
def test_reference_donut_model():
    """models/donut/donut.obj: 1024 vertices, 2048 triangles, no normals and
    no texture coordinates -- generated normals and one dummy coordinate."""
    o = load_obj(os.path.join(REF_MODELS, "donut", "donut.obj"))
    assert o.vertices.shape == (1024, 3) and len(o.meshes) == 1 and o.meshes[0].triangles.shape[0] == 2048
    assert o.texture_vertices.shape[0] == 1 and o.normals_vertices.shape[0] == 1024
    assert _norm(o.vertices).max() == F(100.0) or abs(float(_norm(o.vertices).max()) - 100.0) < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", ["png", "jpg"])
def test_loaded_model_renders_bit_exact(renderer, tmp_path, fmt):
    from draw_amd.obj import upload_obj
    from draw_amd.scene import Camera, Scene
    from parity import assert_same, oracle
    _write_model(tmp_path, fmt)
    obj = load_obj(str(tmp_path / "m.obj"))
    s = Scene(640, 480, Camera(position=(30.0, 40.0, 150.0)), (0.0, 300.0, 300.0), [obj])
    c, w = renderer.render(s)
    assert_same(c.frame, c.depth, w, oracle(s), label="loaded OBJ")
    # the one-call native upload renders the same frame
    _, handles = upload_obj(renderer.ctx, str(tmp_path / "m.obj"))
    from draw_amd.render import frame_uniforms
    cv = renderer.render_device(s, frame_uniforms(s), handles)
    f2, d2, w2 = renderer.read(cv, 640, 480)
    assert np.array_equal(f2, c.frame) and np.array_equal(d2.view(np.uint32), c.depth.view(np.uint32))
    assert np.array_equal(w2, w)

"""Texture files (TextureMap::load_from_file, scene/mod.rs:174-202) through
dh_image_info / dh_image_load, the JPEG path in particular.  The reference
decodes with stb_image, which is not in the reference tree, so byte parity
with it is unpinned; the decoder is pinned instead by
  * closed forms: files written here from ITU-T T.81 with chosen coefficients
    (a flat block decodes to 128 + round(DC * q / 8); other blocks to the
    exact inverse DCT within 1 level);
  * invariants: progressive and restart-marked encodings of the same
    coefficients decode to the same pixels as the plain baseline file;
  * libjpeg (Pillow's decoder): the same files within a few levels, the
    rounding of the two fixed-point pipelines being different;
  * the reference's own texture, models/airplane/11804_Airplane_diff.jpg (a
    progressive file), when the reference tree is present."""
import io
import os

import numpy as np
import pytest

from draw_amd._lib import DrawHipError
from draw_amd.obj import load_obj
from draw_amd.scene import TextureMap

PIL = pytest.importorskip("PIL.Image")
REF_JPG = "/root/reference/models/airplane/11804_Airplane_diff.jpg"

# T.81 figure A.6: natural index of the k-th zig-zag coefficient
ZIGZAG = np.array([0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,
                   6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45,
                   38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63])


class _BitWriter:
    def __init__(self):
        self.out, self.acc, self.n = bytearray(), 0, 0

    def put(self, code, length):
        for i in range(length - 1, -1, -1):
            self.acc = (self.acc << 1) | ((code >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)  # byte stuffing, T.81 F.1.2.3
                self.acc, self.n = 0, 0

    def flush(self):
        while self.n:
            self.put(1, 1)  # pad with 1-bits
        return bytes(self.out)


def _seg(marker, body):
    return bytes([0xFF, marker]) + (len(body) + 2).to_bytes(2, "big") + body


def _jpeg_rgb(coefs, q, restart=0):
    """A baseline JPEG (T.81, Huffman, 8-bit) of three components coded as
    RGB (Adobe marker, transform 0, no JFIF marker): coefs[c, by, bx, 64] are
    the quantized coefficients in natural order, q[64] the one quantisation
    table (natural order).  Huffman tables: DC categories 0-15 as 5-bit codes,
    every AC run/size symbol as an 8-bit code."""
    nc, bh, bw, _ = coefs.shape
    dc_syms = list(range(16))
    ac_syms = [0x00, 0xF0] + [(r << 4) | s for r in range(16) for s in range(1, 11)]
    dc_code = {s: (i, 5) for i, s in enumerate(dc_syms)}
    ac_code = {s: (i, 8) for i, s in enumerate(ac_syms)}

    def dht(tc, syms, length):
        counts = [0] * 16
        counts[length - 1] = len(syms)
        return bytes([tc << 4]) + bytes(counts) + bytes(syms)

    def cat(v):
        return 0 if v == 0 else int(abs(v)).bit_length()

    def bits(v, s):
        return v if v >= 0 else v + (1 << s) - 1

    head = b"\xff\xd8" + _seg(0xEE, b"Adobe\x00\x64\x00\x00\x00\x00\x00")
    head += _seg(0xDB, bytes([0]) + bytes(int(q[ZIGZAG[k]]) for k in range(64)))
    sof = bytes([8]) + (bh * 8).to_bytes(2, "big") + (bw * 8).to_bytes(2, "big") + bytes([nc])
    for c in range(nc):
        sof += bytes([1 + c, 0x11, 0])
    head += _seg(0xC0, sof) + _seg(0xC4, dht(0, dc_syms, 5)) + _seg(0xC4, dht(1, ac_syms, 8))
    if restart:
        head += _seg(0xDD, restart.to_bytes(2, "big"))
    sos = bytes([nc]) + b"".join(bytes([1 + c, 0x00]) for c in range(nc)) + bytes([0, 63, 0])
    head += _seg(0xDA, sos)
    data, w, pred, mcu, rst = b"", _BitWriter(), [0] * nc, 0, 0
    for by in range(bh):
        for bx in range(bw):
            if restart and mcu and mcu % restart == 0:
                data += w.flush() + bytes([0xFF, 0xD0 + rst % 8])
                w, pred, rst = _BitWriter(), [0] * nc, rst + 1
            for c in range(nc):
                z = [int(coefs[c, by, bx, ZIGZAG[k]]) for k in range(64)]
                d = z[0] - pred[c]
                pred[c] = z[0]
                w.put(*dc_code[cat(d)])
                w.put(bits(d, cat(d)), cat(d))
                run = 0
                for k in range(1, 64):
                    if z[k] == 0:
                        run += 1
                        continue
                    while run > 15:
                        w.put(*ac_code[0xF0])
                        run -= 16
                    s = cat(z[k])
                    w.put(*ac_code[(run << 4) | s])
                    w.put(bits(z[k], s), s)
                    run = 0
                if run:
                    w.put(*ac_code[0x00])
            mcu += 1
    return head + data + w.flush() + b"\xff\xd9"


def _idct_exact(coef):
    """The 2-D inverse DCT of T.81 A.3.3 in float64, + 128, rounded and clamped."""
    x = np.arange(8)
    cu = np.where(x == 0, 1 / np.sqrt(2), 1.0)
    basis = cu[None, :] * np.cos((2 * x[:, None] + 1) * x[None, :] * np.pi / 16) / 2  # [x, u]
    f = basis @ coef.reshape(8, 8).astype(np.float64) @ basis.T  # rows y: [y, v] @ [v, u] @ [u, x]
    return np.clip(np.floor(f + 128 + 0.5), 0, 255).astype(np.int64)


def _decode(path):
    return TextureMap.load_from_file(path).img


def _to_image(coefs):
    nc, bh, bw, _ = coefs.shape
    return coefs.reshape(nc, bh, bw, 8, 8).transpose(1, 3, 2, 4, 0).reshape(bh * 8, bw * 8, nc)


def test_flat_blocks_decode_to_the_rounded_dc(tmp_path):
    """A block with only a DC coefficient is flat: 128 + DC*q/8 rounded half
    up (and clamped), which checks the DC prediction, the dequantization and
    the RGB path (no colour conversion)."""
    rng = np.random.default_rng(11)
    q = np.full(64, 1, np.int64)
    q[0] = 7
    coefs = np.zeros((3, 5, 6, 64), np.int64)
    coefs[..., 0] = rng.integers(-200, 200, (3, 5, 6))  # some clamp at 0 or 255
    p = tmp_path / "flat.jpg"
    p.write_bytes(_jpeg_rgb(coefs, q))
    want = np.clip(128 + np.floor((coefs[..., 0] * 7 + 4) / 8), 0, 255).astype(np.int64)
    img = _decode(str(p)).astype(np.int64)
    assert img.shape == (40, 48, 3)
    expect = np.repeat(np.repeat(want.transpose(1, 2, 0), 8, 0), 8, 1)
    assert np.array_equal(img, expect)


@pytest.mark.parametrize("restart", [0, 1, 7])
def test_ac_blocks_match_the_exact_inverse_dct(tmp_path, restart):
    """Random AC coefficients (asymmetric, so a transposed or mis-ordered
    transform fails): the integer IDCT is within 1 level of the exact one on
    every pixel, and the restart-marked file decodes identically."""
    rng = np.random.default_rng(12 + restart)
    q = rng.integers(1, 9, 64)
    coefs = np.zeros((3, 3, 4, 64), np.int64)
    coefs[..., 0] = rng.integers(-60, 60, (3, 3, 4))
    ac = rng.integers(-12, 13, (3, 3, 4, 63)) * (rng.random((3, 3, 4, 63)) < 0.3)
    coefs[..., 1:] = ac
    coefs[0, 1, 2, 1:] = 0      # a flat block among them
    coefs[1, 2, 3, 63] = 9      # the last coefficient
    p = tmp_path / "ac.jpg"
    p.write_bytes(_jpeg_rgb(coefs, q, restart=restart))
    img = _decode(str(p)).astype(np.int64)
    deq = coefs * q
    want = np.zeros((3, 3, 4, 64), np.int64)
    for c in range(3):
        for by in range(3):
            for bx in range(4):
                want[c, by, bx] = _idct_exact(deq[c, by, bx]).ravel()
    diff = np.abs(img - _to_image(want))
    assert diff.max() <= 1 and diff.mean() < 0.2
    if restart:
        p0 = tmp_path / "ac0.jpg"
        p0.write_bytes(_jpeg_rgb(coefs, q))
        assert np.array_equal(_decode(str(p0)), img)


def _smooth(h, w, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([x * 200 // max(w - 1, 1) + 20, y * 200 // max(h - 1, 1) + 30, (x + 2 * y) * 100 // (w + 2 * h)],
                   -1)
    return np.clip(img + rng.integers(-8, 9, img.shape), 0, 255).astype(np.uint8)


def _save(img, **kw):
    b = io.BytesIO()
    PIL.fromarray(img).save(b, "JPEG", **kw)
    return b.getvalue()


@pytest.mark.parametrize("sub", [0, 1, 2])
@pytest.mark.parametrize("hw", [(1, 1), (9, 13), (64, 64), (50, 77)])
def test_progressive_and_restarts_decode_as_baseline(tmp_path, sub, hw):
    """Pillow's progressive script (spectral selection and successive
    approximation, EOB runs, refinement scans) and its restart intervals carry
    the same coefficients as the baseline file: same pixels."""
    img = _smooth(*hw, seed=hw[0] * 100 + hw[1])
    paths = {}
    for name, kw in {"base": {}, "prog": {"progressive": True}, "rst": {"restart_marker_blocks": 2}}.items():
        paths[name] = tmp_path / f"{name}.jpg"
        paths[name].write_bytes(_save(img, quality=85, subsampling=sub, **kw))
    base = _decode(str(paths["base"]))
    assert base.shape == (*hw, 3)
    assert np.array_equal(_decode(str(paths["prog"])), base)
    assert np.array_equal(_decode(str(paths["rst"])), base)


@pytest.mark.parametrize("quality", [50, 90, 100])
@pytest.mark.parametrize("hw", [(1, 1), (7, 5), (64, 64), (100, 131)])
def test_full_chroma_jpeg_matches_libjpeg(tmp_path, quality, hw):
    """4:4:4 files: the integer IDCTs and YCbCr conversions of stb_image and
    libjpeg round differently, by a few levels at most."""
    img = _smooth(*hw, seed=quality)
    p = tmp_path / "t.jpg"
    p.write_bytes(_save(img, quality=quality, subsampling=0))
    ref = np.asarray(PIL.open(p).convert("RGB")).astype(np.int64)
    diff = np.abs(_decode(str(p)).astype(np.int64) - ref)
    assert diff.max() <= 3 and diff.mean() < 0.1


@pytest.mark.parametrize("sub", [1, 2])
@pytest.mark.parametrize("hw", [(32, 48), (50, 77), (100, 131)])
def test_subsampled_jpeg_matches_libjpeg_triangle_filter(tmp_path, sub, hw):
    """4:2:2 and 4:2:0 on smooth images: chroma is upsampled by the triangle
    filter (3/4 nearer + 1/4 further sample) as in libjpeg's fancy upsampling,
    which rounds differently.  (libjpeg switches to box upsampling for chroma
    rows of 2 samples or fewer, so narrow images are not compared.)  stb's
    horizontal-only filter weights the last-but-one output column 3/4 on the
    further sample; that column is compared with the nearer-sample form."""
    img = _smooth(*hw, seed=sub)
    p = tmp_path / "t.jpg"
    p.write_bytes(_save(img, quality=92, subsampling=sub))
    ref = np.asarray(PIL.open(p).convert("RGB")).astype(np.int64)
    got = _decode(str(p)).astype(np.int64)
    diff = np.abs(got - ref)
    if sub == 1:
        col = 2 * ((hw[1] + 1) // 2 - 1)
        assert diff[:, col].max() <= 8
        diff[:, col] = 0
    assert diff.max() <= 4 and diff.mean() < 0.3


def test_cmyk_jpeg_multiplies_by_k(tmp_path):
    """Adobe CMYK (transform 0): each of C, M, Y times K / 255 (exact
    rounding), on the stored samples; Pillow reports them inverted."""
    rng = np.random.default_rng(9)
    cmyk = np.clip(_smooth(24, 40, 9)[..., [0, 1, 2, 0]].astype(np.int64) + rng.integers(0, 60, (24, 40, 4)), 0, 255)
    p = tmp_path / "cmyk.jpg"
    PIL.fromarray(cmyk.astype(np.uint8), "CMYK").save(p, quality=95)
    stored = 255 - np.asarray(PIL.open(p)).astype(np.int64)
    want = np.rint(stored[..., :3] * stored[..., 3:4] / 255.0)
    diff = np.abs(_decode(str(p)).astype(np.int64) - want)
    assert diff.max() <= 4 and diff.mean() < 0.5


def test_image_info_and_load(tmp_path):
    img = _smooth(10, 12, 1)
    p = tmp_path / "t.jpg"
    p.write_bytes(_save(img, quality=90))
    import ctypes as C
    from draw_amd import _lib
    L = _lib.lib()
    w, h, c = C.c_int32(), C.c_int32(), C.c_int32()
    assert L.dh_image_info(str(p).encode(), C.byref(w), C.byref(h), C.byref(c), None, 0) == 0
    assert (w.value, h.value, c.value) == (12, 10, 3)
    small = np.zeros(12 * 10 * 3 - 1, np.uint8)
    rc = L.dh_image_load(str(p).encode(), small.ctypes.data, small.nbytes, C.byref(w), C.byref(h), C.byref(c), None, 0)
    assert _lib.ERRORS[rc] == "DH_E_CAPACITY"


def test_unsupported_and_corrupt_files(tmp_path):
    grey = tmp_path / "grey.jpg"
    PIL.fromarray(np.zeros((8, 8), np.uint8), "L").save(grey)
    with pytest.raises(DrawHipError, match="DH_E_UNSUPPORTED"):
        _decode(str(grey))
    # an arithmetic-coded frame header (SOF9)
    arith = tmp_path / "arith.jpg"
    arith.write_bytes(b"\xff\xd8" + _seg(0xC9, bytes([8, 0, 8, 0, 8, 3, 1, 0x11, 0, 2, 0x11, 0, 3, 0x11, 0])))
    with pytest.raises(DrawHipError, match="DH_E_UNSUPPORTED"):
        _decode(str(arith))
    good = _save(_smooth(40, 40, 2), quality=80, progressive=True)
    # truncated inside the entropy-coded data, and truncated header
    for n in (len(good) // 2, 30):
        t = tmp_path / "trunc.jpg"
        t.write_bytes(good[:n])
        with pytest.raises(DrawHipError):
            _decode(str(t))
    # random byte damage must end in pixels or an error code, never a crash
    rng = np.random.default_rng(4)
    for i in range(200):
        b = bytearray(good)
        for _ in range(rng.integers(1, 8)):
            b[rng.integers(2, len(b))] = rng.integers(0, 256)
        t = tmp_path / "bad.jpg"
        t.write_bytes(bytes(b))
        try:
            _decode(str(t))
        except DrawHipError:
            pass


def test_jpeg_texture_through_the_obj_loader(tmp_path):
    img = _smooth(16, 24, 3)
    (tmp_path / "t.jpg").write_bytes(_save(img, quality=90))
    (tmp_path / "t.mtl").write_text("newmtl M\nKa 1 1 1\nKd 1 1 1\nKs 0 0 0\nmap_Kd t.jpg\nmap_Ka t.jpg\n")
    (tmp_path / "t.obj").write_text("mtllib t.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl M\nf 1 2 3\n")
    o = load_obj(str(tmp_path / "t.obj"))
    want = _decode(str(tmp_path / "t.jpg"))
    assert np.array_equal(o.textures[1].map_kd.img, want) and np.array_equal(o.textures[1].map_ka.img, want)


@pytest.mark.skipif(not os.path.exists(REF_JPG), reason="the reference's model files are not on this machine")
def test_reference_airplane_texture():
    """The reference's own JPEG texture: progressive, 1024x1024, JFIF and
    Adobe markers (transform 1: YCbCr), 4:4:4."""
    img = _decode(REF_JPG).astype(np.int64)
    assert img.shape == (1024, 1024, 3)
    ref = np.asarray(PIL.open(REF_JPG).convert("RGB")).astype(np.int64)
    diff = np.abs(img - ref)
    assert diff.max() <= 3 and diff.mean() < 0.01

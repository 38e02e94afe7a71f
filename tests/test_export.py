"""Frame export (App::export_frame_as, app/mod.rs:316-379): BGRA -> RGBA swap
on the device and an 8-bit RGBA PNG.  The PNG is checked by decoding it with
zlib and the PNG rules (chunk CRCs, IHDR, filter bytes), not by re-running the
writer."""
import ctypes as C
import struct
import zlib

import numpy as np
import pytest

from draw_amd import _lib, workloads
from parity import oracle


def decode_png(data: bytes) -> np.ndarray:
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    W, H, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(H, 1 + 4 * W)
    assert np.all(rows[:, 0] == 0), "this writer uses filter type 0 on every row"
    return rows[:, 1:].reshape(H, W, 4)


@pytest.mark.parametrize("W,H", [(1, 1), (3, 2), (97, 61), (640, 480)])
def test_png_writer_round_trip(tmp_path, W, H):
    rgba = np.random.default_rng(W * H).integers(0, 256, (H, W, 4), dtype=np.uint8)
    path = tmp_path / "f.png"
    assert _lib.lib().dh_write_png(str(path).encode(), rgba.ctypes.data, W, H) == 0
    assert np.array_equal(decode_png(path.read_bytes()), rgba)


def test_png_writer_rejects_bad_arguments(tmp_path):
    L = _lib.lib()
    buf = np.zeros(4, np.uint8)
    assert L.dh_write_png(str(tmp_path / "x.png").encode(), buf.ctypes.data, 0, 1) == -1
    assert L.dh_write_png(str(tmp_path / "no" / "dir.png").encode(), buf.ctypes.data, 1, 1) == \
        -1


@pytest.mark.gpu
def test_export_swaps_channels_of_the_rendered_frame(renderer, tmp_path):
    s = workloads.torus_scene(1)
    cv = renderer.render_device(s)
    rgba = renderer.read_rgba(cv, s.width, s.height)
    ref = oracle(s).frame                       # BGRA, rows top-down
    assert np.array_equal(rgba, ref[:, :, [2, 1, 0, 3]])
    path = tmp_path / "torus.png"
    renderer.export_png(cv, path)
    assert np.array_equal(decode_png(path.read_bytes()), rgba)


def _decode_jpeg(path):
    """Decode with Pillow (an independent libjpeg decoder), as RGB."""
    PIL = pytest.importorskip("PIL.Image")
    with PIL.open(path) as im:
        assert im.format == "JPEG" and im.mode == "RGB"
        return np.asarray(im.convert("RGB"))


def _markers(data: bytes):
    """(marker, segment body) of every segment before the scan, T.81 B.1."""
    assert data[:2] == b"\xff\xd8" and data[-2:] == b"\xff\xd9"
    pos, out = 2, []
    while True:
        assert data[pos] == 0xFF
        m = data[pos + 1]
        (n,) = struct.unpack(">H", data[pos + 2:pos + 4])
        out.append((m, data[pos + 4:pos + 2 + n]))
        pos += 2 + n
        if m == 0xDA:
            return out


def _psnr(a, b):
    mse = np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)
    return 10 * np.log10(255.0 ** 2 / max(mse, 1e-12))


@pytest.mark.parametrize("W,H", [(1, 1), (13, 7), (640, 480)])
def test_jpeg_writer_quality_100_decodes_close(tmp_path, W, H):
    """Quality 100 (the reference's export call: its quality argument is
    width*4, clamped to 100): every quantisation step is 1, no chroma
    subsampling; Pillow's decoder returns the frame within a few levels."""
    s = workloads.torus_scene(1, width=W, height=H)
    rgba = np.ascontiguousarray(oracle(s).frame[:, :, [2, 1, 0, 3]])
    path = tmp_path / "f.jpg"
    # the app's 800-pixel-wide canvas: quality argument 800 * 4
    assert _lib.lib().dh_write_jpg(str(path).encode(), rgba.ctypes.data, W, H, 800 * 4) == 0
    segs = _markers(path.read_bytes())
    dqt = b"".join(b for m, b in segs if m == 0xDB)
    assert len(dqt) == 130 and set(dqt[1:65]) == {1} and set(dqt[66:130]) == {1}
    sof = [b for m, b in segs if m == 0xC0][0]
    assert struct.unpack(">BHHB", sof[:6]) == (8, H, W, 3)
    assert [sof[6 + 3 * k + 1] for k in range(3)] == [0x11, 0x11, 0x11]
    got = _decode_jpeg(path)
    assert got.shape == (H, W, 3)
    err = np.abs(got.astype(int) - rgba[:, :, :3].astype(int))
    assert err.max() <= 4 and err.mean() < 0.8


def test_jpeg_writer_lower_quality_and_stb_quality_rule(tmp_path):
    """Lower quality scales the Annex K tables (IJG rule) and still decodes
    to the picture; 0 means 90 and values above 100 mean 100, as
    stbi_write_jpg treats its quality argument."""
    s = workloads.torus_scene(2)
    rgba = np.ascontiguousarray(oracle(s).frame[:, :, [2, 1, 0, 3]])
    L = _lib.lib()
    sizes = {}
    for q in (0, 50, 90, 100, 3200):
        path = tmp_path / f"q{q}.jpg"
        assert L.dh_write_jpg(str(path).encode(), rgba.ctypes.data, 640, 480, q) == 0
        sizes[q] = path.stat().st_size
        dqt = b"".join(b for m, b in _markers(path.read_bytes()) if m == 0xDB)
        if q == 50:
            assert dqt[1] == 16 and dqt[66] == 17        # Annex K tables unscaled at 50
        assert _psnr(_decode_jpeg(path), rgba[:, :, :3]) > (30 if q == 50 else 38)
    assert (tmp_path / "q0.jpg").read_bytes() == (tmp_path / "q90.jpg").read_bytes()
    assert (tmp_path / "q3200.jpg").read_bytes() == (tmp_path / "q100.jpg").read_bytes()
    assert sizes[50] < sizes[90] < sizes[100]
    assert L.dh_write_jpg(str(tmp_path / "x.jpg").encode(), rgba.ctypes.data, 0, 1, 90) == -1


@pytest.mark.gpu
def test_export_jpg_of_the_rendered_frame(renderer, tmp_path):
    s = workloads.torus_scene(1)
    cv = renderer.render_device(s)
    rgba = renderer.read_rgba(cv, s.width, s.height)
    path = tmp_path / "torus.jpg"
    renderer.export_jpg(cv, path)
    host = tmp_path / "host.jpg"
    assert _lib.lib().dh_write_jpg(str(host).encode(), rgba.ctypes.data, s.width, s.height, 100) == 0
    assert path.read_bytes() == host.read_bytes()

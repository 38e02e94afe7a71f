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

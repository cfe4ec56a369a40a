"""PNG dump of the BGRA8 output image (xrh_png_encode / xrh_write_png, no GPU).

Decoded here with the standard library's zlib and checked against the pixels:
signature, IHDR fields, every chunk's CRC-32, the zlib stream's Adler-32
(zlib checks it while decompressing), filter bytes and the RGB values (the
B8G8R8A8 word's R, G, B bytes; alpha dropped).
"""
from __future__ import annotations

import ctypes as C
import struct
import zlib

import numpy as np
import pytest

from xvk_ray_amd import native as N


def _encode(img: np.ndarray, w: int, h: int) -> bytes:
    img = np.ascontiguousarray(img, np.uint32)
    n = C.c_size_t()
    assert N.host().xrh_png_encode(img.ctypes.data, w, h, None, 0, C.byref(n)) == 0
    buf = (C.c_uint8 * n.value)()
    assert N.host().xrh_png_encode(img.ctypes.data, w, h, buf, n.value, C.byref(n)) == 0
    return bytes(buf)


def _decode(png: bytes):
    assert png[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(png):
        (length,) = struct.unpack(">I", png[pos:pos + 4])
        typ = png[pos + 4:pos + 8]
        data = png[pos + 8:pos + 8 + length]
        (crc,) = struct.unpack(">I", png[pos + 8 + length:pos + 12 + length])
        assert crc == zlib.crc32(typ + data), typ
        chunks.append((typ, data))
        pos += 12 + length
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])  # verifies the Adler-32 too
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert not rows[:, 0].any()  # filter type 0 on every row
    return w, h, rows[:, 1:].reshape(h, w, 3)


@pytest.mark.parametrize("w,h", [(1, 1), (7, 3), (640, 360), (300, 250)])  # 300x250: > 65535 B, several blocks
def test_png_round_trip(w, h):
    rng = np.random.default_rng(w * 1000 + h)
    img = rng.integers(0, 2**32, size=w * h, dtype=np.uint64).astype(np.uint32)
    dw, dh, rgb = _decode(_encode(img, w, h))
    assert (dw, dh) == (w, h)
    b = (img & 0xFF).reshape(h, w)
    g = ((img >> 8) & 0xFF).reshape(h, w)
    r = ((img >> 16) & 0xFF).reshape(h, w)
    assert np.array_equal(rgb[..., 0], r) and np.array_equal(rgb[..., 1], g) and np.array_equal(rgb[..., 2], b)


def test_write_png_file(tmp_path):
    img = np.array([0xFF0000FF, 0xFF00FF00, 0xFFFF0000, 0xFFFFFFFF], np.uint32)  # blue, green, red, white
    path = tmp_path / "out.png"
    assert N.host().xrh_write_png(str(path).encode(), img.ctypes.data, 2, 2) == 0
    _, _, rgb = _decode(path.read_bytes())
    assert rgb.reshape(4, 3).tolist() == [[0, 0, 255], [0, 255, 0], [255, 0, 0], [255, 255, 255]]


def test_png_rejects_bad_input():
    img = np.zeros(4, np.uint32)
    n = C.c_size_t()
    assert N.host().xrh_png_encode(img.ctypes.data, 0, 2, None, 0, C.byref(n)) == N.XR_ERR_INVALID_ARGUMENT
    assert N.host().xrh_png_encode(img.ctypes.data, 2, 2, None, 0, C.byref(n)) == 0
    small = (C.c_uint8 * 8)()
    assert N.host().xrh_png_encode(img.ctypes.data, 2, 2, small, 8, C.byref(n)) == N.XR_ERR_INVALID_ARGUMENT

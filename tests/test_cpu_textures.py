"""Texture images: PNG decoding (xrh_png_decode) and textured glTF loading
(no GPU needed).

PNGs of every colour type are written by an independent encoder (Pillow,
with stored, fixed-Huffman and dynamic-Huffman deflate and adaptive row
filters) and the decoder's RGBA8 is compared with Pillow's own decoding.
"""
from __future__ import annotations

import base64
import ctypes as C
import io
import json
from pathlib import Path

import numpy as np
import pytest

from xvk_ray_amd import native as N

PIL = pytest.importorskip("PIL.Image")
SB2 = Path(__file__).parent / "golden" / "gltf" / "sb2" / "scene.gltf"


def _decode(data: bytes):
    w, h = C.c_uint32(), C.c_uint32()
    rc = N.host().xrh_png_decode(data, len(data), C.byref(w), C.byref(h), None, 0)
    if rc:
        return rc, None
    out = np.zeros(4 * w.value * h.value, np.uint8)
    assert N.host().xrh_png_decode(data, len(data), C.byref(w), C.byref(h), out.ctypes.data, out.nbytes) == 0
    return 0, out.reshape(h.value, w.value, 4)


def _png(img, **kw) -> bytes:
    buf = io.BytesIO()
    img.save(buf, format="PNG", **kw)
    return buf.getvalue()


def _pillow_rgba(data: bytes) -> np.ndarray:
    return np.asarray(PIL.open(io.BytesIO(data)).convert("RGBA"))


def _noise(w, h, c, seed, smooth=False):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 256, size=(h, w, c), dtype=np.uint8)
    if smooth:  # long runs: back-references and non-trivial filters
        a = np.cumsum(a // 32, axis=1).astype(np.uint8)
    return a


@pytest.mark.parametrize("level", [0, 1, 9])
@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P", "P+tRNS", "1", "L+tRNS"])
@pytest.mark.parametrize("size", [(1, 1), (13, 7), (160, 120)])
def test_decode_matches_pillow(mode, size, level):
    w, h = size
    seed = w * 7 + h + level
    if mode in ("RGB", "RGBA", "L", "LA"):
        c = {"RGB": 3, "RGBA": 4, "L": 1, "LA": 2}[mode]
        a = _noise(w, h, c, seed, smooth=(level == 9))
        img = PIL.fromarray(a[..., 0] if c == 1 else a, mode)
        data = _png(img, compress_level=level)
    elif mode.startswith("P"):
        img = PIL.fromarray(_noise(w, h, 1, seed)[..., 0] % 37, "P")
        pal = _noise(37, 1, 3, seed + 1).reshape(-1).tolist()
        img.putpalette(pal)
        kw = {"transparency": bytes(_noise(37, 1, 1, seed + 2).reshape(-1))} if mode == "P+tRNS" else {}
        data = _png(img, compress_level=level, **kw)
    elif mode == "1":
        img = PIL.fromarray((_noise(w, h, 1, seed)[..., 0] > 127).astype(np.uint8) * 255, "L").convert("1")
        data = _png(img, compress_level=level)
    else:  # grey with a transparent key
        a = _noise(w, h, 1, seed)[..., 0] % 4 * 60
        data = _png(PIL.fromarray(a, "L"), compress_level=level, transparency=120)
    rc, got = _decode(data)
    assert rc == 0
    assert np.array_equal(got, _pillow_rgba(data)), mode


def test_decode_16_bit_keeps_high_byte():
    rng = np.random.default_rng(3)
    a = rng.integers(0, 65536, size=(9, 11), dtype=np.uint16)
    data = _png(PIL.fromarray(a))  # uint16 -> a 16-bit greyscale PNG
    rc, got = _decode(data)
    assert rc == 0
    hi = (a >> 8).astype(np.uint8)
    assert np.array_equal(got[..., 0], hi) and np.array_equal(got[..., 1], hi) and (got[..., 3] == 255).all()


def test_decode_rejects_damage():
    data = bytearray(_png(PIL.fromarray(_noise(8, 8, 3, 1), "RGB")))
    assert _decode(bytes(data[:40]))[0] == N.XR_ERR_INVALID_ARGUMENT  # truncated
    bad = bytearray(data)
    bad[-20] ^= 0xFF  # inside the IDAT: chunk CRC
    assert _decode(bytes(bad))[0] == N.XR_ERR_INVALID_ARGUMENT
    assert _decode(b"\xff\xd8\xff\xe0 not a png")[0] == N.XR_ERR_INVALID_ARGUMENT
    assert _decode(bytes(data))[0] == 0


# ------------------------------------------------------------------ glTF ----
def _textured_gltf(tmp: Path) -> tuple[Path, list[np.ndarray]]:
    """One quad, three materials naming textures 0, 1, 2; three images: a file
    uri, a data uri and a bufferView; texture 2 reuses image 0."""
    imgs = [_noise(5, 3, 4, 11), _noise(2, 7, 3, 12), _noise(4, 4, 1, 13)[..., 0]]
    pngs = [_png(PIL.fromarray(imgs[0], "RGBA")), _png(PIL.fromarray(imgs[1], "RGB")),
            _png(PIL.fromarray(imgs[2], "L"))]
    (tmp / "tex a.png").write_bytes(pngs[0])
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32).tobytes()
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32).tobytes()
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32).tobytes()
    blob = pos + uv + idx + pngs[2]
    (tmp / "buf.bin").write_bytes(blob)
    doc = {
        "asset": {"version": "2.0"},
        "scenes": [{"nodes": [0]}],
        "nodes": [{"mesh": 0}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "indices": 2, "material": 1}]}],
        "buffers": [{"uri": "buf.bin", "byteLength": len(blob)}],
        "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 48},
                        {"buffer": 0, "byteOffset": 48, "byteLength": 32},
                        {"buffer": 0, "byteOffset": 80, "byteLength": 24},
                        {"buffer": 0, "byteOffset": 104, "byteLength": len(pngs[2])}],
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
                      {"bufferView": 1, "componentType": 5126, "count": 4, "type": "VEC2"},
                      {"bufferView": 2, "componentType": 5125, "count": 6, "type": "SCALAR"}],
        "images": [{"uri": "tex%20a.png"},
                   {"uri": "data:image/png;base64," + base64.b64encode(pngs[1]).decode()},
                   {"bufferView": 3, "mimeType": "image/png"}],
        "textures": [{"source": 0}, {"source": 1}, {"source": 2}, {"source": 0}],
        "materials": [{"pbrMetallicRoughness": {"baseColorTexture": {"index": k}}} for k in range(4)],
    }
    path = tmp / "textured.gltf"
    path.write_text(json.dumps(doc))
    want = [_pillow_rgba(pngs[0]), _pillow_rgba(pngs[1]), _pillow_rgba(pngs[2]), _pillow_rgba(pngs[0])]
    return path, want


def test_gltf_textures_decoded_in_texture_order(tmp_path):
    path, want = _textured_gltf(tmp_path)
    scene = N.Scene.gltf(path)
    tex = scene.textures()
    assert len(tex) == 4
    for (w, h, px), ref in zip(tex, want):
        assert (h, w) == ref.shape[:2] and np.array_equal(px.reshape(h, w, 4), ref)
    m = scene.arrays()[2]
    assert list(m["base_color_texture_index"]) == [0, 1, 2, 3]
    assert not m["base_color"].any()  # no baseColorFactor: zeros, as the value-initialised AMaterial
    untextured = N.Scene.gltf(path, untextured=True)
    assert untextured.textures() == [] and (untextured.arrays()[2]["base_color_texture_index"] == -1).all()


def test_gltf_non_png_image_is_refused(tmp_path):
    path, _ = _textured_gltf(tmp_path)
    (tmp_path / "tex a.png").write_bytes(b"\xff\xd8\xff\xe0" + bytes(64))  # a JPEG header
    with pytest.raises(N.XrError) as e:
        N.Scene.gltf(path)
    assert "image 0" in str(e.value)
    N.Scene.gltf(path, untextured=True)  # loads without images


def test_reference_sb2_scene():
    """The reference's sb2 scene (src/main.cpp:56-68) with its PNG textures:
    75,277 triangles, every texture equal to Pillow's decoding of its image."""
    scene = N.Scene.gltf(SB2)
    assert scene.triangle_count == 75277
    doc = json.loads(SB2.read_text())
    tex = scene.textures()
    assert len(tex) == len(doc["textures"])
    for (w, h, px), t in zip(tex, doc["textures"]):
        img = doc["images"][t["source"]]
        ref = _pillow_rgba((SB2.parent / img["uri"]).read_bytes())
        assert np.array_equal(px.reshape(h, w, 4), ref)


# ------------------------------------------------- the oracle's sampling ----
def _oracle_with(textures):
    from oracle.binding import OracleScene
    v = np.zeros(3, N.VERTEX_DTYPE)
    v["pos"] = [[0, 0, 1], [1, 0, 1], [0, 1, 1]]
    m = np.zeros(1, N.MATERIAL_DTYPE)
    m["base_color_texture_index"] = 0
    return OracleScene(v, np.arange(3, dtype=np.uint32), m, np.array([0, 0], np.uint32), textures)


def test_oracle_bilinear_closed_forms():
    """Texel centres give the texel exactly (power-of-two sizes: no rounding in
    u * w - 0.5), midpoints the mean of two or four texels, coordinates past
    the edge the edge texel (clamp to edge), bytes read as byte / 255."""
    rng = np.random.default_rng(9)
    W, H = 8, 4
    px = rng.integers(0, 256, size=(H, W, 4), dtype=np.uint8)
    o = _oracle_with([(W, H, px.reshape(-1))])
    t = px[..., :3].astype(np.float64) / 255.0
    for j in range(H):
        for i in range(W):
            got = o.texture_sample(0, (i + 0.5) / W, (j + 0.5) / H)
            assert np.array_equal(got, (px[j, i, :3] / np.float32(255)).astype(np.float32))
    mid = o.texture_sample(0, 3.0 / W, 1.5 / H)  # halfway between texels 2 and 3 of row 1
    assert np.allclose(mid, (t[1, 2] + t[1, 3]) / 2, atol=1e-6)
    quad = o.texture_sample(0, 5.0 / W, 2.0 / H)  # the corner of texels (4..5, 1..2)
    assert np.allclose(quad, (t[1, 4] + t[1, 5] + t[2, 4] + t[2, 5]) / 4, atol=1e-6)
    assert np.allclose(o.texture_sample(0, -3.0, -0.2), t[0, 0], atol=1e-6)
    assert np.allclose(o.texture_sample(0, 7.0, 0.5 / H), t[0, W - 1], atol=1e-6)
    assert np.allclose(o.texture_sample(0, float("nan"), 1e30), t[H - 1, 0], atol=1e-6)
    # a weighted mix at an arbitrary point, against the float64 formula
    u, v = 0.3712, 0.6409
    x, y = u * W - 0.5, v * H - 0.5
    i0, j0 = int(np.floor(x)), int(np.floor(y))
    a, b = x - i0, y - j0
    want = ((1 - a) * (1 - b) * t[j0, i0] + a * (1 - b) * t[j0, i0 + 1] + (1 - a) * b * t[j0 + 1, i0]
            + a * b * t[j0 + 1, i0 + 1])
    assert np.allclose(o.texture_sample(0, u, v), want, atol=2e-6)


def _chunks(data: bytes):
    out, p = [], 8
    while p + 12 <= len(data):
        n = int.from_bytes(data[p:p + 4], "big")
        out.append([data[p + 4:p + 8], bytearray(data[p + 8:p + 8 + n])])
        p += 12 + n
    return out


def _assemble(chunks) -> bytes:
    import zlib
    b = bytearray(b"\x89PNG\r\n\x1a\n")
    for t, d in chunks:
        b += len(d).to_bytes(4, "big") + t + d + (zlib.crc32(t + bytes(d)) & 0xFFFFFFFF).to_bytes(4, "big")
    return bytes(b)


def test_decoder_survives_damaged_chunks():
    """Damaged PNGs whose chunk CRCs are recomputed (so the damage reaches the
    header checks, the inflater and the row filters): header bytes, deflate
    bytes, truncated and extended chunks, and valid zlib streams of damaged
    rows.  Every call returns 0 with an image or an error code, never
    crashes; a decoded image has the size its header states."""
    import zlib
    rng = np.random.default_rng(31)
    seeds = [_png(PIL.fromarray(_noise(13, 7, 4, 1), "RGBA"), compress_level=9),
             _png(PIL.fromarray(_noise(5, 9, 3, 2), "RGB"), compress_level=1),
             _png(PIL.fromarray(_noise(16, 3, 4, 3, smooth=True), "RGBA").convert("P"), compress_level=6),
             _png(PIL.fromarray(_noise(4, 4, 1, 4)[:, :, 0], "L").convert("I;16"))]
    decoded = 0
    for it in range(600):
        ch = _chunks(seeds[it % len(seeds)])
        for _ in range(int(rng.integers(1, 4))):
            c = ch[int(rng.integers(0, len(ch)))]
            d = c[1]
            op = int(rng.integers(0, 4))
            if op == 0 and len(d):
                d[int(rng.integers(0, len(d)))] = int(rng.integers(0, 256))
            elif op == 1 and len(d):
                del d[int(rng.integers(0, len(d))):]
            elif op == 2:
                p = int(rng.integers(0, len(d) + 1))
                d[p:p] = bytes(rng.integers(0, 256, size=int(rng.integers(1, 8)), dtype=np.uint8))
            elif c[0] == b"IDAT":
                try:
                    raw = bytearray(zlib.decompress(bytes(d)))
                except zlib.error:
                    continue
                if raw:
                    raw[int(rng.integers(0, len(raw)))] = int(rng.integers(0, 256))
                c[1] = bytearray(zlib.compress(bytes(raw[:int(rng.integers(0, len(raw) + 1))])))
        data = _assemble(ch)
        rc, img = _decode(data)
        if rc == 0:
            decoded += 1
            w = int.from_bytes(data[16:20], "big")
            h = int.from_bytes(data[20:24], "big")
            assert img.shape == (h, w, 4)
    assert 0 < decoded < 600

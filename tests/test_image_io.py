"""Image writers (rg_image_write, the output stage of the render CLI,
main.cpp:400-410, 479-485).  CPU only: every format is read back by a
reader that does not share code with the writer — the .hdr through the scene
loader's Radiance reader (an environment map of a scene file), PNG with
Python's zlib, BMP and PPM with struct, JPEG through the scene loader's
baseline decoder (itself checked against an independent encoder in
test_textures.py)."""
import json
import struct
import zlib

import numpy as np
import pytest

import rgapi


def ldr_image(h, w, seed, noise=0.01):
    rng = np.random.default_rng(seed)
    img = np.zeros((h, w, 4), np.float32)
    y, x = np.mgrid[0:h, 0:w]
    img[..., 0] = x / max(w - 1, 1)
    img[..., 1] = y / max(h - 1, 1)
    img[..., 2] = 0.5 + 0.4 * np.sin(x * 0.3) * np.cos(y * 0.2)
    img[..., :3] += rng.normal(0, noise, (h, w, 3))
    img[..., 3] = 1
    img[0, 0, :3] = (-0.5, 1.5, np.nan)       # out of range: clamped (NaN -> 0)
    return img


def as_bytes(img):
    v = np.nan_to_num(img[..., :3], nan=0.0)
    return np.where(v > 0, np.where(v >= 1, 255, np.floor(v * np.float32(255.0) + np.float32(0.5))), 0).astype(np.uint8)


def test_hdr_round_trips_through_the_scene_loader(tmp_path):
    rng = np.random.default_rng(1)
    h, w = 13, 29
    img = np.zeros((h, w, 4), np.float32)
    img[..., :3] = np.exp(rng.uniform(-12, 10, (h, w, 3))).astype(np.float32)
    img[2, 3, :3] = 0.0
    img[4, 5, :3] = (1e-40, 0, 0)               # below RGBE range: black
    img[..., 3] = 1
    rgapi.image_write(tmp_path / "env.hdr", img)
    (tmp_path / "s.json").write_text(json.dumps({"mainCamera": {}, "envmap": "env.hdr",
                                                  "o": {"mesh": {"type": "plane"}}}))
    _, d = rgapi.load_json(rgapi.Context(-1), str(tmp_path / "s.json"))
    got = d["env"]["rgb"]
    assert got.shape == (h, w, 3)
    want = img[..., :3]
    # RGBE keeps 8 mantissa bits of the largest channel: each channel is
    # truncated by less than max / 128
    mx = want.max(-1, keepdims=True)
    assert np.all(got <= want * (1 + 1e-6) + 1e-30)
    assert np.all(want - got < mx / 128 + 1e-30)
    assert got[2, 3].max() == 0 and got[4, 5].max() == 0


def test_png_is_a_valid_zlib_rgb_png(tmp_path):
    img = ldr_image(70, 1000, 2)                 # > 65535 raw bytes: several stored blocks
    rgapi.image_write(tmp_path / "a.png", img)
    data = (tmp_path / "a.png").read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body)
        chunks.setdefault(typ, b"")
        chunks[typ] += body
        pos += 12 + n
    w, h, depth, ctype, comp, filt, lace = struct.unpack(">IIBBBBB", chunks[b"IHDR"])
    assert (w, h, depth, ctype, comp, filt, lace) == (1000, 70, 8, 2, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(chunks[b"IDAT"]), np.uint8).reshape(h, 1 + 3 * w)
    assert (raw[:, 0] == 0).all()
    np.testing.assert_array_equal(raw[:, 1:].reshape(h, w, 3), as_bytes(img))
    assert b"IEND" in chunks


def test_bmp_and_ppm(tmp_path):
    img = ldr_image(9, 13, 3)
    rgapi.image_write(tmp_path / "a.bmp", img)
    b = (tmp_path / "a.bmp").read_bytes()
    assert b[:2] == b"BM"
    off, = struct.unpack("<I", b[10:14])
    w, h = struct.unpack("<ii", b[18:26])
    bpp, = struct.unpack("<H", b[28:30])
    assert (w, h, bpp) == (13, 9, 24)
    stride = (3 * w + 3) & ~3
    rows = np.frombuffer(b[off:off + stride * h], np.uint8).reshape(h, stride)[::-1, :3 * w].reshape(h, w, 3)
    np.testing.assert_array_equal(rows[..., ::-1], as_bytes(img))
    rgapi.image_write(tmp_path / "a.ppm", img)
    p = (tmp_path / "a.ppm").read_bytes()
    head = b"P6\n13 9\n255\n"
    assert p.startswith(head)
    np.testing.assert_array_equal(np.frombuffer(p[len(head):], np.uint8).reshape(9, 13, 3), as_bytes(img))


@pytest.mark.parametrize("shape,quality,tol", [((48, 64), 100, 1.5), ((48, 64), 95, 4.0), ((37, 53), 90, 6.0),
                                               ((8, 8), 100, 1.5)])
def test_jpeg_decodes_back(tmp_path, shape, quality, tol):
    h, w = shape
    img = ldr_image(h, w, 4, noise=0.002)
    img[..., :3] = np.clip(np.nan_to_num(img[..., :3]), 0, 1)
    rgapi.image_write(tmp_path / "t.jpg", img, quality=quality)
    data = (tmp_path / "t.jpg").read_bytes()
    assert data[:2] == b"\xff\xd8" and data[-2:] == b"\xff\xd9"
    (tmp_path / "s.json").write_text(json.dumps({"mainCamera": {}, "o": {"mesh": {"type": "plane"},
                                                                          "mat": {"tex": "t.jpg"}}}))
    _, d = rgapi.load_json(rgapi.Context(-1), str(tmp_path / "s.json"))
    got = np.round(d["textures"][0]["rgb"] * 255.0)
    want = as_bytes(img).astype(np.float64)
    err = np.abs(got - want)
    assert err.mean() < tol / 3 and np.percentile(err, 99) < 2 * tol, (err.mean(), np.percentile(err, 99))


def test_unknown_extension_and_unwritable_path(tmp_path):
    img = ldr_image(4, 4, 5)
    with pytest.raises(rgapi.RGError) as e:
        rgapi.image_write(tmp_path / "a.tga", img)
    assert e.value.code == -6
    with pytest.raises(rgapi.RGError) as e:
        rgapi.image_write(tmp_path / "no" / "such" / "a.png", img)
    assert e.value.code == -8

"""Material textures through the scene loader (SceneJsonLoader::readMaterial
`tex` / `texTiling`, sceneloader.cpp:101-127; Texture::loadFromFile,
texture.cpp:177-195) on a host-only context: our baseline JPEG decoder
against images encoded here by an independent baseline encoder (whose own
reconstruction is the expected output), the reference's checkerboard texture,
binary PPM, the UV sets of the primitives, .mesh files and OBJ `vt`, and the
per-object texture binding.  CPU only."""
import os
import struct

import numpy as np
import pytest

import rgapi

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

ZIGZAG = np.array([0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20,
                   13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59,
                   52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63])


def dct_matrix():
    c = np.zeros((8, 8))
    for u in range(8):
        for x in range(8):
            c[u, x] = (np.sqrt(0.5) if u == 0 else 1.0) * np.cos((2 * x + 1) * u * np.pi / 16) / 2
    return c   # F = C f C^T, f = C^T F C (orthonormal with the 1/4 C(u)C(v) factor)


class BitWriter:
    def __init__(self):
        self.out, self.acc, self.n = bytearray(), 0, 0

    def put(self, value, nbits):
        for k in range(nbits - 1, -1, -1):
            self.acc = (self.acc << 1) | ((value >> k) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)
                self.acc, self.n = 0, 0

    def flush(self):
        while self.n:
            self.put(1, 1)


def category(v):
    return 0 if v == 0 else int(abs(v)).bit_length()


def amp_bits(v, s):
    return v if v >= 0 else v + (1 << s) - 1


def encode_jpeg(rgb8, sampling=(1, 1), restart=0, gray=False, quality_scale=1.0):
    """Baseline sequential JPEG of an (H, W, 3) uint8 image with fixed-length
    Huffman codes (DC: 4-bit codes for categories 0-11, AC: 8-bit codes for
    the 162 run/size symbols).  Returns (jpeg bytes, expected decoded RGB):
    the reconstruction from the quantised coefficients with float64 IDCT,
    replicated chroma and the JFIF colour equations."""
    H, W = rgb8.shape[:2]
    rgb = rgb8.astype(np.float64)
    if gray:
        planes = [0.299 * rgb[..., 0] + 0.587 * rgb[..., 1] + 0.114 * rgb[..., 2]]
        samp = [(1, 1)]
    else:
        y = 0.299 * rgb[..., 0] + 0.587 * rgb[..., 1] + 0.114 * rgb[..., 2]
        cb = -0.168736 * rgb[..., 0] - 0.331264 * rgb[..., 1] + 0.5 * rgb[..., 2] + 128
        cr = 0.5 * rgb[..., 0] - 0.418688 * rgb[..., 1] - 0.081312 * rgb[..., 2] + 128
        planes = [y, cb, cr]
        samp = [sampling, (1, 1), (1, 1)]
    hmax = max(h for h, v in samp)
    vmax = max(v for h, v in samp)
    mx, my = -(-W // (8 * hmax)), -(-H // (8 * vmax))
    qy = np.clip(np.round((np.arange(64) % 8 + np.arange(64) // 8 + 2) * 2 * quality_scale), 1, 255).astype(int)
    qc = np.clip(qy + 3, 1, 255)
    qts = [qy.reshape(8, 8), qc.reshape(8, 8)]
    C = dct_matrix()
    comps = []
    for k, p in enumerate(planes):
        h, v = samp[k]
        fw, fh = mx * 8 * hmax, my * 8 * vmax
        full = np.pad(p, ((0, fh - H), (0, fw - W)), mode="edge")
        sx, sy = hmax // h, vmax // v
        sub = full.reshape(fh // sy, sy, fw // sx, sx).mean((1, 3))
        qt = qts[0 if k == 0 else 1]
        bh, bw = sub.shape[0] // 8, sub.shape[1] // 8
        coef = np.zeros((bh, bw, 8, 8), int)
        rec = np.zeros_like(sub)
        for by in range(bh):
            for bx in range(bw):
                blk = sub[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] - 128
                F = C @ blk @ C.T
                q = np.round(F / qt).astype(int)
                coef[by, bx] = q
                r = C.T @ (q * qt) @ C + 128
                rec[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] = np.clip(np.round(r), 0, 255)
        comps.append((h, v, coef, rec, 0 if k == 0 else 1))
    # expected decode: replicate chroma, convert
    up = []
    for h, v, _, rec, _ in comps:
        sx, sy = hmax // h, vmax // v
        up.append(np.repeat(np.repeat(rec, sy, 0), sx, 1)[:H, :W])
    if gray:
        want = np.repeat(up[0][..., None], 3, 2)
    else:
        Y, cb, cr = up[0], up[1] - 128, up[2] - 128
        want = np.stack([Y + 1.402 * cr, Y - 0.344136 * cb - 0.714136 * cr, Y + 1.772 * cb], -1)
        want = np.clip(np.round(want), 0, 255)
    # entropy coding
    ac_syms = [0x00, 0xF0] + [(r << 4) | s for r in range(16) for s in range(1, 11)]
    ac_code = {sym: i for i, sym in enumerate(ac_syms)}
    bw_ = BitWriter()
    preds = [0] * len(comps)
    mcu = 0
    for y in range(my):
        for x in range(mx):
            if restart and mcu and mcu % restart == 0:
                bw_.flush()
                bw_.out += bytes([0xFF, 0xD0 + ((mcu // restart - 1) % 8)])
                preds = [0] * len(comps)
            for k, (h, v, coef, _, _) in enumerate(comps):
                for vv in range(v):
                    for hh in range(h):
                        zz = coef[y * v + vv, x * h + hh].ravel()[ZIGZAG]
                        d = zz[0] - preds[k]
                        preds[k] = zz[0]
                        s = category(d)
                        bw_.put(s, 4)
                        bw_.put(amp_bits(d, s), s)
                        run = 0
                        last = max([i for i in range(1, 64) if zz[i] != 0], default=0)
                        for i in range(1, last + 1):
                            if zz[i] == 0:
                                run += 1
                                continue
                            while run > 15:
                                bw_.put(ac_code[0xF0], 8)
                                run -= 16
                            s = category(zz[i])
                            bw_.put(ac_code[(run << 4) | s], 8)
                            bw_.put(amp_bits(zz[i], s), s)
                            run = 0
                        if last < 63:
                            bw_.put(ac_code[0x00], 8)
            mcu += 1
    bw_.flush()

    def seg(marker, payload):
        return bytes([0xFF, marker]) + struct.pack(">H", len(payload) + 2) + payload

    out = bytearray(b"\xFF\xD8")
    for t, qt in enumerate(qts[:1 if gray else 2]):
        out += seg(0xDB, bytes([t]) + bytes(int(qt.ravel()[ZIGZAG[i]]) for i in range(64)))
    sof = struct.pack(">BHHB", 8, H, W, len(comps))
    for k, (h, v, _, _, tq) in enumerate(comps):
        sof += bytes([k + 1, (h << 4) | v, tq])
    out += seg(0xC0, sof)
    dc_bits = [0, 0, 0, 12] + [0] * 12
    ac_bits = [0] * 7 + [len(ac_syms)] + [0] * 8
    for tc_th, bits, vals in ((0x00, dc_bits, list(range(12))), (0x10, ac_bits, ac_syms),
                              (0x01, dc_bits, list(range(12))), (0x11, ac_bits, ac_syms)):
        if gray and (tc_th & 1):
            continue
        out += seg(0xC4, bytes([tc_th]) + bytes(bits) + bytes(vals))
    if restart:
        out += seg(0xDD, struct.pack(">H", restart))
    sos = bytes([len(comps)])
    for k, (_, _, _, _, tq) in enumerate(comps):
        sos += bytes([k + 1, (tq << 4) | tq])
    out += seg(0xDA, sos + bytes([0, 63, 0]))
    out += bw_.out + b"\xFF\xD9"
    return bytes(out), want.astype(np.uint8)


def load_texture(tmp_path, name, data, tiling=None):
    (tmp_path / name).write_bytes(data)
    mat = {"tex": name}
    if tiling is not None:
        mat["texTiling"] = tiling
    import json
    (tmp_path / "s.json").write_text(json.dumps({"mainCamera": {}, "o": {"mesh": {"type": "plane"}, "mat": mat}}))
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, str(tmp_path / "s.json"))
    return d


def test_rgba_export_is_byte_over_255_and_srgb(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (6, 5, 3), dtype=np.uint8)
    ppm = b"P6\n# test\n5 6\n255\n" + img.tobytes()
    d = load_texture(tmp_path, "t.ppm", ppm, tiling=[3, 2])
    t = d["textures"][0]
    assert t["srgb"] and t["rgb"].shape == (6, 5, 3)
    np.testing.assert_array_equal(t["rgb"], img.astype(np.float32) / np.float32(255.0))
    np.testing.assert_array_equal(d["tiling"][0], [3, 2])
    assert d["tex_of_obj"][0] == 0


@pytest.mark.parametrize("shape,kw", [((16, 16), dict()),
                                      ((21, 37), dict(sampling=(2, 2))),
                                      ((40, 24), dict(sampling=(2, 1), restart=2)),
                                      ((17, 23), dict(gray=True, restart=3)),
                                      ((33, 65), dict(sampling=(2, 2), restart=5, quality_scale=3.0))])
def test_jpeg_decoder_matches_independent_encoder(tmp_path, shape, kw):
    rng = np.random.default_rng(shape[0] * 100 + shape[1])
    H, W = shape
    yy, xx = np.mgrid[0:H, 0:W]
    img = np.stack([xx * 255 / max(W - 1, 1), yy * 255 / max(H - 1, 1), (xx + yy) % 64 * 4], -1)
    img = np.clip(img + rng.normal(0, 12, img.shape), 0, 255).astype(np.uint8)
    data, want = encode_jpeg(img, **kw)
    d = load_texture(tmp_path, "t.jpg", data)
    got = np.round(d["textures"][0]["rgb"] * 255.0).astype(int)
    assert got.shape == want.shape
    # float32 against float64 reconstruction: at most one level apart
    assert np.abs(got - want.astype(int)).max() <= 1, np.abs(got - want.astype(int)).max()


def test_reference_checkerboard_texture_decodes():
    """The reference's resources/textures/checkerboard.jpg (256 x 256,
    greyscale baseline): two levels in a regular checker pattern."""
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, os.path.join(GOLD, "ref_scenes", "sphereArray", "sphereArray.json"))
    floor = 1  # the textured floor object (sphereArray.json: "floor", texTiling [2, 2])
    t = d["textures"][d["tex_of_obj"][floor]]
    rgb = t["rgb"]
    assert rgb.shape == (256, 256, 3) and t["srgb"]
    assert np.array_equal(rgb[..., 0], rgb[..., 1]) and np.array_equal(rgb[..., 1], rgb[..., 2])
    np.testing.assert_array_equal(d["tiling"][floor], [2, 2])
    g = rgb[..., 0]
    lo, hi = np.percentile(g, 10), np.percentile(g, 90)
    assert hi - lo > 0.5                                   # two clearly separated levels
    # the pattern: cells of equal size alternating along rows and columns
    b = g > (lo + hi) / 2
    flips_x = np.count_nonzero(b[128, 1:] != b[128, :-1])
    flips_y = np.count_nonzero(b[1:, 128] != b[:-1, 128])
    assert flips_x == flips_y and flips_x >= 3
    assert abs(b.mean() - 0.5) < 0.05


def test_primitive_and_mesh_uvs(tmp_path):
    """PlaneMesh / CubeMesh texcoords (polygons.cpp:49-55, 90-126), no UVs on
    SphereMesh (uvCount 0), the first UV set of a .mesh and OBJ vt with V
    flipped (objreader.cpp:141-148)."""
    img = np.zeros((2, 2, 3), np.uint8)
    (tmp_path / "t.ppm").write_bytes(b"P6 2 2 255\n" + img.tobytes())
    (tmp_path / "q.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0.25 0.125\nvt 0.75 0.5\nvt 0.5 1\n"
                                    "f 1/1 2/2 3/3\n")
    import json
    sc = {"mainCamera": {}, "_materials": {"t": {"tex": "t.ppm"}},
          "p": {"mesh": {"type": "plane"}, "mat": "t"},
          "c": {"mesh": {"type": "cube"}, "mat": "t"},
          "s": {"mesh": {"type": "sphere", "stacks": 2, "slices": 3}, "mat": "t"},
          "o": {"mat": "t", "mesh": "q.obj"},      # mat first: the OBJ object inherits it
          "o2": {"mesh": "q.obj", "mat": "t"}}     # mesh first: the default material at that point
    (tmp_path / "s.json").write_text(json.dumps(sc))
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, str(tmp_path / "s.json"))
    uv = d["uv6"]
    sizes = [int((d["obj"] == k).sum()) for k in range(len(d["objects"]))]
    assert sizes == [2, 12, 12, 1, 1]
    np.testing.assert_array_equal(uv[0], [0, 0, 0, 1, 1, 0])
    np.testing.assert_array_equal(uv[1], [0, 1, 1, 1, 1, 0])
    for f in range(6):
        np.testing.assert_array_equal(uv[2 + 2 * f], [0, 0, 1, 0, 1, 1])
        np.testing.assert_array_equal(uv[3 + 2 * f], [0, 0, 1, 1, 0, 1])
    assert not uv[14:26].any()
    np.testing.assert_allclose(uv[26], [0.25, 0.875, 0.75, 0.5, 0.5, 0.0])
    np.testing.assert_array_equal(d["tex_of_obj"], [0, 0, 0, 0, -1])


def test_mesh_file_first_uv_set():
    """Sphere.mesh (flags 0xA: normals, no texcoords) leaves zero UVs;
    a written .mesh with one UV set carries it per corner."""
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, os.path.join(GOLD, "ref_scenes", "sphereArray", "sphereArray.json"))
    assert "uv6" in d
    meshes = [k for k in range(len(d["objects"])) if int((d["obj"] == k).sum()) == 528]
    assert meshes and not d["uv6"][d["obj"] == meshes[0]].any()


def test_unreadable_texture_is_dropped_with_a_warning(tmp_path):
    d = load_texture(tmp_path, "t.png", b"\x89PNG not decoded here")
    assert "textures" not in d
    assert any("t.png" in w for w in d["warnings"])
    d = load_texture(tmp_path, "t.jpg", b"\xFF\xD8\xFF\xC2\x00\x02")   # progressive / truncated
    assert "textures" not in d and any("JPEG" in w or "jpeg" in w for w in d["warnings"])

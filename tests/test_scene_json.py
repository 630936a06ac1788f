"""Scene ingest (rg_scene_load_json) on a host-only context: the reference's
JSON schema, primitive meshes, transforms, materials, media, camera and the
binary .mesh format.  CPU only."""
import os
import struct

import numpy as np
import pytest

import rgapi

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def loaded():
    ctx = rgapi.Context(-1)
    sc, d = rgapi.load_json(ctx, os.path.join(GOLD, "room_scene.json"))
    return ctx, sc, d


def tris_of(d, k):
    return d["v9"][d["obj"] == k].reshape(-1, 3, 3)


def test_object_order_and_counts(loaded):
    _, _, d = loaded
    # light, floor, left, back (2 each), box 12, ball 2*8*12, mesh 528; hidden skipped
    sizes = [int((d["obj"] == k).sum()) for k in range(len(d["objects"]))]
    assert sizes == [2, 2, 2, 2, 12, 192, 528]


def test_mesh_file_matches_independent_parse(loaded):
    _, _, d = loaded
    raw = open(os.path.join(GOLD, "ref_scenes", "sphereArray", "Sphere.mesh"), "rb").read()
    tag, ver, flags, length = struct.unpack("<IHHI", raw[:12])
    vcount = struct.unpack("<I", raw[12:16])[0]
    assert (tag, ver, flags, length, vcount) == (0x6873656D, 0x102, 0xA, 80, 1584)
    verts = np.frombuffer(raw, np.float32, 3 * vcount, length).reshape(-1, 3)
    want = verts * 0.5 + np.array([0, -2.5, 1.5 - 5.0], np.float32)  # camera at z = 5
    got = tris_of(d, 6).reshape(-1, 3)
    np.testing.assert_allclose(got, want, atol=1e-6)


def test_transforms_and_camera_view(loaded):
    _, _, d = loaded
    light = tris_of(d, 0).reshape(-1, 3)
    # plane scaled 1.5 in x/z, flipped by 180 deg about X, at y = 2.95, camera at z = 5
    assert np.allclose(light[:, 1], 2.95, atol=1e-6)
    assert np.isclose(light[:, 0].min(), -0.75) and np.isclose(light[:, 0].max(), 0.75)
    assert np.isclose(light[:, 2].min(), -0.75 - 5.0) and np.isclose(light[:, 2].max(), 0.75 - 5.0)
    n = d["n9"][d["obj"] == 0].reshape(-1, 3)
    assert np.allclose(n, [0, -1, 0], atol=1e-6)          # faces down after the flip
    left = d["n9"][d["obj"] == 2].reshape(-1, 3)
    assert np.allclose(left, [1, 0, 0], atol=1e-6)        # -90 deg about Z turns +Y into +X
    back = d["n9"][d["obj"] == 3].reshape(-1, 3)
    assert np.allclose(back, [0, 0, 1], atol=1e-6)        # +90 deg about X turns +Y into +Z
    assert np.isclose(d["camera"]["fov_deg"], 60) and np.isclose(d["camera"]["exposure"], 1.5)


def test_rotated_cube_normals_are_unit_and_rotated(loaded):
    _, _, d = loaded
    n = d["n9"][d["obj"] == 4].reshape(-1, 3)
    assert np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-6)
    c, s = np.cos(np.radians(40)), np.sin(np.radians(40))
    # the +X face normal after 40 deg about Y (x' = x c + z s, z' = -x s + z c)
    assert np.isclose(n, [c, 0, -s], atol=1e-5).all(axis=1).any()


def test_materials_and_media(loaded):
    _, _, d = loaded
    objs = d["objects"]
    assert np.allclose(objs[0]["color"], [1, 1, 240 / 255]) and objs[0]["emission"] == 10
    assert np.allclose(objs[1]["color"], 204 / 255) and objs[1]["roughness"] == 0  # "wall"
    assert objs[2]["roughness"] == np.float32(0.5)                                 # default
    assert objs[4]["transparency"] == 1 and objs[4]["interior_medium"] == 0
    assert objs[5]["refraction"] == np.float32(0.9) and objs[5]["refraction_ratio"] == np.float32(1.5)
    assert all(o["interior_medium"] == -1 for i, o in enumerate(objs) if i != 4)


def test_bad_files_report_io_errors():
    ctx = rgapi.Context(-1)
    with pytest.raises(rgapi.RGError) as e:
        rgapi.load_json(ctx, "/nonexistent/scene.json")
    assert e.value.code == -8
    bad = os.path.join(GOLD, "..", "..", "_bad_scene.json")
    with open(bad, "w") as f:
        f.write("{ \"a\": [1, 2 }")
    try:
        with pytest.raises(rgapi.RGError) as e:
            rgapi.load_json(ctx, bad)
        assert e.value.code == -8 and "line" in str(e.value)
    finally:
        os.remove(bad)


def test_object_medium_exported(loaded):
    _, _, d = loaded
    assert d["global_medium"] == -1 and len(d["media"]) == 1
    m = d["media"][0]
    assert np.allclose(m["sigma_a"], 0.1) and np.allclose(m["sigma_s"], 0.4) and m["density"] == 1


def _rgbe(c):
    v = max(c)
    if v < 1e-32:
        return [0, 0, 0, 0]
    m, e = np.frexp(v)
    s = m * 256.0 / v
    return [int(c[0] * s), int(c[1] * s), int(c[2] * s), int(e + 128)]


def _write_hdr(path, img, rle):
    h, w, _ = img.shape
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
        for y in range(h):
            px = np.array([_rgbe(img[y, x]) for x in range(w)], np.uint8)
            if not rle:
                f.write(px.tobytes())
                continue
            f.write(bytes([2, 2, w >> 8, w & 255]))
            for c in range(4):
                ch = px[:, c]
                x = 0
                while x < w:                      # runs of equal bytes, else literals
                    r = 1
                    while x + r < w and r < 127 and ch[x + r] == ch[x]:
                        r += 1
                    if r >= 3:
                        f.write(bytes([128 + r, ch[x]]))
                        x += r
                    else:
                        n = min(w - x, 128)
                        f.write(bytes([n]) + ch[x:x + n].tobytes())
                        x += n


@pytest.mark.parametrize("rle", [False, True])
def test_envmap_hdr_and_root_medium(tmp_path, rle):
    rng = np.random.default_rng(5)
    img = rng.uniform(0, 4, (6, 16, 3)).astype(np.float32)
    img[2, :8] = 1.5                              # a run for the RLE encoder
    _write_hdr(tmp_path / "sky.hdr", img, rle)
    (tmp_path / "s.json").write_text("""{
      "envmap": { "texture": "sky.hdr", "intensity": 2.5, "rotation": 30 },
      "medium": { "sigma_s": [0.01, 0.02, 0.03], "g": 0.3 },
      "quad": { "mesh": { "type": "plane" } } }""")
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, str(tmp_path / "s.json"))
    env = d["env"]
    assert env["rgb"].shape == img.shape and env["intensity"] == 2.5 and env["rotation_deg"] == 30
    # RGBE keeps 8 mantissa bits of the largest channel: decode = m * 2^(e-136)
    want = np.zeros_like(img)
    for y in range(6):
        for x in range(16):
            r, g, b, e = _rgbe(img[y, x])
            want[y, x] = np.array([r, g, b], np.float32) * np.float32(2.0 ** (e - 136))
    np.testing.assert_array_equal(env["rgb"], want)
    assert d["global_medium"] == 0 and np.isclose(d["media"][0]["g"], 0.3)
    assert np.allclose(d["media"][0]["sigma_s"], [0.01, 0.02, 0.03])


def test_inactive_medium_dropped_and_skips_warn(tmp_path):
    (tmp_path / "s.json").write_text("""{
      "a": { "mesh": { "type": "cone" }, "medium": { "sigma_a": [0, 0, 0] } },
      "b": { "mesh": ["x.fbx"], "mat": "nosuch" },
      "c": { "mesh": { "type": "cube" }, "mat": { "color": "#f80" } } }""")
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, str(tmp_path / "s.json"))
    assert len(d["media"]) == 0 and len(d["objects"]) == 1
    w = "\n".join(d["warnings"])
    assert "cone" in w and "x.fbx" in w and "nosuch" in w
    # 3-digit hex is read as one number times 2 (sceneloader.cpp:184-186):
    # 0xf80 * 2 = 0x001F00
    np.testing.assert_array_equal(d["objects"][0]["color"], np.float32([0, 31 / 255, 0]))


def test_rgx_cli_argument_errors():
    import subprocess
    exe = os.path.join(rgapi.HERE, "rgx")
    if not os.path.exists(exe):
        pytest.skip("rgx not built")
    r = subprocess.run([exe, "render", os.path.join(GOLD, "room_scene.json"), "-d", "4"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "unknown shading system 4" in r.stderr
    r = subprocess.run([exe, "render", "x.json", "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "unknown argument" in r.stdout


REF = os.path.join(GOLD, "ref_scenes")


def test_reference_cube_room_scene():
    """The reference's own cubeRoom.json: six planes and two rotated cubes
    (36 triangles), camera at z = 5.6 with a 70 degree field of view."""
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, os.path.join(REF, "cubeRoom", "cubeRoom.json"))
    assert d["warnings"] == []
    assert len(d["obj"]) == 36 and len(d["objects"]) == 8
    assert np.isclose(d["camera"]["fov_deg"], 70)
    floor = tris_of(d, 2).reshape(-1, 3)
    assert np.allclose(floor[:, 1], -3) and np.allclose(sorted(set(np.round(floor[:, 2], 4))), [-8.6, -2.6])
    objs = d["objects"]
    assert objs[2]["glossy"] == np.float32(0.7) and objs[2]["metallic"] == 1 and objs[2]["roughness"] == 0.5
    np.testing.assert_allclose(objs[3]["color"], [1, 0x55 / 255, 0x55 / 255])  # left wall "#ff5555"
    assert objs[0]["emission"] == 10
    # top light: 180 degrees about X faces down
    assert np.allclose(d["n9"][d["obj"] == 0].reshape(-1, 3), [0, -1, 0], atol=1e-6)


def test_reference_sphere_array_scene():
    """sphereArray.json: five Sphere.mesh instances, a camera tilted by -14
    degrees about X with depth of field, and a floor textured with the
    reference's checkerboard.jpg (copied beside the scene as data)."""
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, os.path.join(REF, "sphereArray", "sphereArray.json"))
    assert len(d["obj"]) == 2 + 2 + 5 * 528
    assert not any("checkerboard" in w for w in d["warnings"])
    assert len(d["textures"]) == 1 and (d["tex_of_obj"] >= 0).sum() == 1
    cam = d["camera"]
    assert np.isclose(cam["depth_of_field"], 6.2) and np.isclose(cam["aperture"], 5.6)
    # render space = R_x(+14 deg) (world - location): the floor plane's corners
    a = np.radians(14.0)
    R = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    corners = np.array([[sx * 6, -0.5, sz * 6] for sx in (-1, 1) for sz in (-1, 1)]) - [0, 2.2, 6.0]
    want = corners @ R.T
    got = tris_of(d, 1).reshape(-1, 3)
    for w in want:
        assert np.min(np.linalg.norm(got - w, axis=1)) < 1e-4
    obj3 = d["objects"][4]
    assert np.isclose(obj3["refraction"], 0.8) and np.isclose(obj3["roughness"], 0.01)


def test_reference_medium_and_aniso_scenes():
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, os.path.join(REF, "mediumTest", "waterCube.json"))
    m = d["media"][0]
    assert np.allclose(m["sigma_a"], [1.5, 0.4, 0.2]) and m["density"] == 2
    cube = [o for o in d["objects"] if o["interior_medium"] == 0]
    assert len(cube) == 1 and cube[0]["refraction"] == 1 and np.isclose(cube[0]["refraction_ratio"], 1.33)
    assert any("sky_sun.hdr" in w for w in d["warnings"])
    _, d = rgapi.load_json(ctx, os.path.join(REF, "mediumTest", "cloudHomo.json"))
    assert np.isclose(d["media"][0]["g"], 0.6) and d["objects"][1]["transparency"] == 1
    _, d = rgapi.load_json(ctx, os.path.join(REF, "anisoTest", "anisoTest.json"))
    assert np.isclose(d["objects"]["anisotropy"].max(), 0.95)


def test_reference_bokeh_and_ibl_scenes():
    """bokehTest.json: an 8-blade iris rotated 10 degrees, depth of field
    4.5, twelve sphere emitters (SphereMesh defaults 16 x 32: 1024 triangles
    each); iblTest.json: a checkerboard floor tiled 12 x 12 (the reference's
    texture, decoded) and a dispersive glass ball (chromaDispersion 0.03)."""
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, os.path.join(REF, "bokehTest", "bokehTest.json"))
    cam = d["camera"]
    assert cam["aperture_blades"] == 8 and np.isclose(cam["aperture_rotation_deg"], 10)
    assert np.isclose(cam["depth_of_field"], 4.5) and np.isclose(cam["aperture"], 2.8)
    sizes = [int((d["obj"] == k).sum()) for k in range(len(d["objects"]))]
    assert sizes.count(1024) == 13 and sizes[:2] == [2, 2]
    assert (d["objects"]["emission"] > 0).sum() >= 13
    _, d = rgapi.load_json(ctx, os.path.join(REF, "iblTest", "iblTest.json"))
    floor = int(np.nonzero(d["tex_of_obj"] >= 0)[0][0])
    np.testing.assert_array_equal(d["tiling"][floor], [12, 12])
    assert d["textures"][0]["rgb"].shape == (256, 256, 3)
    glass = d["objects"][d["objects"]["refraction"] == 1]
    assert len(glass) == 1 and np.isclose(glass[0]["chroma_dispersion"], 0.03)
    assert not any("dispersion" in w for w in d["warnings"])


def test_reference_suzanne_scene_and_unquoted_keys():
    """suzanne.json: Suzanne.mesh (checked against an independent struct
    parse) on a floor textured with the reference's wood.jpg (1024 x 1024,
    4:2:0 with restart intervals, decoded by our JPEG decoder); test/bokeh.json
    uses bare identifiers as keys (`stacks: 16`), as the reference's
    JavaScript-style scene files may."""
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, os.path.join(REF, "suzanne", "suzanne.json"))
    raw = open(os.path.join(REF, "suzanne", "Suzanne.mesh"), "rb").read()
    tag, ver, flags, length = struct.unpack("<IHHI", raw[:12])
    vcount, uvcount, icount = struct.unpack("<III", raw[12:24])
    assert tag == 0x6873656D
    ntri = icount // 3 if icount else vcount // 3
    sizes = [int((d["obj"] == k).sum()) for k in range(len(d["objects"]))]
    assert sizes == [2, ntri]
    wood = d["textures"][d["tex_of_obj"][0]]["rgb"]
    assert wood.shape == (1024, 1024, 3)
    mean = wood.reshape(-1, 3).mean(0)
    assert mean[0] > mean[1] > mean[2]            # a warm brown wood
    assert wood.std() > 0.02                        # with grain, not flat
    _, d = rgapi.load_json(ctx, os.path.join(REF, "test", "bokeh.json"))
    sizes = [int((d["obj"] == k).sum()) for k in range(len(d["objects"]))]
    assert sizes[0] == 2 * 16 * 32

def _rgbe_decode(path):
    """Independent Radiance RLE decoder (new-style scanlines)."""
    raw = open(path, "rb").read()
    pos = raw.index(b"\n\n") + 2
    eol = raw.index(b"\n", pos)
    _, h, _, w = raw[pos:eol].split()
    h, w = int(h), int(w)
    pos = eol + 1
    out = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        assert raw[pos] == 2 and raw[pos + 1] == 2
        pos += 4
        for c in range(4):
            x = 0
            while x < w:
                n = raw[pos]
                pos += 1
                if n > 128:
                    out[y, x:x + n - 128, c] = raw[pos]
                    pos += 1
                    x += n - 128
                else:
                    out[y, x:x + n, c] = np.frombuffer(raw[pos:pos + n], np.uint8)
                    pos += n
                    x += n
    e = out[..., 3].astype(np.int32)
    f = np.where(e > 0, np.ldexp(np.float32(1.0), e - 136), 0).astype(np.float32)
    return out[..., :3].astype(np.float32) * f[..., None]


@pytest.mark.skipif(not os.path.exists("/root/reference/resources/scenes"), reason="reference checkout absent")
def test_reference_scene_in_place_with_environment():
    """Loaded where it lies in the reference checkout, waterCube.json finds
    its run-length-encoded sky_sun.hdr; the decoded map equals an
    independent decoder's."""
    ctx = rgapi.Context(-1)
    _, d = rgapi.load_json(ctx, "/root/reference/resources/scenes/mediumTest/waterCube.json")
    assert d["warnings"] == []
    env = d["env"]
    assert env["rgb"].shape == (512, 1024, 3) and np.isclose(env["intensity"], 1.5)
    want = _rgbe_decode("/root/reference/resources/textures/sky_sun.hdr")
    np.testing.assert_array_equal(env["rgb"], want)

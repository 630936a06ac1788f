"""Output stage on the GPU through the C ABI (rg_postprocess,
rg_postprocess_device) against the oracle's restatement of
RayRenderer::applyPostProcess + applyTonemapGamma.

Bar: the bloomed HDR is bit-identical when curve = 1 (every operation is a
binary32 add, multiply or divide in the reference's order); with curve != 1
it goes through powf, whose last bit may differ between the host library and
the device library: relative 1e-5.  The LDR image goes through powf(x,
1/2.2): absolute 2e-6 (about 30 ulp at 1.0)."""
import sys

import numpy as np
import pytest

import oracle
import rgapi
import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return rgapi.Context(0)


def field(h, w, seed, n_bright=40):
    rng = np.random.default_rng(seed)
    x = np.zeros((h, w, 4), np.float32)
    x[..., :3] = rng.uniform(0, 0.8, (h, w, 3))
    x[..., 3] = 1.0
    ys, xs = rng.integers(0, h, n_bright), rng.integers(0, w, n_bright)
    x[ys, xs, :3] = rng.uniform(2, 500, (n_bright, 3))
    return x


@pytest.mark.parametrize("h,w", [(1, 1), (7, 3), (131, 257), (1080, 1920)])
def test_bloom_and_tonemap_match_oracle(ctx, h, w):
    say(f"[post] bloom + tonemap {w}x{h}")
    x = field(h, w, 10 + h)
    b = rgapi.default_bloom()
    hdr, ldr = rgapi.postprocess(ctx, x, bloom=b)
    want = oracle.bloom(x)
    assert hdr.tobytes() == want.tobytes()
    np.testing.assert_allclose(ldr, oracle.tonemap(want), rtol=0, atol=2e-6)
    assert np.array_equal(ldr[..., 3], x[..., 3])


@pytest.mark.parametrize("kw", [dict(threshold=0.5, strength=2.5, radius=0.08, size_aspect=0.3),
                                dict(threshold=0.0, radius=0.0, size_aspect=1.0),
                                dict(size_aspect=0.01, radius=0.2),       # aspect clamped to 0.02
                                dict(threshold=-3.0, radius=-1.0)])       # clamps of the reference
def test_bloom_settings_match_oracle(ctx, kw):
    say(f"[post] bloom settings {kw}")
    x = field(96, 160, 5)
    hdr, _ = rgapi.postprocess(ctx, x, bloom=rgapi.default_bloom(**kw), want_ldr=False)
    assert hdr.tobytes() == oracle.bloom(x, **kw).tobytes()


def test_bloom_curve_within_powf_tolerance(ctx):
    x = field(64, 80, 6)
    hdr, _ = rgapi.postprocess(ctx, x, bloom=rgapi.default_bloom(curve=2.0), want_ldr=False)
    np.testing.assert_allclose(hdr, oracle.bloom(x, curve=2.0), rtol=1e-5, atol=1e-7)


def test_tonemap_only_leaves_hdr_and_matches(ctx):
    x = field(50, 70, 7)
    hdr, ldr = rgapi.postprocess(ctx, x, bloom=None)
    assert np.array_equal(hdr, x)
    np.testing.assert_allclose(ldr, oracle.tonemap(x), rtol=0, atol=2e-6)


def test_device_entry_point_equals_host_entry_point(ctx):
    import torch

    x = field(240, 320, 8)
    want, want_ldr = rgapi.postprocess(ctx, x, bloom=rgapi.default_bloom())
    d_hdr = torch.from_numpy(x.copy()).to("cuda:0")
    d_ldr = torch.empty_like(d_hdr)
    torch.cuda.synchronize()
    rgapi.postprocess_device(ctx, 320, 240, d_hdr.data_ptr(), rgapi.default_bloom(), d_ldr.data_ptr())
    ctx.synchronize()
    assert np.array_equal(d_hdr.cpu().numpy(), want)
    assert np.array_equal(d_ldr.cpu().numpy(), want_ldr)
    # a second call on smaller images reuses the grown buffers
    y = field(30, 20, 9)
    hdr, _ = rgapi.postprocess(ctx, y, bloom=rgapi.default_bloom(), want_ldr=False)
    assert hdr.tobytes() == oracle.bloom(y).tobytes()


def test_rendered_frame_postprocess_matches_oracle(ctx):
    # the cube room's emitter (radiance 15, clamped to 10 per sample) is the
    # bright source the bloom spreads
    rs = rgapi.RenderScene(ctx, scenes.cube_room())
    out = rs.render(rgapi.default_settings(width=128, height=96, spp=8, max_depth=5))
    hdr, ldr = rgapi.postprocess(ctx, out["hdr"], bloom=rgapi.default_bloom(threshold=0.8))
    want = oracle.bloom(out["hdr"], threshold=0.8)
    assert hdr.tobytes() == want.tobytes()
    assert (hdr[..., :3] >= out["hdr"][..., :3]).all()
    assert hdr[..., :3].sum() > out["hdr"][..., :3].sum()
    np.testing.assert_allclose(ldr, oracle.tonemap(want), rtol=0, atol=2e-6)


def test_bad_arguments_rejected(ctx):
    with pytest.raises(rgapi.RGError):
        rgapi.postprocess(ctx, np.zeros((0, 4, 4), np.float32))
    host = rgapi.Context(-1)
    with pytest.raises(rgapi.RGError) as e:
        rgapi.postprocess(host, np.zeros((2, 2, 4), np.float32))
    assert e.value.code == -2


def test_cli_writes_the_postprocessed_image(ctx, tmp_path):
    """rgx render ... -enpp true -o x.ppm: the 8-bit image is the rounded
    LDR of rg_render + rg_postprocess with the same settings."""
    import os
    import subprocess

    exe = os.path.join(rgapi.HERE, "rgx")
    scene = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "room_scene.json")
    out = tmp_path / "x.ppm"
    say("[post] rgx render -enpp")
    r = subprocess.run([exe, "render", scene, "-r", "64,48", "-s", "4", "-enpp", "true", "-blth", "0.7",
                        "-o", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "rendering... 100%" in r.stdout   # the progress line, as main.cpp prints it
    raw = out.read_bytes()
    head = b"P6\n64 48\n255\n"
    assert raw.startswith(head)
    img = np.frombuffer(raw[len(head):], np.uint8).reshape(48, 64, 3)
    sc, _ = rgapi.load_json(ctx, scene)
    sc.build()
    st = rgapi.default_settings(width=64, height=48, spp=4)
    hdr = sc.render(st)["hdr"]
    _, ldr = rgapi.postprocess(ctx, hdr, bloom=rgapi.default_bloom(threshold=0.7))
    want = np.floor(np.clip(ldr[..., :3], 0, 1) * 255.0 + 0.5).astype(np.uint8)
    assert np.array_equal(img, want)

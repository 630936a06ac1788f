"""Point lights of traceLight (rayrenderer.cpp:2205-2293) in the path-tracing
restatement (oracle/rg_oracle_pt.c), CPU only:

* an omni light over a diffuse plane, seen through pixel centres: every
  pixel reads albedo x min(1, colour x emission x cos(theta) / r^2), the
  light's term of traceLight times the diffuse albedo, computed here from
  the camera geometry; with the light bright enough every channel clamps
  to 1 and the pixel reads the albedo;
* a spot: inside the inner edge of the smoothstep (dot >= cos + 0.1) the
  omni falloff with dot(l, n_light) in place of cos(theta), outside the
  outer edge (dot <= cos - 0.1) nothing, and monotone in between;
* two lights: traceLight picks one uniformly and returns its whole value,
  so the image converges to the mean of the two one-light images (not their
  sum, not half of it);
* a light behind an opaque wall gives nothing; one behind a transparent
  (transparency >= 0.01, refraction <= 0.1) wall shines through (the
  environment rays' occluders, rayrenderer.cpp:2211-2215);
* the loader turns an emissive object without meshes into a point light at
  the object's render-space origin.
"""
import json
import os

import numpy as np
import pytest

import oracle
import rgapi
import scenes

THREADS = 8


def centre_hits(W, H, fov_deg):
    """Pixel-centre rays of the pinhole camera (rayrenderer.cpp:97-132,
    860-934 without jitter) hitting the plane y = -1: points and mask."""
    vs = 2.0 * np.tan(np.radians(fov_deg) * 0.5) / H
    x, y = np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5)
    d = np.stack([(x - W / 2) * vs, -(y - H / 2) * vs, -np.ones_like(x)], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    ok = d[..., 1] < -1e-3
    t = np.where(ok, -1.0 / np.where(ok, d[..., 1], -1.0), 0.0)
    return d * t[..., None], ok


def omni_expected(p, light_pos, color, emission, albedo):
    lr = np.asarray(light_pos, np.float64) - p
    r2 = (lr * lr).sum(-1)
    cos = lr[..., 1] / np.sqrt(r2)   # plane normal +Y
    val = np.clip(np.asarray(color)[None, None, :] * (emission * np.maximum(cos, 0) / r2)[..., None], 0, 1)
    return np.asarray(albedo)[None, None, :] * val


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("emission", [2.0, 1e6])
def test_omni_light_over_a_diffuse_plane(mode, emission):
    W, H = 48, 32
    albedo, color = (0.8, 0.6, 0.4), (1.0, 0.5, 0.25)
    sc = scenes.point_light_plane(height=1.0, emission=emission, color=color, albedo=albedo)
    s, _ = oracle.render(sc, W, H, 4, 3, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    img = s / 4
    p, ok = centre_hits(W, H, 60.0)
    ok &= p[..., 2] < -1.0   # on the quad (z in [-7, -1], |x| <= 4)
    ok &= (p[..., 2] > -7.0) & (np.abs(p[..., 0]) < 4.0)
    want = omni_expected(p, (0.0, 0.0, -3.0), color, emission, albedo)
    assert ok.sum() > 300
    np.testing.assert_allclose(img[ok], want[ok], rtol=2e-4, atol=1e-7)
    if emission > 1e5:
        np.testing.assert_allclose(img[ok], np.broadcast_to(albedo, img[ok].shape), rtol=1e-6)
    # rows above the horizon see no geometry and no environment: backColor
    np.testing.assert_array_equal(img[:H // 2 - 1], np.broadcast_to(np.float32(oracle.BACK_COLOR[:3]),
                                                                     img[:H // 2 - 1].shape))


@pytest.mark.parametrize("mode", [0, 1])
def test_spot_light_cone(mode):
    W, H = 64, 48
    spot = 14.0                                  # cosf(RADIAN_TO_DEGREE(7)) = 0.4946
    c = float(np.cos(np.float32(np.float32(spot * 0.5) * np.float32(180.0 / np.pi)), dtype=np.float32))
    sc = scenes.point_light_plane(height=1.0, emission=0.5, color=(1, 1, 1), albedo=(1, 1, 1))
    sc["point_lights"]["normal"] = (0.0, 1.0, 0.0)    # dot(l, n_light) > 0 below the light
    sc["point_lights"]["spot_range"] = spot
    s, _ = oracle.render(sc, W, H, 2, 3, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    img = (s / 2)[..., 0].astype(np.float64)
    p, ok = centre_hits(W, H, 60.0)
    ok &= (p[..., 2] > -7.0) & (p[..., 2] < -1.0) & (np.abs(p[..., 0]) < 4.0)
    lr = np.array([0.0, 0.0, -3.0]) - p
    r = np.linalg.norm(lr, axis=-1)
    dl = lr[..., 1] / r                          # dot(l, n_light) = cos(theta) for this light
    inner = ok & (dl >= c + 0.1 + 1e-4)
    outer = ok & (dl <= c - 0.1 - 1e-4)
    mid = ok & ~inner & ~outer
    assert inner.sum() > 50 and outer.sum() > 50 and mid.sum() > 20
    np.testing.assert_allclose(img[inner], np.minimum(1.0, 0.5 * dl[inner] / r[inner] ** 2), rtol=2e-4)
    assert (img[outer] == 0).all()
    # between the edges: below the omni value, above zero, rising with dot
    full = np.minimum(1.0, 0.5 * dl / r ** 2)
    assert (img[mid] <= full[mid] * (1 + 1e-4)).all() and (img[mid] > 0).mean() > 0.9
    frac = img[mid] / (full[mid])
    order = np.argsort(dl[mid])
    assert np.all(np.diff(frac[order]) > -1e-3)


@pytest.mark.parametrize("mode", [0, 1])
def test_two_lights_pick_one_uniformly(mode):
    W, H, n = 24, 16, 4096
    a = scenes.point_light_plane(emission=1.0, color=(1, 1, 1), albedo=(0.5, 0.5, 0.5))
    b = {**a, "point_lights": a["point_lights"].copy()}
    b["point_lights"]["pos"] = (1.5, 0.2, -4.0)
    both = {**a, "point_lights": np.concatenate([a["point_lights"], b["point_lights"]])}
    ia = oracle.render(a, W, H, 2, 3, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)[0] / 2
    ib = oracle.render(b, W, H, 2, 3, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)[0] / 2
    s, q = oracle.render(both, W, H, n, 3, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    m = s.astype(np.float64) / n
    var = np.maximum(q / n - m * m, 0) / n
    want = 0.5 * (ia + ib)
    lit = (want > 1e-3) & (ia != ib)    # pixels the two lights light differently
    # the float32 sum of 4096 samples carries up to ~3e-4 of rounding
    z = (m - want)[lit] / np.sqrt(var[lit] + (3e-4 * want[lit]) ** 2)
    assert np.mean(np.abs(z) > 4) < 0.01
    assert abs(m[lit].sum() / want[lit].sum() - 1) < 0.01


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("wall,lit", [(dict(), False), (dict(transparency=0.5), True),
                                      (dict(transparency=0.5, refraction=0.5), False)])
def test_point_light_occluders(mode, wall, lit):
    """A sheet between the light (y = 1.5) and the floor, above the camera
    (y = 0.5), so the camera still sees the floor."""
    W, H = 32, 24
    sc = scenes.point_light_plane(height=2.5, emission=3.0, color=(1, 1, 1), albedo=(1, 1, 1))
    b = scenes.Builder()
    for i in range(len(sc["objects"])):
        b.add_object(sc["objects"][i])
    b.v, b.n, b.o = [sc["v9"]], [sc["n9"]], [sc["obj"]]
    w = b.add_object(scenes.material(**wall))
    b.quad(w, (-6, 0.5, 0), (6, 0.5, 0), (6, 0.5, -9), (-6, 0.5, -9), normal=(0, -1, 0))
    occl = b.finish(camera=sc["camera"], point_lights=sc["point_lights"])
    base, _ = oracle.render(sc, W, H, 2, 3, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    got, _ = oracle.render(occl, W, H, 2, 3, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    p, ok = centre_hits(W, H, 60.0)
    floor = ok & (p[..., 2] > -7.0) & (p[..., 2] < -1.0) & (np.abs(p[..., 0]) < 4.0)
    assert floor.sum() > 100 and (base[floor] > 0).all()
    if lit:
        np.testing.assert_array_equal(got[floor], base[floor])
    else:
        assert (got[floor] == 0).all()


def test_loader_reads_point_lights(tmp_path):
    sc = {
        "mainCamera": {"location": [0, 0, 0], "angle": [0, 0, 0], "fieldOfView": 50},
        "floor": {"mesh": {"type": "plane"}, "location": [0, -1, -3], "scale": [4, 1, 4]},
        "bulb": {"location": [0.5, 1.0, -3.0], "mat": {"emission": 2.5, "color": [1.0, 0.5, 0.25]}},
        "spot": {"location": [-1, 2, -4], "angle": [0, 0, 0], "mat": {"emission": 1.0, "spotRange": 14}},
        "dark": {"location": [3, 3, 3], "mat": {"emission": 0.0}},
    }
    p = os.path.join(tmp_path, "pl.json")
    with open(p, "w") as f:
        json.dump(sc, f)
    ctx = rgapi.Context(-1)
    h, d = rgapi.load_json(ctx, p)
    pl = d["point_lights"]
    assert len(pl) == 2
    np.testing.assert_allclose(pl[0]["pos"], (0.5, 1.0, -3.0), atol=1e-6)
    np.testing.assert_allclose(pl[0]["color"], (1.0, 0.5, 0.25), atol=1e-6)
    assert pl[0]["emission"] == np.float32(2.5) and pl[0]["spot_range"] == 0.0
    np.testing.assert_allclose(pl[1]["pos"], (-1.0, 2.0, -4.0), atol=1e-6)
    assert pl[1]["spot_range"] == np.float32(14.0)
    # angle (0, 0): normalize(cos 0 sin 0, cos 0 cos 0, sin 0) = +Y
    np.testing.assert_allclose(pl[1]["normal"], (0.0, 1.0, 0.0), atol=1e-6)


def test_single_lobe_matches_recursive_point_lit_room():
    """The GPU's single-lobe estimator (mode 1) and the reference's call
    structure (mode 0) agree per pixel on the point-lit room with an
    environment (the diffuse events' three NEE terms)."""
    sc = scenes.point_lit_room(env_value=0.2)
    W = H = 24
    n = 512
    a, aq = oracle.render(sc, W, H, n, 5, mode=0, threads=THREADS, seed=31)
    b, bq = oracle.render(sc, W, H, n, 5, mode=1, threads=THREADS, seed=32)
    ma, mb = a / n, b / n
    va = np.maximum(aq / n - ma * ma, 0) / n
    vb = np.maximum(bq / n - mb * mb, 0) / n
    z = (ma - mb) / np.sqrt(va + vb + 1e-12)
    assert np.mean(np.abs(z) > 3.5) < 0.01
    assert abs(ma.mean() / mb.mean() - 1) < 0.01

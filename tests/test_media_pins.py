"""Pins of the heterogeneous and emissive media and heat haze of the
path-tracing restatement (oracle/rg_oracle_pt.c; medium.cpp:125-472,
rayrenderer.cpp:1596-1864, 2066-2203), CPU only.

* value-noise fBm: continuous across lattice cells with zero slope at the
  lattice points (smoothstep weights), mid-edge / face / cell centres equal
  the means of 2 / 4 / 8 corners, octaves combine as sum g^k n(2^k f p) /
  sum g^k, and densityAt = max(0, amplitude n + bias) of the offset point;
* delta tracking: the fraction of walks that pass x without a collision is
  exp(-int_0^x sigma_t density ds) (dense quadrature of the density field);
  a constant field behaves as the homogeneous medium of that density (same
  image expectation), scatter weight sigma_s density / sigma_t_max included;
* emission: the cone profile at its peak, off its ends and beyond its
  radius; the path tube's colour blend and radial power falloff; the
  stratified emission integral's mean equals the quadrature of sigma_e Tr,
  and constant emission gives the closed form exactly;
* volume-light NEE: an emissive cone with no bounding mesh lights a
  diffuse plane by albedo / pi x int sigma_e(axis) cos / r^2 along its axis
  (equiangular sampling, its pdf, r^2 and the 1/pi); a constant-emission
  volume, sampled at a point, contributes next to nothing, as there;
* heat haze: no bend without amplitude, unit directions, and for small
  amplitudes the direction change equals the integral of the normal
  gradient's component across the ray, linear in the amplitude;
* the single-lobe estimator (mode 1) against the recursive restatement
  (mode 0) on a scene with all of them.
"""
import numpy as np
import pytest

import oracle
import scenes

THREADS = 8


def noise_medium(**kw):
    base = dict(sigma_s=(1.0, 1.0, 1.0), density_field=1, noise_octaves=1, noise_frequency=1.0,
                noise_amplitude=1.0, noise_bias=0.0)
    base.update(kw)
    return scenes.medium(**base)


# --- value noise / fBm --------------------------------------------------------
def test_value_noise_cells():
    m = noise_medium()
    rng = np.random.default_rng(3)
    cells = rng.integers(-20, 20, (200, 3)).astype(np.float32)
    f = lambda p: oracle.medium_eval(m, "density", p).astype(np.float64)
    c = f(cells)
    assert (c >= 0).all() and (c < 1).all() and np.unique(c).size > 150
    # continuity and zero slope at the lattice points, from either side
    for ax in range(3):
        e = np.zeros(3, np.float32)
        e[ax] = 1e-3
        assert np.abs(f(cells + e) - c).max() < 1e-5
        assert np.abs(f(cells - e) - c).max() < 1e-5
    # fade(1/2) = 1/2: mid-edge, face and cell centres are corner means
    corner = lambda d: f(cells + np.asarray(d, np.float32))
    mid = f(cells + np.float32([0.5, 0, 0]))
    np.testing.assert_allclose(mid, 0.5 * (c + corner([1, 0, 0])), atol=1e-6)
    face = f(cells + np.float32([0.5, 0.5, 0]))
    np.testing.assert_allclose(face, 0.25 * (c + corner([1, 0, 0]) + corner([0, 1, 0]) + corner([1, 1, 0])),
                               atol=1e-6)
    ctr = f(cells + np.float32([0.5, 0.5, 0.5]))
    want = sum(corner([i, j, k]) for i in (0, 1) for j in (0, 1) for k in (0, 1)) / 8
    np.testing.assert_allclose(ctr, want, atol=1e-6)


def test_fbm_octaves_amplitude_bias_offset():
    rng = np.random.default_rng(4)
    p = rng.uniform(-5, 5, (500, 3)).astype(np.float32)
    f0, g, lac = 1.6, 0.5, 2.0
    one = [oracle.medium_eval(noise_medium(noise_frequency=f0 * lac ** k), "density", p).astype(np.float64)
           for k in range(4)]
    want = sum(g ** k * one[k] for k in range(4)) / sum(g ** k for k in range(4))
    got = oracle.medium_eval(noise_medium(noise_frequency=f0, noise_octaves=4, noise_gain=g,
                                          noise_lacunarity=lac), "density", p)
    np.testing.assert_allclose(got, want, atol=2e-6)
    biased = oracle.medium_eval(noise_medium(noise_frequency=f0, noise_octaves=4, noise_amplitude=1.4,
                                             noise_bias=-0.25), "density", p)
    np.testing.assert_allclose(biased, np.maximum(0, 1.4 * want - 0.25), atol=3e-6)
    assert (biased == 0).any()                  # the bias carves empty pockets
    off = np.float32([0.3, -1.2, 2.5])
    shifted = oracle.medium_eval(noise_medium(noise_offset=tuple(off)), "density", p + off)
    np.testing.assert_allclose(shifted, oracle.medium_eval(noise_medium(), "density", p), atol=2e-6)


# --- delta tracking -------------------------------------------------------------
def test_delta_tracking_survival_is_the_transmittance():
    m = noise_medium(sigma_a=(0.2, 0.2, 0.2), sigma_s=(0.6, 0.6, 0.6), noise_octaves=3, noise_frequency=1.3,
                     noise_amplitude=1.4, noise_bias=-0.2)
    o, d, L = np.float32([0.1, 0.2, 0.3]), np.float32([0.6, 0.0, 0.8]), 6.0
    n = 200_000
    t = oracle.delta_track(m, o, d, L, n)
    assert ((t == -1) | ((t > 0) & (t < L))).all()
    s = np.linspace(0, L, 6001)
    den = oracle.medium_eval(m, "density", o[None, :] + s[:, None].astype(np.float32) * d[None, :]).astype(
        np.float64)
    hero = 0.8
    tau = np.concatenate([[0], np.cumsum(0.5 * (den[1:] + den[:-1]) * np.diff(s))]) * hero
    for x in (0.5, 1.5, 3.0, 5.9):
        surv = np.mean((t == -1) | (t > x))
        want = np.exp(-np.interp(x, s, tau))
        assert abs(surv - want) < 4 * np.sqrt(want * (1 - want) / n) + 2e-4, (x, surv, want)


def fog_scene(med):
    b = scenes.Builder()
    fl = b.add_object(scenes.material(color=(0.6, 0.6, 0.6)))
    b.quad(fl, (-3, -1.0, -1), (3, -1.0, -1), (3, -1.0, -7), (-3, -1.0, -7), normal=(0, 1, 0))
    light = b.add_object(scenes.material(color=(1.0, 1.0, 1.0), emission=8.0))
    b.quad(light, (-0.5, 1.5, -3.5), (0.5, 1.5, -3.5), (0.5, 1.5, -2.5), (-0.5, 1.5, -2.5), normal=(0, -1, 0))
    return b.finish(camera={"fov_deg": 60.0, "aperture": 1.8}, media=np.array([med], scenes.MEDIUM_DTYPE),
                    global_medium=0, env=scenes.constant_env(0.2))


@pytest.mark.parametrize("bias", [1.0, 0.5])
def test_constant_density_field_is_the_homogeneous_medium(bias):
    """amplitude 0: density = bias everywhere, accepted at every tentative
    collision; the image's expectation is that of the homogeneous medium of
    density x bias (grey coefficients)."""
    het = scenes.medium(sigma_a=(0.05,) * 3, sigma_s=(0.25,) * 3, g=0.3, density_field=1, noise_amplitude=0.0,
                        noise_bias=bias)
    hom = scenes.medium(sigma_a=(0.05,) * 3, sigma_s=(0.25,) * 3, g=0.3, density=bias)
    W, H, n = 24, 16, 1024
    a, aq = oracle.render(fog_scene(het), W, H, n, 5, mode=1, threads=THREADS, seed=41)
    b, bq = oracle.render(fog_scene(hom), W, H, n, 5, mode=1, threads=THREADS, seed=42)
    ma, mb = a / n, b / n
    z = (ma - mb) / np.sqrt(np.maximum(aq / n - ma ** 2, 0) / n + np.maximum(bq / n - mb ** 2, 0) / n + 1e-12)
    assert np.mean(np.abs(z) > 3.5) < 0.01
    assert abs(ma.mean() / mb.mean() - 1) < 0.01


# --- emission profiles ------------------------------------------------------------
def test_cone_profile_points():
    inner, outer, I, den = np.array([0.6, 0.7, 1.0]), np.array([1.0, 0.4, 0.05]), 25.0, 0.8
    m = scenes.medium(emission_mode=1, cone_origin=(1.0, 2.0, -3.0), cone_axis=(0.0, 0.0, -2.0), cone_length=2.0,
                      cone_radius=0.4, cone_intensity=I, cone_peak_axial=0.25, cone_peak_sharpness=4.0,
                      density=den)
    o, ax = np.array([1.0, 2.0, -3.0]), np.array([0.0, 0.0, -1.0])
    side = np.array([1.0, 0.0, 0.0])
    pts = np.array([o + ax * 0.5,                       # the peak, on the axis
                    o + ax * 0.5 + side * 0.4 * np.sqrt(0.5),   # v^2 = 1/2 at the peak
                    o + ax * 1.5,                       # u = 0.75
                    o - ax * 0.1, o + ax * 2.1,         # beyond the ends
                    o + ax * 0.5 + side * 0.41], np.float32)    # beyond the radius
    e = oracle.medium_eval(m, "emission", pts).astype(np.float64)
    np.testing.assert_allclose(e[0], inner * I * den, rtol=1e-6)
    np.testing.assert_allclose(e[1], (inner * 0.25 + outer * 0.75) * 0.25 * I * den, rtol=1e-5)
    t = np.exp(-((0.75 - 0.25) * 4.0) ** 2)
    np.testing.assert_allclose(e[2], (inner * t + outer * (1 - t)) * t * I * den, rtol=1e-5)
    assert (e[3:] == 0).all()


def test_path_profile_points():
    m = scenes.medium(emission_mode=2, path=[((0, 0, -2), 0.4, 0.0), ((2, 0, -2), 0.2, 1.0)],
                      path_inner=(1.0, 0.5, 0.2), path_outer=(0.1, 0.2, 0.9), path_intensity=3.0,
                      path_falloff_power=2.0, density=0.5)
    inner, outer = np.array([1.0, 0.5, 0.2]), np.array([0.1, 0.2, 0.9])
    pts = np.array([[0, 0, -2], [1, 0, -2], [1, 0.15, -2], [1, 0.31, -2], [-0.3, 0, -2], [2.3, 0.0, -2]], np.float32)
    e = oracle.medium_eval(m, "emission", pts).astype(np.float64)
    np.testing.assert_allclose(e[0], inner * 3.0 * 0.5, rtol=1e-6)
    mid = 0.5 * (inner + outer)
    np.testing.assert_allclose(e[1], mid * 3.0 * 0.5, rtol=1e-6)
    np.testing.assert_allclose(e[2], mid * (1 - 0.15 / 0.3) ** 2 * 3.0 * 0.5, rtol=1e-5)
    assert (e[3] == 0).all()                           # outside the lerped radius 0.3
    # beyond the ends: the clamped end point's tube
    np.testing.assert_allclose(e[4], inner * (1 - 0.3 / 0.4) ** 2 * 1.5, rtol=1e-5)
    assert (e[5] == 0).all()


@pytest.mark.parametrize("mode", [1, 2])
def test_emission_integral_mean_is_the_quadrature(mode):
    if mode == 1:
        m = scenes.medium(emission_mode=1, sigma_a=(0.3, 0.5, 0.7), cone_origin=(0, 0, 0), cone_axis=(1, 0, 0),
                          cone_length=2.0, cone_radius=0.5, cone_samples=4)
    else:
        m = scenes.medium(emission_mode=2, sigma_a=(0.3, 0.5, 0.7), path=[((0.2, 0, 0), 0.5, 0.0),
                                                                         ((1.8, 0.1, 0), 0.3, 1.0)],
                          path_samples=6)
    o, d, L = np.float32([-0.2, 0.05, 0.0]), np.float32([1.0, 0.0, 0.0]), 2.4
    n = 100_000
    est = oracle.emission_integral(m, o, d, L, n).astype(np.float64)
    s = np.linspace(0, L, 20001)
    se = oracle.medium_eval(m, "emission", o[None] + s[:, None].astype(np.float32) * d[None]).astype(np.float64)
    tr = np.exp(-np.outer(s, [0.3, 0.5, 0.7]))
    want = np.trapezoid(se * tr, s, axis=0)
    np.testing.assert_allclose(est.mean(0), want, rtol=4 * (est.std(0) / np.sqrt(n)).max() / want.min() + 2e-3)
    # constant emission: the closed form, every draw
    c = scenes.medium(sigma_a=(0.3, 0.5, 0.7), sigma_e=(2.0, 1.0, 0.5), density=1.5)
    ce = oracle.emission_integral(c, o, d, L, 4).astype(np.float64)
    st = np.array([0.3, 0.5, 0.7]) * 1.5
    np.testing.assert_allclose(ce, np.broadcast_to(np.array([2.0, 1.0, 0.5]) * 1.5 * (1 - np.exp(-st * L)) / st,
                                                   (4, 3)), rtol=1e-5)


# --- volume-light NEE -------------------------------------------------------------
def invisible_flame_plane(mode_const=False):
    """A diffuse plane y = -1 and an emissive cone medium no object contains:
    it lights the plane only through volume NEE."""
    b = scenes.Builder()
    fl = b.add_object(scenes.material(color=(0.5, 0.5, 0.5)))
    b.quad(fl, (-4, -1.0, -1), (4, -1.0, -1), (4, -1.0, -7), (-4, -1.0, -7), normal=(0, 1, 0))
    if mode_const:
        med = scenes.medium(sigma_e=(3.0, 3.0, 3.0))
    else:
        med = scenes.medium(emission_mode=1, cone_origin=(-1.0, 0.0, -4.0), cone_axis=(1.0, 0.0, 0.0),
                            cone_length=2.0, cone_radius=0.3, cone_intensity=4.0, cone_peak_axial=0.3,
                            cone_peak_sharpness=2.0)
    sc = b.finish(camera={"fov_deg": 60.0, "aperture": 1.8}, media=np.array([med], scenes.MEDIUM_DTYPE))
    sc["volume_lights"] = np.array([(0, (0.0, 0.0, -4.0))], scenes.VOLUME_LIGHT_DTYPE)
    return sc, med


@pytest.mark.parametrize("mode", [0, 1])
def test_volume_nee_is_the_line_integral(mode):
    from test_point_lights import centre_hits
    sc, med = invisible_flame_plane()
    W, H, n = 32, 24, 4096
    s, q = oracle.render(sc, W, H, n, 2, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    img = s.astype(np.float64) / n
    var = np.maximum(q / n - img ** 2, 0) / n
    p, ok = centre_hits(W, H, 60.0)
    ok &= (p[..., 2] > -7.0) & (p[..., 2] < -1.0) & (np.abs(p[..., 0]) < 4.0)
    tt = np.linspace(0, 2.0, 4001)
    axis_pts = np.stack([-1.0 + tt, np.zeros_like(tt), np.full_like(tt, -4.0)], -1)
    se = oracle.medium_eval(med, "emission", axis_pts.astype(np.float32)).astype(np.float64)
    pix = p[ok]
    lr = axis_pts[None, :, :] - pix[:, None, :]
    r2 = (lr ** 2).sum(-1)
    cos = np.maximum(lr[..., 1] / np.sqrt(r2), 0)
    want = 0.5 / np.pi * np.trapezoid(se[None, :, :] * (cos / r2)[..., None], tt, axis=1)
    got = img[ok]
    z = (got - want) / np.sqrt(var[ok] + (3e-4 * want) ** 2 + 1e-14)
    assert np.mean(np.abs(z) > 4) < 0.01
    assert abs(got.sum() / want.sum() - 1) < 0.01
    # a constant-emission volume is sampled at a point with pdf ~ 1e6: ~0
    sc2, _ = invisible_flame_plane(mode_const=True)
    s2, _ = oracle.render(sc2, W, H, 64, 2, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    assert (s2[ok] / 64).max() < 1e-5


# --- heat haze ----------------------------------------------------------------------
def test_heat_haze_bend():
    o, d = np.float32([0.2, -0.1, 0.0]), np.float32([0.0, 0.0, -1.0])
    m0 = scenes.medium(heat_haze=1, ior_amplitude=0.0)
    bo, bd = oracle.bend(m0, o, d, 2.0)
    np.testing.assert_array_equal(bd, d)
    np.testing.assert_allclose(bo, o + 2.0 * d, atol=1e-7)
    defl = []
    for amp in (1e-3, 2e-3):
        m = scenes.medium(heat_haze=1, ior_amplitude=amp, ior_frequency=3.0, ior_march_steps=64)
        bo, bd = oracle.bend(m, o, d, 2.0)
        assert abs(np.linalg.norm(bd) - 1) < 1e-6
        defl.append(bd.astype(np.float64) - d)
        # first order: d' - d = int grad_perp(n) / n ds along the straight ray
        s = (np.arange(64) + 0.5) / 64 * 2.0
        pts = o[None] + s[:, None].astype(np.float32) * d[None]
        h = np.float32(0.5 / 3.0)
        grad = np.stack([(oracle.medium_eval(m, "ior", pts + e) - oracle.medium_eval(m, "ior", pts - e)) / (2 * h)
                         for e in np.eye(3, dtype=np.float32) * h], -1).astype(np.float64)
        perp = grad - np.outer(grad @ d, d)
        want = perp.sum(0) * (2.0 / 64)
        np.testing.assert_allclose(defl[-1], want, atol=0.05 * np.abs(want).max() + 1e-7)
    np.testing.assert_allclose(defl[1], 2 * defl[0], rtol=0.02, atol=1e-6)   # z: second order
    n = oracle.medium_eval(scenes.medium(heat_haze=1, ior_amplitude=0.01), "ior",
                           np.random.default_rng(1).uniform(-3, 3, (2000, 3)).astype(np.float32))
    assert (n >= 1 - 0.01 - 1e-6).all() and (n <= 1 + 0.01 + 1e-6).all() and n.std() > 1e-3


# --- all of them: single-lobe vs recursive ---------------------------------------------
def test_single_lobe_matches_recursive_media():
    sc = scenes.media_room()
    W, H, n = 32, 24, 1024
    a, aq = oracle.render(sc, W, H, n, 6, mode=0, threads=THREADS, seed=51)
    b, bq = oracle.render(sc, W, H, n, 6, mode=1, threads=THREADS, seed=52)
    ma, mb = a / n, b / n
    va = np.maximum(aq / n - ma * ma, 0) / n
    vb = np.maximum(bq / n - mb * mb, 0) / n
    z = (ma - mb) / np.sqrt(va + vb + (3e-4 * ma) ** 2 + 1e-12)
    assert np.mean(np.abs(z) > 3.5) < 0.015
    assert abs(ma.mean() / mb.mean() - 1) < 0.01


# --- the loader -----------------------------------------------------------------------
def test_loader_reads_and_bakes_media(tmp_path):
    import json
    import os

    import rgapi
    sc = {
        "mainCamera": {"location": [0, 0, 4], "angle": [0, 0, 0], "fieldOfView": 45},
        "medium": {"heatHaze": True, "iorAmplitude": 0.01},        # a heat-haze global medium
        "floor": {"mesh": {"type": "plane"}, "location": [0, -1, 0], "scale": [4, 1, 4]},
        "flame": {
            "location": [-0.2, 0.4, 0], "scale": [1.6, 0.7, 0.7], "mesh": {"type": "cube"},
            "mat": {"transparency": 1.0},
            "medium": {"emissionMode": "cone", "coneOrigin": [0.6, 0.4, 0], "coneAxis": [-2, 0, 0],
                       "coneLength": 1.6, "coneRadius": 0.35, "coneIntensity": 25, "coneEmissionSamples": 8,
                       "densityField": "fbm", "noiseFrequency": 2.5, "noiseOctaves": 3, "noiseBias": -0.1}},
        "local": {
            "location": [1, 0, 0], "scale": [2, 2, 2], "mesh": {"type": "cube"}, "mat": {"transparency": 1.0},
            "medium": {"emissionMode": "cone", "coneOrigin": [0.5, 0, 0], "coneAxis": [0, 1, 0],
                       "coneFollowObject": True, "coneIntensity": 3}},
        "tube": {
            "mesh": {"type": "cube"}, "mat": {"transparency": 1.0},
            "medium": {"emissionMode": "path", "pathIntensity": 2.0, "pathFollowObject": False,
                       "pathPoints": [{"p": [0, 0, 0], "radius": 0.2, "t": 0}, {"p": [1, 1, 0], "t": 1}]}},
        "glow": {"location": [0, 2, -1], "mesh": {"type": "cube"}, "mat": {"transparency": 1.0},
                 "medium": {"emission": [1, 2, 3]}},
        "hidden": {"visible": False, "location": [5, 0, 0],
                   "inner": {"location": [0, 1, 0],
                             "medium": {"emissionMode": "cone", "coneIntensity": 1}}},
        "dud": {"mesh": {"type": "cube"}, "medium": {"sigma_a": [0, 0, 0]}},
    }
    p = os.path.join(tmp_path, "m.json")
    with open(p, "w") as f:
        json.dump(sc, f)
    h, d = rgapi.load_json(rgapi.Context(-1), p)
    assert any("heat-haze global medium dropped" in w for w in d["warnings"])
    assert d["global_medium"] == -1
    med = d["media"]
    # heat-haze global (kept in the table, unused), flame, local, tube, glow, hidden inner; the dud is dropped
    assert len(med) == 6
    f = med[1]
    assert f["emission_mode"] == 1 and f["cone_samples"] == 8 and f["density_field"] == 1
    assert f["noise_octaves"] == 3 and np.isclose(f["noise_frequency"], 2.5) and np.isclose(f["noise_bias"], -0.1)
    np.testing.assert_allclose(f["cone_origin"], (0.6, 0.4, -4.0), atol=1e-6)   # world -> render
    np.testing.assert_allclose(f["cone_axis"], (-1.0, 0.0, 0.0), atol=1e-6)     # normalised
    loc = med[2]
    # follows its object: (0.5, 0, 0) local -> 2 * 0.5 + 1 = 2 world -> (2, 0, -4)
    np.testing.assert_allclose(loc["cone_origin"], (2.0, 0.0, -4.0), atol=1e-6)
    np.testing.assert_allclose(loc["cone_axis"], (0.0, 1.0, 0.0), atol=1e-6)
    t = med[3]
    assert t["emission_mode"] == 2 and t["n_path"] == 2
    np.testing.assert_allclose(t["path"][1]["p"], (1.0, 1.0, -4.0), atol=1e-6)
    assert np.isclose(t["path"][1]["radius"], 0.1) and np.isclose(t["path"][0]["radius"], 0.2)
    vl = d["volume_lights"]
    # the flame (cone), the local cone, the constant glow and the hidden cone; not the path tube
    assert [int(v["medium"]) for v in vl] == [1, 2, 4, 5]
    np.testing.assert_allclose(vl[0]["pos"], (-0.2, 0.4, -4.0), atol=1e-6)      # the object's location
    np.testing.assert_allclose(vl[2]["pos"], (0.0, 2.0, -5.0), atol=1e-6)
    np.testing.assert_allclose(vl[3]["pos"], (0.0, 1.0, -4.0), atol=1e-6)       # own location, not composed

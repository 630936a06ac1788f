"""Shading systems 0-3 (-d 0..3, rayrenderer.cpp:59-77) in the restatement
(oracle/rg_oracle_pt.c), CPU only, through closed forms the providers fix:

* RaySimpleShaderProvider: 0.1 + max(dot(normalize(1, 1, 1), n), 0) for the
  interpolated normal, whatever the material;
* calcAO: the mean cosine of cosine-weighted directions is 2/3 under an open
  sky, and (2/3) cos^3(alpha) under a disc ceiling of angular radius alpha
  (the occluded cone removes 2 (1 - cos^3 alpha) / 3); its occluders are the
  triangles with transparency < 0.01, emissive ones included, within
  RAY_MAX_DISTANCE (100) for system 1 and 2.0 for system 3;
* lambertTraceLights: the sum of every point light's tracePointLight (not one
  picked), one area light's traceAreaLight (the light-strategy MIS part of a
  Lambert NEE, checked against quadrature over the light), plus worldColor;
  times 0.3 and the albedo;
* LambertShaderProvider's transparency: colour x (direct + 0.5) x (1 - t) +
  t x tracePath behind, the continuation shading even an invisible object;
  emission reads colour x emission; enableColorSampling off drops the albedo;
* an absorbing global medium routes the eye ray through tracePath: a sample
  reaches the surface with probability exp(-sigma d);
* the C ABI refuses an unknown system and a non-finite worldColor.
"""
import numpy as np
import pytest

import oracle
import rgapi
import scenes

f32 = np.float32


def simple_value(n):
    """0.1 + max(dot(normalize(1, 1, 1), n), 0) in binary32, the provider's order."""
    one = f32(1.0)
    inv = f32(1.0) / np.sqrt(f32(3.0), dtype=f32)
    lx = ly = lz = one * inv
    n = np.asarray(n, f32)
    d = lx * n[0] + ly * n[1] + lz * n[2]
    return f32(0.1) + max(d, f32(0.0))


def floor_scene(floor=None, extra=None, point_lights=None, media=None, global_medium=-1, normal=None):
    """A large floor z = -1 facing the camera at the origin; extra(b) adds
    objects.  A 1 x 1 image without antialiasing sees (0, 0, -1)."""
    b = scenes.Builder()
    fl = b.add_object(floor if floor is not None else scenes.material(color=(0.8, 0.6, 0.4)))
    b.quad(fl, (-50, -50, -1), (50, -50, -1), (50, 50, -1), (-50, 50, -1), normal=normal)
    if extra is not None:
        extra(b)
    return b.finish(camera={"fov_deg": 60.0, "exposure": 1.0, "aperture": 1.8, "depth_of_field": 0.0},
                    media=media, global_medium=global_medium, point_lights=point_lights)


def disc(b, mat, z, r, segs=256):
    """A disc of radius r at height z facing -z (a triangle fan)."""
    obj = b.add_object(mat)
    a = np.linspace(0.0, 2.0 * np.pi, segs + 1)
    for k in range(segs):
        p0 = (0.0, 0.0, z)
        p1 = (r * np.cos(a[k + 1]), r * np.sin(a[k + 1]), z)
        p2 = (r * np.cos(a[k]), r * np.sin(a[k]), z)
        b.v.append(np.asarray(p0 + p1 + p2, np.float32))
        b.n.append(np.tile(np.float32([0, 0, -1]), 3))
        b.o.append(obj)
        b.uv.append(np.zeros(6, np.float32))


def centre(sc, shader, spp=1, ao=0, **kw):
    """Mean and standard error of the single pixel of a 1 x 1 image."""
    s, q = oracle.render(sc, 1, 1, spp, 5, mode=0, threads=1, antialias=False, firefly_clamp=0.0, shader=shader,
                         ao_samples=ao, **kw)
    m = s[0, 0].astype(np.float64) / spp
    var = np.maximum(q[0, 0].astype(np.float64) / spp - m * m, 0.0)
    return m, np.sqrt(var / spp)


@pytest.mark.parametrize("normal", [(0.0, 0.0, 1.0), (0.3, -0.5, 0.8), (-0.6, -0.6, 0.52)])
@pytest.mark.parametrize("emission", [0.0, 3.0])
def test_simple_shader_closed_form(normal, emission):
    n = np.asarray(normal, np.float32)
    n = n / np.sqrt(np.float32((n * n).sum()))
    sc = floor_scene(floor=scenes.material(color=(0.2, 0.9, 0.4), emission=emission), normal=n)
    m, _ = centre(sc, 0, spp=3)
    want = simple_value(n.astype(np.float32))
    np.testing.assert_allclose(m, [want] * 3, rtol=2e-7)
    if np.dot(n, [1, 1, 1]) < 0:
        assert m[0] == pytest.approx(0.1, rel=1e-7)


def test_simple_shader_background_and_visibility():
    """A miss reads backColor (no environment); an invisible first hit too."""
    sc = floor_scene(floor=scenes.material(visible=0))
    m, _ = centre(sc, 0, spp=2)
    np.testing.assert_allclose(m, oracle.BACK_COLOR[:3], rtol=1e-7)
    m, _ = centre(sc, 0, spp=2, back_color=(0.3, 0.2, 0.1))
    np.testing.assert_allclose(m, [0.3, 0.2, 0.1], rtol=1e-7)


N_AO = 65536


def ao_expect(cos_a):
    """Mean and per-ray standard deviation of cos x [unoccluded] under
    cosine-weighted sampling with the cone theta < alpha occluded."""
    e = 2.0 / 3.0 * cos_a ** 3
    e2 = 0.5 * cos_a ** 4
    return e, np.sqrt(e2 - e * e)


def test_ao_open_sky_is_two_thirds():
    m, _ = centre(floor_scene(), 1, ao=N_AO)
    e, sd = ao_expect(1.0)
    assert abs(m[0] - e) < 5 * sd / np.sqrt(N_AO)
    assert m[0] == m[1] == m[2]


@pytest.mark.parametrize("height,radius", [(1.5, 1.5), (1.2, 3.6), (3.0, 1.0)])
def test_ao_disc_ceiling_closed_form(height, radius):
    sc = floor_scene(extra=lambda b: disc(b, scenes.material(), -1.0 + height, radius))
    m, _ = centre(sc, 1, ao=N_AO)
    e, sd = ao_expect(height / np.hypot(height, radius))
    assert abs(m[0] - e) < 5 * sd / np.sqrt(N_AO) + 2e-4   # the 256-gon is 1e-4 smaller than the disc


def test_ao_occluders_and_distance():
    """Transparent ceilings do not occlude, emissive opaque ones do; system 1
    looks 100 units out, system 3 two."""
    e45, sd = ao_expect(np.cos(np.pi / 4))
    tol = 5 * sd / np.sqrt(N_AO) + 2e-4
    clear = floor_scene(extra=lambda b: disc(b, scenes.material(transparency=0.5), 0.5, 1.5))
    m, _ = centre(clear, 1, ao=N_AO)
    assert abs(m[0] - 2.0 / 3.0) < 0.005
    lit = floor_scene(extra=lambda b: disc(b, scenes.material(emission=5.0), 0.5, 1.5))
    m, _ = centre(lit, 1, ao=N_AO)
    assert abs(m[0] - e45) < tol
    far = floor_scene(extra=lambda b: disc(b, scenes.material(), 149.0, 150.0))
    m, _ = centre(far, 1, ao=N_AO)
    assert abs(m[0] - 2.0 / 3.0) < 0.005              # 150 > RAY_MAX_DISTANCE
    # system 3 without lights: albedo x (w x 0.2 + w x 0.8 x AO), w = 0.3 x
    # worldColor.  A ceiling 3 up is beyond its 2.0; one 1.5 up occludes the
    # directions that reach it within 2.0, cos(theta) >= 0.75 (the disc's own
    # edge, cos = 0.707, lies farther out)
    high = floor_scene(extra=lambda b: disc(b, scenes.material(), 2.0, 3.0))
    low = floor_scene(extra=lambda b: disc(b, scenes.material(), 0.5, 1.5))
    e_low, sd_low = ao_expect(0.75)
    w = np.float32(oracle.WORLD_COLOR) * f32(0.3)
    alb = np.array([0.8, 0.6, 0.4])
    for sc, ao, t in ((high, 2.0 / 3.0, 0.005), (low, e_low, 5 * sd_low / np.sqrt(N_AO))):
        m, _ = centre(sc, 3, ao=N_AO)
        want = alb * (w * 0.2 + w * 0.8 * ao)
        np.testing.assert_allclose(m, want, rtol=0, atol=t * 0.8 * w.max() * 1.01 + 1e-7)
    m1, _ = centre(high, 1, ao=N_AO)
    e3, sd3 = ao_expect(3.0 / np.hypot(3.0, 3.0))
    assert abs(m1[0] - e3) < 5 * sd3 / np.sqrt(N_AO) + 2e-4   # system 1 sees it


def point_value(light_pos, color, emission, p=(0.0, 0.0, -1.0)):
    """tracePointLight of an omni light seen from p on the floor (n = +z), binary32."""
    lr = np.asarray(light_pos, f32) - np.asarray(p, f32)
    l = np.sqrt(f32((lr * lr).sum()))
    dto = lr[2] / l
    da = np.power(l, f32(-2.0))
    lum = f32(emission) * abs(dto) * da
    return np.clip(np.asarray(color, f32) * lum, 0, 1).astype(f32)


@pytest.mark.parametrize("shader", [2, 3])
def test_lambert_sums_every_point_light(shader):
    """Two omni lights: the pixel reads albedo x 0.3 x (v1 + v2 + worldColor)
    -- their sum, where traceLight would pick one.  System 3 without
    occluders in reach of its AO (the floor alone) scales the light by
    0.2 + 0.8 x AO, AO near 2/3."""
    lights = [scenes.point_light((0.0, 0.0, 0.2), color=(1.0, 0.5, 0.25), emission=0.9),
              scenes.point_light((0.5, -0.3, -0.4), color=(0.3, 0.6, 1.0), emission=0.5)]
    sc = floor_scene(point_lights=lights)
    v1 = point_value((0.0, 0.0, 0.2), (1.0, 0.5, 0.25), 0.9)
    v2 = point_value((0.5, -0.3, -0.4), (0.3, 0.6, 1.0), 0.5)
    light = ((v1 + v2) + np.float32(oracle.WORLD_COLOR)) * f32(0.3)
    alb = np.float32([0.8, 0.6, 0.4])
    if shader == 2:
        m, _ = centre(sc, 2, spp=2)
        np.testing.assert_allclose(m, light * alb, rtol=2e-6)
    else:
        m, _ = centre(sc, 3, ao=N_AO)
        _, sd = ao_expect(1.0)
        want = (light * 0.2 + light * 0.8 * (2.0 / 3.0)) * alb
        np.testing.assert_allclose(m, want, rtol=0, atol=5 * sd / np.sqrt(N_AO) * 0.8 * light.max())


def test_lambert_area_light_matches_quadrature():
    """One emissive square over the floor: E[traceAreaLight] = the integral
    over the light of Le cos_o cos_l / (pi r^2) x w_light, w_light the power
    heuristic against the cosine pdf (rayrenderer.cpp:1916-1974)."""
    h, a, Le = 1.2, 1.0, 4.0

    def light(b):
        o = b.add_object(scenes.material(color=(1.0, 1.0, 1.0), emission=Le))
        z = -1.0 + h
        b.quad(o, (-a / 2, -a / 2, z), (-a / 2, a / 2, z), (a / 2, a / 2, z), (a / 2, -a / 2, z))

    sc = floor_scene(extra=light)
    assert sc["n9"][-1][2] < 0    # the light faces the floor
    spp = 16384
    m, se = centre(sc, 2, spp=spp)
    k = 400
    u = (np.arange(k) + 0.5) / k * a - a / 2
    X, Y = np.meshgrid(u, u)
    r2 = X * X + Y * Y + h * h
    co = h / np.sqrt(r2)
    cl = co
    A = a * a
    pl = r2 / (cl * A)
    pb = co / np.pi
    w = pl * pl / (pl * pl + pb * pb)
    E = (Le * co * cl / (np.pi * r2) * w).sum() * (a / k) ** 2
    alb = np.array([0.8, 0.6, 0.4])
    want = alb * 0.3 * (E + np.asarray(oracle.WORLD_COLOR))
    assert np.all(np.abs(m - want) < 4 * se + 1e-4 * want), (m, want, se)


def test_lambert_transparency_closed_form():
    """A pane (transparency 0.4) at z = -1 over an opaque floor at z = -2, one
    omni light in front: c1 x (L1 + 0.5) x 0.6 + 0.4 x c2 x L2, L = 0.3 x
    (light + worldColor); the pane does not stop the light's shadow ray
    (environment occluders).  The continuation shades the floor even when it
    is invisible; with an invisible pane the eye reads the background."""
    c1, c2 = (0.9, 0.3, 0.2), (0.2, 0.5, 0.9)
    lp, lc, le = (0.0, 0.0, -0.5), (1.0, 0.9, 0.8), 0.05

    def build(floor_visible=1, pane_visible=1):
        b = scenes.Builder()
        fl = b.add_object(scenes.material(color=c2, visible=floor_visible))
        b.quad(fl, (-50, -50, -2), (50, -50, -2), (50, 50, -2), (-50, 50, -2))
        pn = b.add_object(scenes.material(color=c1, transparency=0.4, visible=pane_visible))
        b.quad(pn, (-50, -50, -1), (50, -50, -1), (50, 50, -1), (-50, 50, -1))
        return b.finish(camera={"fov_deg": 60.0}, point_lights=[scenes.point_light(lp, color=lc, emission=le)])

    world = np.float32(oracle.WORLD_COLOR)
    L1 = (point_value(lp, lc, le, (0, 0, -1)) + world) * f32(0.3)
    L2 = (point_value(lp, lc, le, (0, 0, -2)) + world) * f32(0.3)
    want = np.float32(c1) * (L1 + f32(0.5)) * f32(0.6) + (np.float32(c2) * L2) * f32(0.4)
    for fv in (1, 0):
        m, _ = centre(build(floor_visible=fv), 2, spp=2)
        np.testing.assert_allclose(m, want, rtol=2e-6)
    m, _ = centre(build(pane_visible=0), 2, spp=2)
    np.testing.assert_allclose(m, oracle.BACK_COLOR[:3], rtol=1e-7)
    # colour sampling off: the opaque floor reads L2 alone, the pane keeps its colour
    m, _ = centre(build(), 2, spp=2, color_sampling=False)
    np.testing.assert_allclose(m, np.float32(c1) * (L1 + f32(0.5)) * f32(0.6) + L2 * f32(0.4), rtol=2e-6)


@pytest.mark.parametrize("shader", [2, 3])
def test_lambert_emission_and_world_color(shader):
    sc = floor_scene(floor=scenes.material(color=(0.5, 0.25, 1.0), emission=3.0))
    m, _ = centre(sc, shader, spp=2, ao=4)
    np.testing.assert_allclose(m, [1.5, 0.75, 3.0], rtol=1e-7)
    dark = floor_scene()
    m, _ = centre(dark, shader, spp=2, ao=4, world_color=(0.0, 0.0, 0.0))
    assert (m == 0).all()
    m, _ = centre(dark, 2, spp=2, world_color=(1.0, 2.0, 3.0))
    np.testing.assert_allclose(m, np.float32([0.8, 0.6, 0.4]) * np.float32([1, 2, 3]) * f32(0.3), rtol=1e-6)


def test_absorbing_global_medium_routes_through_trace_path():
    """Pure absorption sigma_a = 0.3 (grey), surface at distance 1: tracePath
    returns the provider's value when the free flight passes the surface
    (probability exp(-0.3)), zero otherwise (sigma_s = 0: the scatter weight
    is zero)."""
    fog = scenes.medium(sigma_a=(0.3, 0.3, 0.3), sigma_s=(0.0, 0.0, 0.0))
    sc = floor_scene(media=[fog], global_medium=0)
    spp = 4096
    m, _ = centre(sc, 0, spp=spp)
    v = float(simple_value((0.0, 0.0, 1.0)))
    p = np.exp(-0.3)
    sd = v * np.sqrt(p * (1 - p) / spp)
    assert abs(m[0] - v * p) < 5 * sd
    assert m[0] == m[1] == m[2]


def test_c_abi_refuses_unknown_systems():
    ctx = rgapi.Context(-1)
    rs = rgapi.RenderScene(ctx, floor_scene())
    for bad in (dict(shader=4), dict(shader=-1), dict(shader=6), dict(world_color=(0.0, float("nan"), 0.0))):
        with pytest.raises(rgapi.RGError) as e:
            rs.render(rgapi.default_settings(width=4, height=4, spp=1, **bad))
        assert e.value.code == -1, bad
    with pytest.raises(rgapi.RGError) as e:   # a valid system reaches the device check
        rs.render(rgapi.default_settings(width=4, height=4, spp=1, shader=3))
    assert e.value.code == -2
    st = rgapi.default_settings()
    assert st.shader == 5
    np.testing.assert_array_equal(np.float32(list(st.world_color)), np.float32(oracle.WORLD_COLOR))

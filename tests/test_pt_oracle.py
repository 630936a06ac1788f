"""Pins of the path-tracing restatement (oracle/rg_oracle_pt.c) against closed
forms, independent quadrature and each other.  CPU only.

* recursive (reference call structure, every Mix lobe evaluated) vs
  single-lobe (the GPU's estimator): same expectation, checked per pixel with
  a z-test on the two estimates' own variances;
* environment furnace: a diffuse sphere of albedo rho under a uniform
  environment L reflects rho * L (NEE + BSDF sampling with MIS must sum to 1);
* area-light direct illumination of a diffuse plane against a dense
  quadrature of rho/pi * Le * cos cos' / r^2 over the light;
* GGX furnace: a metallic white sphere under a uniform environment returns
  L times the directional albedo of the separable-Smith GGX lobe, computed
  here by quadrature;
* absorbing medium: transmittance exp(-sigma_a * d) through a slab;
* chromatic dispersion: per band, a dispersive plane's Fresnel reflectance
  at its band's index of refraction (Schlick); a white dispersive glass
  sphere under a uniform environment returns L in every band, with the
  second moment of a single 3x band mask (3 L^2);
* A-Trous: constant images are fixed points, and with the edge-stopping terms
  switched off the filter is the B3-spline a-trous convolution.
"""
import numpy as np
import pytest

import oracle
import scenes

THREADS = 8


def test_single_lobe_matches_recursive_cube_room():
    sc = scenes.cube_room()
    W = H = 24
    n = 512
    a, aq = oracle.render(sc, W, H, n, 5, mode=0, threads=THREADS, seed=11)
    b, bq = oracle.render(sc, W, H, n, 5, mode=1, threads=THREADS, seed=12)
    ma, mb = a / n, b / n
    va = np.maximum(aq / n - ma * ma, 0) / n
    vb = np.maximum(bq / n - mb * mb, 0) / n
    z = (ma - mb) / np.sqrt(va + vb + 1e-12)
    assert np.mean(np.abs(z) > 3.5) < 0.01
    assert abs(ma.mean() / mb.mean() - 1) < 0.01


def test_single_lobe_matches_recursive_materials():
    sc = scenes.bsdf_frame(grid=3, stacks=10, slices=14)
    W, H = 36, 24
    n = 512
    a, aq = oracle.render(sc, W, H, n, 8, mode=0, threads=THREADS, seed=21)
    b, bq = oracle.render(sc, W, H, n, 8, mode=1, threads=THREADS, seed=22)
    ma, mb = a / n, b / n
    va = np.maximum(aq / n - ma * ma, 0) / n
    vb = np.maximum(bq / n - mb * mb, 0) / n
    z = (ma - mb) / np.sqrt(va + vb + 1e-12)
    assert np.mean(np.abs(z) > 3.5) < 0.02
    assert abs(ma.mean() / mb.mean() - 1) < 0.02


def sphere_under_env(mat, env_value=1.0, stacks=48, slices=64):
    b = scenes.Builder()
    o = b.add_object(mat)
    b.sphere(o, (0.0, 0.0, -4.0), 1.0, stacks, slices)
    return b.finish(camera={"fov_deg": 40.0, "aperture": 1.8}, env=scenes.constant_env(env_value, 128, 64))


@pytest.mark.parametrize("mode", [0, 1])
def test_environment_furnace_diffuse(mode):
    rho = 0.5
    sc = sphere_under_env(scenes.material(color=(rho, rho, rho)))
    W = H = 16
    n = 2048
    s, _ = oracle.render(sc, W, H, n, 8, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    c = s[6:10, 6:10] / n  # centre pixels: on the sphere, facing the camera
    assert abs(c.mean() / rho - 1) < 0.01
    corner = s[0, 0] / n  # misses the sphere: sees the environment
    assert np.allclose(corner, 1.0, rtol=1e-5)


def lambert_form_factor_radiance(p, nrm, quad, Le, rho, m=200):
    """rho/pi * Le * integral over the light of cos cos' / r^2 dA (midpoint rule)."""
    q = [np.asarray(v, np.float64) for v in quad]
    e1, e2 = q[1] - q[0], q[3] - q[0]
    ln = np.cross(e1, e2)
    area = np.linalg.norm(ln)
    ln /= area
    uu, vv = np.meshgrid((np.arange(m) + 0.5) / m, (np.arange(m) + 0.5) / m)
    pts = q[0] + uu[..., None] * e1 + vv[..., None] * e2
    d = pts - p
    r2 = (d * d).sum(-1)
    dn = d / np.sqrt(r2)[..., None]
    cos_p = np.clip(dn @ nrm, 0, None)
    cos_l = np.clip(-(dn @ ln), 0, None)
    return rho / np.pi * Le * (cos_p * cos_l / r2).mean() * area


@pytest.mark.parametrize("mode", [0, 1])
def test_area_light_direct_closed_form(mode):
    rho, Le = 0.6, 12.0
    b = scenes.Builder()
    fl = b.add_object(scenes.material(color=(rho, rho, rho)))
    b.quad(fl, (-20, -1, 20), (20, -1, 20), (20, -1, -20), (-20, -1, -20), normal=(0, 1, 0))
    li = b.add_object(scenes.material(color=(1, 1, 1), emission=Le))
    quad = [(-0.3, 0.5, -3.3), (0.3, 0.5, -3.3), (0.3, 0.5, -2.7), (-0.3, 0.5, -2.7)]
    b.quad(li, *quad, normal=(0, -1, 0))
    sc = b.finish(camera={"fov_deg": 30.0, "aperture": 1.8})
    W, H = 8, 8
    n = 4096
    s, _ = oracle.render(sc, W, H, n, 4, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    # pixel (4, 6): ray through its centre hits the floor at a known point
    x, y = 4, 6
    vs = 2 * np.tan(np.radians(15.0)) / H
    d = np.array([(x + 0.5 - W / 2) * vs, -(y + 0.5 - H / 2) * vs, -1.0])
    t = -1.0 / d[1]
    p = d * t
    want = lambert_form_factor_radiance(p, np.array([0, 1.0, 0]), quad, Le, rho)
    got = s[y, x].mean() / n
    assert abs(got / want - 1) < 0.02, (got, want)


def ggx_albedo_separable(alpha, cos_v, n=400):
    """integral over the hemisphere of D(h) G1(v) G1(l) / (4 cos_v) (F = 1)."""
    v = np.array([np.sqrt(1 - cos_v ** 2), 0.0, cos_v])

    def g1(w):
        tan2 = (1 - w[..., 2] ** 2) / np.maximum(w[..., 2] ** 2, 1e-30)
        return np.where(w[..., 2] > 0, 2 / (1 + np.sqrt(1 + alpha ** 2 * tan2)), 0.0)

    th = (np.arange(n) + 0.5) / n * (np.pi / 2)
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * (2 * np.pi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    L = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1)
    Hh = L + v
    Hh /= np.linalg.norm(Hh, axis=-1, keepdims=True)
    den = (Hh[..., 0] ** 2 + Hh[..., 1] ** 2) / alpha ** 2 + Hh[..., 2] ** 2
    D = 1 / (np.pi * alpha ** 2 * den ** 2)
    f = D * g1(v[None, None]) * g1(L) / (4 * cos_v)
    dw = np.sin(T) * (np.pi / 2 / n) * (2 * np.pi / (2 * n))
    return (f * dw).sum()


@pytest.mark.parametrize("mode", [0, 1])
def test_ggx_furnace(mode):
    rough = 0.6
    sc = sphere_under_env(scenes.material(color=(1, 1, 1), glossy=1.0, metallic=1.0, roughness=rough))
    W = H = 16
    n = 4096
    s, _ = oracle.render(sc, W, H, n, 8, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    got = s[7:9, 7:9].mean() / n
    # centre pixels see the sphere almost head-on
    want = ggx_albedo_separable(rough * rough, 0.994)
    assert abs(got / want - 1) < 0.02, (got, want)


@pytest.mark.parametrize("mode", [0, 1])
def test_absorbing_medium_transmittance(mode):
    """Camera inside a purely absorbing global medium looking at an emitter:
    E * exp(-sigma * d).  (A medium inside a closed transparency-1 mesh is not
    left again in the reference: the silhouette fix of tracePath,
    rayrenderer.cpp:1842-1846, flips the exit face's shading normal, so
    TransparencyShader re-enters the interior medium; the restatement keeps
    that behaviour.)"""
    sigma, E, dist = 0.7, 3.0, 2.0
    b = scenes.Builder()
    li = b.add_object(scenes.material(color=(1, 1, 1), emission=E))
    b.quad(li, (-50, -50, -dist), (50, -50, -dist), (50, 50, -dist), (-50, 50, -dist), normal=(0, 0, 1))
    med = np.zeros(1, scenes.MEDIUM_DTYPE)
    med[0]["sigma_a"] = sigma
    med[0]["density"] = 1.0
    sc = b.finish(camera={"fov_deg": 2.0, "aperture": 1.8}, media=med, global_medium=0)
    n = 8192
    s, _ = oracle.render(sc, 2, 2, n, 8, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    got = s.mean() / n
    want = E * np.exp(-sigma * dist)
    assert abs(got / want - 1) < 0.03, (got, want)


def schlick(cos_t, ior):
    r0 = ((1 - ior) / (1 + ior)) ** 2
    return r0 + (1 - r0) * (1 - cos_t) ** 5


@pytest.mark.parametrize("mode", [0, 1])
def test_dispersion_fresnel_per_band(mode):
    """A smooth dispersive refractive plane seen at 45 degrees under a
    uniform environment L, a black plane behind it: only the Fresnel-reflected
    branch reaches the environment, so band c returns L * F(cos 45, ior_c)
    with ior = n (1 - s), n, n (1 + s) for red, green, blue (bsdf.cpp:394-420:
    the band is drawn with probability 1/3 and masked by 3)."""
    n_ior, s, L = 1.5, 0.2, 1.0
    b = scenes.Builder()
    g = b.add_object(scenes.material(color=(1, 1, 1), refraction=1.0, refraction_ratio=n_ior, roughness=0.0,
                                     chroma_dispersion=s))
    c = np.array([0.0, 0.0, -5.0])
    t1, t2 = np.array([1.0, 0, 0]), np.array([0, 1.0, -1.0]) / np.sqrt(2)
    b.quad(g, c - 2 * t1 - 2 * t2, c + 2 * t1 - 2 * t2, c + 2 * t1 + 2 * t2, c - 2 * t1 + 2 * t2,
           normal=np.array([0, 1.0, 1.0]) / np.sqrt(2))
    k = b.add_object(scenes.material(color=(0, 0, 0)))
    b.quad(k, (-50, -50, -8), (50, -50, -8), (50, 50, -8), (-50, 50, -8), normal=(0, 0, 1))
    sc = b.finish(camera={"fov_deg": 1.0, "aperture": 1.8}, env=scenes.constant_env(L, 64, 32))
    W = H = 16
    n = 4096
    sm, _ = oracle.render(sc, W, H, n, 4, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    got = sm.reshape(-1, 3).mean(0) / n
    ct = np.cos(np.pi / 4)
    want = L * np.array([schlick(ct, n_ior * (1 - s)), schlick(ct, n_ior), schlick(ct, n_ior * (1 + s))])
    # about 3,600 red, 15,000 green and 29,000 blue reflections: <= 1.7% noise
    np.testing.assert_allclose(got, want, rtol=0.08)


@pytest.mark.parametrize("mode", [0, 1])
def test_dispersion_white_furnace(mode):
    """A smooth white dispersive glass sphere under a uniform environment L
    loses nothing, so every band returns L; the band drawn at the first
    interface is kept at the second, so a sample is 3 L in one band and 0 in
    the others and the second moment is 3 L^2 (a band drawn again at each
    interface would give 9 L^2)."""
    sc = sphere_under_env(scenes.material(color=(1, 1, 1), refraction=1.0, refraction_ratio=1.5, roughness=0.0,
                                          chroma_dispersion=0.1))
    W = H = 16
    n = 4096
    s, q = oracle.render(sc, W, H, n, 12, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    m1 = s[5:11, 5:11].reshape(-1, 3).mean(0) / n
    m2 = q[5:11, 5:11].reshape(-1, 3).mean(0) / n
    np.testing.assert_allclose(m1, 1.0, rtol=0.03)
    # Russian roulette reweights the few paths with three or more events
    np.testing.assert_allclose(m2, 3.0, rtol=0.06)


def test_medium_is_not_left_through_a_transparent_exit_face():
    sigma = 0.7
    b = scenes.Builder()
    box = b.add_object(scenes.material(color=(1, 1, 1), transparency=1.0, interior_medium=0))
    b.box(box, (-3, -3, -4.0), (3, 3, -2.0))
    med = np.zeros(1, scenes.MEDIUM_DTYPE)
    med[0]["sigma_a"] = sigma
    med[0]["density"] = 1.0
    sc = b.finish(camera={"fov_deg": 10.0, "aperture": 1.8}, media=med, env=scenes.constant_env(1.0, 64, 32))
    for mode in (0, 1):
        s, _ = oracle.render(sc, 4, 4, 64, 8, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
        assert s.max() == 0.0  # absorbed: the ray keeps the medium after the box


# --- A-Trous denoiser (rayrenderer.cpp:2536-2772) ----------------------------
def guides(h, w):
    nrm = np.zeros((h, w, 4), np.float32)
    nrm[..., :3] = (0.5, 0.5, 1.0)  # encoded n*0.5+0.5 of (0, 0, 1)
    nrm[..., 3] = 1.0
    alb = np.ones((h, w, 4), np.float32)
    dep = np.full((h, w), 0.5, np.float32)
    return nrm, alb, dep


def test_denoise_constant_is_fixed_point():
    h, w = 20, 24
    img = np.zeros((h, w, 4), np.float32)
    img[..., :3] = (0.3, 0.5, 0.7)
    img[..., 3] = 1
    nrm, alb, dep = guides(h, w)
    out = oracle.denoise(img, nrm, alb, dep)
    np.testing.assert_allclose(out[..., :3], img[..., :3], rtol=2e-6)


def b3_atrous(img, levels):
    """Edge-stopping switched off: normalised separable B3 a-trous."""
    k = np.array([1, 4, 6, 4, 1], np.float64) / 16
    h, w = img.shape[:2]
    cur = img.astype(np.float64)
    for l in range(levels):
        s = 1 << l
        acc = np.zeros_like(cur)
        wsum = np.zeros((h, w, 1))
        for ky in range(5):
            for kx in range(5):
                dy, dx = (ky - 2) * s, (kx - 2) * s
                ys, xs = np.arange(h) + dy, np.arange(w) + dx
                vy, vx = (ys >= 0) & (ys < h), (xs >= 0) & (xs < w)
                wgt = k[ky] * k[kx] * (vy[:, None] & vx[None, :])[..., None]
                acc += wgt * cur[np.clip(ys, 0, h - 1)][:, np.clip(xs, 0, w - 1)]
                wsum += wgt
        cur = acc / wsum
    return cur


def test_denoise_reduces_to_b3_atrous():
    rng = np.random.default_rng(3)
    h, w = 40, 36
    img = np.zeros((h, w, 4), np.float32)
    img[..., :3] = rng.uniform(0.2, 0.4, (h, w, 3))  # smooth enough to dodge the firefly clamp
    img[..., 3] = 1
    nrm, alb, dep = guides(h, w)
    out = oracle.denoise(img, nrm, alb, dep, levels=3, sigma_c=1e6, sigma_n=0.0, sigma_d=1e9)
    want = b3_atrous(img[..., :3], 3)
    np.testing.assert_allclose(out[..., :3], want, rtol=2e-5, atol=1e-6)


def test_adaptive_driver_concentrates_samples():
    sc = scenes.cube_room()
    sm, sq, nt = oracle.render_adaptive(sc, 96, 64, 64, 5, base=4, threshold=0.12, threads=THREADS)
    assert nt.min() >= 4 and nt.max() <= 64
    assert len(np.unique(nt)) > 1          # some tiles stopped early
    assert all(v % 4 == 0 for v in nt)     # passes of adaptiveBaseSamples


def textured_plane(tex, srgb, tiling=(1.0, 1.0)):
    """A diffuse plane facing the camera (z = -3, x, y in [-1, 1]) under a
    uniform environment, UV (0, 0) at its top-left corner and (1, 1) at the
    bottom-right, carrying texture `tex`."""
    b = scenes.Builder()
    t = b.add_texture(tex, srgb=srgb)
    o = b.add_object(scenes.material(color=(0.8, 0.8, 0.8)), texture=t, tiling=tiling)
    b.quad(o, (-1, 1, -3), (-1, -1, -3), (1, -1, -3), (1, 1, -3), normal=(0, 0, 1),
           uv=((0, 0), (0, 1), (1, 1), (1, 0)))
    return b.finish(camera={"fov_deg": 30.0, "aperture": 1.8}, env=scenes.constant_env(1.0, 64, 32))


def srgb_to_linear_f64(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("srgb", [False, True])
def test_texture_furnace_quadrants(mode, srgb):
    """Under a uniform environment of radiance 1 a diffuse surface reflects
    exactly its albedo (the furnace above), so every pixel of the plane reads
    colour x texel: the 2x2 texture's top-left texel on the top-left
    quadrant, and so on (nearest texel, V down, sRGB decoded)."""
    tex = np.array([[[0.9, 0.2, 0.1], [0.1, 0.8, 0.3]],
                    [[0.4, 0.4, 0.9], [0.7, 0.05, 0.6]]], np.float32)
    sc = textured_plane(tex, srgb)
    W = H = 40
    n = 512
    s, _ = oracle.render(sc, W, H, n, 8, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    img = s / n
    lin = srgb_to_linear_f64(tex) if srgb else tex.astype(np.float64)
    # the plane spans the middle of the view (30 degree fov at distance 3)
    for (r0, c0), (ty, tx) in (((8, 8), (0, 0)), ((8, 26), (0, 1)), ((26, 8), (1, 0)), ((26, 26), (1, 1))):
        q = img[r0:r0 + 6, c0:c0 + 6].reshape(-1, 3).mean(0)
        np.testing.assert_allclose(q, 0.8 * lin[ty, tx], rtol=0.02)


def test_texture_tiling_repeats_and_wraps():
    """texTiling (2, 2) repeats the texture twice across the plane; a UV
    of exactly 1 wraps to texel 0 (the fold into [0, 1))."""
    tex = np.array([[[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]], np.float32)   # 1 x 2: red | green
    sc = textured_plane(tex, False, tiling=(2.0, 1.0))
    W = H = 48
    s, _ = oracle.render(sc, W, H, 64, 8, mode=1, threads=THREADS, antialias=False, firefly_clamp=0.0)
    row = (s / 64)[24]
    # the plane fills the view: pixel centre c looks at x = (c + 0.5 - 24) *
    # 2 tan(15 deg) * 3 / 48, u = (x + 1) / 2, texel = floor(frac(2u) * 2)
    half = 3.0 * np.tan(np.radians(15.0))
    x = (np.arange(W) + 0.5 - W / 2) * (2 * half / W)
    su = 2.0 * (x + 1.0) / 2.0
    f = su - np.floor(su)
    want = np.floor(f * 2).astype(int)
    away = np.abs(f * 2 - np.round(f * 2)) > 0.1          # not at a band edge
    got = np.argmax(row, axis=1)
    assert away.sum() > 30 and np.array_equal(got[away], want[away])
    assert set(want[away]) == {0, 1} and (np.diff(want[away]) != 0).sum() >= 3   # four bands


def test_dispersion_band_left_by_the_eye_hit_mix_branch():
    """A transparent, refractive, dispersive pane seen directly: the
    reference evaluates its Mix branch, whose refraction lobe draws a
    dispersion band and leaves it in the BSDFParam, before the transparency
    branch traces on with that band (rayrenderer.cpp:2889-2894; bsdf.cpp:
    402-407).  The single-lobe estimator must carry such a band on its
    transparency path to have the recursive restatement's expectation: with a
    banded environment behind a dispersive sphere the difference is large
    (over half the pixel channels beyond 3.5 sigma without it)."""
    sc = scenes.dispersive_window()
    W = H = 16
    n = 8192
    a, aq = oracle.render(sc, W, H, n, 8, mode=0, threads=THREADS, seed=1, firefly_clamp=0.0)
    b, bq = oracle.render(sc, W, H, n, 8, mode=1, threads=THREADS, seed=2, firefly_clamp=0.0)
    ma, mb = a / n, b / n
    va = np.maximum(aq / n - ma * ma, 0) / n
    vb = np.maximum(bq / n - mb * mb, 0) / n
    z = (ma - mb) / np.sqrt(va + vb + 1e-12)
    assert np.mean(np.abs(z) > 3.5) <= 0.02
    assert abs(ma.mean() / mb.mean() - 1) < 0.005

"""More pins of the path-tracing restatement (oracle/rg_oracle_pt.c): each
restated function that the reference's arithmetic fixes is checked against
a closed form or an independent computation that a plausible mistake (a
dropped term, a flipped sign, a transposed axis) would fail.  CPU only.

* Henyey-Greenstein: sampled cos(theta) about wo follows the textbook HG
  density with mean g (g in {-0.7, 0.3, 0.8}), azimuths are uniform, and
  phasePdf integrates to one and is largest along wo for g > 0;
* scattering-medium furnace: a camera inside a chromatic, partly absorbing
  global medium under a uniform environment returns, per channel, the closed
  form of the reference's estimator: escape (1 - p_c) per segment,
  single-scattering albedo rho_c p_c per scatter, and in-medium environment
  NEE that keeps only directions toward the incoming side weighted by
  phasePdf(-dir, .), i.e. the HG mass F(g) of the hemisphere around -dir;
* tile noise: synthetic per-pixel sums with known mean and deviation give the
  mean of max-channel sigma / (m sqrt n), dark pixels skipped, 1e9 below two
  samples, 0 for an all-dark tile, ragged border tiles;
* A-Trous: one level on step-edge images with one edge-stopping term at a
  time (colour, normal, depth) equals the float64 weighted average with
  exp(-dL^2 / (2 sc^2)), max(n.n', 0)^sn and exp(-|dz| / (sd step));
* non-constant environment: sampled directions land in each texel with the
  probability the pdf integrates to there, the pdf of a sampled direction is
  envmapPdf of it, and an upward diffuse plane under a sky of two
  hemispheres reflects rho * L_up (NEE + BSDF sampling with MIS);
* GGX delta mirror: a roughness-0 metallic plane under a uniform sky returns
  Schlick F(cos i) * L;
* anisotropic GGX: the furnace at normal incidence against float64
  quadrature of the anisotropic separable-Smith albedo (alpha_x, alpha_y
  from roughness^2 (1 -/+ anisotropy)).
"""
import numpy as np
import pytest

import oracle
import scenes

THREADS = 8


# --- Henyey-Greenstein -------------------------------------------------------
def hg_mu_density(g, mu):
    """Textbook HG density of mu = cos(theta) (integrates to 1 over [-1, 1])."""
    return 0.5 * (1 - g * g) / (1 + g * g - 2 * g * mu) ** 1.5


@pytest.mark.parametrize("g", [-0.7, 0.3, 0.8])
def test_hg_sampling_follows_the_hg_density(g):
    rng = np.random.default_rng(5)
    wo = np.array([0.3, -0.5, 0.81], np.float64)
    wo /= np.linalg.norm(wo)
    n = 400_000
    d = oracle.phase_sample(g, wo.astype(np.float32), rng.random((n, 2), dtype=np.float32)).astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    mu = d @ wo
    # the mean cosine of HG(g) is g
    assert abs(mu.mean() - g) < 4 * mu.std() / np.sqrt(n) + 1e-3
    edges = np.linspace(-1, 1, 41)
    hist, _ = np.histogram(mu, edges)
    fine = np.linspace(-1, 1, 40 * 200 + 1)
    mids = 0.5 * (fine[1:] + fine[:-1])
    want = (hg_mu_density(g, mids) * (fine[1] - fine[0])).reshape(40, 200).sum(1) * n
    big = want > 200
    z = (hist[big] - want[big]) / np.sqrt(want[big])
    assert np.abs(z).max() < 5, np.abs(z).max()
    # azimuth about wo is uniform
    t1 = np.cross(wo, [1.0, 0, 0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(wo, t1)
    phi = np.arctan2(d @ t2, d @ t1)
    h, _ = np.histogram(phi, np.linspace(-np.pi, np.pi, 17))
    assert np.abs(h / (n / 16) - 1).max() < 0.03


@pytest.mark.parametrize("g", [-0.7, 0.0, 0.3, 0.8])
def test_hg_pdf_normalised_and_matches_sampling(g):
    rng = np.random.default_rng(9)
    wo = np.array([0.0, 0.6, -0.8], np.float32)
    n = 1_000_000
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    r = np.sqrt(1 - z * z)
    wi = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1).astype(np.float32)
    p = oracle.phase_pdf(g, wo, wi).astype(np.float64)
    # uniform-sphere Monte Carlo of the integral of the pdf
    assert abs(p.mean() * 4 * np.pi - 1) < 0.02
    # as a function of mu = wo . wi it is the HG density / (2 pi)
    mu = wi.astype(np.float64) @ wo.astype(np.float64)
    np.testing.assert_allclose(p * 2 * np.pi, hg_mu_density(g, mu), rtol=2e-4, atol=1e-6)
    if g > 0:
        at = oracle.phase_pdf(g, wo, np.stack([wo, -wo]))
        assert at[0] > at[1]


# --- scattering medium furnace ----------------------------------------------
def hg_hemisphere_mass(g):
    """HG(g) probability of mu > 0 (float64 quadrature of the density)."""
    mu = (np.arange(200_000) + 0.5) / 200_000
    return hg_mu_density(g, mu).mean()


def medium_furnace_scene(g, sigma_a, sigma_s, env=1.0):
    b = scenes.Builder()
    o = b.add_object(scenes.material())
    # one tiny triangle far behind the camera: the scene needs geometry, no
    # ray ever meets it
    b.quad(o, (0, 0, 900), (1e-4, 0, 900), (1e-4, 1e-4, 900), (0, 1e-4, 900), normal=(0, 0, -1))
    med = np.zeros(1, scenes.MEDIUM_DTYPE)
    med[0]["sigma_a"] = sigma_a
    med[0]["sigma_s"] = sigma_s
    med[0]["g"] = g
    med[0]["density"] = 1.0
    return b.finish(camera={"fov_deg": 5.0, "aperture": 1.8}, media=med, global_medium=0,
                    env=scenes.constant_env(env, 256, 128))


def medium_furnace_expected(g, sigma_a, sigma_s, depth, dist=100.0):
    """Per channel: sum over segments k <= depth of (rho p)^k (1 - p) (escape)
    plus F(g) * sum over scatters k < depth of (rho p)^(k+1) (NEE)."""
    sa, ss = np.asarray(sigma_a, np.float64), np.asarray(sigma_s, np.float64)
    st = sa + ss
    p = 1 - np.exp(-st * dist)
    a = ss / st * p
    k = np.arange(depth + 1)[:, None]
    esc = (a[None] ** k * (1 - p)[None]).sum(0)
    nee = hg_hemisphere_mass(g) * (a[None] ** (k[:depth] + 1)).sum(0)
    return esc + nee


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("g", [-0.5, 0.0, 0.6])
def test_scattering_medium_furnace(mode, g):
    sa, ss, depth = (0.002, 0.0, 0.004), (0.006, 0.01, 0.008), 6
    sc = medium_furnace_scene(g, sa, ss)
    n = 8192
    s, _ = oracle.render(sc, 4, 4, n, depth, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0,
                         back_color=(0, 0, 0))
    got = s.reshape(-1, 3).mean(0) / n
    want = medium_furnace_expected(g, sa, ss, depth)
    np.testing.assert_allclose(got, want, rtol=0.02)
    # F(g) separates the signs: the NEE term alone differs by > 10% between g and -g
    if g != 0:
        assert abs(hg_hemisphere_mass(g) / hg_hemisphere_mass(-g) - 1) > 0.5


# --- tile noise --------------------------------------------------------------
def test_tile_noise_closed_form():
    rng = np.random.default_rng(2)
    W, H, n = 70, 40, 16     # tiles 3 x 2, ragged on the right and bottom
    m = rng.uniform(0.05, 2.0, (H, W, 3))
    sd = rng.uniform(0.0, 1.0, (H, W, 3))
    m[5:9, 3:20] = rng.uniform(0, 9e-4, (4, 17, 3))    # dark pixels: skipped
    s = (m * n).astype(np.float32)
    q = ((sd * sd + m * m) * n).astype(np.float32)
    for ty in range(2):
        for tx in range(3):
            y0, x0 = 32 * ty, 32 * tx
            mm = s[y0:y0 + 32, x0:x0 + 32].astype(np.float64) / n
            var = np.maximum(q[y0:y0 + 32, x0:x0 + 32].astype(np.float64) / n - mm * mm, 0)
            rsem = np.sqrt(var / n) / np.maximum(mm, 1e-3)
            keep = mm.max(-1) >= 1e-3
            want = rsem.max(-1)[keep].mean()
            got = oracle.tile_noise(s, q, tx, ty, n)
            assert abs(got / want - 1) < 1e-4, (tx, ty, got, want)
    assert oracle.tile_noise(s, q, 0, 0, 1) == 1e9
    dark = np.zeros_like(s)
    assert oracle.tile_noise(dark, dark, 1, 1, n) == 0.0


# --- A-Trous edge-stopping terms ---------------------------------------------
def one_level(img, nrm_vec, dep, sc, sn, sd):
    """float64 one-level a-trous (step 1) with the reference's three weights."""
    k = np.array([1, 4, 6, 4, 1], np.float64) / 16
    h, w = img.shape[:2]
    lum = img @ np.array([0.2126, 0.7152, 0.0722])
    out = np.zeros_like(img)
    for y in range(h):
        for x in range(w):
            acc = np.zeros(3)
            ws = 0.0
            for ky in range(5):
                for kx in range(5):
                    yy, xx = y + ky - 2, x + kx - 2
                    if not (0 <= yy < h and 0 <= xx < w):
                        continue
                    wc = np.exp(-(lum[yy, xx] - lum[y, x]) ** 2 / (2 * sc * sc))
                    wn = max(float(nrm_vec[yy, xx] @ nrm_vec[y, x]), 0.0) ** sn
                    wd = np.exp(-abs(dep[yy, xx] - dep[y, x]) / sd)
                    wt = k[ky] * k[kx] * wc * wn * wd
                    acc += wt * img[yy, xx]
                    ws += wt
            out[y, x] = acc / ws
    return out


@pytest.mark.parametrize("term", ["colour", "normal", "depth"])
def test_atrous_edge_stopping_term(term):
    h, w = 12, 14
    img = np.full((h, w, 3), 0.25)
    img[:, 7:] = (0.4, 0.35, 0.3)        # a step edge in every case
    nv = np.tile([0.0, 0.0, 1.0], (h, w, 1))
    dep = np.full((h, w), 0.3)
    sc, sn, sd = 1e6, 0.0, 1e9
    if term == "colour":
        sc = 0.05
    elif term == "normal":
        nv[:, 7:] = (0.6, 0.0, 0.8)
        sn = 6.0
    else:
        dep[:, 7:] = 0.45
        sd = 0.1
    rgba = np.concatenate([img, np.ones((h, w, 1))], -1).astype(np.float32)
    nrm = np.concatenate([nv * 0.5 + 0.5, np.ones((h, w, 1))], -1).astype(np.float32)
    alb = np.ones((h, w, 4), np.float32)
    out = oracle.denoise(rgba, nrm, alb, dep.astype(np.float32), levels=1, sigma_c=sc, sigma_n=sn, sigma_d=sd)
    want = one_level(img, (nrm[..., :3].astype(np.float64) * 2 - 1), dep.astype(np.float32).astype(np.float64),
                     sc, sn, sd)
    np.testing.assert_allclose(out[..., :3], want, rtol=3e-5, atol=1e-6)
    # the term does stop the filter at the edge: the pixel next to it keeps
    # most of its own value
    plain = one_level(img, nv, dep, 1e6, 0.0, 1e9)
    assert np.abs(want[:, 6] - img[:, 6]).max() < 0.5 * np.abs(plain[:, 6] - img[:, 6]).max()


# --- non-constant environment -----------------------------------------------
def env_scene(rgb, rotation=0.0):
    b = scenes.Builder()
    o = b.add_object(scenes.material())
    b.quad(o, (0, 0, 900), (1e-4, 0, 900), (1e-4, 1e-4, 900), (0, 1e-4, 900), normal=(0, 0, -1))
    return b.finish(camera={"fov_deg": 30.0, "aperture": 1.8},
                    env={"rgb": np.asarray(rgb, np.float32), "intensity": 1.0, "rotation_deg": rotation})


def test_env_sampling_matches_its_pdf():
    rng = np.random.default_rng(4)
    H, W = 16, 32
    rgb = rng.uniform(0.05, 1.0, (H, W, 3)).astype(np.float32)
    rgb[3, 20] = (40.0, 35.0, 30.0)                      # a sun
    rgb[H - 1] = 0.0                                     # black bottom row
    sc = env_scene(rgb)
    n = 600_000
    u = rng.random((n, 2), dtype=np.float32)
    dirs, pdfs, rad, qp, qr = oracle.env_probe(sc, u, np.zeros((0, 3), np.float32))
    d = dirs.astype(np.float64)
    # equirect cell of each direction (u = 0.5 + atan2(z, x) / 2pi,
    # v = 0.5 - asin(y) / pi; no rotation)
    uu = 0.5 + np.arctan2(d[:, 2], d[:, 0]) / (2 * np.pi)
    vv = 0.5 - np.arcsin(np.clip(d[:, 1], -1, 1)) / np.pi
    cx = np.clip((uu * W).astype(int), 0, W - 1)
    cy = np.clip((vv * H).astype(int), 0, H - 1)
    counts = np.bincount(cy * W + cx, minlength=H * W).reshape(H, W)
    # probability of each texel from the pdf: pdf is constant per texel in
    # direction space up to the sin factor; integrate it over the texel's
    # solid angle on a fine grid of the same formula's inputs
    lum = np.maximum(0, rgb @ np.array([0.2126, 0.7152, 0.0722]))
    vc = (np.arange(H) + 0.5) / H
    wrow = lum * np.sin(np.pi * vc)[:, None]
    mass = wrow / wrow.sum()
    big = mass * n > 400
    z = (counts[big] - mass[big] * n) / np.sqrt(mass[big] * n)
    assert np.abs(z).max() < 5, np.abs(z).max()
    assert counts[H - 1].sum() == 0                      # black row never sampled
    # the pdf of a sampled direction: its texel's probability over the
    # texel's solid angle 2 pi^2 sin(pi v) / (W H), with the sine taken at the
    # row centre (the sampler's density within a row differs from it by
    # sin(pi v) / sin(pi v_row), which the histogram above covers)
    want_pdf = mass[cy, cx] * W * H / (2 * np.pi ** 2 * np.sin(np.pi * vc[cy]))
    # (a direction on a texel edge may round into the neighbouring texel here)
    assert np.mean(np.abs(pdfs / want_pdf - 1) > 2e-3) < 1e-4
    # envmapPdf of the sampled direction is the sampler's pdf; the radiance
    # seen along it is the texel's
    _, _, _, qp2, qr2 = oracle.env_probe(sc, np.zeros((0, 2), np.float32), dirs[:5000])
    np.testing.assert_allclose(qp2, pdfs[:5000], rtol=2e-3)
    np.testing.assert_allclose(rad[:5000], rgb[cy[:5000], cx[:5000]], rtol=1e-6)


def test_env_rows_run_from_zenith_to_nadir_and_rotation_turns_the_sky():
    H, W = 8, 16
    rgb = np.zeros((H, W, 3), np.float32)
    rgb[0] = 1.0                           # top row: zenith
    rgb[H - 1] = 2.0                       # bottom row: nadir
    rgb[4, 8] = 5.0                        # just below the horizon, u = 8.5 / 16 (+x side)
    # (exactly -y has v = 1, which the texture fold wraps to row 0)
    q = np.array([[0, 1, 0], [0.01, -1, 0], [1, -0.1, 0.05]], np.float32)
    _, _, _, _, r0 = oracle.env_probe(env_scene(rgb), np.zeros((0, 2), np.float32), q)
    assert r0[0, 0] == 1.0 and r0[1, 0] == 2.0 and r0[2, 0] == 5.0
    # a 90 degree rotation moves that texel a quarter turn in azimuth: it is
    # seen along +z or -z, no longer along +x
    _, _, _, _, r90 = oracle.env_probe(env_scene(rgb, 90.0), np.zeros((0, 2), np.float32),
                                       np.array([[1, -0.1, 0.05], [0.05, -0.1, 1], [0.05, -0.1, -1]], np.float32))
    assert r90[0, 0] == 0.0 and max(r90[1, 0], r90[2, 0]) == 5.0


@pytest.mark.parametrize("mode", [0, 1])
def test_two_hemisphere_sky_on_a_diffuse_plane(mode):
    """A level diffuse floor of albedo rho under a sky of L_up above the
    horizon and L_dn below sees only the upper hemisphere: rho * L_up.  The
    camera looks along -z; the bottom rows of a tall view see the floor
    (y = -1), the top rows the sky above the horizon."""
    rho, Lup, Ldn = 0.5, 2.0, 0.25
    H, W = 32, 64
    rgb = np.full((H, W, 3), Ldn, np.float32)
    rgb[: H // 2] = Lup
    b = scenes.Builder()
    fl = b.add_object(scenes.material(color=(rho, rho, rho)))
    b.quad(fl, (-300, -1, 300), (300, -1, 300), (300, -1, -300), (-300, -1, -300), normal=(0, 1, 0))
    sc = b.finish(camera={"fov_deg": 20.0, "aperture": 1.8},
                  env={"rgb": rgb, "intensity": 1.0, "rotation_deg": 0.0})
    n = 2048
    s_, _ = oracle.render(sc, 8, 64, n, 6, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    img = s_ / n
    np.testing.assert_allclose(img[40:].reshape(-1, 3).mean(0), rho * Lup, rtol=0.02)
    np.testing.assert_allclose(img[:20], Lup, rtol=1e-6)
    # the same floor under the sky turned upside down reads rho * L_dn
    sc["env"] = {"rgb": rgb[::-1].copy(), "intensity": 1.0, "rotation_deg": 0.0}
    s2, _ = oracle.render(sc, 8, 64, n, 6, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    np.testing.assert_allclose((s2 / n)[40:].reshape(-1, 3).mean(0), rho * Ldn, rtol=0.02)


# --- GGX delta mirror and anisotropy ------------------------------------------
def tilted_plane_scene(mat, incidence_deg, env=1.0):
    """A large plane whose normal makes incidence_deg with the camera's -z
    view axis (rotation about x), seen through a narrow field of view."""
    b = scenes.Builder()
    o = b.add_object(mat)
    a = np.radians(incidence_deg)
    nrm = np.array([0.0, np.sin(a), np.cos(a)])
    t1 = np.array([1.0, 0, 0])
    t2 = np.cross(nrm, t1)
    c = np.array([0.0, 0.0, -5.0])
    b.quad(o, c - 3 * t1 - 3 * t2, c + 3 * t1 - 3 * t2, c + 3 * t1 + 3 * t2, c - 3 * t1 + 3 * t2, normal=nrm)
    return b.finish(camera={"fov_deg": 1.0, "aperture": 1.8}, env=scenes.constant_env(env, 128, 64))


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("metallic,incidence", [(1.0, 60.0), (0.0, 75.0)])
def test_ggx_delta_mirror_is_schlick(mode, metallic, incidence):
    alb = np.array([0.5, 0.7, 0.9])
    mat = scenes.material(color=tuple(alb), glossy=1.0, roughness=0.0, metallic=metallic)
    sc = tilted_plane_scene(mat, incidence, env=1.5)
    s, q = oracle.render(sc, 4, 4, 64, 4, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    got = s / 64
    F0 = alb if metallic >= 1 else alb * metallic + 0.04 * (1 - metallic)
    # each pixel's centre ray meets the plane at its own incidence
    a = np.radians(incidence)
    nrm = np.array([0.0, np.sin(a), np.cos(a)])
    vs = 2 * np.tan(np.radians(0.5)) / 4
    ys, xs = np.mgrid[0:4, 0:4]
    d = np.stack([(xs + 0.5 - 2) * vs, -(ys + 0.5 - 2) * vs, -np.ones((4, 4))], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    ci = -(d @ nrm)
    want = (F0[None, None] + (1 - F0[None, None]) * (1 - ci[..., None]) ** 5) * 1.5
    np.testing.assert_allclose(got, want, rtol=2e-5)
    # a delta lobe: every sample returns the same value
    np.testing.assert_allclose(q / 64, (s / 64) ** 2, rtol=1e-3)


def ggx_aniso_albedo_normal(ax, ay, n=700):
    """float64 quadrature of D(h) G1(v) G1(l) / (4 cos_v) over l for v = n
    (normal incidence), anisotropic GGX with separable Smith."""
    th = (np.arange(n) + 0.5) / n * (np.pi / 2)
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * (2 * np.pi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    L = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1)
    v = np.array([0.0, 0.0, 1.0])
    Hh = L + v
    Hh /= np.linalg.norm(Hh, axis=-1, keepdims=True)
    den = Hh[..., 0] ** 2 / ax ** 2 + Hh[..., 1] ** 2 / ay ** 2 + Hh[..., 2] ** 2
    D = 1 / (np.pi * ax * ay * den ** 2)
    lam = 0.5 * (np.sqrt(1 + (ax ** 2 * L[..., 0] ** 2 + ay ** 2 * L[..., 1] ** 2) / L[..., 2] ** 2) - 1)
    g1l = 1 / (1 + lam)
    f = D * 1.0 * g1l / 4
    dw = np.sin(T) * (np.pi / 2 / n) * (2 * np.pi / (2 * n))
    return (f * dw).sum()


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("aniso", [-0.95, 0.6])
def test_ggx_anisotropic_furnace(mode, aniso):
    rough = 0.7
    mat = scenes.material(color=(1, 1, 1), glossy=1.0, metallic=1.0, roughness=rough, anisotropy=aniso)
    sc = tilted_plane_scene(mat, 0.0)
    n = 16384
    s, _ = oracle.render(sc, 2, 2, n, 8, mode=mode, threads=THREADS, antialias=False, firefly_clamp=0.0)
    got = s.mean() / n
    r2 = rough * rough
    ax, ay = max(1e-3, r2 * (1 - aniso)), max(1e-3, r2 * (1 + aniso))
    want = ggx_aniso_albedo_normal(ax, ay)
    iso = ggx_aniso_albedo_normal(r2, r2)
    assert abs(want / iso - 1) > 0.01            # anisotropy changes the albedo measurably
    assert abs(got / want - 1) < 0.01, (got, want, iso)

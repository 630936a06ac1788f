"""Pins of the output-stage restatement (oracle/rg_oracle_post.c): the HDR
bloom (RayRenderer::applyPostProcess, rayrenderer.cpp:598-697) and the
Reinhard + gamma tonemap (applyTonemapGamma, :2614-2647), against closed
forms and independent float64 computations.  CPU only."""
import math

import numpy as np
import pytest

import oracle


def img(h, w, rgb, a=1.0):
    out = np.zeros((h, w, 4), np.float32)
    out[..., :3] = rgb
    out[..., 3] = a
    return out


# --- tonemap ----------------------------------------------------------------
def test_tonemap_grey_is_reinhard_then_gamma():
    for c in (0.01, 0.25, 1.0, 3.0, 40.0):
        out = oracle.tonemap(img(1, 1, [c, c, c], a=0.5))[0, 0]
        want = (c / (1.0 + c)) ** (1.0 / 2.2)   # luma of grey c is c
        assert out[:3] == pytest.approx([want] * 3, rel=2e-6)
        assert out[3] == 0.5                      # alpha copied


def test_tonemap_black_and_negative():
    out = oracle.tonemap(img(1, 2, [0, 0, 0]))
    assert np.all(out[..., :3] == 0)
    neg = img(1, 1, [-1.0, -1.0, -1.0])
    assert np.all(oracle.tonemap(neg)[..., :3] == 0)   # L <= 1e-6 -> black


def test_tonemap_saturated_colour_normalised_by_peak():
    # pure red 10: L = 2.126, Reinhard scale = 1 / 3.126, red 3.2 > 1 -> peak
    # normalisation gives exactly (1, 0, 0)
    out = oracle.tonemap(img(1, 1, [10.0, 0.0, 0.0]))[0, 0]
    assert out[0] == 1.0 and out[1] == 0.0 and out[2] == 0.0
    # a colour whose mapped peak stays below 1 keeps its hue ratios
    c = np.array([0.6, 0.3, 0.1])
    L = 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]
    m = c * (L / (1 + L)) / L
    out = oracle.tonemap(img(1, 1, c))[0, 0]
    assert out[:3] == pytest.approx(m ** (1 / 2.2), rel=3e-6)


# --- bloom ------------------------------------------------------------------
def bright_field(h, w, seed):
    rng = np.random.default_rng(seed)
    x = img(h, w, 0.0)
    x[..., :3] = rng.uniform(0, 0.5, (h, w, 3))
    ys, xs = rng.integers(0, h, 12), rng.integers(0, w, 12)
    x[ys, xs, :3] = rng.uniform(5, 200, (12, 3))
    return x


def test_bloom_strength_zero_and_dim_images_unchanged():
    x = bright_field(40, 56, 1)
    assert np.array_equal(oracle.bloom(x, strength=0.0), x)
    dim = img(33, 47, 0.4)
    assert np.array_equal(oracle.bloom(dim), dim)   # L = 0.4 < threshold 1


def test_bloom_identity_resample_adds_excess_energy():
    # size_aspect 1 and radius 0: the glow buffer is the image itself, no blur,
    # the upsample is the identity, so bloom adds strength * c * (L - th) / L
    x = bright_field(20, 30, 2)
    th, k = 1.0, 0.75
    out = oracle.bloom(x, threshold=th, strength=k, radius=0.0, size_aspect=1.0)
    c = x[..., :3].astype(np.float64)
    L = 0.2126 * c[..., 0] + 0.7152 * c[..., 1] + 0.0722 * c[..., 2]
    s = np.where(L > th, (L - th) / np.maximum(L, 1e-30), 0.0)
    want = c + k * c * s[..., None]
    assert np.allclose(out[..., :3], want, rtol=1e-6, atol=1e-7)
    assert np.array_equal(out[..., 3], x[..., 3])
    # luma added to a bright pixel is exactly the excess above the threshold
    i = np.unravel_index(np.argmax(L), L.shape)
    Lo = 0.2126 * out[i][0] + 0.7152 * out[i][1] + 0.0722 * out[i][2]
    assert Lo - L[i] == pytest.approx(k * (L[i] - th), rel=1e-5)


def test_bloom_curve_sharpens_knee():
    x = img(1, 1, [2.0, 2.0, 2.0])   # L = 2, excess ratio 0.5
    out = oracle.bloom(x, threshold=1.0, strength=1.0, radius=0.0, size_aspect=1.0, curve=2.0)
    assert out[0, 0, 0] == pytest.approx(2.0 + 2.0 * 0.25, rel=1e-6)


def test_bloom_downsample_is_box_mean_at_integer_ratio():
    # a threshold of 0 keeps all energy; with aspect 0.5 on even sizes each
    # glow pixel is the mean of a 2x2 block; radius 0 skips the blur and the
    # bilinear upsample of factor 2 is then checked at the block centres'
    # neighbourhood through the total energy of the glow
    rng = np.random.default_rng(3)
    h, w = 16, 24
    x = img(h, w, 0.0)
    x[..., :3] = rng.uniform(0.5, 4.0, (h, w, 3))
    out = oracle.bloom(x, threshold=0.0, strength=1.0, radius=0.0, size_aspect=0.5)
    glow = out[..., :3].astype(np.float64) - x[..., :3]
    box = x[..., :3].astype(np.float64).reshape(h // 2, 2, w // 2, 2, 3).mean((1, 3))
    # bilinear x2 upsample with aligned centres, clamped edges, of the box means
    def up(a, n):
        src = (np.arange(n) + 0.5) / 2 - 0.5
        src = np.clip(src, 0, None)
        i0 = np.minimum(np.floor(src).astype(int), a.shape[0] - 1)
        i1 = np.minimum(i0 + 1, a.shape[0] - 1)
        t = src - i0
        return a[i0] * (1 - t)[:, None, None] + a[i1] * t[:, None, None]
    want = up(np.swapaxes(up(box, h), 0, 1), w)
    want = np.swapaxes(want, 0, 1)
    assert np.allclose(glow, want, rtol=1e-5, atol=1e-5)


def gauss(sigma):
    r = max(1, math.ceil(sigma * 3.0))
    k = np.exp(-np.arange(-r, r + 1, dtype=np.float64) ** 2 / (2 * sigma * sigma))
    return k / k.sum()


def test_bloom_blur_of_a_point_is_twice_convolved_gaussian():
    # aspect 1 (no resampling), a single bright pixel far from the borders:
    # the glow is the separable kernel (g * g) (x) (g * g) of sigma / sqrt2 per
    # pass, scaled by the pixel's excess
    h = w = 121
    x = img(h, w, 0.0)
    x[60, 60, :3] = 10.0
    radius = 0.05                        # sigma_down = 0.05 * 121 = 6.05
    out = oracle.bloom(x, threshold=1.0, strength=1.0, radius=radius, size_aspect=1.0)
    sd = radius * w
    g = gauss(sd * 0.70710678)
    g2 = np.convolve(g, g)
    r2 = (len(g2) - 1) // 2
    kern = np.outer(g2, g2)
    glow = out[..., 0].astype(np.float64) - x[..., 0]
    glow[60, 60] = out[60, 60, 0] - 10.0
    excess = 10.0 * (10.0 - 1.0) / 10.0
    want = np.zeros((h, w))
    want[60 - r2:60 + r2 + 1, 60 - r2:60 + r2 + 1] = excess * kern
    assert np.abs(glow - want).max() < 1e-5 * excess
    assert glow.sum() == pytest.approx(excess, rel=1e-5)     # energy preserved


def test_upsample_convention_reproduces_constant_and_linear_glow():
    # constant glow -> constant halo at any ratio (convention pin)
    x = img(30, 40, 3.0)
    out = oracle.bloom(x, threshold=1.0, strength=1.0, radius=0.0, size_aspect=0.25)
    glow = out[..., :3] - x[..., :3]
    assert np.allclose(glow, 3.0 * 2.0 / 3.0, rtol=1e-6)

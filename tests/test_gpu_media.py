"""Heterogeneous and emissive media and heat haze on the GPU through the C
ABI, against the oracle (medium.cpp:125-472; rayrenderer.cpp:1596-1864,
2066-2203):

* the media room (an fBm cloud, a noisy cone flame that is also an NEE
  volume light, a path-tube glow and a heat-haze volume in front of stripes,
  under an area light and a dim environment): per pixel against the
  single-lobe restatement, the batched engine against the single-job path
  bit for bit, and converged against the recursive restatement;
* volume-light NEE from a cone no object contains against the line integral
  albedo / pi x int sigma_e cos / r^2 along its axis.
"""
import sys

import numpy as np
import pytest

import oracle
import rgapi
import scenes

from test_gpu_pt_extra import CHUNK, chunked, converged, rel_diff
from test_point_lights import centre_hits
from test_media_pins import invisible_flame_plane

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return rgapi.Context(0)


def test_media_room_matches_single_lobe_restatement(ctx):
    sc = scenes.media_room()
    rs = rgapi.RenderScene(ctx, sc)
    W, H, spp = 64, 48, 32
    say("[f4] media room 64x48x32 vs oracle twin")
    out = rs.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=6))
    want, _ = oracle.render(sc, W, H, spp, 6, mode=1, threads=8)
    r = rel_diff(out["sum"], want)
    say(f"[f4]   frac > 1e-3: {np.mean(r > 1e-3):.5f}, median {np.median(r):.2e}")
    assert np.mean(r > 1e-3) < 0.01
    assert np.median(r) < 1e-5
    assert abs(out["sum"].mean() / want.mean() - 1) < 2e-3
    st = dict(width=W, height=H, spp=4, max_depth=6)
    outs, _ = rgapi.render_batch(ctx, [rs], [rgapi.make_job(0, **st)])
    assert np.array_equal(outs[0]["hdr"], rs.render(rgapi.default_settings(**st))["hdr"])


def test_media_room_converges_to_recursive_restatement(ctx):
    sc = scenes.media_room()
    rs = rgapi.RenderScene(ctx, sc)
    W, H, depth = 16, 12, 6
    n_gpu, n_cpu = 131072, 32768
    say(f"[f4] convergence {W}x{H}: gpu {n_gpu} spp vs recursive {n_cpu} spp")

    def gpu_chunk(seed):
        o = rs.render(rgapi.default_settings(width=W, height=H, spp=CHUNK, max_depth=depth, seed=seed,
                                             firefly_clamp=0.0))
        return o["sum"], o["sumsq"]

    g, gq = chunked(gpu_chunk, n_gpu, 0x500)
    c, cq = chunked(lambda seed: oracle.render(sc, W, H, CHUNK, depth, mode=0, threads=8, seed=seed,
                                               firefly_clamp=0.0), n_cpu, 0x600)
    frac, rmse, rmse_sys = converged(g, gq, n_gpu, c, cq, n_cpu)
    say(f"[f4]   |z| > 3 on {frac:.4f} of pixel channels, rel RMSE {rmse:.4f} (noise removed {rmse_sys:.4f})")
    assert frac <= 0.02
    assert rmse_sys <= 0.01
    assert rmse <= 0.03


def test_volume_nee_line_integral_on_the_gpu(ctx):
    sc, med = invisible_flame_plane()
    rs = rgapi.RenderScene(ctx, sc)
    W, H, n = 32, 24, 4096
    out = rs.render(rgapi.default_settings(width=W, height=H, spp=n, max_depth=2, antialias=0, firefly_clamp=0.0))
    img = out["sum"].astype(np.float64) / n
    var = np.maximum(out["sumsq"] / n - img ** 2, 0) / n
    p, ok = centre_hits(W, H, 60.0)
    ok &= (p[..., 2] > -7.0) & (p[..., 2] < -1.0) & (np.abs(p[..., 0]) < 4.0)
    tt = np.linspace(0, 2.0, 4001)
    axis_pts = np.stack([-1.0 + tt, np.zeros_like(tt), np.full_like(tt, -4.0)], -1)
    se = oracle.medium_eval(med, "emission", axis_pts.astype(np.float32)).astype(np.float64)
    lr = axis_pts[None, :, :] - p[ok][:, None, :]
    r2 = (lr ** 2).sum(-1)
    cos = np.maximum(lr[..., 1] / np.sqrt(r2), 0)
    want = 0.5 / np.pi * np.trapezoid(se[None, :, :] * (cos / r2)[..., None], tt, axis=1)
    z = (img[ok] - want) / np.sqrt(var[ok] + (3e-4 * want) ** 2 + 1e-14)
    assert np.mean(np.abs(z) > 4) < 0.01
    assert abs(img[ok].sum() / want.sum() - 1) < 0.01

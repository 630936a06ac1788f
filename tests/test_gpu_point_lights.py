"""Point lights (traceLight, rayrenderer.cpp:2205-2293) on the GPU through
the C ABI, against the oracle:

* the cube room with its ceiling quad, an omni and a spot light and a
  constant environment (a diffuse event casts up to three NEE rays: area,
  point, environment): per pixel against the single-lobe restatement, and
  the batched engine renders the same image as the single-job path;
* an omni light over a diffuse plane against the closed form
  albedo x min(1, colour x emission x cos / r^2) at pixel centres;
* converged against the recursive restatement of the reference.
"""
import sys

import numpy as np
import pytest

import oracle
import rgapi
import scenes

from test_gpu_pt_extra import CHUNK, chunked, converged, rel_diff
from test_point_lights import centre_hits, omni_expected

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return rgapi.Context(0)


def test_point_lit_room_matches_single_lobe_restatement(ctx):
    sc = scenes.point_lit_room(env_value=0.2)
    rs = rgapi.RenderScene(ctx, sc)
    W, H, spp = 64, 48, 32
    say("[pl] point-lit cube room 64x48x32 vs oracle twin")
    out = rs.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=6))
    want, _ = oracle.render(sc, W, H, spp, 6, mode=1, threads=8)
    r = rel_diff(out["sum"], want)
    say(f"[pl]   frac > 1e-3: {np.mean(r > 1e-3):.5f}")
    assert np.mean(r > 1e-3) < 0.01
    assert np.median(r) < 1e-5
    assert abs(out["sum"].mean() / want.mean() - 1) < 2e-3
    # the point lights add light: the same room without them is darker
    dark = {k: v for k, v in sc.items() if k != "point_lights"}
    off = rgapi.RenderScene(ctx, dark).render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=6))
    assert out["sum"].mean() > 1.05 * off["sum"].mean()
    st = dict(width=W, height=H, spp=4, max_depth=6)
    outs, _ = rgapi.render_batch(ctx, [rs], [rgapi.make_job(0, **st)])
    assert np.array_equal(outs[0]["hdr"], rs.render(rgapi.default_settings(**st))["hdr"])


def test_omni_light_closed_form_on_the_gpu(ctx):
    W, H = 48, 32
    albedo, color, emission = (0.8, 0.6, 0.4), (1.0, 0.5, 0.25), 2.0
    sc = scenes.point_light_plane(height=1.0, emission=emission, color=color, albedo=albedo)
    rs = rgapi.RenderScene(ctx, sc)
    out = rs.render(rgapi.default_settings(width=W, height=H, spp=4, max_depth=3, antialias=0, firefly_clamp=0.0))
    img = out["hdr"][..., :3]
    p, ok = centre_hits(W, H, 60.0)
    ok &= (p[..., 2] > -7.0) & (p[..., 2] < -1.0) & (np.abs(p[..., 0]) < 4.0)
    want = omni_expected(p, (0.0, 0.0, -3.0), color, emission, albedo)
    np.testing.assert_allclose(img[ok], want[ok], rtol=3e-4, atol=1e-7)


def test_point_lit_room_converges_to_recursive_restatement(ctx):
    sc = scenes.point_lit_room(env_value=0.2)
    rs = rgapi.RenderScene(ctx, sc)
    W, H, depth = 16, 12, 6
    n_gpu, n_cpu = 131072, 32768
    say(f"[pl] convergence {W}x{H}: gpu {n_gpu} spp vs recursive {n_cpu} spp")

    def gpu_chunk(seed):
        o = rs.render(rgapi.default_settings(width=W, height=H, spp=CHUNK, max_depth=depth, seed=seed,
                                             firefly_clamp=0.0))
        return o["sum"], o["sumsq"]

    g, gq = chunked(gpu_chunk, n_gpu, 0x300)
    c, cq = chunked(lambda seed: oracle.render(sc, W, H, CHUNK, depth, mode=0, threads=8, seed=seed,
                                               firefly_clamp=0.0), n_cpu, 0x400)
    frac, rmse, rmse_sys = converged(g, gq, n_gpu, c, cq, n_cpu)
    say(f"[pl]   |z| > 3 on {frac:.4f} of pixel channels, rel RMSE {rmse:.4f} (noise removed {rmse_sys:.4f})")
    assert frac <= 0.02
    assert rmse_sys <= 0.01
    assert rmse <= 0.03

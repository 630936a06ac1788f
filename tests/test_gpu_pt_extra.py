"""GPU path tracer on the branches and settings the BSDF frame leaves out,
through the C ABI, against the oracle:

* delta mirrors, anisotropic GGX (both signs, rotated brush), a
  Henyey-Greenstein medium (g = 0.6) and a non-constant sky with a sun:
  per pixel against the single-lobe restatement, and converged against the
  recursive restatement of the reference (per-pixel 3 sigma, rel. RMSE);
* dispersive glass converged against the recursive restatement;
* the scattering-medium furnace closed form on the GPU;
* a global medium without environment: rays that cross the fog without a
  hit and gather nothing read backColor (default and user-set), on the
  single-job and batched paths;
* enableColorSampling off.
"""
import sys

import numpy as np
import pytest

import oracle
import rgapi
import scenes

from test_pt_oracle_pins import medium_furnace_expected, medium_furnace_scene

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return rgapi.Context(0)


def rel_diff(a, b, floor=1e-3):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


# Both renderers sum a pixel's samples in binary32, in sample order, as the
# reference does; over 4096 samples of a constant value (a pixel that sees
# only sky) that sum is off by up to ~2.4e-4 relative, far above the Monte
# Carlo error of such a pixel.  The estimates are therefore built from
# 4096-sample chunks added in float64, and the z-test allows that rounding.
CHUNK = 4096
SUM_ROUNDING = 3e-4


def chunked(render_chunk, n, seed0):
    s = q = 0.0
    for k in range(n // CHUNK):
        a, b = render_chunk(seed0 + k)
        s = s + a.astype(np.float64)
        q = q + b.astype(np.float64)
    return s, q


def converged(g_sum, g_sq, n_g, c_sum, c_sq, n_c):
    mg, mc = g_sum / n_g, c_sum / n_c
    vg = np.maximum(g_sq / n_g - mg * mg, 0) / n_g
    vc = np.maximum(c_sq / n_c - mc * mc, 0) / n_c
    z = (mg - mc) / np.sqrt(vg + vc + (SUM_ROUNDING * mg) ** 2 + (SUM_ROUNDING * mc) ** 2 + 1e-12)
    rmse = np.sqrt(np.mean((mg - mc) ** 2)) / mc.mean()
    # E[(mg - mc)^2] = bias^2 + vg + vc: the systematic part of the RMSE with
    # the two estimates' own Monte Carlo variance taken out
    rmse_sys = np.sqrt(max(np.mean((mg - mc) ** 2 - vg - vc), 0.0)) / mc.mean()
    return np.mean(np.abs(z) > 3), rmse, rmse_sys


def test_extra_materials_match_single_lobe_restatement(ctx):
    sc = scenes.materials_extra()
    rs = rgapi.RenderScene(ctx, sc)
    W, H, spp = 72, 48, 32
    say("[pt+] mirrors / anisotropic GGX / HG medium / sky 72x48x32 vs oracle twin")
    out = rs.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=8))
    want, _ = oracle.render(sc, W, H, spp, 8, mode=1, threads=8)
    r = rel_diff(out["sum"], want)
    say(f"[pt+]   frac > 1e-3: {np.mean(r > 1e-3):.5f}")
    assert np.mean(r > 1e-3) < 0.01
    assert np.median(r) < 1e-5
    assert abs(out["sum"].mean() / want.mean() - 1) < 2e-3
    # the batched engine renders the same image
    st = dict(width=W, height=H, spp=4, max_depth=8)
    single = rs.render(rgapi.default_settings(**st))
    outs, _ = rgapi.render_batch(ctx, [rs], [rgapi.make_job(0, **st)])
    assert np.array_equal(outs[0]["hdr"], single["hdr"])


def test_dispersive_window_matches_single_lobe_restatement(ctx):
    """The eye hit's Mix branch hands its dispersion band to the
    transparency path (see test_pt_oracle): per pixel against the twin, and
    the batched engine equals the single-job path."""
    sc = scenes.dispersive_window()
    rs = rgapi.RenderScene(ctx, sc)
    W, H, spp = 48, 48, 32
    out = rs.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=8))
    want, _ = oracle.render(sc, W, H, spp, 8, mode=1, threads=8)
    r = rel_diff(out["sum"], want)
    assert np.mean(r > 1e-3) < 0.01
    assert abs(out["sum"].mean() / want.mean() - 1) < 2e-3
    st = dict(width=W, height=H, spp=4, max_depth=8)
    outs, _ = rgapi.render_batch(ctx, [rs], [rgapi.make_job(0, **st)])
    assert np.array_equal(outs[0]["hdr"], rs.render(rgapi.default_settings(**st))["hdr"])


@pytest.mark.parametrize("which", ["extra", "dispersion", "window"])
def test_converges_to_recursive_restatement(ctx, which):
    """GPU single-lobe estimator vs the recursive restatement (every Mix lobe
    evaluated, the reference's call structure) at high sample counts."""
    if which == "extra":
        sc = scenes.materials_extra()
        depth = 8
    elif which == "dispersion":
        sc = scenes.bsdf_frame(grid=3, stacks=16, slices=24, dispersion=0.05)
        depth = 8
    else:
        sc = scenes.dispersive_window()
        depth = 8
    rs = rgapi.RenderScene(ctx, sc)
    W, H = 16, 12
    n_gpu, n_cpu = 131072, 32768
    say(f"[pt+] convergence {which} {W}x{H}: gpu {n_gpu} spp vs recursive {n_cpu} spp")
    def gpu_chunk(seed):
        o = rs.render(rgapi.default_settings(width=W, height=H, spp=CHUNK, max_depth=depth, seed=seed,
                                             firefly_clamp=0.0))
        return o["sum"], o["sumsq"]

    g, gq = chunked(gpu_chunk, n_gpu, 0x100)
    c, cq = chunked(lambda seed: oracle.render(sc, W, H, CHUNK, depth, mode=0, threads=8, seed=seed,
                                               firefly_clamp=0.0), n_cpu, 0x200)
    frac, rmse, rmse_sys = converged(g, gq, n_gpu, c, cq, n_cpu)
    say(f"[pt+]   |z| > 3 on {frac:.4f} of pixel channels, rel RMSE {rmse:.4f} (noise removed {rmse_sys:.4f})")
    assert frac <= 0.02
    assert rmse_sys <= 0.01
    assert rmse <= 0.03


def test_medium_furnace_closed_form_on_the_gpu(ctx):
    g_, sa, ss, depth = 0.6, (0.002, 0.0, 0.004), (0.006, 0.01, 0.008), 6
    sc = medium_furnace_scene(g_, sa, ss)
    rs = rgapi.RenderScene(ctx, sc)
    n = 16384
    st = rgapi.default_settings(width=8, height=8, spp=n, max_depth=depth, antialias=0, firefly_clamp=0.0)
    for i in range(3):
        st.back_color[i] = 0.0
    out = rs.render(st)
    got = out["sum"].reshape(-1, 3).mean(0) / n
    np.testing.assert_allclose(got, medium_furnace_expected(g_, sa, ss, depth), rtol=0.01)


@pytest.mark.parametrize("back", [None, (0.1, 0.3, 0.5)])
def test_fog_without_environment_reads_back_color(ctx, back):
    """Absorbing global fog, no environment: an eye ray that meets nothing
    gathers nothing and reads backColor exactly; the rest matches the twin.
    Then a scattering fog against the twin, single-job and batched."""
    kw = {} if back is None else {"back_color": back}
    bc = np.asarray(oracle.BACK_COLOR[:3] if back is None else back, np.float32)
    W, H, spp = 64, 48, 16
    sc = scenes.fog_room(scatter=0.0, absorb=0.2)
    rs = rgapi.RenderScene(ctx, sc)
    st = rgapi.default_settings(width=W, height=H, spp=spp, max_depth=6)
    if back is not None:
        for i in range(3):
            st.back_color[i] = back[i]
    say(f"[pt+] absorbing fog, back colour {bc.tolist()}")
    out = rs.render(st)
    want, _ = oracle.render(sc, W, H, spp, 6, mode=1, threads=8, **kw)
    miss = np.all(want == bc * np.float32(spp), axis=-1) | np.all(np.isclose(want, bc * spp, rtol=1e-6), axis=-1)
    assert miss.mean() > 0.2
    np.testing.assert_allclose(out["hdr"][miss][:, :3], np.broadcast_to(bc, (miss.sum(), 3)), rtol=1e-6)
    r = rel_diff(out["sum"], want)
    assert np.mean(r > 1e-4) < 0.002
    sc2 = scenes.fog_room()
    rs2 = rgapi.RenderScene(ctx, sc2)
    out2 = rs2.render(st)
    want2, _ = oracle.render(sc2, W, H, spp, 6, mode=1, threads=8, **kw)
    r2 = rel_diff(out2["sum"], want2)
    say(f"[pt+] scattering fog: frac > 1e-3 {np.mean(r2 > 1e-3):.5f}")
    assert np.mean(r2 > 1e-3) < 0.01
    assert abs(out2["sum"].mean() / want2.mean() - 1) < 2e-3
    job = rgapi.make_job(0, width=W, height=H, spp=spp, max_depth=6)
    if back is not None:
        for i in range(3):
            job.settings.back_color[i] = back[i]
    outs, _ = rgapi.render_batch(ctx, [rs2], [job])
    assert np.array_equal(outs[0]["hdr"], out2["hdr"])


def test_color_sampling_off_matches_restatement(ctx):
    """enableColorSampling = false: diffuse albedo 1, glossy F0 from the
    untextured colour, depth-cap light term 0 (bsdf.cpp:68-73, 185-188;
    rayrenderer.cpp:2839-2856)."""
    sc = scenes.textured_room()
    rs = rgapi.RenderScene(ctx, sc)
    W, H, spp = 64, 48, 16
    on = rs.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=5))
    out = rs.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=5, color_sampling=0))
    want, _ = oracle.render(sc, W, H, spp, 5, mode=1, threads=8, color_sampling=False)
    r = rel_diff(out["sum"], want)
    assert np.mean(r > 1e-4) < 0.002
    assert out["sum"].mean() > 1.2 * on["sum"].mean()        # white albedo everywhere is brighter

"""Wavefront path tracer on the GPU through the C ABI (rg_render), against the
oracle: per pixel against the single-lobe restatement that draws the same
random numbers, statistically against the recursive restatement of the
reference, and for the denoiser and the adaptive driver."""
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


def rel_diff(a, b, floor=1e-3):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def test_cube_room_matches_single_lobe_restatement(ctx):
    sc = scenes.cube_room()
    rs = rgapi.RenderScene(ctx, sc)
    W = H = 64
    spp = 64
    say("[pt] cube room 64x64x64 vs oracle twin")
    out = rs.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=5))
    want, wq = oracle.render(sc, W, H, spp, 5, mode=1, threads=8)
    r = rel_diff(out["sum"], want)
    assert np.mean(r > 1e-4) < 0.001, np.sort(r.ravel())[-10:]
    assert r.max() < 0.05
    np.testing.assert_allclose(out["hdr"][..., :3], want / spp, rtol=1e-4, atol=1e-6)
    assert (out["hdr"][..., 3] == 1).all()
    assert (out["nspp"] == spp).all()


def test_materials_media_env_match_single_lobe_restatement(ctx):
    sc = scenes.bsdf_frame(grid=3, stacks=16, slices=24)
    rs = rgapi.RenderScene(ctx, sc)
    W, H = 72, 48
    spp = 32
    say("[pt] materials/media/env 72x48x32 vs oracle twin")
    out = rs.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=8))
    want, _ = oracle.render(sc, W, H, spp, 8, mode=1, threads=8)
    r = rel_diff(out["sum"], want)
    # transcendental functions differ in the last ulp between libm and the
    # device library, which can flip a branch on a rare path
    assert np.mean(r > 1e-3) < 0.01, np.mean(r > 1e-3)
    assert abs(out["sum"].mean() / want.mean() - 1) < 2e-3


def test_converges_to_recursive_restatement(ctx):
    """GPU single-lobe estimator vs the recursive restatement of the
    reference at high spp: per-pixel 3-sigma and relative RMSE <= 1%."""
    sc = scenes.cube_room()
    rs = rgapi.RenderScene(ctx, sc)
    W = H = 16
    n_gpu, n_cpu = 65536, 16384
    say("[pt] convergence 16x16 gpu 64k spp vs recursive 16k spp")
    g = rs.render(rgapi.default_settings(width=W, height=H, spp=n_gpu, max_depth=5, seed=101))
    c, cq = oracle.render(sc, W, H, n_cpu, 5, mode=0, threads=8, seed=202)
    mg, mc = g["sum"] / n_gpu, c / n_cpu
    vg = np.maximum(g["sumsq"] / n_gpu - mg * mg, 0) / n_gpu
    vc = np.maximum(cq / n_cpu - mc * mc, 0) / n_cpu
    z = (mg - mc) / np.sqrt(vg + vc + 1e-12)
    assert np.mean(np.abs(z) > 3) < 0.02
    rmse = np.sqrt(np.mean((mg - mc) ** 2)) / mc.mean()
    say(f"[pt] rel RMSE {rmse:.4f}")
    assert rmse <= 0.01


def test_denoiser_matches_restatement(ctx):
    sc = scenes.cube_room()
    rs = rgapi.RenderScene(ctx, sc)
    W, H = 96, 80
    st = rgapi.default_settings(width=W, height=H, spp=8, max_depth=5)
    noisy = rs.render(st)["hdr"]
    st.denoise = 1
    den = rs.render(st)["hdr"]
    nrm, alb, dep = oracle.aov(sc, W, H)
    want = oracle.denoise(noisy, nrm, alb, dep)
    np.testing.assert_allclose(den, want, rtol=2e-5, atol=2e-6)
    # the filter removes noise: neighbouring-pixel differences shrink
    assert np.abs(np.diff(den[..., :3], axis=1)).mean() < 0.5 * np.abs(np.diff(noisy[..., :3], axis=1)).mean()


def test_adaptive_driver_matches_restatement(ctx):
    sc = scenes.cube_room()
    rs = rgapi.RenderScene(ctx, sc)
    W, H = 96, 64
    st = rgapi.default_settings(width=W, height=H, spp=64, max_depth=5, adaptive=1, adaptive_base=4,
                                adaptive_threshold=0.12)
    out = rs.render(st)
    sm, sq, nt = oracle.render_adaptive(sc, W, H, 64, 5, base=4, threshold=0.12, mode=1, threads=8)
    say(f"[pt] adaptive tile spp gpu {out['nspp'].tolist()} oracle {nt.tolist()}")
    assert (out["nspp"] == nt).all()
    assert len(np.unique(nt)) > 1
    r = rel_diff(out["sum"], sm)
    assert np.mean(r > 1e-4) < 0.002
    # HDR is each pixel's mean over its tile's own sample count
    per_px = np.repeat(np.repeat(nt.reshape((H + 31) // 32, (W + 31) // 32), 32, 0), 32, 1)[:H, :W]
    np.testing.assert_allclose(out["hdr"][..., :3], out["sum"] / per_px[..., None], rtol=1e-5, atol=1e-7)


def test_cube_room_full_size_and_samples_in_flight(ctx):
    sc = scenes.cube_room()
    rs = rgapi.RenderScene(ctx, sc)
    a = rs.render(rgapi.default_settings(width=256, height=256, spp=16, max_depth=5))
    b = rs.render(rgapi.default_settings(width=256, height=256, spp=16, max_depth=5, samples_in_flight=3))
    assert np.isfinite(a["hdr"]).all()
    # a different split of the samples into launches gives the same image
    np.testing.assert_allclose(a["hdr"], b["hdr"], rtol=1e-5, atol=1e-6)
    assert 0.05 < a["hdr"][..., :3].mean() < 1.0


def test_render_errors_and_cancel(ctx):
    sc = scenes.cube_room()
    s = rgapi.Scene(ctx, sc["v9"])
    s.build()
    st = rgapi.default_settings(width=8, height=8, spp=1, max_depth=5)
    with pytest.raises(rgapi.RGError):
        rgapi.check(rgapi.lib().rg_render(ctx.h, s.h, rgapi.ctypes.byref(st), None, None, None, None))
    c2 = rgapi.Context(0)
    rs = rgapi.RenderScene(c2, sc)
    # as the reference (rayrenderer.cpp:537), a render clears an earlier cancel
    c2.cancel()
    want = rs.render(st)
    # cancelled from the progress callback, on the rendering thread
    big = rgapi.default_settings(width=64, height=64, spp=64, max_depth=5, samples_in_flight=4)
    seen = []

    def cancel_at_first(p):
        seen.append(p)
        c2.cancel()
    c2.set_progress(cancel_at_first)
    with pytest.raises(rgapi.RGError) as e:
        rs.render(big)
    assert e.value.code == -5 and len(seen) == 1 and 0.0 < seen[0] < 1.0
    c2.set_progress(None)
    again = rs.render(st)
    assert np.array_equal(want["hdr"], again["hdr"])


def test_progress_callback(ctx):
    """Progress reaches 1.0 through non-decreasing fractions of the sample
    budget, one per finished generation, and does not change the image."""
    rs = rgapi.RenderScene(ctx, scenes.cube_room())
    st = rgapi.default_settings(width=48, height=32, spp=20, max_depth=5, samples_in_flight=3)
    plain = rs.render(st)
    seen = []
    ctx.set_progress(seen.append)
    try:
        out = rs.render(st)
        ada = rgapi.default_settings(width=48, height=32, spp=32, max_depth=5, adaptive=1,
                                     adaptive_base=4, adaptive_threshold=0.05)
        seen_ada_start = len(seen)
        rs.render(ada)
    finally:
        ctx.set_progress(None)
    assert np.array_equal(plain["hdr"], out["hdr"])
    first = seen[:seen_ada_start]
    # 7 generations of 3, 3, ..., 2 samples: the first six are reported when
    # the next is queued, then 1.0
    np.testing.assert_allclose(first, [3 / 20, 6 / 20, 9 / 20, 12 / 20, 15 / 20, 18 / 20, 1.0], rtol=1e-6)
    second = seen[seen_ada_start:]
    assert second[-1] == 1.0 and all(a < b for a, b in zip(second, second[1:]))


def test_loaded_json_scene_matches_single_lobe_restatement(ctx):
    """rg_scene_load_json -> rg_bvh_build -> rg_render on the GPU against the
    oracle twin given the loader's flattened arrays (planes, rotated cube
    with a medium, refractive sphere, a .mesh file)."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "room_scene.json")
    sc, d = rgapi.load_json(ctx, path)
    sc.build()
    W, H, spp = 64, 48, 32
    say("[pt] loaded JSON room 64x48x32 vs oracle twin")
    out = sc.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=6))
    want, _ = oracle.render(d, W, H, spp, 6, mode=1, threads=8)
    r = rel_diff(out["sum"], want)
    assert np.mean(r > 1e-3) < 0.01, np.mean(r > 1e-3)
    assert abs(out["sum"].mean() / want.mean() - 1) < 2e-3
    assert want.mean() > 0.01


# name, spp, depth, largest fraction of pixels off by more than 1e-3: cubeRoom's
# big cube stands on the floor with coplanar faces, where a last-ulp
# difference of a transcendental function picks the other face; every other
# scene is held to the materials scene's bar
REF_SCENES = [("cubeRoom/cubeRoom.json", 32, 6, 0.10),
              ("sphereArray/sphereArray.json", 16, 6, 0.01),
              ("mediumTest/waterCube.json", 16, 8, 0.01),
              ("mediumTest/cloudHomo.json", 16, 8, 0.01),
              ("anisoTest/anisoTest.json", 16, 6, 0.01),
              ("bokehTest/bokehTest.json", 16, 6, 0.01),
              ("iblTest/iblTest.json", 16, 6, 0.01),
              ("suzanne/suzanne.json", 16, 8, 0.01)]


@pytest.mark.parametrize("name,spp,depth,bar", REF_SCENES)
def test_reference_scene_files_match_single_lobe_restatement(ctx, name, spp, depth, bar):
    """The reference's own sample scenes (loaded by rg_scene_load_json, with
    the checkerboard and wood floor textures; environment maps absent, so
    iblTest and suzanne see only the background colour and their own
    surfaces) rendered on the GPU against the oracle twin on the loader's
    flattened arrays: anisoTest runs the anisotropic GGX lobes, cloudHomo a
    Henyey-Greenstein medium with g = 0.6."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_scenes", name)
    sc, d = rgapi.load_json(ctx, path)
    sc.build()
    W, H = 64, 48
    say(f"[pt] reference scene {name} {W}x{H}x{spp} vs oracle twin")
    out = sc.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=depth))
    want, _ = oracle.render(d, W, H, spp, depth, mode=1, threads=8)
    r = rel_diff(out["sum"], want)
    say(f"[pt]   frac > 1e-3: {np.mean(r > 1e-3):.4f}, mean ratio {out['sum'].mean() / want.mean():.5f}")
    assert np.mean(r > 1e-3) < bar
    assert np.median(r) < 1e-5
    assert abs(out["sum"].mean() / want.mean() - 1) < 5e-3
    assert want.mean() > 1e-3


def test_eight_wide_tree_renders_identically(ctx):
    """The 8-wide tree visits the binary tree's leaves in the same order as
    the 4-wide one, so the whole path tracer's output is bit-identical."""
    sc = scenes.bsdf_frame(grid=3, stacks=16, slices=24)
    st = rgapi.default_settings(width=48, height=32, spp=8, max_depth=6)
    say("[pt] BSDF frame, 4-wide vs 8-wide BVH")
    a = rgapi.RenderScene(ctx, sc, width=4).render(st)
    b = rgapi.RenderScene(ctx, sc, width=8).render(st)
    assert np.array_equal(a["sum"], b["sum"]) and a["rays"] == b["rays"]


def test_textured_room_matches_single_lobe_restatement(ctx):
    """Albedo textures (diffuse floor and back wall in sRGB with tiling, a
    glossy metallic box whose F0 is tinted by a linear texture) on the GPU
    against the oracle twin: the texel lookups are integer/float-exact, so
    the bar is the untextured cube room's."""
    sc = scenes.textured_room()
    rs = rgapi.RenderScene(ctx, sc)
    W, H, spp = 64, 48, 32
    say("[pt] textured room 64x48x32 vs oracle twin")
    out = rs.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=5))
    want, _ = oracle.render(sc, W, H, spp, 5, mode=1, threads=8)
    r = rel_diff(out["sum"], want)
    assert np.mean(r > 1e-4) < 0.002, np.mean(r > 1e-4)
    assert abs(out["sum"].mean() / want.mean() - 1) < 1e-4
    plain = rgapi.RenderScene(ctx, scenes.cube_room()).render(
        rgapi.default_settings(width=W, height=H, spp=spp, max_depth=5))
    assert not np.allclose(out["sum"], plain["sum"], rtol=0.05)     # the textures change the image


def test_texture_furnace_on_the_gpu(ctx):
    """The oracle's texture furnace pin (a diffuse plane under a uniform sky
    reads colour x texel per quadrant) rendered on the GPU."""
    b = scenes.Builder()
    tex = np.array([[[0.9, 0.2, 0.1], [0.1, 0.8, 0.3]], [[0.4, 0.4, 0.9], [0.7, 0.05, 0.6]]], np.float32)
    t = b.add_texture(tex, srgb=True)
    o = b.add_object(scenes.material(color=(0.8, 0.8, 0.8)), texture=t)
    b.quad(o, (-1, 1, -3), (-1, -1, -3), (1, -1, -3), (1, 1, -3), normal=(0, 0, 1),
           uv=((0, 0), (0, 1), (1, 1), (1, 0)))
    sc = b.finish(camera={"fov_deg": 30.0, "aperture": 1.8}, env=scenes.constant_env(1.0, 64, 32))
    rs = rgapi.RenderScene(ctx, sc)
    st = rgapi.default_settings(width=40, height=40, spp=256, max_depth=8, antialias=0, firefly_clamp=0.0)
    out = rs.render(st)
    want, _ = oracle.render(sc, 40, 40, 256, 8, mode=1, threads=8, antialias=False, firefly_clamp=0.0)
    np.testing.assert_allclose(out["sum"], want, rtol=1e-4, atol=1e-5)
    lin = np.where(tex <= 0.04045, tex / 12.92, ((tex + 0.055) / 1.055) ** 2.4)
    q = out["hdr"][8:14, 8:14, :3].reshape(-1, 3).mean(0)
    np.testing.assert_allclose(q, 0.8 * lin[0, 0], rtol=0.02)


def test_textured_batch_equals_single_job(ctx):
    rs = rgapi.RenderScene(ctx, scenes.textured_room())
    st = dict(width=48, height=40, spp=6, max_depth=5)
    single = rs.render(rgapi.default_settings(**st))
    outs, _ = rgapi.render_batch(ctx, [rs], [rgapi.make_job(0, **st)])
    assert np.array_equal(outs[0]["hdr"], single["hdr"])


def test_dispersive_glass_matches_single_lobe_restatement(ctx):
    """Refractive spheres with chromatic dispersion (bsdf.cpp:394-420): the
    band drawn at a path's first dispersive interface is kept in the path
    state across events; per pixel against the restatement, and the batched
    engine's image equals the single-job path's."""
    sc = scenes.bsdf_frame(grid=3, stacks=16, slices=24, dispersion=0.05)
    assert (sc["objects"]["chroma_dispersion"] > 0).sum() >= 2
    rs = rgapi.RenderScene(ctx, sc)
    W, H = 72, 48
    spp = 32
    say("[pt] dispersive glass 72x48x32 vs oracle twin")
    out = rs.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=8))
    want, _ = oracle.render(sc, W, H, spp, 8, mode=1, threads=8)
    r = rel_diff(out["sum"], want)
    assert np.mean(r > 1e-3) < 0.01, np.mean(r > 1e-3)
    assert abs(out["sum"].mean() / want.mean() - 1) < 2e-3
    plain = rgapi.RenderScene(ctx, scenes.bsdf_frame(grid=3, stacks=16, slices=24))
    base = plain.render(rgapi.default_settings(width=W, height=H, spp=spp, max_depth=8))
    assert not np.array_equal(base["sum"], out["sum"])
    st = dict(width=W, height=H, spp=8, max_depth=8)
    single = rs.render(rgapi.default_settings(**st))
    outs, _ = rgapi.render_batch(ctx, [rs], [rgapi.make_job(0, **st)])
    assert np.array_equal(outs[0]["hdr"], single["hdr"])


def test_render_into_page_locked_buffer(ctx):
    """rg_host_alloc'd output (rgapi.host_zeros) receives the same image as
    pageable memory, over repeated renders into the same buffer."""
    rs = rgapi.RenderScene(ctx, scenes.cube_room())
    st = rgapi.default_settings(width=40, height=24, spp=4, max_depth=5)
    want = rs.render(st, hdr_only=True)["hdr"]
    buf = rgapi.host_zeros((24, 40, 4))
    for _ in range(2):
        got = rs.render(st, hdr_only=True, hdr_out=buf)["hdr"]
        assert got is buf
        assert np.array_equal(buf, want)

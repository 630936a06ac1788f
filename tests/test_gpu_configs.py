"""The configured workloads that only the benchmark ran before, as parity
tests on the GPU through the C ABI:

* full BSDF frame (25 spheres of 80,000 triangles, GGX / glossy / refraction
  materials, medium box, area light and constant environment, depth 8,
  adaptive sampling at threshold 1e-3 with base 4, 5-level A-Trous): on a
  3 x 3 grid at 96 x 64 every stage against the oracle (tile sample counts,
  per-pixel sums against the single-lobe restatement, the denoised image
  against the oracle's denoiser on the same noisy image and guides); at the
  configured 1920 x 1080 and 2,000,014 triangles by properties: finite,
  non-negative, per-tile counts in passes of the base up to 64, and the
  adaptive image's mean equal to an independent 4-spp render's;
* serving stream: the four resident scenes of the served workload (cube room,
  two full BSDF frames, the 1.31M-triangle GGX icosphere with the medium box)
  and a seeded Poisson stream of requests drawn like the batch jobs; every
  served image equals rg_render's for the same request bit for bit.
"""
import sys
import time

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


def test_bsdf_frame_adaptive_and_denoise_match_restatement(ctx):
    sc = scenes.bsdf_frame(grid=3, stacks=16, slices=24)
    rs = rgapi.RenderScene(ctx, sc)
    W, H, spp, depth = 96, 64, 64, 8
    say("[cfg2] 3x3 BSDF frame 96x64, adaptive 1e-3 base 4 + 5-level A-Trous vs oracle")
    st = rgapi.default_settings(width=W, height=H, spp=spp, max_depth=depth, adaptive=1, adaptive_base=4,
                                adaptive_threshold=1e-3)
    noisy = rs.render(st)
    sm, sq, nt = oracle.render_adaptive(sc, W, H, spp, depth, base=4, threshold=1e-3, mode=1, threads=8)
    assert (noisy["nspp"] == nt).all()
    r = np.abs(noisy["sum"] - sm) / np.maximum(np.abs(sm), 1e-3)
    assert np.mean(r > 1e-3) < 0.01
    assert abs(noisy["sum"].mean() / sm.mean() - 1) < 2e-3
    # a looser threshold stops some tiles early, identically on both sides
    st2 = rgapi.default_settings(width=W, height=H, spp=spp, max_depth=depth, adaptive=1, adaptive_base=4,
                                 adaptive_threshold=0.12)
    a2 = rs.render(st2)
    _, _, nt2 = oracle.render_adaptive(sc, W, H, spp, depth, base=4, threshold=0.12, mode=1, threads=8)
    assert (a2["nspp"] == nt2).all() and len(np.unique(nt2)) > 1
    st.denoise = 1
    den = rs.render(st)
    nrm, alb, dep = oracle.aov(sc, W, H)
    want = oracle.denoise(noisy["hdr"], nrm, alb, dep, levels=5)
    np.testing.assert_allclose(den["hdr"], want, rtol=2e-5, atol=2e-6)


def test_full_bsdf_frame_properties(ctx):
    t0 = time.perf_counter()
    sc = scenes.bsdf_frame()
    rs = rgapi.RenderScene(ctx, sc)
    assert len(sc["v9"]) == 2_000_014
    say(f"[cfg2] full BSDF frame 1920x1080, 2,000,014 tris (set-up {time.perf_counter() - t0:.1f}s)")
    W, H = 1920, 1080
    st = rgapi.default_settings(width=W, height=H, spp=64, max_depth=8, adaptive=1, adaptive_base=4,
                                adaptive_threshold=1e-3)
    ada = rs.render(st)
    hdr = ada["hdr"]
    assert np.isfinite(hdr).all() and (hdr >= 0).all()
    n = ada["nspp"]
    assert len(n) == ((W + 31) // 32) * ((H + 31) // 32)
    assert (n % 4 == 0).all() and n.min() >= 4 and n.max() == 64
    low = rs.render(rgapi.default_settings(width=W, height=H, spp=4, max_depth=8, seed=0x4444))
    ratio = hdr[..., :3].mean() / low["hdr"][..., :3].mean()
    say(f"[cfg2]   tiles at 64 spp: {np.mean(n == 64):.3f}, mean ratio adaptive / 4 spp {ratio:.4f}")
    assert abs(ratio - 1) < 0.01
    st.denoise = 1
    den = rs.render(st)["hdr"]
    assert np.isfinite(den).all() and (den >= 0).all()
    assert abs(den[..., :3].mean() / hdr[..., :3].mean() - 1) < 0.03
    # the denoiser smooths: neighbouring-pixel differences shrink
    assert np.abs(np.diff(den[..., :3], axis=1)).mean() < 0.7 * np.abs(np.diff(hdr[..., :3], axis=1)).mean()


def draw_requests(n, n_scenes, seed):
    """Requests like the batch workload: scene uniform, W and H log-uniform
    in [128, 1024], spp U[4, 256], depth U[4, 8], own seed and job index."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        w, h = (int(np.exp(rng.uniform(np.log(128), np.log(1024)))) for _ in range(2))
        out.append((int(rng.integers(n_scenes)), dict(width=w, height=h, spp=int(rng.integers(4, 257)),
                                                       max_depth=int(rng.integers(4, 9)),
                                                       seed=(seed + i) & 0xFFFFFFFF, job=i)))
    return out


def test_served_stream_over_the_four_resident_scenes(ctx):
    t0 = time.perf_counter()
    resident = [rgapi.RenderScene(ctx, scenes.cube_room()),
                rgapi.RenderScene(ctx, scenes.bsdf_frame()),
                rgapi.RenderScene(ctx, scenes.bsdf_frame(seed=0xB5DF0003)),
                rgapi.RenderScene(ctx, scenes.ico_ggx())]
    assert len(resident[3].src["v9"]) == 1_310_736
    say(f"[cfg4] four resident scenes ready in {time.perf_counter() - t0:.1f}s")
    spec = draw_requests(20, 4, 0xC4C40001)
    assert {s for s, _ in spec} == {0, 1, 2, 3}
    rng = np.random.default_rng(0x5E7E0001)
    gaps = rng.exponential(0.02, len(spec))
    srv = rgapi.Server(ctx, resident, hbm_budget=8 << 30)
    try:
        tickets = []
        for (s, kw), gap in zip(spec, gaps):
            time.sleep(gap)
            tickets.append(srv.submit(rgapi.make_job(s, **kw)))
        served = []
        for i, t in enumerate(tickets):
            info, out = srv.wait(t, timeout=300.0)
            assert info is not None and info.status == 0, f"request {i}"
            assert info.submit_s <= info.admit_s <= info.done_s
            served.append(out["hdr"])
        st = srv.stats()
        assert st["requests"] == len(spec) and st["fb_bytes"] == 0
    finally:
        srv.close()
    say(f"[cfg4] {len(spec)} requests served in {st['generations']} generations; comparing with rg_render")
    for i, (s, kw) in enumerate(spec):
        o = resident[s].render(rgapi.default_settings(**kw), hdr_only=True)
        assert np.array_equal(o["hdr"], served[i]), f"request {i} (scene {s}, {kw}) differs"

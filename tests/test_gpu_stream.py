"""Batched and served rendering (rg_render_batch, rg_server_*, rg_autotune)
on the GPU.  Every job's image must equal what the single-job path
(rg_render) returns for the same scene and settings: the engine draws the
same random numbers per sample and sums each pixel's samples in the same
order and grouping, so the comparison is exact."""
import sys
import time

import numpy as np
import pytest

import rgapi
import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return rgapi.Context(0)


@pytest.fixture(scope="module")
def resident(ctx):
    room = rgapi.RenderScene(ctx, scenes.cube_room())
    frame = rgapi.RenderScene(ctx, scenes.bsdf_frame(grid=3, stacks=16, slices=24))
    return [room, frame]


# uneven resolutions, sample counts, depths, seeds; one job with an explicit
# samples-in-flight count so the per-pixel grouping is not trivial
JOBS = [
    (0, dict(width=64, height=48, spp=8, max_depth=5)),
    (1, dict(width=96, height=64, spp=5, max_depth=8)),
    (0, dict(width=33, height=17, spp=3, max_depth=4, seed=0x1234)),
    (1, dict(width=128, height=96, spp=16, max_depth=6, samples_in_flight=3)),
    (0, dict(width=200, height=150, spp=1, max_depth=5, job=7)),
    (1, dict(width=17, height=200, spp=7, max_depth=4, firefly_clamp=0.0)),
    # depth 1 beside depth 8: this job's paths must stop after their own
    # max_depth + 2 events although the generation runs 10
    (1, dict(width=80, height=60, spp=4, max_depth=1)),
]


def make_jobs():
    return [rgapi.make_job(s, **kw) for s, kw in JOBS]


@pytest.fixture(scope="module")
def single(resident):
    """rg_render of every job, one at a time."""
    outs = []
    for s, kw in JOBS:
        outs.append(resident[s].render(rgapi.default_settings(**kw)))
    return outs


def assert_same(got, want, what):
    assert np.array_equal(got["hdr"], want["hdr"]), (what, np.abs(got["hdr"] - want["hdr"]).max())
    if "sum" in got:
        assert np.array_equal(got["sum"], want["sum"]), (what, np.abs(got["sum"] - want["sum"]).max())


def test_batch_matches_single_job_path(ctx, resident, single):
    say("[stream] batch of 7 uneven jobs vs rg_render")
    outs, stats = rgapi.render_batch(ctx, resident, make_jobs())
    for i, (o, w) in enumerate(zip(outs, single)):
        assert_same(o, w, f"job {i}")
    want_samples = sum(kw["width"] * kw["height"] * kw["spp"] for _, kw in JOBS)
    assert int(stats[2]) == want_samples
    assert int(stats[0]) == sum(w["rays"] for w in single)
    assert int(stats[1]) == sum(w["shadow_rays"] for w in single)


@pytest.mark.parametrize("tune", [
    # generations of one chunk each, bounce loops replayed from graphs
    dict(pool_paths=4096, stack=256, blocks_per_cu=4, iters_per_graph=1),
    # several jobs' chunks per generation, direct launches, one lane
    dict(pool_paths=1 << 16, blocks_per_cu=16, lanes=1),
    dict(pool_paths=1 << 20, blocks_per_cu=16, iters_per_graph=1, lanes=2),
    dict(pool_paths=1 << 18, lanes=4),
])
def test_tuning_knobs_do_not_change_images(ctx, resident, single, tune):
    say(f"[stream] batch with tune {tune}")
    outs, _ = rgapi.render_batch(ctx, resident, make_jobs(), tune=rgapi.make_tune(**tune))
    for i, (o, w) in enumerate(zip(outs, single)):
        assert_same(o, w, f"job {i} {tune}")


def test_server_alone_equals_mixed(ctx, resident, single):
    """A request renders the same image alone and among others that arrive
    while it is on the GPU, with admission limited by a small budget."""
    jobs = make_jobs()
    sums = [kw["width"] * kw["height"] * (kw["spp"] * 12 + 40) for _, kw in JOBS]
    budget = max(sums) * 2 + (1 << 20)      # room for about two jobs at a time
    srv = rgapi.Server(ctx, resident, hbm_budget=budget)
    try:
        say("[stream] server: request 3 alone")
        t = srv.submit(jobs[3], want_sum=True)
        info, alone = srv.wait(t, timeout=60.0)
        assert info is not None and info.status == 0
        assert_same(alone, single[3], "alone")
        say("[stream] server: all seven, staggered")
        tickets = []
        for i, j in enumerate(jobs):
            tickets.append(srv.submit(j, want_sum=True))
            time.sleep(0.002 * i)
        for i, t in enumerate(tickets):
            info, out = srv.wait(t, timeout=120.0)
            assert info is not None, f"request {i} timed out"
            assert info.status == 0
            assert info.submit_s <= info.admit_s <= info.done_s
            assert info.samples == JOBS[i][1]["width"] * JOBS[i][1]["height"] * JOBS[i][1]["spp"]
            assert_same(out, single[i], f"mixed {i}")
        st = srv.stats()
        assert st["requests"] == len(JOBS) + 1 and st["fb_bytes"] == 0 and st["budget"] == budget
        assert st["generations"] >= 2
    finally:
        srv.close()


def test_server_rejects_bad_requests(ctx, resident):
    srv = rgapi.Server(ctx, resident, hbm_budget=1 << 20)
    try:
        with pytest.raises(rgapi.RGError) as e:
            srv.submit(rgapi.make_job(0, width=1024, height=1024, spp=64))
        assert e.value.code == -7                      # RG_E_BUDGET
        with pytest.raises(rgapi.RGError) as e:
            srv.submit(rgapi.make_job(0, width=16, height=16, spp=4, adaptive=1))
        assert e.value.code == -6                      # RG_E_UNSUPPORTED
        with pytest.raises(rgapi.RGError) as e:
            srv.submit(rgapi.make_job(5, width=16, height=16, spp=4))
        assert e.value.code == -1                      # RG_E_INVALID
        t = srv.submit(rgapi.make_job(0, width=16, height=16, spp=4))
        info, out = srv.wait(t, timeout=60.0)
        assert info.status == 0 and out["hdr"].shape == (16, 16, 4)
    finally:
        srv.close()


def test_autotune_returns_a_working_setting(ctx, resident, single):
    probe = [rgapi.make_job(0, width=128, height=96, spp=4, max_depth=5),
             rgapi.make_job(1, width=128, height=96, spp=4, max_depth=5)]
    say("[stream] autotune on two small probe jobs")
    t = rgapi.autotune(ctx, resident, probe)
    d = t.as_dict()
    say(f"[stream] tuned: {d}")
    assert d["pool_paths"] in (1 << 23, 1 << 24, 1 << 25) and d["stack"] in (32, 64, 128, 256)
    assert d["iters_per_graph"] in (0, 1) and d["blocks_per_cu"] in (0, 4, 8, 16) and d["lanes"] in (1, 2, 3, 4)
    outs, _ = rgapi.render_batch(ctx, resident, make_jobs()[:2], tune=t)
    for i in range(2):
        assert_same(outs[i], single[i], f"tuned job {i}")


def test_batch_on_eight_wide_trees(ctx, single):
    """The batched engine on scenes built with 8-wide nodes gives the same
    images; scenes of different widths cannot share an engine."""
    r8 = [rgapi.RenderScene(ctx, scenes.cube_room(), width=8),
          rgapi.RenderScene(ctx, scenes.bsdf_frame(grid=3, stacks=16, slices=24), width=8)]
    outs, _ = rgapi.render_batch(ctx, r8, make_jobs()[:3])
    for i in range(3):
        assert_same(outs[i], single[i], f"8-wide job {i}")
    r4 = rgapi.RenderScene(ctx, scenes.cube_room(), width=4)
    with pytest.raises(rgapi.RGError) as e:
        rgapi.render_batch(ctx, [r4, r8[1]], make_jobs()[:1])
    assert e.value.code == -6


def test_cancel_and_destroy_with_work_in_flight():
    """A server destroyed with requests queued or rendering returns and leaves
    the device usable; rg_cancel cancels the batch in flight (a batch clears
    an earlier cancel when it starts, as rg_render does) and fails a
    server's queued and running requests, after which it serves again."""
    import threading

    c2 = rgapi.Context(0)
    room = rgapi.RenderScene(c2, scenes.cube_room())
    big = rgapi.make_job(0, width=1024, height=1024, spp=1024, max_depth=5)
    small = rgapi.make_job(0, width=32, height=32, spp=2)
    srv = rgapi.Server(c2, [room])
    try:
        say("[stream] destroy with requests in flight")
        for _ in range(4):
            srv.submit(big)
    finally:
        srv.close()
    out = room.render(rgapi.default_settings(width=32, height=32, spp=2, max_depth=3))
    assert np.isfinite(out["hdr"]).all()

    say("[stream] cancel before a batch, then during one")
    c2.cancel()
    outs, _ = rgapi.render_batch(c2, [room], [small])
    assert np.isfinite(outs[0]["hdr"]).all()
    timer = threading.Timer(0.3, c2.cancel)
    timer.start()
    try:
        with pytest.raises(rgapi.RGError) as e:
            rgapi.render_batch(c2, [room], [big] * 3, want_sum=False)
        assert e.value.code == -5
    finally:
        timer.join()

    say("[stream] cancel a server's running requests, then serve again")
    srv = rgapi.Server(c2, [room])
    try:
        tickets = [srv.submit(big) for _ in range(3)]
        time.sleep(0.2)
        c2.cancel()
        for t in tickets:
            with pytest.raises(rgapi.RGError) as e:
                srv.wait(t, timeout=60.0)
            assert e.value.code == -5
        t = srv.submit(small)
        info, out = srv.wait(t, timeout=60.0)
        assert info.status == 0 and np.isfinite(out["hdr"]).all()
    finally:
        srv.close()


def test_server_queue_limit_refuses_beyond_the_outstanding_work(ctx, resident, single):
    """With a limit on the samples submitted and not done, requests beyond it
    are refused with RG_E_BUSY at submission; the accepted ones render as
    they do alone, and the outstanding count returns to zero."""
    jobs = make_jobs()
    sam = [kw["width"] * kw["height"] * kw["spp"] for _, kw in JOBS]
    limit = max(sam) + min(sam)
    srv = rgapi.Server(ctx, resident, hbm_budget=1 << 30)
    try:
        srv.set_queue_limit(limit)
        accepted, refused = [], 0
        for _ in range(3):
            for i, j in enumerate(jobs):
                try:
                    accepted.append((i, srv.submit(j)))
                except rgapi.RGError as e:
                    assert e.code == -9
                    refused += 1
        say(f"[stream] queue limit {limit}: {len(accepted)} accepted, {refused} refused")
        assert refused > 0 and len(accepted) >= 3
        for i, t in accepted:
            info, out = srv.wait(t, timeout=120.0)
            assert info is not None and info.status == 0
            assert np.array_equal(out["hdr"], single[i]["hdr"]), f"request {i}"
        q = srv.queue()
        assert q == {"outstanding": 0, "limit": limit, "rejected": refused}
        srv.set_queue_limit(0)
        tickets = [srv.submit(j) for j in jobs]
        for t in tickets:
            info, _ = srv.wait(t, timeout=120.0)
            assert info.status == 0
    finally:
        srv.close()

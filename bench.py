#!/usr/bin/env python3
"""Headline benchmark: closest-hit traversal throughput (Mrays/s) on the
1,310,720-triangle displaced icosphere with 16,777,216 incoherent rays,
inputs resident in HBM.

One step = one rg_trace_closest launch over the whole 16M-ray batch.  The
timed region is bracketed by torch.cuda.synchronize(); kernel durations are
taken with HIP events on the stream the kernel runs on.  The CPU baseline is
the oracle (the strict-FP C restatement of the reference traversal) on a
time-bounded prefix of the same rays, on the host's cores.

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "raygen-renderer_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

PEAK_HBM_GBS = 8000.0
N_RAYS = 1 << 24
LEVEL = 8
RAY_KEY = 0x5EED0001
MESH_SEED = 0x1C05EED


def host_threads():
    """Threads for the CPU baselines: the CPUs this process may run on,
    capped at 16 (the host share of one GPU on the boxes this runs on)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_desc(threads, fast_math=False):
    build = "gcc -O3 -ffp-contract=off -fno-fast-math"
    if fast_math:
        build += " (value); gcc -O3 -ffast-math -funroll-loops -march=x86-64-v3 (value_fast_math)"
    return {"threads": threads, "host_cpus": os.cpu_count(), "cpu_model": cpu_model(), "oracle_build": build}


def host_has_avx2_fma():
    """The fast-math oracle build targets x86-64-v3 (AVX2, FMA)."""
    try:
        with open("/proc/cpuinfo") as f:
            flags = next((ln.split(":", 1)[1].split() for ln in f if ln.startswith("flags")), [])
    except OSError:
        return False
    return "avx2" in flags and "fma" in flags


def cpu_baseline(v9, rays_np, seconds, threads):
    """Oracle traversal on a bounded prefix of the same rays.  Also returns
    the mean binary-tree node pops and triangle tests per ray, which fix the
    algorithmic bytes per ray."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle

    t0 = time.perf_counter()
    nodes, perm = oracle.bvh_build(v9)
    build_s = time.perf_counter() - t0
    chunk = 1 << 17
    done, busy = 0, 0.0
    pops = tests = 0
    while done < len(rays_np) and busy < seconds:
        r = rays_np[done: done + chunk]
        t0 = time.perf_counter()
        _, cnt = oracle.trace_closest(nodes, perm, v9, r, counts=True, threads=threads)
        busy += time.perf_counter() - t0
        pops += int(cnt[:, 0].sum())
        tests += int(cnt[:, 1].sum())
        done += len(r)
    # the same traversal built with the reference's compiler flags
    # (fast-math, build/linux/makefile:3), on the same rays and threads
    fdone, fbusy = 0, 0.0
    while fdone < done and fbusy < seconds / 2 and host_has_avx2_fma():
        r = rays_np[fdone: fdone + chunk]
        t0 = time.perf_counter()
        oracle.trace_closest(nodes, perm, v9, r, threads=threads, fast_math=True)
        fbusy += time.perf_counter() - t0
        fdone += len(r)
    # the reference's own default is one render thread (rayrenderer.h:44)
    n1 = min(len(rays_np), 1 << 17)
    t0 = time.perf_counter()
    oracle.trace_closest(nodes, perm, v9, rays_np[:n1], threads=1)
    one = n1 / (time.perf_counter() - t0) / 1e6
    return {
        "value": done / busy / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {done} of the {len(rays_np)} rays, closest-hit, oracle binary-tree "
                  f"traversal with {threads} threads (4096-ray chunks, atomic work claiming); "
                  f"BVH build {build_s:.2f}s excluded",
        "value_1_thread": round(one, 3),
        "sample_1_thread": f"first {n1} rays, the same traversal on 1 thread",
        "value_fast_math": round(fdone / fbusy / 1e6, 3) if fbusy else None,
        "sample_fast_math": f"first {fdone} rays, the same traversal built with the reference's flags "
                            f"(gcc -O3 -ffast-math -funroll-loops -march=x86-64-v3), {threads} threads",
        "bvh_build_s": round(build_s, 3),
        "host": host_desc(threads, fast_math=fbusy > 0),
    }, pops / done, tests / done


def bench_cube_room(ctx, stream, steps, cpu_seconds, do_cpu):
    """Plain cube room (36 triangles), 256x256, 16 spp, depth 5, NEE+MIS, AA,
    no denoise: frame time, samples/s and rays/s of the whole render call."""
    import numpy as np
    import torch

    import rgapi
    import scenes

    sc = scenes.cube_room()
    rs = rgapi.RenderScene(ctx, sc)
    st = rgapi.default_settings(width=256, height=256, spp=16, max_depth=5, antialias=1, seed=0xC0DE0000)
    # the HDR image goes to a page-locked host buffer, as render() leaves it
    hdr = rgapi.host_zeros((256, 256, 4))
    out = rs.render(st, hdr_only=True, hdr_out=hdr)  # warm-up (buffers, upload)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = rs.render(st, hdr_only=True, hdr_out=hdr)
    wall = (time.perf_counter() - t0) / steps
    res = {"workload": "cube room 36 tris, 256x256, 16 spp, depth 5, NEE+MIS, AA, no denoise",
           "ms_per_frame": round(wall * 1e3, 3),
           "samples_per_s": round(out["samples"] / wall, 1),
           "mrays_per_s": round((out["rays"] + out["shadow_rays"]) / wall / 1e6, 2),
           "rays_per_frame": out["rays"], "shadow_rays_per_frame": out["shadow_rays"],
           "mean_radiance": [round(float(x), 5) for x in out["hdr"][..., :3].mean((0, 1))]}
    if do_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = host_threads()
        spp, t, done = 1, 0.0, 0
        while t < cpu_seconds and done < 16:
            t0 = time.perf_counter()
            oracle.render(sc, 256, 256, spp, 5, mode=0, threads=threads, sample_start=done)
            t += time.perf_counter() - t0
            done += spp
        res["cpu_baseline"] = {"value": round(256 * 256 * done / t, 1), "unit": "samples/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{done} spp of the 256x256 frame, recursive restatement of the "
                                         f"reference shader (all Mix lobes), {threads} threads",
                               "host": host_desc(threads)}
    return res


def bench_bsdf_frame(ctx, do_cpu, cpu_seconds):
    """Full BSDF frame: 5x5 spheres x 80k triangles, GGX/refraction, area
    light + constant environment, homogeneous medium box, 1920x1080, 64 spp,
    depth 8, adaptive (threshold 1e-3), 5-level A-Trous."""
    import numpy as np
    import torch

    import rgapi
    import scenes

    t0 = time.perf_counter()
    sc = scenes.bsdf_frame(grid=5, stacks=200, slices=200)
    t1 = time.perf_counter()
    rs = rgapi.RenderScene(ctx, sc)
    t2 = time.perf_counter()
    rs.build()  # again: the GPU builder alone, the triangles already on the device
    build_s = time.perf_counter() - t2
    st = rgapi.default_settings(width=1920, height=1080, spp=64, max_depth=8, adaptive=1, adaptive_base=4,
                                adaptive_threshold=1e-3, denoise=1, denoise_levels=5, seed=0xB5DF0002)
    warm = rgapi.default_settings(width=1920, height=1080, spp=1, max_depth=8)
    hdr = rgapi.host_zeros((1080, 1920, 4))  # page-locked: the image arrives by DMA
    rs.render(warm, sums=False, hdr_out=hdr)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    out = rs.render(st, sums=False, hdr_out=hdr)  # the HDR image and the tile sample counts to the host
    wall = time.perf_counter() - t3
    # samples taken: each 32x32 tile's sample count times its pixels (the
    # path tracer also traces the later passes of an adaptive window for
    # tiles that stop inside it; those are counted in samples_traced)
    tw, th = 1920, 1080
    taken = 0
    for t, k in enumerate(out["nspp"]):
        x0, y0 = (t % ((tw + 31) // 32)) * 32, (t // ((tw + 31) // 32)) * 32
        taken += int(k) * min(32, tw - x0) * min(32, th - y0)
    res = {"workload": "BSDF frame: 25 spheres x 80k tris (2,000,014 tris), GGX/refraction/medium/env, "
                       "1920x1080, 64 spp, depth 8, adaptive 1e-3, 5-level A-Trous",
           "tris": int(len(sc["v9"])),
           "frame_s": round(wall, 4),
           "samples": taken,
           "samples_traced": out["samples"],
           "samples_per_s": round(taken / wall, 1),
           "mrays_per_s": round((out["rays"] + out["shadow_rays"]) / wall / 1e6, 2),
           "rays": out["rays"], "shadow_rays": out["shadow_rays"],
           "mean_spp": round(float(taken) / (1920 * 1080), 2),
           "scene_gen_s": round(t1 - t0, 2), "scene_setup_s": round(t2 - t1, 3),
           "bvh_rebuild_s": round(build_s, 4)}
    res["postprocess"] = bench_postprocess(ctx, out["hdr"], do_cpu)
    if do_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = host_threads()
        # the oracle builds its BVH inside every render call: a 1-pixel render
        # measures that build, which is subtracted and reported apart
        t0 = time.perf_counter()
        oracle.render(sc, 1, 1, 1, 8, mode=0, threads=1)
        build = time.perf_counter() - t0
        W, H, spp, t, done = 384, 216, 1, 0.0, 0
        while t < cpu_seconds and done < 64:
            t0 = time.perf_counter()
            oracle.render(sc, W, H, spp, 8, mode=0, threads=threads, sample_start=done)
            t += time.perf_counter() - t0 - build
            done += spp
        res["cpu_baseline"] = {"value": round(W * H * done / t, 1), "unit": "samples/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{W}x{H} at {done} spp of the same scene (no adaptive pass, no "
                                         f"denoise), recursive restatement, {threads} threads; the oracle's "
                                         f"BVH build ({build:.2f}s per call) subtracted",
                               "bvh_build_s": round(build, 3), "host": host_desc(threads)}
    return res


def bench_postprocess(ctx, hdr, do_cpu, reps=20):
    """Output stage of the 1080p frame on device images: HDR bloom with the
    reference's defaults + Reinhard/gamma tonemap (rg_postprocess_device),
    timed with HIP events on the context's stream.  Algorithmic bytes: the
    HDR image read by the glow downsample (16 B/px), read and written by the
    composite (32 B/px) and the LDR image written (16 B/px) = 64 B/px."""
    import numpy as np
    import torch

    import rgapi

    h, w = hdr.shape[:2]
    d_src = torch.from_numpy(np.ascontiguousarray(hdr)).to("cuda:0")
    d_hdr = d_src.clone()
    d_ldr = torch.empty_like(d_hdr)
    b = rgapi.default_bloom()
    stream = torch.cuda.current_stream()
    rgapi.postprocess_device(ctx, w, h, d_hdr.data_ptr(), b, d_ldr.data_ptr())   # warm-up
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        d_hdr.copy_(d_src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        rgapi.postprocess_device(ctx, w, h, d_hdr.data_ptr(), b, d_ldr.data_ptr())
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    ms.sort()
    med = ms[len(ms) // 2]
    res = {"workload": f"bloom (threshold 1, strength 1, radius 0.03, size 0.15) + tonemap, {w}x{h}",
           "ms": round(med, 4), "mpix_per_s": round(w * h / (med / 1e3) / 1e6, 1),
           "algorithmic_gb_s": round(64 * w * h / (med / 1e3) / 1e9, 1)}
    if do_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        t0 = time.perf_counter()
        oracle.tonemap(oracle.bloom(hdr))
        t = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(w * h / t / 1e6, 2), "unit": "Mpix/s", "cores": 1, "kind": "port",
                               "sample": f"the same {w}x{h} frame through the oracle's bloom + tonemap, 1 thread"}
    return res


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def draw_jobs(n, n_scenes, seed):
    """n render jobs drawn from one seed: scene uniform over the resident
    scenes, width and height each log-uniform in [128, 1024], spp uniform in
    [4, 256], depth uniform in [4, 8], a distinct Philox key per job."""
    import math

    import numpy as np

    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        scene = int(rng.integers(0, n_scenes))
        w = int(round(math.exp(rng.uniform(math.log(128), math.log(1024)))))
        h = int(round(math.exp(rng.uniform(math.log(128), math.log(1024)))))
        spp = int(rng.integers(4, 257))
        depth = int(rng.integers(4, 9))
        jobs.append((scene, dict(width=w, height=h, spp=spp, max_depth=depth, seed=(seed + i) & 0xFFFFFFFF, job=i)))
    return jobs


def bench_batch(ctx, resident, n_jobs=64, seed=0xC3C30000):
    """64 jobs (cube room or BSDF frame; uneven resolution, spp, depth) as one
    rg_render_batch over a path pool the autotuner sized, against the same
    jobs one at a time through rg_render."""
    import numpy as np

    import rgapi

    spec = draw_jobs(n_jobs, len(resident), seed)
    jobs = [rgapi.make_job(s, **kw) for s, kw in spec]
    samples = sum(kw["width"] * kw["height"] * kw["spp"] for _, kw in spec)
    # warm both paths on a small job
    resident[0].render(rgapi.default_settings(width=64, height=64, spp=2), hdr_only=True)
    t0 = time.perf_counter()
    # probe jobs from the same distribution (another seed), capped at 64 spp so
    # every candidate is timed on ~10^8 samples in well under a second
    probe = []
    for s, kw in draw_jobs(8, len(resident), seed ^ 0x9E3779B9):
        kw["spp"] = min(kw["spp"], 64)
        probe.append(rgapi.make_job(s, **kw))
    tune = rgapi.autotune(ctx, resident, probe)
    tune_s = time.perf_counter() - t0
    log("autotune", tune.as_dict(), f"{tune_s:.1f}s")
    t0 = time.perf_counter()
    outs, stats = rgapi.render_batch(ctx, resident, jobs, tune=tune, want_sum=False)
    batch_s = time.perf_counter() - t0
    log(f"batch {batch_s:.2f}s")
    single_s, rays, same, differ = 0.0, 0, 0, []
    for i, (s, kw) in enumerate(spec):
        st = rgapi.default_settings(**kw)
        t0 = time.perf_counter()
        o = resident[s].render(st, hdr_only=True)
        single_s += time.perf_counter() - t0
        rays += o["rays"] + o["shadow_rays"]
        eq = np.array_equal(o["hdr"], outs[i]["hdr"])
        same += int(eq)
        if not eq:
            differ.append(i)
    log(f"one at a time {single_s:.2f}s")
    brays = int(stats[0]) + int(stats[1])
    return {"workload": f"{n_jobs} jobs (cube room 36 tris / BSDF frame 2,000,014 tris), W,H log-U[128,1024], "
                        f"spp U[4,256], depth U[4,8], seed {seed:#x}",
            "jobs": n_jobs, "samples": samples,
            "batched_s": round(batch_s, 3), "one_at_a_time_s": round(single_s, 3),
            "speedup": round(single_s / batch_s, 3),
            "batched_samples_per_s": round(samples / batch_s, 1),
            "one_at_a_time_samples_per_s": round(samples / single_s, 1),
            "batched_mrays_per_s": round(brays / batch_s / 1e6, 2),
            "one_at_a_time_mrays_per_s": round(rays / single_s / 1e6, 2),
            "rays_equal": brays == rays, "images_bit_equal": f"{same}/{n_jobs}", "images_differing": differ,
            "tune": tune.as_dict(), "autotune_s": round(tune_s, 2)}


def bench_batch_small(ctx, resident, n_jobs=256, seed=0xC5C50000):
    """Many small jobs (preview renders: 128x128, 4 spp, depth 5, alternating
    the two scenes): batched against one at a time.  Each job alone keeps
    only 65,536 paths in flight, a fraction of what the GPU holds, so this is
    where packing jobs into shared generations pays; not one of the
    configured workloads."""
    import numpy as np

    import rgapi

    spec = [(i % len(resident), dict(width=128, height=128, spp=4, max_depth=5, seed=(seed + i) & 0xFFFFFFFF, job=i))
            for i in range(n_jobs)]
    jobs = [rgapi.make_job(s, **kw) for s, kw in spec]
    rgapi.render_batch(ctx, resident, jobs[:8], want_sum=False)  # warm
    t0 = time.perf_counter()
    outs, _ = rgapi.render_batch(ctx, resident, jobs, want_sum=False)
    batch_s = time.perf_counter() - t0
    single_s, same = 0.0, 0
    for i, (s, kw) in enumerate(spec):
        t0 = time.perf_counter()
        o = resident[s].render(rgapi.default_settings(**kw), hdr_only=True)
        single_s += time.perf_counter() - t0
        same += int(np.array_equal(o["hdr"], outs[i]["hdr"]))
    samples = n_jobs * 128 * 128 * 4
    return {"workload": f"{n_jobs} preview jobs 128x128, 4 spp, depth 5, cube room / BSDF frame alternating",
            "batched_s": round(batch_s, 4), "one_at_a_time_s": round(single_s, 4),
            "speedup": round(single_s / batch_s, 3), "batched_samples_per_s": round(samples / batch_s, 1),
            "images_bit_equal": f"{same}/{n_jobs}"}


def bench_serve(ctx, resident, n_req, rates, seed=0xC4C40000, budget=32 << 30, slo_s=1.0):
    """Poisson request stream over four resident scenes: served by rg_server
    (continuous batching, admission by framebuffer footprint, overload
    policy: a request that would take the outstanding work past slo_s of the
    plain path's throughput is refused with RG_E_BUSY) at several arrival
    rates, with p50/p99 latency of the accepted requests; against the plain
    path one request at a time, which also renders every request once so each
    served image is compared with its plain-path image."""
    import hashlib
    import queue
    import threading

    import numpy as np

    import rgapi

    spec = draw_jobs(n_req, len(resident), seed)
    jobs = [rgapi.make_job(s, **kw) for s, kw in spec]
    sam = [kw["width"] * kw["height"] * kw["spp"] for _, kw in spec]
    # plain path, one at a time, on every request; a digest of each image
    plain_s, digest = 0.0, []
    for i, (s, kw) in enumerate(spec):
        t0 = time.perf_counter()
        o = resident[s].render(rgapi.default_settings(**kw), hdr_only=True)
        plain_s += time.perf_counter() - t0
        digest.append(hashlib.blake2b(o["hdr"].tobytes(), digest_size=16).digest())
    plain_sps = sum(sam) / plain_s
    lam_plain = plain_sps / float(np.mean(sam))
    log(f"plain {plain_sps / 1e6:.1f} Msamples/s over {n_req} requests -> {lam_plain:.2f} req/s")
    probe = []
    for s, kw in draw_jobs(8, len(resident), seed ^ 0x9E3779B9):
        kw["spp"] = min(kw["spp"], 64)
        probe.append(rgapi.make_job(s, **kw))
    t0 = time.perf_counter()
    tune = rgapi.autotune(ctx, resident, probe)
    log("autotune", tune.as_dict(), f"{time.perf_counter() - t0:.1f}s")
    limit = int(slo_s * plain_sps)
    res = {"workload": f"{n_req} requests per rate, drawn like the batch jobs over 4 resident scenes "
                       f"(cube room, two BSDF frames, GGX icosphere 1,310,736 tris + medium box), "
                       f"Poisson arrivals, HBM budget {budget >> 30} GiB",
           "plain_samples_per_s": round(plain_sps, 1), "plain_requests": n_req,
           "slo_p99_s": slo_s, "queue_limit_samples": limit,
           "overload_policy": "refuse (RG_E_BUSY) a request that would take the samples submitted and not done "
                              "past slo_p99_s x plain_samples_per_s",
           "tune": tune.as_dict(), "rates": []}
    rng = np.random.default_rng(seed ^ 0x5A5A)
    for f in rates:
        lam = f * lam_plain
        gaps = rng.exponential(1.0 / lam, n_req)
        srv = rgapi.Server(ctx, resident, hbm_budget=budget, tune=tune)
        srv.set_queue_limit(limit)
        todo = queue.Queue()
        lat, spans, checked = [], [None, None], {"same": 0, "differ": [], "failed": []}

        def consume():
            while True:
                item = todo.get()
                if item is None:
                    return
                i, t = item
                info, out = srv.wait(t, timeout=600.0)
                if info is None or info.status != 0:
                    checked["failed"].append(i)
                    continue
                lat.append(info.done_s - info.submit_s)
                spans[0] = info.submit_s if spans[0] is None else min(spans[0], info.submit_s)
                spans[1] = info.done_s if spans[1] is None else max(spans[1], info.done_s)
                if hashlib.blake2b(out["hdr"].tobytes(), digest_size=16).digest() == digest[i]:
                    checked["same"] += 1
                else:
                    checked["differ"].append(i)

        th = threading.Thread(target=consume)
        th.start()
        accepted, refused, acc_samples = 0, 0, 0
        t_next = time.perf_counter()
        for i, j in enumerate(jobs):
            t_next += gaps[i]
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            try:
                todo.put((i, srv.submit(j)))
                accepted += 1
                acc_samples += sam[i]
            except rgapi.RGError as e:
                if e.code != -9:
                    raise
                refused += 1
        todo.put(None)
        th.join()
        st = srv.stats()
        tn = srv.tune().as_dict()
        srv.close()
        if checked["failed"]:
            raise RuntimeError(f"requests {checked['failed'][:5]} failed")
        # a served image that is not bit-identical: its relative RMSE against
        # the plain path's image, rendered again
        rmse = 0.0
        for i in checked["differ"]:
            s_, kw = spec[i]
            a = resident[s_].render(rgapi.default_settings(**kw), hdr_only=True)["hdr"][..., :3]
            srv2 = rgapi.Server(ctx, resident, hbm_budget=budget, tune=tune)
            info, out = srv2.wait(srv2.submit(jobs[i]), timeout=600.0)
            srv2.close()
            b = out["hdr"][..., :3]
            rmse = max(rmse, float(np.sqrt(np.mean((a - b) ** 2)) / max(float(a.mean()), 1e-12)))
        span = spans[1] - spans[0]
        lat = np.array(lat)
        r = {"rate_factor": f, "arrival_rate_req_s": round(lam, 3),
             "accepted": accepted, "refused": refused,
             "served_samples_per_s": round(acc_samples / span, 1),
             "vs_plain": round(acc_samples / span / plain_sps, 3),
             "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
             "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
             "makespan_s": round(span, 3),
             "images_bit_equal_plain": f"{checked['same']}/{accepted}", "max_rel_rmse_vs_plain": rmse,
             "graph_replays": st["graphs"], "tune": tn}
        log("rate", r)
        res["rates"].append(r)
    # the highest rate served within the SLO with at most 1% of requests refused
    ok = [r for r in res["rates"] if r["p99_latency_s"] <= slo_s and r["refused"] <= 0.01 * n_req]
    res["capacity_at_slo"] = ({"rate_factor": max(r["rate_factor"] for r in ok),
                               "served_samples_per_s": max(r["served_samples_per_s"] for r in ok)}
                              if ok else None)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-render", action="store_true", help="traversal microbenchmark only")
    ap.add_argument("--serve-requests", type=int, default=2000, help="requests per arrival rate")
    ap.add_argument("--legs", default="cube,frame,batch,serve",
                    help="comma-separated workloads run after the traversal headline: cube, frame, batch, serve")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device-to-device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (multi-GPU sharding of rays / tiles / jobs "
              "with a replicated BVH is designed but not built).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import rgapi

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    ctx = rgapi.Context(0)
    ctx.set_stream(stream.cuda_stream)

    t0 = time.perf_counter()
    v9 = rgapi.gen_icosphere(LEVEL, 0.05, MESH_SEED)
    gen_s = time.perf_counter() - t0
    # the BVH: built on the GPU (the first build also loads its kernels),
    # timed warm, and the host builder's time for the same tree beside it
    t0 = time.perf_counter()
    sc = rgapi.Scene(ctx, v9)
    sc.build(bins=16, leaf=4, width=4)
    build_first_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    sc.build(bins=16, leaf=4, width=4, builder=rgapi.BUILDER_HOST)
    build_host_s = time.perf_counter() - t0
    builds = []
    for _ in range(3):
        fresh = rgapi.Scene(ctx, v9)  # a new scene: its first build uploads the triangles
        t0 = time.perf_counter()
        fresh.build(bins=16, leaf=4, width=4, builder=rgapi.BUILDER_DEVICE)
        builds.append(time.perf_counter() - t0)
        fresh.close()
    sc.build(bins=16, leaf=4, width=4, builder=rgapi.BUILDER_DEVICE)
    build_s = min(builds)

    n = N_RAYS
    d_rays = torch.empty(n * 8, dtype=torch.float32, device=dev)
    d_hits = torch.empty(n * 4, dtype=torch.float32, device=dev)
    d_occ = torch.empty(n, dtype=torch.uint8, device=dev)
    ctx.gen_rays_sphere(sc.bounds(), 1.2, RAY_KEY, 0, d_rays.data_ptr(), n)

    for _ in range(args.warmup):
        sc.trace_closest(d_rays.data_ptr(), d_hits.data_ptr(), n)
    torch.cuda.synchronize()

    # two HIP events on the launch stream bracket the back-to-back launches
    # (an event between launches costs ~40 us of idle GPU per step)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        sc.trace_closest(d_rays.data_ptr(), d_hits.data_ptr(), n)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    avg_ms = ev0.elapsed_time(ev1) / args.steps
    mrays = n * args.steps / wall / 1e6

    # any-hit variant (secondary line items)
    sc.trace_any(d_rays.data_ptr(), d_occ.data_ptr(), n)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(5):
        sc.trace_any(d_rays.data_ptr(), d_occ.data_ptr(), n)
    b.record(stream)
    torch.cuda.synchronize()
    any_ms = a.elapsed_time(b) / 5

    hits = d_hits.view(n, 4)[:, 0].view(torch.int32)
    hit_frac = float((hits != -1).float().mean().item())

    cpu = None
    pops = tests = None
    if not args.no_cpu_baseline:
        rays_np = d_rays.view(n, 8)[: 1 << 22].cpu().numpy().view(rgapi.RAY_DTYPE).reshape(-1)
        cpu, pops, tests = cpu_baseline(v9, rays_np, args.cpu_seconds, host_threads())

    # algorithmic bytes per ray: the reference traversal's node pops and
    # triangle tests over all 16M rays (tests/golden/c1_bytes_per_ray.json);
    # the CPU leg's own counts on its sample when the fixture is absent
    npr_src = None
    fx = os.path.join(ROOT, "tests", "golden", "c1_bytes_per_ray.json")
    if os.path.exists(fx):
        with open(fx) as f:
            fj = json.load(f)
        pops, tests, npr_src = fj["node_pops_per_ray"], fj["tri_tests_per_ray"], "tests/golden/c1_bytes_per_ray.json"
    elif pops is not None:
        npr_src = "CPU leg sample"
    roof = None
    if pops is not None:
        b_ray = 32 + 16 + 32 * pops + 36 * tests
        achieved = b_ray * n / (avg_ms / 1e3) / 1e9
        traffic = l2 = limiter = None
        tf = os.path.join(ROOT, "profiles", "c1_closest_traffic.json")
        if os.path.exists(tf):
            with open(tf) as f:
                tj = json.load(f)
            traffic, l2, limiter = tj.get("hbm_bytes_per_launch"), tj.get("l2_hit_rate"), tj.get("limiter")
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 5), "traffic": traffic, "l2_hit_rate": l2,
                "limiter": limiter,
                "bytes_per_ray": round(b_ray, 2), "node_pops_per_ray": round(pops, 3),
                "tri_tests_per_ray": round(tests, 3), "bytes_per_ray_source": npr_src,
                "kernel": "k_trace_ps<64, 4, closest>", "kernel_avg_ms": round(avg_ms, 4)}

    extra = {}
    if args.copy_calibration:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        extra["copy_calibration_bytes"] = 1 << 30
    legs = set() if args.skip_render else {x for x in args.legs.split(",") if x}
    if "cube" in legs:
        log("cube room")
        extra["cube_room"] = bench_cube_room(ctx, stream, max(3, args.steps // 2), 8.0, not args.no_cpu_baseline)
    if "frame" in legs:
        log("bsdf frame")
        extra["bsdf_frame"] = bench_bsdf_frame(ctx, not args.no_cpu_baseline, 8.0)
    if legs & {"batch", "serve"}:
        import scenes

        log("resident scenes")
        room = rgapi.RenderScene(ctx, scenes.cube_room())
        frame = rgapi.RenderScene(ctx, scenes.bsdf_frame(grid=5, stacks=200, slices=200))
        if "batch" in legs:
            log("batch")
            extra["batch"] = bench_batch(ctx, [room, frame], n_jobs=args.batch_jobs)
            log("batch of small jobs")
            extra["batch_small"] = bench_batch_small(ctx, [room, frame])
        if "serve" in legs:
            frame2 = rgapi.RenderScene(ctx, scenes.bsdf_frame(grid=5, stacks=200, slices=200, seed=0xB5DF0003))
            ico = rgapi.RenderScene(ctx, scenes.ico_ggx())
            log("serve")
            extra["serve"] = bench_serve(ctx, [room, frame, frame2, ico], args.serve_requests, (0.5, 1.0, 2.0))

    out = {
        "metric": "closest-hit traversal throughput, 1.31M-tri displaced icosphere, 16M incoherent rays",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (generated mesh and Philox4x32-10 rays)",
        "config": {"workload": "traversal microbench: level-8 displaced icosphere (1,310,720 tris), "
                               "SAH 16 bins leaf<=4, 16,777,216 rays origins U(1.2 bbox) dirs uniform sphere, "
                               "closest-hit (prim, t, u, v)",
                   "tris": len(v9), "rays": n, "bvh": "binary SAH collapsed to 4-wide, exact binary visit order"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "anyhit_mrays_s": round(n / (any_ms / 1e3) / 1e6, 2),
        "hit_fraction": round(hit_frac, 4),
        "bvh_build_s": round(build_s, 4),
        "bvh_build": {"device_s": round(build_s, 4), "device_first_call_s": round(build_first_s, 4),
                      "host_s": round(build_host_s, 4),
                      "what": "rg_bvh_build: binary SAH tree + 4-wide collapse + triangle records in HBM, "
                              "from host triangles (includes their upload); device and host builders give "
                              "the same nodes"},
        "mesh_gen_s": round(gen_s, 3),
        "scene_device_mb": round(sc.device_bytes() / 2**20, 1),
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

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
    return {
        "value": done / busy / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {done} of the {len(rays_np)} rays, closest-hit, oracle binary-tree "
                  f"traversal with {threads} threads (4096-ray chunks, atomic work claiming); "
                  f"BVH build {build_s:.2f}s excluded",
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
    out = rs.render(st)  # warm-up (buffers, upload)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = rs.render(st)
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
        threads = max(1, min(16, os.cpu_count() or 1))
        spp, t, done = 1, 0.0, 0
        while t < cpu_seconds and done < 16:
            t0 = time.perf_counter()
            oracle.render(sc, 256, 256, spp, 5, mode=0, threads=threads, sample_start=done)
            t += time.perf_counter() - t0
            done += spp
        res["cpu_baseline"] = {"value": round(256 * 256 * done / t, 1), "unit": "samples/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{done} spp of the 256x256 frame, recursive restatement of the "
                                         f"reference shader (all Mix lobes), {threads} threads"}
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
    st = rgapi.default_settings(width=1920, height=1080, spp=64, max_depth=8, adaptive=1, adaptive_base=4,
                                adaptive_threshold=1e-3, denoise=1, denoise_levels=5, seed=0xB5DF0002)
    warm = rgapi.default_settings(width=1920, height=1080, spp=1, max_depth=8)
    rs.render(warm)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    out = rs.render(st)
    wall = time.perf_counter() - t3
    res = {"workload": "BSDF frame: 25 spheres x 80k tris (2,000,014 tris), GGX/refraction/medium/env, "
                       "1920x1080, 64 spp, depth 8, adaptive 1e-3, 5-level A-Trous",
           "tris": int(len(sc["v9"])),
           "frame_s": round(wall, 4),
           "samples": out["samples"],
           "samples_per_s": round(out["samples"] / wall, 1),
           "mrays_per_s": round((out["rays"] + out["shadow_rays"]) / wall / 1e6, 2),
           "mean_spp": round(float(out["samples"]) / (1920 * 1080), 2),
           "scene_gen_s": round(t1 - t0, 2), "bvh_build_upload_s": round(t2 - t1, 2)}
    if do_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = max(1, min(16, os.cpu_count() or 1))
        t0 = time.perf_counter()
        oracle.render(sc, 192, 108, 2, 8, mode=0, threads=threads)
        t = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(192 * 108 * 2 / t, 1), "unit": "samples/s", "cores": threads,
                               "kind": "port",
                               "sample": f"192x108 at 2 spp of the same scene (includes the oracle's own BVH "
                                         f"build), recursive restatement, {threads} threads"}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-render", action="store_true", help="traversal microbenchmark only")
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
    t0 = time.perf_counter()
    sc = rgapi.Scene(ctx, v9)
    sc.build(bins=16, leaf=4, width=4)
    build_s = time.perf_counter() - t0

    n = N_RAYS
    d_rays = torch.empty(n * 8, dtype=torch.float32, device=dev)
    d_hits = torch.empty(n * 4, dtype=torch.float32, device=dev)
    d_occ = torch.empty(n, dtype=torch.uint8, device=dev)
    ctx.gen_rays_sphere(sc.bounds(), 1.2, RAY_KEY, 0, d_rays.data_ptr(), n)

    for _ in range(args.warmup):
        sc.trace_closest(d_rays.data_ptr(), d_hits.data_ptr(), n)
    torch.cuda.synchronize()

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t_wall = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        sc.trace_closest(d_rays.data_ptr(), d_hits.data_ptr(), n)
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    kern_ms = [a.elapsed_time(b) for a, b in ev]
    avg_ms = sum(kern_ms) / len(kern_ms)
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
        threads = max(1, min(16, os.cpu_count() or 1))
        cpu, pops, tests = cpu_baseline(v9, rays_np, args.cpu_seconds, threads)

    roof = None
    if pops is not None:
        b_ray = 32 + 16 + 32 * pops + 36 * tests
        achieved = b_ray * n / (avg_ms / 1e3) / 1e9
        traffic = None
        tf = os.path.join(ROOT, "profiles", "c1_closest_traffic.json")
        if os.path.exists(tf):
            with open(tf) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 5), "traffic": traffic,
                "bytes_per_ray": round(b_ray, 2), "node_pops_per_ray": round(pops, 3),
                "tri_tests_per_ray": round(tests, 3), "kernel": "k_closest",
                "kernel_avg_ms": round(avg_ms, 4)}

    extra = {}
    if args.copy_calibration:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        extra["copy_calibration_bytes"] = 1 << 30
    if not args.skip_render:
        extra["cube_room"] = bench_cube_room(ctx, stream, max(3, args.steps // 2), 8.0, not args.no_cpu_baseline)
        extra["bsdf_frame"] = bench_bsdf_frame(ctx, not args.no_cpu_baseline, 8.0)

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
        "bvh_build_s": round(build_s, 3),
        "mesh_gen_s": round(gen_s, 3),
        "scene_device_mb": round(sc.device_bytes() / 2**20, 1),
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

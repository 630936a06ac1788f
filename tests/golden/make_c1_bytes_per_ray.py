"""Writes tests/golden/c1_bytes_per_ray.json: the mean node pops and
triangle tests per ray of the reference's binary-tree traversal
(TriangleBVH::intersectClosest, bvh.cpp:233-285, restated by the oracle) over
all 16,777,216 rays of the traversal microbenchmark (1,310,720-triangle
displaced icosphere, Philox4x32-10 key 0x5EED0001).  They fix the
algorithmic bytes per ray B_ray = 32 (ray) + 16 (hit) + 32 n_node + 36 n_tri
that the benchmark's roofline uses.  The rays come from the library's device
generator (the benchmark's own rays), so this runs on a machine with the GPU:
    python tests/golden/make_c1_bytes_per_ray.py
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "raygen-renderer_amd"), os.path.join(ROOT, "oracle")]


def main():
    import numpy as np
    import torch

    import oracle
    import rgapi

    ctx = rgapi.Context(0)
    v9 = rgapi.gen_icosphere(8, 0.05, 0x1C05EED)
    sc = rgapi.Scene(ctx, v9)
    sc.build(bins=16, leaf=4, width=4)
    n = 1 << 24
    d = torch.empty(n * 8, dtype=torch.float32, device="cuda:0")
    ctx.gen_rays_sphere(sc.bounds(), 1.2, 0x5EED0001, 0, d.data_ptr(), n)
    torch.cuda.synchronize()
    rays = d.view(n, 8).cpu().numpy().view(rgapi.RAY_DTYPE).reshape(-1)
    nodes, perm = oracle.bvh_build(v9)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    pops = tests = 0
    t0 = time.perf_counter()
    for a in range(0, n, 1 << 21):
        _, cnt = oracle.trace_closest(nodes, perm, v9, rays[a:a + (1 << 21)], counts=True, threads=threads)
        pops += int(cnt[:, 0].sum(dtype=np.uint64))
        tests += int(cnt[:, 1].sum(dtype=np.uint64))
        print(f"{a + (1 << 21)} rays", flush=True)
    out = {"workload": "1,310,720-tri level-8 displaced icosphere (amp 0.05, seed 0x1C05EED), SAH 16 bins leaf 4; "
                       "16,777,216 rays, Philox4x32-10 key 0x5EED0001, origins U(1.2 bbox), dirs uniform sphere",
           "rays": n, "node_pops": pops, "tri_tests": tests,
           "node_pops_per_ray": pops / n, "tri_tests_per_ray": tests / n,
           "bytes_per_ray": 32 + 16 + 32 * pops / n + 36 * tests / n,
           "traversal": "oracle or_trace_closest (binary tree, stack of 64, near child by split-axis sign)",
           "cpu_s": round(time.perf_counter() - t0, 2)}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_bytes_per_ray.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

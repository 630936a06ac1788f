"""The GPU BVH builder (rg_bvh_build with RG_BUILDER_DEVICE) through the C ABI:

* the binary tree and the permutation equal the oracle's restatement of
  TriangleBVH::build (bvh.cpp:76-231) node for node and bit for bit, on
  meshes around the size where ranges move from the level kernels to the
  per-wave builder, degenerate inputs (coincident centroids, flat boxes,
  signed zeros, NaN coordinates, a far triangle that makes one partition chain run through a
  whole range, more than 65535 coincident triangles) and a tessellated
  sphere;
* the wide trees (4- and 8-wide) and triangle records equal the host path's:
  closest and any hits of the same rays are bit-identical, the footprints
  and traversal stack depths agree;
* on the 1,310,720-triangle icosphere of the traversal benchmark the device
  tree equals the host builder's node for node (the host builder equals the
  oracle on it in the CPU suite).
"""
import sys
import time
import zlib

import numpy as np
import pytest

import oracle
import rgapi

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return rgapi.Context(0)


def soup(n, seed, spread=0.05):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (n, 1, 3))
    return (c + rng.normal(0, spread, (n, 3, 3))).astype(np.float32).reshape(n, 9)


def signed_zero(n, seed):
    rng = np.random.default_rng(seed)
    v = soup(n, seed).reshape(n, 3, 3)
    v[:, :, 0] = np.where(rng.random((n, 3)) < 0.5, np.float32(0.0), np.float32(-0.0))
    v[:, :, 2] = np.where(rng.random((n, 3)) < 0.3, np.float32(-0.0), v[:, :, 2])
    return v.reshape(n, 9)


def far_first(n, seed):
    """One triangle far away, then a cluster: the first split puts the far
    one right and everything else left, so its partition chain visits
    every left position of the range."""
    far = np.array([[50, 50, 50, 51, 50, 50, 50, 51, 50]], np.float32)
    return np.concatenate([far, soup(n - 1, seed)], axis=0)


def nan_coords(n, seed):
    v = soup(n, seed)
    v[7, 1] = np.nan
    v[1000:1004, 3:6] = np.nan
    v[2000, :] = np.nan
    v[2500:2900, :] = np.nan
    return v


def flat(n, seed):
    v = soup(n, seed).reshape(n, 3, 3)
    v[:, :, 2] = 0.25
    return v.reshape(n, 9)


def stacked(n):
    return np.repeat(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32), n, axis=0)


CASES = {
    "soup1": lambda: soup(1, 1),
    "soup4": lambda: soup(4, 2),
    "soup5": lambda: soup(5, 3),
    "soup255": lambda: soup(255, 13),
    "soup256": lambda: soup(256, 14),
    "soup257": lambda: soup(257, 15),
    "soup511": lambda: soup(511, 4),
    "soup512": lambda: soup(512, 5),
    "soup513": lambda: soup(513, 6),
    "soup1025": lambda: soup(1025, 7),
    "soup3000": lambda: soup(3000, 8),
    "soup40000": lambda: soup(40000, 9, spread=0.01),
    "ico5": lambda: rgapi.gen_icosphere(5),
    "signed_zero": lambda: signed_zero(3000, 10),
    "nan_coords": lambda: nan_coords(3000, 16),
    "far_first": lambda: far_first(6000, 11),
    "flat_z": lambda: flat(2500, 12),
    "coincident9": lambda: stacked(9),
    "coincident2000": lambda: stacked(2000),
    "coincident70000": lambda: stacked(70000),
}


def rays_for(v9, n, seed):
    lo = v9.reshape(-1, 3).min(axis=0)
    hi = v9.reshape(-1, 3).max(axis=0)
    rng = np.random.default_rng(seed)
    c, e = (lo + hi) / 2, np.maximum(hi - lo, 1e-3)
    o = c + (rng.random((n, 3)) - 0.5) * 1.4 * e
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return rgapi.make_rays(o, d)


@pytest.mark.parametrize("case", sorted(CASES))
def test_device_tree_equals_oracle(ctx, case):
    v9 = np.ascontiguousarray(CASES[case](), dtype=np.float32)
    say(f"[build] {case}: {len(v9)} triangles, device builder")
    sc = rgapi.Scene(ctx, v9)
    sc.build(builder=rgapi.BUILDER_DEVICE)
    n_p, p_p = sc.export2()
    n_o, p_o = oracle.bvh_build(v9)
    assert len(n_p) == len(n_o)
    assert n_p.tobytes() == n_o.tobytes()
    assert (p_p == p_o).all()


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("case", ["soup513", "soup3000", "ico5", "far_first", "signed_zero", "coincident2000"])
def test_device_wide_tree_traces_like_host(ctx, case, width):
    v9 = np.ascontiguousarray(CASES[case](), dtype=np.float32)
    say(f"[build] {case}: {len(v9)} triangles, width {width}, device vs host traversal")
    dev = rgapi.Scene(ctx, v9)
    dev.build(width=width, builder=rgapi.BUILDER_DEVICE)
    host = rgapi.Scene(ctx, v9)
    host.build(width=width, builder=rgapi.BUILDER_HOST)
    # the device build also keeps its binary tree and permutation
    assert dev.device_bytes() == host.device_bytes() + 32 * len(dev.export2()[0]) + 4 * len(v9)
    rays = rays_for(v9, 40000, zlib.crc32(case.encode()) & 0xFFFF)
    a, b = dev.trace_closest_host(rays), host.trace_closest_host(rays)
    assert a.tobytes() == b.tobytes()
    assert (dev.trace_any_host(rays) == host.trace_any_host(rays)).all()
    for pred in (rgapi.PRED_AREA, rgapi.PRED_ENV):
        assert (dev.trace_any_host(rays, pred) == host.trace_any_host(rays, pred)).all()


def test_render_scene_with_device_tree_equals_host_tree(ctx):
    """Occluder bits present at the build, and rewritten after it (shading
    set again on a device-built scene rewrites its triangle records on the
    device): renders equal the host-built scene's bit for bit."""
    import scenes
    sc = scenes.bsdf_frame(grid=2, stacks=40, slices=40)
    st = rgapi.default_settings(width=64, height=48, spp=4, max_depth=6)
    n9 = np.ascontiguousarray(sc["n9"], np.float32)
    obj = np.ascontiguousarray(sc["obj"], np.uint32)
    objs = np.ascontiguousarray(sc["objects"])
    outs = []
    for builder, reshade in ((rgapi.BUILDER_DEVICE, False), (rgapi.BUILDER_DEVICE, True), (rgapi.BUILDER_HOST, False)):
        rs = rgapi.RenderScene(ctx, sc)
        rs.build(builder=builder)
        if reshade:
            rgapi.check(rgapi.lib().rg_scene_set_shading(rs.h, rgapi._ptr(n9), rgapi._ptr(obj), rgapi._ptr(objs),
                                                         len(objs)))
        outs.append(rs.render(st))
    assert np.array_equal(outs[0]["hdr"], outs[2]["hdr"])
    assert np.array_equal(outs[1]["hdr"], outs[2]["hdr"])


def test_icosphere_1_3m_device_equals_host(ctx):
    v9 = rgapi.gen_icosphere(8)
    say(f"[build] icosphere {len(v9)} triangles")
    dev = rgapi.Scene(ctx, v9)
    dev.build(builder=rgapi.BUILDER_DEVICE)  # first call loads the kernels
    t0 = time.perf_counter()
    dev.build(builder=rgapi.BUILDER_DEVICE)
    t_dev = time.perf_counter() - t0
    host = rgapi.Scene(ctx, v9)
    t0 = time.perf_counter()
    host.build(builder=rgapi.BUILDER_HOST)
    t_host = time.perf_counter() - t0
    say(f"[build]   device {t_dev * 1e3:.1f} ms, host {t_host * 1e3:.1f} ms")
    n_d, p_d = dev.export2()
    n_h, p_h = host.export2()
    assert n_d.tobytes() == n_h.tobytes()
    assert (p_d == p_h).all()
    assert dev.device_bytes() == host.device_bytes() + 32 * len(n_d) + 4 * len(v9)
    rays = rays_for(v9, 1 << 18, 0xB1D)
    assert dev.trace_closest_host(rays).tobytes() == host.trace_closest_host(rays).tobytes()


def clusters(n, seed):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-5, 5, (12, 1, 3))
    which = rng.integers(0, 12, n)
    c = centers[which] + rng.normal(0, rng.uniform(0.01, 0.5, (n, 1, 1)), (n, 1, 3))
    return (c + rng.normal(0, 0.02, (n, 3, 3))).astype(np.float32).reshape(n, 9)


def sorted_line(n, seed):
    """Triangles along x in ascending order with a few far ones interleaved:
    every level's partition moves long runs of elements."""
    rng = np.random.default_rng(seed)
    x = np.sort(rng.uniform(0, 100, n))
    c = np.stack([x, rng.normal(0, 0.1, n), rng.normal(0, 0.1, n)], 1)[:, None, :]
    v = c + rng.normal(0, 0.05, (n, 3, 3))
    far = rng.random(n) < 0.01
    v[far] += np.array([0.0, 40.0, 0.0])
    return v.astype(np.float32).reshape(n, 9)


def mixed_scales(n, seed):
    """Small triangles on a sphere, a few huge ones spanning it, duplicates."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    v = d[:, None, :] + rng.normal(0, 0.01, (n, 3, 3))
    big = rng.random(n) < 0.002
    v[big] = rng.uniform(-3, 3, (int(big.sum()), 3, 3))
    dup = rng.random(n) < 0.05
    v[dup] = v[0]
    return v.astype(np.float32).reshape(n, 9)


@pytest.mark.parametrize("maker,n,seed", [
    (soup, 50000, 21), (soup, 150000, 22), (clusters, 80000, 23), (clusters, 30001, 24),
    (sorted_line, 60000, 25), (sorted_line, 9000, 26), (mixed_scales, 70000, 27), (mixed_scales, 4097, 28),
])
def test_device_tree_equals_host_builder_on_random_meshes(ctx, maker, n, seed):
    v9 = np.ascontiguousarray(maker(n, seed), dtype=np.float32)
    say(f"[build] {maker.__name__} n={n} seed={seed}")
    dev = rgapi.Scene(ctx, v9)
    dev.build(builder=rgapi.BUILDER_DEVICE)
    host = rgapi.Scene(ctx, v9)
    host.build(builder=rgapi.BUILDER_HOST)
    n_d, p_d = dev.export2()
    n_h, p_h = host.export2()
    assert n_d.tobytes() == n_h.tobytes()
    assert (p_d == p_h).all()

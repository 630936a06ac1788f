"""C-ABI library: loads, exports every declared symbol, host-only contexts,
product BVH == oracle BVH node for node.  CPU only (no device calls)."""
import os
import re

import numpy as np
import pytest

import oracle
import rgapi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(rg_\w+)\s*\(", txt, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = rgapi.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_only_context_refuses_device_work():
    ctx = rgapi.Context(-1)
    sc = rgapi.Scene(ctx, rgapi.gen_icosphere(1))
    sc.build()
    with pytest.raises(rgapi.RGError) as e:
        sc.trace_closest(0, 0, 0)
    assert e.value.code == -2
    with pytest.raises(rgapi.RGError):
        sc.trace_closest_host(rgapi.make_rays([[0, 0, 5]], [[0, 0, -1]]))


def test_bad_parameters_rejected():
    ctx = rgapi.Context(-1)
    sc = rgapi.Scene(ctx, rgapi.gen_icosphere(1))
    for kw in ({"bins": 1}, {"leaf": 0}, {"width": 3}):
        with pytest.raises(rgapi.RGError):
            sc.build(**kw)
    with pytest.raises(rgapi.RGError):
        sc.export2()  # not built yet


def soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (n, 1, 3))
    return (c + rng.normal(0, 0.05, (n, 3, 3))).astype(np.float32).reshape(n, 9)


def signed_zero(n, seed):
    rng = np.random.default_rng(seed)
    v = soup(n, seed).reshape(n, 3, 3)
    v[:, :, 0] = np.where(rng.random((n, 3)) < 0.5, np.float32(0.0), np.float32(-0.0))
    v[:, :, 2] = np.where(rng.random((n, 3)) < 0.3, np.float32(-0.0), v[:, :, 2])
    return v.reshape(n, 9)


def nan_coords(n, seed):
    v = soup(n, seed)
    v[7, 1] = np.nan
    v[100:104, 3:6] = np.nan
    v[200, :] = np.nan
    return v


CASES = {
    "ico4": lambda: rgapi.gen_icosphere(4),
    "soup1": lambda: soup(1, 1),
    "soup4": lambda: soup(4, 2),
    "soup5": lambda: soup(5, 3),
    "soup3000": lambda: soup(3000, 4),
    "coincident": lambda: np.repeat(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32), 7, axis=0),
    # coordinates of +0.0 and -0.0 mixed: box min/max ties keep the later
    # operand's sign, as fminf/fmaxf do (bvh.cpp:36-51)
    "signed_zero": lambda: signed_zero(400, 8),
    # NaN coordinates take no part in a box (fminf/fmaxf skip a NaN operand)
    "nan_coords": lambda: nan_coords(300, 9),
    "flat_z": lambda: np.concatenate([soup(500, 5)[:, :2], np.zeros((500, 1), np.float32),
                                      soup(500, 6)[:, 3:5], np.zeros((500, 1), np.float32),
                                      soup(500, 7)[:, 6:8], np.zeros((500, 1), np.float32)], axis=1),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_product_bvh_equals_oracle(case):
    v9 = np.ascontiguousarray(CASES[case](), dtype=np.float32)
    ctx = rgapi.Context(-1)
    sc = rgapi.Scene(ctx, v9)
    sc.build()
    n_p, p_p = sc.export2()
    n_o, p_o = oracle.bvh_build(v9)
    assert n_p.tobytes() == n_o.tobytes()
    assert (p_p == p_o).all()


def test_builder_choice_checked():
    ctx = rgapi.Context(-1)
    sc = rgapi.Scene(ctx, soup(10, 1))
    with pytest.raises(rgapi.RGError) as e:
        sc.build(builder=rgapi.BUILDER_DEVICE)  # a host-only context has no device
    assert e.value.code == -2
    with pytest.raises(rgapi.RGError):
        sc.build(builder=3)
    sc.build(builder=rgapi.BUILDER_HOST)
    n_p, p_p = sc.export2()
    n_o, p_o = oracle.bvh_build(sc.v9)
    assert n_p.tobytes() == n_o.tobytes()


def test_empty_scene_builds():
    ctx = rgapi.Context(-1)
    sc = rgapi.Scene(ctx, np.zeros((0, 9), np.float32))
    sc.build()
    nodes, perm = sc.export2()
    assert len(nodes) == 0


def test_icosphere_counts_and_watertight():
    for level in (0, 2, 3):
        v9 = rgapi.gen_icosphere(level)
        assert v9.shape == (20 * 4 ** level, 9)
        tri = v9.reshape(-1, 3, 3)
        edges = {}
        for t in tri:
            for a, b in ((0, 1), (1, 2), (2, 0)):
                key = tuple(sorted((tuple(t[a]), tuple(t[b]))))
                edges[key] = edges.get(key, 0) + 1
        assert set(edges.values()) == {2}
        r = np.linalg.norm(tri, axis=2)
        assert (r > 0.95 - 1e-6).all() and (r < 1.05 + 1e-6).all()
        # outward winding: face normal points away from the origin
        n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
        assert (np.einsum("ij,ij->i", n, tri.mean(axis=1)) > 0).all()

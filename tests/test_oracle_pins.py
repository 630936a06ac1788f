"""Pins of the CPU restatement (oracle/) against published vectors, closed
forms, brute force and structural invariants.  CPU only."""
from fractions import Fraction

import numpy as np
import pytest

import oracle


# --- Philox4x32-10: Random123 published known-answer vectors ---------------
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_kat(ctr, key, expect):
    assert [int(x) for x in oracle.philox(ctr, key)] == expect


# --- Möller–Trumbore (raycommon.cpp:151-235): hand-built closed forms -------
V1, V2, V3 = (0, 0, 0), (1, 0, 0), (0, 1, 0)
DOWN = (0, 0, -1)


def test_mt_known_hit_asymmetric():
    ok, t, u, v = oracle.mt_closest(V1, V2, V3, (0.25, 0.5, 2.0), DOWN)
    assert ok and t == 2.0 and u == 0.25 and v == 0.5   # u along v2-v1, v along v3-v1


def test_mt_edge_and_vertex_inclusive():
    assert oracle.mt_closest(V1, V2, V3, (0.5, 0.5, 1.0), DOWN)[0]      # u + v == 1
    assert oracle.mt_closest(V1, V2, V3, (0.0, 0.0, 1.0), DOWN)[0]      # vertex
    assert not oracle.mt_closest(V1, V2, V3, (0.6, 0.6, 1.0), DOWN)[0]  # outside
    assert not oracle.mt_closest(V1, V2, V3, (-0.1, 0.5, 1.0), DOWN)[0]


def test_mt_parallel_and_behind():
    assert not oracle.mt_closest(V1, V2, V3, (0.2, 0.2, 0.0), (1, 0, 0))[0]
    assert not oracle.mt_closest(V1, V2, V3, (0.2, 0.2, -1.0), DOWN)[0]   # t = -1
    assert not oracle.mt_any(V1, V2, V3, (0.2, 0.2, -1.0), DOWN, 1e30)[0]


def test_mt_back_face_hits_both_variants():
    ok, t, _, _ = oracle.mt_closest(V1, V2, V3, (0.2, 0.2, -1.0), (0, 0, 1))
    assert ok and t == 1.0
    assert oracle.mt_any(V1, V2, V3, (0.2, 0.2, -1.0), (0, 0, 1), 2.0)[0]


def test_mt_strict_reject_of_current_t():
    # closest: t >= current is rejected (first found wins exact ties)
    assert not oracle.mt_closest(V1, V2, V3, (0.2, 0.2, 1.0), DOWN, tcur=1.0)[0]
    assert oracle.mt_closest(V1, V2, V3, (0.2, 0.2, 1.0), DOWN, tcur=np.nextafter(np.float32(1), np.float32(2)))[0]
    # any: accepts only 1e-6 < t < maxt
    assert not oracle.mt_any(V1, V2, V3, (0.2, 0.2, 1.0), DOWN, 1.0)[0]
    assert oracle.mt_any(V1, V2, V3, (0.2, 0.2, 1.0), DOWN, 1.0001)[0]
    assert not oracle.mt_any(V1, V2, V3, (0.2, 0.2, 5e-7), DOWN, 1.0)[0]


def test_mt_determinant_thresholds_differ():
    # det = s^2: closest rejects |det| < 1e-8, any rejects |det| < 1e-6
    s = 3e-4  # det ~ 9e-8
    a, b, c = (0, 0, 0), (s, 0, 0), (0, s, 0)
    o = (s / 4, s / 4, 1.0)
    assert oracle.mt_closest(a, b, c, o, DOWN)[0]
    assert not oracle.mt_any(a, b, c, o, DOWN, 10.0)[0]
    s = 5e-5  # det ~ 2.5e-9
    a, b, c = (0, 0, 0), (s, 0, 0), (0, s, 0)
    assert not oracle.mt_closest(a, b, c, (s / 4, s / 4, 1.0), DOWN)[0]


# --- Halton LDS (raycommon.cpp:21-92) ---------------------------------------
def test_radical_inverse_exact_rationals():
    for base in (2, 3, 5, 7):
        for i in range(1, 200):
            digits, k, r = [], i, Fraction(0)
            while k:
                digits.append(k % base)
                k //= base
            for j, dgt in enumerate(digits):
                r += Fraction(dgt, base ** (j + 1))
            got = oracle.radical_inverse(i, base)
            assert abs(Fraction(got) - r) <= Fraction(4, 2 ** 24), (base, i)


@pytest.mark.parametrize("dim", [0, 1, 2, 5, 15])
def test_lds_stratified_per_pixel(dim):
    # samples 0..b^2-1 of one pixel fall one per interval of width 1/b^2
    # after undoing the per-pixel Cranley-Patterson shift
    base = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53][dim]
    n = base * base
    x, y = 17, 5
    vals = np.array([oracle.lds_sample(x, y, s, dim) for s in range(n - 1)], dtype=np.float64)
    shift = (oracle.hash_pixel(dim, 0, oracle.hash_pixel(x, y, 0)) >> 8) / 2.0 ** 24
    base_pts = np.mod(vals - shift, 1.0)
    # radical inverses of 1..b^2-1 are exactly {j / b^2 : j = 1..b^2-1}
    cells = np.rint(base_pts * n).astype(int) % n
    assert sorted(cells.tolist()) == list(range(1, n))


def test_lds_pixels_decorrelated():
    a = np.array([oracle.lds_sample(3, 4, 0, d) for d in range(16)])
    b = np.array([oracle.lds_sample(4, 4, 0, d) for d in range(16)])
    assert not np.allclose(a, b)
    assert ((a >= 0) & (a < 1)).all()


# --- BVH build (bvh.cpp:76-231): invariants ---------------------------------
def soup(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 1, 3))
    return (c + rng.normal(0, 0.05, (n, 3, 3))).astype(np.float32).reshape(n, 9)


def check_invariants(nodes, perm, v9):
    n = len(v9)
    seen = np.zeros(n, dtype=int)
    tri = v9.reshape(-1, 3, 3)
    lo_t, hi_t = tri.min(axis=1), tri.max(axis=1)
    stack = [0]
    while stack:
        k = stack.pop()
        nd = nodes[k]
        if nd["count"] > 0:
            idx = perm[nd["first_or_left"]: nd["first_or_left"] + nd["count"]]
            seen[idx] += 1
            assert (lo_t[idx] >= nd["bmin"]).all() and (hi_t[idx] <= nd["bmax"]).all()
            if nd["count"] > 4:  # forced leaf: no improving split existed
                pass
        else:
            l = nd["first_or_left"]
            assert l + 1 < len(nodes) and l > k
            for c in (l, l + 1):
                assert (nodes[c]["bmin"] >= nd["bmin"]).all() and (nodes[c]["bmax"] <= nd["bmax"]).all()
            stack += [l, l + 1]
    assert (seen == 1).all()
    assert sorted(perm.tolist()) == list(range(n))


@pytest.mark.parametrize("n,seed", [(1, 0), (4, 1), (5, 2), (100, 3), (3000, 4)])
def test_bvh_invariants(n, seed):
    v9 = soup(n, seed)
    nodes, perm = oracle.bvh_build(v9)
    check_invariants(nodes, perm, v9)


def test_bvh_coincident_centroids_forced_leaf():
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], dtype=np.float32)
    v9 = np.repeat(tri, 9, axis=0)
    nodes, perm = oracle.bvh_build(v9)
    assert len(nodes) == 1 and nodes[0]["count"] == 9


def test_bvh_empty():
    nodes, perm = oracle.bvh_build(np.zeros((0, 9), np.float32))
    assert len(nodes) == 0


# --- traversal == brute force (bvh.cpp:233-285, bvh.h:63-101) ---------------
def random_rays(n, seed, box=2.0):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, dtype=oracle.RAY_DTYPE)
    r["o"] = rng.uniform(-box, box, (n, 3))
    d = rng.normal(size=(n, 3))
    d[rng.random(n) < 0.05, 0] = 0.0  # axis-parallel components: 1/d -> 1e30 branch
    r["d"] = d
    r["tmax"] = 1e6
    return r


@pytest.mark.parametrize("n,seed", [(1, 0), (7, 1), (500, 2), (4000, 3)])
def test_closest_equals_brute_force(n, seed):
    v9 = soup(n, seed)
    nodes, perm = oracle.bvh_build(v9)
    rays = random_rays(3000, seed + 100)
    h, tie = oracle.trace_closest(nodes, perm, v9, rays, ties=True)
    hb, tieb = oracle.brute_closest(v9, rays, ties=True)
    same_t = h["t"] == hb["t"]
    assert same_t.all()
    nt = tieb == 0
    assert (h["prim"][nt] == hb["prim"][nt]).all()
    assert (h["u"][nt] == hb["u"][nt]).all() and (h["v"][nt] == hb["v"][nt]).all()
    if n >= 500:
        assert (h["prim"] != oracle.MISS).mean() > 0.02


def test_any_equals_brute_force():
    v9 = soup(2000, 9)
    nodes, perm = oracle.bvh_build(v9)
    rays = random_rays(3000, 10)
    rays["tmax"] = np.random.default_rng(11).uniform(0.01, 3.0, len(rays))
    assert (oracle.trace_any(nodes, perm, v9, rays) == oracle.brute_any(v9, rays)).all()


def test_closest_counts_are_positive():
    v9 = soup(1000, 5)
    nodes, perm = oracle.bvh_build(v9)
    rays = random_rays(100, 6)
    _, cnt = oracle.trace_closest(nodes, perm, v9, rays, counts=True)
    assert (cnt[:, 0] >= 1).all()


def test_closest_threads_match_single():
    v9 = soup(3000, 12)
    nodes, perm = oracle.bvh_build(v9)
    rays = random_rays(20000, 13)
    a = oracle.trace_closest(nodes, perm, v9, rays, threads=1)
    b = oracle.trace_closest(nodes, perm, v9, rays, threads=4)
    assert a.tobytes() == b.tobytes()

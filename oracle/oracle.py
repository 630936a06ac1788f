"""ctypes wrapper of liboracle.so — the CPU restatement of the reference's
hot path.  TEST INFRASTRUCTURE ONLY: imported by tests/, by
__graft_entry__.smoke() and by bench.py's CPU-baseline leg as the checker;
the product library never calls it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

RAY_DTYPE = np.dtype([("o", "<f4", (3,)), ("tmin", "<f4"), ("d", "<f4", (3,)), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
NODE2_DTYPE = np.dtype([("bmin", "<f4", (3,)), ("first_or_left", "<u4"), ("bmax", "<f4", (3,)),
                        ("count", "<u2"), ("axis", "<u2")])
MISS = 0xFFFFFFFF

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, i32, i64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_int64
        sig = {
            "or_philox4x32_10": (None, [vp, vp, vp]),
            "or_hash_pixel": (u32, [i32, i32, u32]),
            "or_radical_inverse": (ctypes.c_float, [u32, i32]),
            "or_lds_sample": (ctypes.c_float, [i32, i32, i32, i32]),
            "or_mt_closest": (i32, [vp, vp, vp, vp, vp, ctypes.c_float, vp, vp, vp]),
            "or_mt_any": (i32, [vp, vp, vp, vp, vp, ctypes.c_float, vp]),
            "or_bvh_build": (i64, [vp, u32, vp, vp]),
            "or_trace_closest": (None, [vp, i64, vp, vp, vp, u64, vp, vp, vp, i32]),
            "or_trace_any": (None, [vp, i64, vp, vp, vp, u64, vp, i32]),
            "or_brute_closest": (None, [vp, u32, vp, u64, vp, vp]),
            "or_brute_any": (None, [vp, u32, vp, u64, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def hash_pixel(x, y, salt):
    return lib().or_hash_pixel(x, y, salt)


def radical_inverse(i, base):
    return lib().or_radical_inverse(i, base)


def lds_sample(x, y, s, dim):
    return lib().or_lds_sample(x, y, s, dim)


def mt_closest(v1, v2, v3, o, d, tmin=0.0, tcur=1e6):
    a = [np.asarray(x, dtype=np.float32) for x in (v1, v2, v3, o, d)]
    t = np.array([tcur], dtype=np.float32)
    u = np.zeros(1, dtype=np.float32)
    v = np.zeros(1, dtype=np.float32)
    r = lib().or_mt_closest(*[_p(x) for x in a], tmin, _p(t), _p(u), _p(v))
    return bool(r), float(t[0]), float(u[0]), float(v[0])


def mt_any(v1, v2, v3, o, d, maxt):
    a = [np.asarray(x, dtype=np.float32) for x in (v1, v2, v3, o, d)]
    t = np.zeros(1, dtype=np.float32)
    r = lib().or_mt_any(*[_p(x) for x in a], maxt, _p(t))
    return bool(r), float(t[0])


def bvh_build(v9):
    v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(-1, 9)
    n = len(v9)
    nodes = np.zeros(max(2 * n, 1), dtype=NODE2_DTYPE)
    perm = np.zeros(max(n, 1), dtype=np.uint32)
    nn = lib().or_bvh_build(_p(v9), n, _p(nodes), _p(perm))
    return nodes[:nn].copy(), perm[:n].copy()


def trace_closest(nodes, perm, v9, rays, counts=False, ties=False, threads=1):
    v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    hits = np.zeros(len(rays), dtype=HIT_DTYPE)
    cnt = np.zeros((len(rays), 2), dtype=np.uint32) if counts else None
    tie = np.zeros(len(rays), dtype=np.uint8) if ties else None
    lib().or_trace_closest(_p(nodes), len(nodes), _p(v9), _p(perm), _p(rays), len(rays),
                           _p(hits), _p(cnt), _p(tie), threads)
    out = [hits]
    if counts:
        out.append(cnt)
    if ties:
        out.append(tie)
    return out[0] if len(out) == 1 else tuple(out)


def trace_any(nodes, perm, v9, rays, threads=1):
    v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    occ = np.zeros(len(rays), dtype=np.uint8)
    lib().or_trace_any(_p(nodes), len(nodes), _p(v9), _p(perm), _p(rays), len(rays), _p(occ), threads)
    return occ


def brute_closest(v9, rays, ties=False):
    v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    hits = np.zeros(len(rays), dtype=HIT_DTYPE)
    tie = np.zeros(len(rays), dtype=np.uint8)
    lib().or_brute_closest(_p(v9), len(v9), _p(rays), len(rays), _p(hits), _p(tie))
    return (hits, tie) if ties else hits


def brute_any(v9, rays):
    v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    occ = np.zeros(len(rays), dtype=np.uint8)
    lib().or_brute_any(_p(v9), len(v9), _p(rays), len(rays), _p(occ))
    return occ


def u01(x):
    """Top 24 bits of a uint32 as a float in [0, 1)."""
    return (np.asarray(x, dtype=np.uint32) >> 8).astype(np.float32) * np.float32(1.0 / 16777216.0)


def gen_rays_sphere(box, scale, key, first, n):
    """Restatement of the microbenchmark's ray generator (Philox4x32-10,
    counter = (index lo, index hi, block, 0)).  Directions use numpy's
    float32 sin/cos, so they agree with the device generator to a few ulp."""
    box = np.asarray(box, dtype=np.float32)
    c = (box[:3] + box[3:]) * np.float32(0.5)
    e = (box[3:] - box[:3]) * np.float32(scale)
    rays = np.zeros(n, dtype=RAY_DTYPE)
    for i in range(n):
        idx = first + i
        a = philox([idx & 0xFFFFFFFF, idx >> 32, 0, 0], [key, 0])
        b = philox([idx & 0xFFFFFFFF, idx >> 32, 1, 0], [key, 0])
        ua = u01(a)
        o = c + (ua[:3] - np.float32(0.5)) * e
        z = np.float32(1.0) - np.float32(2.0) * ua[3]
        r = np.sqrt(np.maximum(np.float32(0.0), np.float32(1.0) - z * z))
        phi = np.float32(6.28318530717958647692) * u01(b[0])
        rays[i]["o"] = o
        rays[i]["d"] = (r * np.cos(phi), r * np.sin(phi), z)
        rays[i]["tmax"] = np.inf
    return rays

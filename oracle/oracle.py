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
# the same sources under the reference's compiler flags (fast-math): for
# timing the CPU baseline only, never for checking results
FAST_LIB_PATH = os.path.join(HERE, "liboracle_fast.so")

RAY_DTYPE = np.dtype([("o", "<f4", (3,)), ("tmin", "<f4"), ("d", "<f4", (3,)), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
NODE2_DTYPE = np.dtype([("bmin", "<f4", (3,)), ("first_or_left", "<u4"), ("bmax", "<f4", (3,)),
                        ("count", "<u2"), ("axis", "<u2")])
MISS = 0xFFFFFFFF

_lib = None
_fast = None


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built (make -C oracle)")
    L = ctypes.CDLL(path)
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
    return L


def lib():
    global _lib
    if _lib is None:
        _lib = _load(LIB_PATH)
    return _lib


def fast_lib():
    """The fast-math build: CPU-baseline timing only."""
    global _fast
    if _fast is None:
        _fast = _load(FAST_LIB_PATH)
    return _fast


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


def trace_closest(nodes, perm, v9, rays, counts=False, ties=False, threads=1, fast_math=False):
    """fast_math: the fast-math build, for timing only."""
    v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    hits = np.zeros(len(rays), dtype=HIT_DTYPE)
    cnt = np.zeros((len(rays), 2), dtype=np.uint32) if counts else None
    tie = np.zeros(len(rays), dtype=np.uint8) if ties else None
    (fast_lib() if fast_math else lib()).or_trace_closest(_p(nodes), len(nodes), _p(v9), _p(perm), _p(rays), len(rays),
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


# --------------------------------------------------------------------------
# path tracer
# --------------------------------------------------------------------------
OBJECT_DTYPE = np.dtype([("color", "<f4", (3,)), ("glossy", "<f4"), ("roughness", "<f4"),
                         ("transparency", "<f4"), ("refraction", "<f4"), ("refraction_ratio", "<f4"),
                         ("metallic", "<f4"), ("anisotropy", "<f4"), ("aniso_rotation", "<f4"),
                         ("emission", "<f4"), ("interior_medium", "<i4"), ("visible", "<u4"),
                         ("chroma_dispersion", "<f4"), ("pad", "<u4")])
POINT_LIGHT_DTYPE = np.dtype([("pos", "<f4", (3,)), ("normal", "<f4", (3,)), ("color", "<f4", (3,)),
                              ("emission", "<f4"), ("spot_range", "<f4"), ("pad", "<u4")])
MEDIUM_DTYPE = np.dtype([("sigma_a", "<f4", (3,)), ("sigma_s", "<f4", (3,)), ("sigma_e", "<f4", (3,)),
                         ("g", "<f4"), ("density", "<f4"),
                         ("emission_mode", "<i4"), ("cone_origin", "<f4", (3,)), ("cone_axis", "<f4", (3,)),
                         ("cone_length", "<f4"), ("cone_radius", "<f4"), ("cone_inner", "<f4", (3,)),
                         ("cone_outer", "<f4", (3,)), ("cone_intensity", "<f4"), ("cone_peak_axial", "<f4"),
                         ("cone_peak_sharpness", "<f4"), ("cone_samples", "<i4"),
                         ("path", [("p", "<f4", (3,)), ("radius", "<f4"), ("t", "<f4")], (16,)),
                         ("n_path", "<u4"), ("path_inner", "<f4", (3,)), ("path_outer", "<f4", (3,)),
                         ("path_intensity", "<f4"), ("path_falloff_power", "<f4"), ("path_samples", "<i4"),
                         ("density_field", "<i4"), ("noise_frequency", "<f4"), ("noise_octaves", "<i4"),
                         ("noise_gain", "<f4"), ("noise_lacunarity", "<f4"), ("noise_amplitude", "<f4"),
                         ("noise_bias", "<f4"), ("noise_offset", "<f4", (3,)),
                         ("heat_haze", "<i4"), ("ior_amplitude", "<f4"), ("ior_frequency", "<f4"),
                         ("ior_octaves", "<i4"), ("ior_gain", "<f4"), ("ior_lacunarity", "<f4"),
                         ("ior_march_steps", "<i4"), ("ior_offset", "<f4", (3,)), ("ior_falloff", "<f4")])
VOLUME_LIGHT_DTYPE = np.dtype([("medium", "<u4"), ("pos", "<f4", (3,))])


class _Cam(ctypes.Structure):
    _fields_ = [("fov_deg", ctypes.c_float), ("exposure", ctypes.c_float), ("depth_of_field", ctypes.c_float),
                ("aperture", ctypes.c_float), ("aperture_blades", ctypes.c_int32),
                ("aperture_rotation_deg", ctypes.c_float), ("view_far", ctypes.c_float), ("pad", ctypes.c_float)]


class _Env(ctypes.Structure):
    _fields_ = [("w", ctypes.c_uint32), ("h", ctypes.c_uint32), ("rgb", ctypes.c_void_p),
                ("intensity", ctypes.c_float), ("rotation_deg", ctypes.c_float)]


class _Tex(ctypes.Structure):
    _fields_ = [("w", ctypes.c_uint32), ("h", ctypes.c_uint32), ("rgb", ctypes.c_void_p), ("srgb", ctypes.c_int32),
                ("pad", ctypes.c_uint32)]


class _Desc(ctypes.Structure):
    _fields_ = [("v9", ctypes.c_void_p), ("n9", ctypes.c_void_p), ("obj_of_tri", ctypes.c_void_p),
                ("ntris", ctypes.c_uint32), ("objs", ctypes.c_void_p), ("nobj", ctypes.c_uint32),
                ("media", ctypes.c_void_p), ("nmedia", ctypes.c_uint32), ("global_medium", ctypes.c_int32),
                ("env", _Env), ("cam", _Cam), ("tex", ctypes.c_void_p), ("ntex", ctypes.c_uint32),
                ("tex_of_obj", ctypes.c_void_p), ("tiling", ctypes.c_void_p), ("uv6", ctypes.c_void_p),
                ("plights", ctypes.c_void_p), ("nplights", ctypes.c_uint32),
                ("vlights", ctypes.c_void_p), ("nvlights", ctypes.c_uint32)]


class _Settings(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("spp", ctypes.c_uint32),
                ("max_depth", ctypes.c_uint32), ("antialias", ctypes.c_int32), ("firefly_clamp", ctypes.c_float),
                ("seed", ctypes.c_uint32), ("job", ctypes.c_uint32), ("sample_start", ctypes.c_uint32),
                ("color_sampling", ctypes.c_int32), ("back_color", ctypes.c_float * 4),
                ("shader", ctypes.c_int32), ("world_color", ctypes.c_float * 3), ("ao_samples", ctypes.c_uint32)]


BACK_COLOR = (1.0 * 0.2, 0.95 * 0.2, 0.9 * 0.2, 0.0)   # RendererSettings::backColor default (rayrenderer.h:152)
# RendererSettings::worldColor default (1, 0.95, 0.9) * 0.1 (rayrenderer.h:151), binary32 products
WORLD_COLOR = tuple(float(np.float32(c) * np.float32(0.1)) for c in (1.0, 0.95, 0.9))


def _settings(width, height, spp, max_depth, antialias, firefly_clamp, seed, job, sample_start,
              color_sampling=True, back_color=None, shader=5, world_color=None, ao_samples=0):
    s = _Settings(width, height, spp, max_depth, 1 if antialias else 0, firefly_clamp, seed & 0xFFFFFFFF, job,
                  sample_start)
    s.color_sampling = 1 if color_sampling else 0
    bc = BACK_COLOR if back_color is None else tuple(back_color) + (0.0,) * (4 - len(back_color))
    for i in range(4):
        s.back_color[i] = bc[i]
    s.shader = shader
    wc = WORLD_COLOR if world_color is None else tuple(world_color)
    for i in range(3):
        s.world_color[i] = wc[i]
    s.ao_samples = ao_samples
    return s


def _desc(scene, keep):
    arrs = {}
    for k, dt in (("v9", np.float32), ("n9", np.float32), ("obj", np.uint32)):
        arrs[k] = np.ascontiguousarray(scene[k], dtype=dt)
    objs = np.ascontiguousarray(scene["objects"], dtype=OBJECT_DTYPE)
    media = np.ascontiguousarray(scene.get("media", np.zeros(0, MEDIUM_DTYPE)), dtype=MEDIUM_DTYPE)
    keep += [arrs, objs, media]
    d = _Desc()
    d.v9 = arrs["v9"].ctypes.data
    d.n9 = arrs["n9"].ctypes.data
    d.obj_of_tri = arrs["obj"].ctypes.data
    d.ntris = len(arrs["obj"])
    d.objs = objs.ctypes.data
    d.nobj = len(objs)
    d.media = media.ctypes.data if len(media) else None
    d.nmedia = len(media)
    d.global_medium = int(scene.get("global_medium", -1))
    env = scene.get("env")
    if env is not None:
        rgb = np.ascontiguousarray(env["rgb"], dtype=np.float32)
        keep.append(rgb)
        d.env = _Env(rgb.shape[1], rgb.shape[0], rgb.ctypes.data, float(env.get("intensity", 1.0)),
                     float(env.get("rotation_deg", 0.0)))
    texs = scene.get("textures")
    if texs:
        rgbs = [np.ascontiguousarray(t["rgb"], np.float32) for t in texs]
        tab = (_Tex * len(texs))(*[_Tex(r.shape[1], r.shape[0], r.ctypes.data, 1 if t.get("srgb", True) else 0, 0)
                                   for r, t in zip(rgbs, texs)])
        too = np.ascontiguousarray(scene["tex_of_obj"], np.int32)
        til = np.ascontiguousarray(scene["tiling"], np.float32)
        uv6 = np.ascontiguousarray(scene["uv6"], np.float32)
        keep += [rgbs, tab, too, til, uv6]
        d.tex = ctypes.addressof(tab)
        d.ntex = len(texs)
        d.tex_of_obj = too.ctypes.data
        d.tiling = til.ctypes.data
        d.uv6 = uv6.ctypes.data
    vl = scene.get("volume_lights")
    if vl is not None and len(vl):
        vla = np.ascontiguousarray(vl, dtype=VOLUME_LIGHT_DTYPE)
        keep.append(vla)
        d.vlights = vla.ctypes.data
        d.nvlights = len(vla)
    pl = scene.get("point_lights")
    if pl is not None and len(pl):
        pla = np.ascontiguousarray(pl, dtype=POINT_LIGHT_DTYPE)
        keep.append(pla)
        d.plights = pla.ctypes.data
        d.nplights = len(pla)
    cam = scene["camera"]
    d.cam = _Cam(cam.get("fov_deg", 75.0), cam.get("exposure", 1.0), cam.get("depth_of_field", 0.0),
                 cam.get("aperture", 1.8), cam.get("aperture_blades", 0), cam.get("aperture_rotation_deg", 0.0),
                 cam.get("view_far", 50.0), 0.0)
    return d


def render(scene, width, height, spp, max_depth=5, mode=0, threads=8, antialias=True, firefly_clamp=10.0,
           seed=0xC0DE0000, job=0, sample_start=0, active_tile=None, color_sampling=True, back_color=None,
           shader=5, world_color=None, ao_samples=0):
    """Returns (sum, sumsq), each (H, W, 3) float32, over the given samples.
    shader selects the shading system (0-3, 5); ao_samples is calcAO's
    sample count (settings.samples, the whole render's spp; 0 = spp)."""
    L = lib()
    if not hasattr(L, "_render_sig"):
        L.or_render.restype = ctypes.c_int
        L.or_render.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_void_p]
        L.or_render_one.restype = None
        L.or_render_one.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_void_p]
        L._render_sig = True
    keep = []
    d = _desc(scene, keep)
    s = _settings(width, height, spp, max_depth, antialias, firefly_clamp, seed, job, sample_start, color_sampling,
                  back_color, shader, world_color, ao_samples)
    sm = np.zeros((height, width, 3), np.float32)
    sq = np.zeros((height, width, 3), np.float32)
    at = None if active_tile is None else np.ascontiguousarray(active_tile, dtype=np.uint8)
    L.or_render(ctypes.byref(d), ctypes.byref(s), mode, threads, _p(at), _p(sm), _p(sq))
    return sm, sq


def aov(scene, width, height, back_color=None):
    """Primary-hit guide buffers: (normal_rgba, albedo_rgba, depth)."""
    L = lib()
    L.or_aov.restype = None
    L.or_aov.argtypes = [ctypes.c_void_p] * 5
    keep = []
    d = _desc(scene, keep)
    s = _settings(width, height, 1, 1, True, 10.0, 0, 0, 0, back_color=back_color)
    nrm = np.zeros((height, width, 4), np.float32)
    alb = np.zeros((height, width, 4), np.float32)
    dep = np.zeros((height, width), np.float32)
    L.or_aov(ctypes.byref(d), ctypes.byref(s), _p(nrm), _p(alb), _p(dep))
    return nrm, alb, dep


def denoise(noisy_rgba, nrm, alb, dep, levels=5, sigma_c=0.4, sigma_n=128.0, sigma_d=0.1, intensity=1.0):
    L = lib()
    L.or_denoise.restype = None
    f = ctypes.c_float
    L.or_denoise.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int, f, f, f, f,
                                                                                      ctypes.c_void_p]
    h, w = noisy_rgba.shape[:2]
    out = np.zeros((h, w, 4), np.float32)
    arrs = [np.ascontiguousarray(a, np.float32) for a in (noisy_rgba, nrm, alb, dep)]
    L.or_denoise(w, h, *[_p(a) for a in arrs], levels, sigma_c, sigma_n, sigma_d, intensity, _p(out))
    return out


def render_adaptive(scene, width, height, spp, max_depth=5, base=4, threshold=0.02, mode=1, threads=8,
                    antialias=True, firefly_clamp=10.0, seed=0xC0DE0000, job=0, color_sampling=True,
                    back_color=None, shader=5, world_color=None):
    """renderAdaptive restated: returns (sum, sumsq, nspp per 32x32 tile)."""
    L = lib()
    L.or_render_adaptive.restype = ctypes.c_int
    L.or_render_adaptive.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                     ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    keep = []
    d = _desc(scene, keep)
    s = _settings(width, height, spp, max_depth, antialias, firefly_clamp, seed, job, 0, color_sampling, back_color,
                  shader, world_color, spp)
    sm = np.zeros((height, width, 3), np.float32)
    sq = np.zeros((height, width, 3), np.float32)
    nt = np.zeros(((width + 31) // 32) * ((height + 31) // 32), np.uint32)
    L.or_render_adaptive(ctypes.byref(d), ctypes.byref(s), mode, threads, base, threshold, _p(sm), _p(sq), _p(nt))
    return sm, sq, nt


def bloom(hdr, threshold=1.0, strength=1.0, radius=0.03, size_aspect=0.15, curve=1.0):
    """or_bloom on a copy of hdr (H x W x 4 float32); returns the bloomed image."""
    L = lib()
    f = ctypes.c_float
    L.or_bloom.restype = None
    L.or_bloom.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, f, f, f, f, f]
    out = np.ascontiguousarray(hdr, dtype=np.float32).copy()
    h, w = out.shape[:2]
    L.or_bloom(_p(out), w, h, threshold, strength, radius, size_aspect, curve)
    return out


def tonemap(hdr):
    """or_tonemap: H x W x 4 float32 HDR -> LDR in [0, 1]."""
    L = lib()
    L.or_tonemap.restype = None
    L.or_tonemap.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    src = np.ascontiguousarray(hdr, dtype=np.float32)
    h, w = src.shape[:2]
    out = np.empty_like(src)
    L.or_tonemap(_p(src), w, h, _p(out))
    return out


def phase_sample(g, wo, u2):
    """or_phase_sample: directions sampled about wo for the (u1, u2) rows of u2."""
    L = lib()
    L.or_phase_sample.restype = None
    L.or_phase_sample.argtypes = [ctypes.c_float, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    wo = np.ascontiguousarray(wo, np.float32)
    u2 = np.ascontiguousarray(u2, np.float32).reshape(-1, 2)
    out = np.zeros((len(u2), 3), np.float32)
    L.or_phase_sample(g, _p(wo), len(u2), _p(u2), _p(out))
    return out


def phase_pdf(g, wo, wi):
    L = lib()
    L.or_phase_pdf.restype = None
    L.or_phase_pdf.argtypes = [ctypes.c_float, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    wo = np.ascontiguousarray(wo, np.float32)
    wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    out = np.zeros(len(wi), np.float32)
    L.or_phase_pdf(g, _p(wo), len(wi), _p(wi), _p(out))
    return out


def env_probe(scene, u2, qdirs):
    """or_env_probe: (dirs, pdfs, radiance) of sampled directions and
    (pdf, radiance) of the query directions of scene's environment."""
    L = lib()
    L.or_env_probe.restype = None
    L.or_env_probe.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    keep = []
    d = _desc(scene, keep)
    u2 = np.ascontiguousarray(u2, np.float32).reshape(-1, 2)
    q = np.ascontiguousarray(qdirs, np.float32).reshape(-1, 3)
    dirs = np.zeros((len(u2), 3), np.float32)
    pdfs = np.zeros(len(u2), np.float32)
    rad = np.zeros((len(u2), 3), np.float32)
    qp = np.zeros(len(q), np.float32)
    qr = np.zeros((len(q), 3), np.float32)
    L.or_env_probe(ctypes.byref(d), len(u2), _p(u2), _p(dirs), _p(pdfs), _p(rad), len(q), _p(q), _p(qp), _p(qr))
    return dirs, pdfs, rad, qp, qr


def tile_noise(sum_rgb, sumsq_rgb, tx, ty, n):
    """or_tile_noise of tile (tx, ty) of (H, W, 3) sums over n samples."""
    L = lib()
    L.or_tile_noise.restype = ctypes.c_float
    L.or_tile_noise.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int]
    a = np.ascontiguousarray(sum_rgb, np.float32)
    b = np.ascontiguousarray(sumsq_rgb, np.float32)
    h, w = a.shape[:2]
    return float(L.or_tile_noise(_p(a), _p(b), w, h, tx, ty, n))


def _medium_arg(med):
    m = np.ascontiguousarray(np.atleast_1d(med)[:1], dtype=MEDIUM_DTYPE)
    return m, _p(m)


def medium_eval(med, op, pts):
    """or_medium_eval: op "density" (densityAt), "emission" (emissionAt,
    RGB) or "ior" (iorAt) of one medium record at render-space points."""
    L = lib()
    L.or_medium_eval.restype = None
    L.or_medium_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    m, mp = _medium_arg(med)
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    code = {"density": 0, "emission": 1, "ior": 2}[op]
    out = np.zeros((len(pts), 3) if code == 1 else len(pts), np.float32)
    L.or_medium_eval(mp, code, len(pts), _p(pts), _p(out))
    return out


def delta_track(med, o, d, max_t, n, seed=1):
    """or_delta_track_probe: the collision distance of n independent
    delta-tracking walks along o + t d (-1 when none before max_t)."""
    L = lib()
    L.or_delta_track_probe.restype = None
    L.or_delta_track_probe.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                       ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p]
    m, mp = _medium_arg(med)
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    out = np.zeros(n, np.float32)
    L.or_delta_track_probe(mp, _p(o), _p(d), max_t, n, seed, _p(out))
    return out


def emission_integral(med, o, d, max_t, n, seed=1):
    """or_emission_integral_probe: n independent estimates (RGB) of the
    emission integral over [0, max_t] of o + t d."""
    L = lib()
    L.or_emission_integral_probe.restype = None
    L.or_emission_integral_probe.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                             ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p]
    m, mp = _medium_arg(med)
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    out = np.zeros((n, 3), np.float32)
    L.or_emission_integral_probe(mp, _p(o), _p(d), max_t, n, seed, _p(out))
    return out


def bend(med, o, d, length):
    """or_bend_probe: (origin, direction) of the ray bent over length."""
    L = lib()
    L.or_bend_probe.restype = None
    L.or_bend_probe.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p]
    m, mp = _medium_arg(med)
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    out = np.zeros(6, np.float32)
    L.or_bend_probe(mp, _p(o), _p(d), length, _p(out))
    return out[:3], out[3:]

"""ctypes binding of librg.so (the C ABI declared in include/rg.h).

The library is loaded from this directory; a missing or unloadable library
raises RGError — there is no fallback path.  Device buffers are passed as
integer addresses (e.g. ``torch.Tensor.data_ptr()``); host buffers as numpy
arrays.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librg.so")

RAY_DTYPE = np.dtype([("o", "<f4", (3,)), ("tmin", "<f4"), ("d", "<f4", (3,)), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
NODE2_DTYPE = np.dtype([("bmin", "<f4", (3,)), ("first_or_left", "<u4"), ("bmax", "<f4", (3,)),
                        ("count", "<u2"), ("axis", "<u2")])
MISS = 0xFFFFFFFF

PRED_ALL, PRED_AREA, PRED_ENV = 0, 1, 2


class RGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rg error {code}: {msg}")
        self.code = code


class BvhParams(ctypes.Structure):
    _fields_ = [("bins", ctypes.c_uint32), ("leaf", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("builder", ctypes.c_uint32)]


BUILDER_AUTO, BUILDER_HOST, BUILDER_DEVICE = 0, 1, 2


_lib = None

# void (*)(float progress, void *user)
PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_float, ctypes.c_void_p)


def lib():
    """Load librg.so once; raise RGError when it is not built."""
    global _lib
    if _lib is None:
        # torch ships its own libamdhip64 (same SONAME); when torch is present
        # it is loaded first so that librg binds to that runtime and the
        # process holds a single HIP runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RGError(-1, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, i32, fp = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float
        pp = ctypes.POINTER(ctypes.c_void_p)
        sig = {
            "rg_create": (i32, [i32, pp]),
            "rg_destroy": (None, [vp]),
            "rg_last_error": (ctypes.c_char_p, []),
            "rg_set_stream": (i32, [vp, vp]),
            "rg_set_progress": (i32, [vp, vp, vp]),
            "rg_synchronize": (i32, [vp]),
            "rg_cancel": (None, [vp]),
            "rg_scene_create": (i32, [vp, vp, u32, pp]),
            "rg_scene_destroy": (None, [vp]),
            "rg_bvh_build": (i32, [vp, vp, ctypes.POINTER(BvhParams)]),
            "rg_bvh_export2": (i32, [vp, vp, u64, vp, ctypes.POINTER(u64)]),
            "rg_scene_device_bytes": (i32, [vp, ctypes.POINTER(u64)]),
            "rg_scene_bounds": (i32, [vp, vp]),
            "rg_trace_closest": (i32, [vp, vp, vp, vp, u64]),
            "rg_trace_any": (i32, [vp, vp, vp, vp, u64, i32]),
            "rg_trace_closest_host": (i32, [vp, vp, vp, vp, u64]),
            "rg_trace_any_host": (i32, [vp, vp, vp, vp, u64, i32]),
            "rg_gen_icosphere": (i32, [u32, fp, u32, vp]),
            "rg_gen_rays_sphere": (i32, [vp, vp, fp, u32, u64, vp, u64]),
            "rg_object_default": (None, [vp]),
            "rg_settings_default": (None, [vp]),
            "rg_scene_set_shading": (i32, [vp, vp, vp, vp, u32]),
            "rg_scene_set_media": (i32, [vp, vp, u32, i32]),
            "rg_scene_set_env": (i32, [vp, vp]),
            "rg_scene_set_point_lights": (i32, [vp, vp, u32]),
            "rg_scene_set_volume_lights": (i32, [vp, vp, u32]),
            "rg_scene_export_volume_lights": (i32, [vp, vp, vp]),
            "rg_medium_default": (None, [vp]),
            "rg_scene_export_point_lights": (i32, [vp, vp, vp]),
            "rg_scene_set_camera": (i32, [vp, vp]),
            "rg_render": (i32, [vp, vp, vp, vp, vp, vp, vp]),
            "rg_render_stats": (i32, [vp, vp]),
            "rg_host_alloc": (i32, [ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
            "rg_host_free": (None, [vp]),
            "rg_scene_load_json": (i32, [vp, ctypes.c_char_p, pp]),
            "rg_scene_warnings": (i32, [vp, ctypes.c_char_p, u64]),
            "rg_scene_get_camera": (i32, [vp, vp]),
            "rg_scene_export": (i32, [vp, vp, vp, vp, vp, vp, vp]),
            "rg_scene_export_media": (i32, [vp, vp, vp, vp, vp, vp]),
            "rg_render_batch": (i32, [vp, vp, ctypes.c_uint32, vp, ctypes.c_uint32, vp, vp, vp, vp]),
            "rg_autotune": (i32, [vp, vp, ctypes.c_uint32, vp, ctypes.c_uint32, vp]),
            "rg_server_create": (i32, [vp, vp, ctypes.c_uint32, u64, vp, pp]),
            "rg_server_submit": (i32, [vp, vp, vp, vp, vp]),
            "rg_server_wait": (i32, [vp, u64, ctypes.c_double, vp]),
            "rg_server_stats": (i32, [vp, vp]),
            "rg_server_get_tune": (i32, [vp, vp]),
            "rg_server_set_queue_limit": (i32, [vp, u64]),
            "rg_server_queue": (i32, [vp, vp]),
            "rg_server_destroy": (i32, [vp]),
            "rg_bloom_default": (None, [vp]),
            "rg_scene_set_textures": (i32, [vp, vp, u32, vp, vp, vp]),
            "rg_scene_export_textures": (i32, [vp, vp, vp, vp, vp, vp, vp]),
            "rg_postprocess": (i32, [vp, u32, u32, vp, vp, vp]),
            "rg_postprocess_device": (i32, [vp, u32, u32, vp, vp, vp]),
            "rg_postprocess_device_noalloc": (i32, [vp, u32, u32, vp, vp, vp]),
            "rg_image_write": (i32, [ctypes.c_char_p, u32, u32, vp, i32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RGError(rc, lib().rg_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """rg_ctx: device >= 0 for a HIP device, -1 for a host-only context."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().rg_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def set_stream(self, stream_ptr):
        check(lib().rg_set_stream(self.h, ctypes.c_void_p(stream_ptr or None)))

    def synchronize(self):
        check(lib().rg_synchronize(self.h))

    def cancel(self):
        lib().rg_cancel(self.h)

    def set_progress(self, fn):
        """rg_set_progress: fn(progress: float) on the rendering thread, or
        None to remove it."""
        if fn is None:
            self._progress = None
            check(lib().rg_set_progress(self.h, None, None))
            return
        self._progress = PROGRESS_FN(lambda p, user: fn(float(p)))
        check(lib().rg_set_progress(self.h, ctypes.cast(self._progress, ctypes.c_void_p), None))

    def close(self):
        if self.h:
            lib().rg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def gen_rays_sphere(self, box, scale, key, first, d_rays, n):
        b = np.ascontiguousarray(box, dtype=np.float32)
        check(lib().rg_gen_rays_sphere(self.h, _ptr(b), scale, key, first, ctypes.c_void_p(d_rays), n))


class Scene:
    """rg_scene: triangle soup (n, 9) float32 in render space."""

    def __init__(self, ctx, v9):
        self.ctx = ctx
        self.v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(-1, 9)
        h = ctypes.c_void_p()
        check(lib().rg_scene_create(ctx.h, _ptr(self.v9), len(self.v9), ctypes.byref(h)))
        self.h = h

    @property
    def n_tris(self):
        return len(self.v9)

    def build(self, bins=16, leaf=4, width=4, builder=BUILDER_AUTO):
        p = BvhParams(bins, leaf, width, builder)
        check(lib().rg_bvh_build(self.ctx.h, self.h, ctypes.byref(p)))

    def export2(self):
        n = ctypes.c_uint64()
        check(lib().rg_bvh_export2(self.h, None, 0, None, ctypes.byref(n)))
        nodes = np.zeros(n.value, dtype=NODE2_DTYPE)
        perm = np.zeros(self.n_tris, dtype=np.uint32)
        check(lib().rg_bvh_export2(self.h, _ptr(nodes), n.value, _ptr(perm), ctypes.byref(n)))
        return nodes, perm

    def device_bytes(self):
        b = ctypes.c_uint64()
        check(lib().rg_scene_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def bounds(self):
        b = np.zeros(6, dtype=np.float32)
        check(lib().rg_scene_bounds(self.h, _ptr(b)))
        return b

    def trace_closest(self, d_rays, d_hits, n):
        check(lib().rg_trace_closest(self.ctx.h, self.h, ctypes.c_void_p(d_rays), ctypes.c_void_p(d_hits), n))

    def trace_any(self, d_rays, d_occ, n, pred=PRED_ALL):
        check(lib().rg_trace_any(self.ctx.h, self.h, ctypes.c_void_p(d_rays), ctypes.c_void_p(d_occ), n, pred))

    def trace_closest_host(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        check(lib().rg_trace_closest_host(self.ctx.h, self.h, _ptr(rays), _ptr(hits), len(rays)))
        return hits

    def trace_any_host(self, rays, pred=PRED_ALL):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        occ = np.zeros(len(rays), dtype=np.uint8)
        check(lib().rg_trace_any_host(self.ctx.h, self.h, _ptr(rays), _ptr(occ), len(rays), pred))
        return occ

    def render(self, settings, hdr_only=False, sums=True, hdr_out=None):
        """rg_render of this scene (shading set and BVH built) with host
        outputs; hdr_only skips the sums and tile sample counts, sums=False
        the sums only; hdr_out (H, W, 4) float32, e.g. from host_zeros,
        receives the HDR image instead of a new array."""
        W, H = settings.width, settings.height
        if hdr_out is not None:
            assert hdr_out.shape == (H, W, 4) and hdr_out.dtype == np.float32 and hdr_out.flags.c_contiguous
            hdr = hdr_out
        else:
            hdr = np.zeros((H, W, 4), np.float32)
        sm = sq = nt = None
        if not hdr_only:
            if sums:
                sm = np.zeros((H, W, 3), np.float32)
                sq = np.zeros((H, W, 3), np.float32)
            nt = np.zeros(((W + 31) // 32) * ((H + 31) // 32), np.uint32)
        check(lib().rg_render(self.ctx.h, self.h, ctypes.byref(settings), _ptr(hdr), _ptr(sm), _ptr(sq), _ptr(nt)))
        stats = np.zeros(3, np.uint64)
        check(lib().rg_render_stats(self.ctx.h, _ptr(stats)))
        out = {"hdr": hdr, "rays": int(stats[0]), "shadow_rays": int(stats[1]), "samples": int(stats[2])}
        if not hdr_only:
            out["nspp"] = nt
            if sums:
                out.update({"sum": sm, "sumsq": sq})
        return out

    def close(self):
        if self.h:
            lib().rg_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gen_icosphere(level, amp=0.05, seed=0x1C05EED):
    """Displaced icosphere, (20*4**level, 9) float32."""
    n = 20 * 4 ** level
    out = np.zeros((n, 9), dtype=np.float32)
    check(lib().rg_gen_icosphere(level, amp, seed, _ptr(out)))
    return out


def make_rays(o, d, tmin=0.0, tmax=np.inf):
    """Pack origins/directions (n,3) into the 32-byte ray records."""
    o = np.asarray(o, dtype=np.float32).reshape(-1, 3)
    d = np.asarray(d, dtype=np.float32).reshape(-1, 3)
    r = np.zeros(len(o), dtype=RAY_DTYPE)
    r["o"] = o
    r["d"] = d
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


# --------------------------------------------------------------------------
# path tracing
# --------------------------------------------------------------------------
class Camera(ctypes.Structure):
    _fields_ = [("fov_deg", ctypes.c_float), ("exposure", ctypes.c_float), ("depth_of_field", ctypes.c_float),
                ("aperture", ctypes.c_float), ("aperture_blades", ctypes.c_int32),
                ("aperture_rotation_deg", ctypes.c_float), ("view_far", ctypes.c_float), ("pad", ctypes.c_float)]


class Env(ctypes.Structure):
    _fields_ = [("w", ctypes.c_uint32), ("h", ctypes.c_uint32), ("rgb", ctypes.c_void_p),
                ("intensity", ctypes.c_float), ("rotation_deg", ctypes.c_float)]


class Settings(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("spp", ctypes.c_uint32),
                ("max_depth", ctypes.c_uint32), ("antialias", ctypes.c_int32), ("firefly_clamp", ctypes.c_float),
                ("seed", ctypes.c_uint32), ("job", ctypes.c_uint32), ("adaptive", ctypes.c_int32),
                ("adaptive_base", ctypes.c_uint32), ("adaptive_threshold", ctypes.c_float),
                ("denoise", ctypes.c_int32), ("denoise_levels", ctypes.c_uint32), ("sigma_color", ctypes.c_float),
                ("sigma_normal", ctypes.c_float), ("sigma_depth", ctypes.c_float),
                ("denoise_intensity", ctypes.c_float), ("samples_in_flight", ctypes.c_uint32),
                ("color_sampling", ctypes.c_int32), ("cull_back_face", ctypes.c_int32),
                ("back_color", ctypes.c_float * 4), ("shader", ctypes.c_int32), ("world_color", ctypes.c_float * 3)]


def default_settings(**kw):
    s = Settings()
    lib().rg_settings_default(ctypes.byref(s))
    for k, v in kw.items():
        if not hasattr(s, k):
            raise KeyError(k)
        setattr(s, k, v)
    return s


def make_camera(d):
    return Camera(d.get("fov_deg", 75.0), d.get("exposure", 1.0), d.get("depth_of_field", 0.0),
                  d.get("aperture", 1.8), d.get("aperture_blades", 0), d.get("aperture_rotation_deg", 0.0),
                  d.get("view_far", 50.0), 0.0)


class Texture(ctypes.Structure):
    """rg_texture: RGB floats as stored (LDR byte / 255), srgb flag."""
    _fields_ = [("w", ctypes.c_uint32), ("h", ctypes.c_uint32), ("rgb", ctypes.c_void_p), ("srgb", ctypes.c_int32),
                ("pad", ctypes.c_uint32)]


class _HostBlock:
    """Owner of one rg_host_alloc block; freed when the arrays over it are."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        check(lib().rg_host_alloc(nbytes, ctypes.byref(p)))
        self.p = p.value

    def __del__(self):
        if self.p:
            lib().rg_host_free(ctypes.c_void_p(self.p))
            self.p = None


def host_zeros(shape, dtype=np.float32):
    """A zeroed numpy array in page-locked host memory (rg_host_alloc):
    rg_render's copies into it go by DMA, without staging."""
    dtype = np.dtype(dtype)
    nbytes = int(np.prod(shape)) * dtype.itemsize
    if nbytes == 0:
        return np.zeros(shape, dtype)
    blk = _HostBlock(nbytes)
    buf = (ctypes.c_byte * nbytes).from_address(blk.p)
    buf._owner = blk
    a = np.frombuffer(buf, dtype=dtype).reshape(shape)
    a[...] = 0
    return a


class RenderScene(Scene):
    """A scene dict (see scenes.py) made resident on the device: geometry,
    BVH, shading, media, environment and camera."""

    def __init__(self, ctx, sc, width=4):
        super().__init__(ctx, sc["v9"])
        self.src = sc
        self._keep = []
        n9 = np.ascontiguousarray(sc["n9"], np.float32)
        obj = np.ascontiguousarray(sc["obj"], np.uint32)
        objs = np.ascontiguousarray(sc["objects"])
        assert objs.dtype.itemsize == 64
        media = np.ascontiguousarray(sc.get("media", np.zeros(0, objs.dtype)))
        self._keep += [n9, obj, objs, media]
        check(lib().rg_scene_set_media(self.h, _ptr(media) if len(media) else None, len(media),
                                       int(sc.get("global_medium", -1))))
        check(lib().rg_scene_set_shading(self.h, _ptr(n9), _ptr(obj), _ptr(objs), len(objs)))
        vl = sc.get("volume_lights")
        if vl is not None and len(vl):
            from scenes import VOLUME_LIGHT_DTYPE
            vla = np.ascontiguousarray(vl, dtype=VOLUME_LIGHT_DTYPE)
            self._keep.append(vla)
            check(lib().rg_scene_set_volume_lights(self.h, _ptr(vla), len(vla)))
        env = sc.get("env")
        if env is not None:
            rgb = np.ascontiguousarray(env["rgb"], np.float32)
            self._keep.append(rgb)
            e = Env(rgb.shape[1], rgb.shape[0], rgb.ctypes.data, float(env.get("intensity", 1.0)),
                    float(env.get("rotation_deg", 0.0)))
            check(lib().rg_scene_set_env(self.h, ctypes.byref(e)))
        texs = sc.get("textures")
        if texs:
            rgbs = [np.ascontiguousarray(t["rgb"], np.float32) for t in texs]
            tab = (Texture * len(texs))(*[Texture(r.shape[1], r.shape[0], r.ctypes.data,
                                                  1 if t.get("srgb", True) else 0, 0) for r, t in zip(rgbs, texs)])
            too = np.ascontiguousarray(sc["tex_of_obj"], np.int32)
            til = np.ascontiguousarray(sc["tiling"], np.float32)
            uv6 = np.ascontiguousarray(sc["uv6"], np.float32)
            self._keep += [rgbs, tab, too, til, uv6]
            check(lib().rg_scene_set_textures(self.h, tab, len(texs), _ptr(too), _ptr(til), _ptr(uv6)))
        pl = sc.get("point_lights")
        if pl is not None and len(pl):
            from scenes import POINT_LIGHT_DTYPE
            pla = np.ascontiguousarray(pl, dtype=POINT_LIGHT_DTYPE)
            self._keep.append(pla)
            check(lib().rg_scene_set_point_lights(self.h, _ptr(pla), len(pla)))
        cam = make_camera(sc["camera"])
        check(lib().rg_scene_set_camera(self.h, ctypes.byref(cam)))
        self.build(width=width)

def load_json(ctx, path):
    """rg_scene_load_json: returns (rg_scene handle wrapper, scene dict with
    the flattened arrays, camera and warnings)."""
    h = ctypes.c_void_p()
    check(lib().rg_scene_load_json(ctx.h, path.encode(), ctypes.byref(h)))
    nt, no = ctypes.c_uint32(), ctypes.c_uint32()
    check(lib().rg_scene_export(h, None, None, None, None, ctypes.byref(nt), ctypes.byref(no)))
    from scenes import OBJECT_DTYPE
    v9 = np.zeros((nt.value, 9), np.float32)
    n9 = np.zeros((nt.value, 9), np.float32)
    obj = np.zeros(nt.value, np.uint32)
    objs = np.zeros(no.value, OBJECT_DTYPE)
    check(lib().rg_scene_export(h, _ptr(v9), _ptr(n9), _ptr(obj), _ptr(objs), ctypes.byref(nt), ctypes.byref(no)))
    cam = Camera()
    check(lib().rg_scene_get_camera(h, ctypes.byref(cam)))
    buf = ctypes.create_string_buffer(8192)
    lib().rg_scene_warnings(h, buf, len(buf))
    nm, gm, env = ctypes.c_uint32(), ctypes.c_int32(), Env()
    check(lib().rg_scene_export_media(h, None, ctypes.byref(nm), ctypes.byref(gm), ctypes.byref(env), None))
    from scenes import MEDIUM_DTYPE
    media = np.zeros(nm.value, MEDIUM_DTYPE)
    env_rgb = np.zeros((env.h, env.w, 3), np.float32)
    check(lib().rg_scene_export_media(h, _ptr(media), ctypes.byref(nm), ctypes.byref(gm), None, _ptr(env_rgb)))
    sc = Scene.__new__(Scene)
    sc.ctx, sc.h = ctx, h
    sc.v9 = v9
    cam_d = {f: getattr(cam, f) for f, _ in Camera._fields_ if f != "pad"}
    d = {"v9": v9, "n9": n9, "obj": obj, "objects": objs, "camera": cam_d, "media": media,
         "global_medium": gm.value, "warnings": [w for w in buf.value.decode().split("\n") if w]}
    if env.w and env.h:
        d["env"] = {"rgb": env_rgb, "intensity": env.intensity, "rotation_deg": env.rotation_deg}
    nvl = ctypes.c_uint32()
    check(lib().rg_scene_export_volume_lights(h, None, ctypes.byref(nvl)))
    if nvl.value:
        from scenes import VOLUME_LIGHT_DTYPE
        vla = np.zeros(nvl.value, VOLUME_LIGHT_DTYPE)
        check(lib().rg_scene_export_volume_lights(h, _ptr(vla), ctypes.byref(nvl)))
        d["volume_lights"] = vla
    npl = ctypes.c_uint32()
    check(lib().rg_scene_export_point_lights(h, None, ctypes.byref(npl)))
    if npl.value:
        from scenes import POINT_LIGHT_DTYPE
        pla = np.zeros(npl.value, POINT_LIGHT_DTYPE)
        check(lib().rg_scene_export_point_lights(h, _ptr(pla), ctypes.byref(npl)))
        d["point_lights"] = pla
    ntex = ctypes.c_uint32()
    check(lib().rg_scene_export_textures(h, ctypes.byref(ntex), None, None, None, None, None))
    if ntex.value:
        tab = (Texture * ntex.value)()
        check(lib().rg_scene_export_textures(h, ctypes.byref(ntex), tab, None, None, None, None))
        texels = np.zeros(sum(3 * t.w * t.h for t in tab), np.float32)
        too = np.zeros(no.value, np.int32)
        til = np.zeros((no.value, 2), np.float32)
        uv6 = np.zeros((nt.value, 6), np.float32)
        check(lib().rg_scene_export_textures(h, ctypes.byref(ntex), tab, _ptr(texels), _ptr(too), _ptr(til),
                                             _ptr(uv6)))
        texs, at = [], 0
        for t in tab:
            m = 3 * t.w * t.h
            texs.append({"rgb": texels[at:at + m].reshape(t.h, t.w, 3), "srgb": bool(t.srgb)})
            at += m
        d.update(textures=texs, tex_of_obj=too, tiling=til, uv6=uv6)
    return sc, d


# ---------------------------------------------------------------------------
# many jobs: rg_render_batch, rg_autotune, rg_server_*
# ---------------------------------------------------------------------------
TIMEOUT = 1


class Job(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("settings", Settings)]


class Tune(ctypes.Structure):
    _fields_ = [("pool_paths", ctypes.c_uint32), ("stack", ctypes.c_uint32), ("lanes", ctypes.c_uint32),
                ("blocks_per_cu", ctypes.c_uint32), ("iters_per_graph", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_ if f != "pad"}


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("pad", ctypes.c_uint32), ("samples", ctypes.c_uint64),
                ("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double)]


def make_job(scene, **kw):
    j = Job()
    j.scene = scene
    j.settings = default_settings(**kw)
    return j


def make_tune(**kw):
    t = Tune()
    for k, v in kw.items():
        setattr(t, k, v)
    return t


def _scene_array(scenes):
    arr = (ctypes.c_void_p * len(scenes))(*[s.h for s in scenes])
    return arr


def render_batch(ctx, scenes, jobs, tune=None, want_sum=True):
    """rg_render_batch: returns (list of {"hdr", "sum"}, stats[4])."""
    jarr = (Job * len(jobs))(*jobs)
    outs = []
    hp = (ctypes.c_void_p * len(jobs))()
    sp = (ctypes.c_void_p * len(jobs))()
    for i, j in enumerate(jobs):
        W, H = j.settings.width, j.settings.height
        o = {"hdr": np.zeros((H, W, 4), np.float32)}
        hp[i] = o["hdr"].ctypes.data
        if want_sum:
            o["sum"] = np.zeros((H, W, 3), np.float32)
            sp[i] = o["sum"].ctypes.data
        outs.append(o)
    stats = np.zeros(4, np.uint64)
    sarr = _scene_array(scenes)
    check(lib().rg_render_batch(ctx.h, sarr, len(scenes), jarr, len(jobs), ctypes.byref(tune) if tune else None,
                                hp, sp if want_sum else None, _ptr(stats)))
    return outs, stats


def autotune(ctx, scenes, probe=None):
    t = Tune()
    sarr = _scene_array(scenes)
    parr = (Job * len(probe))(*probe) if probe else None
    check(lib().rg_autotune(ctx.h, sarr, len(scenes), parr, len(probe) if probe else 0, ctypes.byref(t)))
    return t


class Server:
    """rg_server_*: scenes stay resident; submit() queues a request and
    returns its ticket, wait() returns (info, outputs)."""

    def __init__(self, ctx, scenes, hbm_budget=0, tune=None):
        self.ctx = ctx
        self.scenes = list(scenes)
        h = ctypes.c_void_p()
        check(lib().rg_server_create(ctx.h, _scene_array(self.scenes), len(self.scenes), hbm_budget,
                                     ctypes.byref(tune) if tune else None, ctypes.byref(h)))
        self.h = h
        self.outs = {}

    def submit(self, job, want_hdr=True, want_sum=False):
        W, H = job.settings.width, job.settings.height
        o = {}
        if want_hdr:
            o["hdr"] = np.zeros((H, W, 4), np.float32)
        if want_sum:
            o["sum"] = np.zeros((H, W, 3), np.float32)
        t = ctypes.c_uint64()
        check(lib().rg_server_submit(self.h, ctypes.byref(job), _ptr(o["hdr"]) if want_hdr else None,
                                     _ptr(o["sum"]) if want_sum else None, ctypes.byref(t)))
        self.outs[t.value] = o
        return t.value

    def wait(self, ticket, timeout=-1.0):
        info = RequestInfo()
        rc = lib().rg_server_wait(self.h, ticket, timeout, ctypes.byref(info))
        if rc == TIMEOUT:
            return None, None
        check(rc)
        return info, self.outs.pop(ticket)

    def stats(self):
        out = np.zeros(8, np.uint64)
        check(lib().rg_server_stats(self.h, _ptr(out)))
        keys = ["requests", "samples", "rays", "shadow_rays", "generations", "fb_bytes", "budget", "graphs"]
        return dict(zip(keys, (int(x) for x in out)))

    def set_queue_limit(self, max_outstanding_samples):
        check(lib().rg_server_set_queue_limit(self.h, int(max_outstanding_samples)))

    def queue(self):
        out = np.zeros(3, np.uint64)
        check(lib().rg_server_queue(self.h, _ptr(out)))
        return {"outstanding": int(out[0]), "limit": int(out[1]), "rejected": int(out[2])}

    def tune(self):
        t = Tune()
        check(lib().rg_server_get_tune(self.h, ctypes.byref(t)))
        return t

    def close(self):
        if getattr(self, "h", None):
            lib().rg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Bloom(ctypes.Structure):
    """rg_bloom: RendererSettings' bloom knobs (rayrenderer.h:137-144)."""
    _fields_ = [("threshold", ctypes.c_float), ("strength", ctypes.c_float), ("radius", ctypes.c_float),
                ("size_aspect", ctypes.c_float), ("curve", ctypes.c_float), ("pad", ctypes.c_float)]


def default_bloom(**kw):
    b = Bloom()
    lib().rg_bloom_default(ctypes.byref(b))
    for k, v in kw.items():
        setattr(b, k, v)
    return b


def postprocess(ctx, hdr, bloom=None, want_ldr=True):
    """Output stage on host images (RayRenderer::applyPostProcess +
    applyTonemapGamma): returns (bloomed HDR, LDR or None); hdr (H x W x 4
    float32) is not modified."""
    out = np.ascontiguousarray(hdr, dtype=np.float32).copy()
    h, w = out.shape[:2]
    ldr = np.empty_like(out) if want_ldr else None
    check(lib().rg_postprocess(ctx.h, w, h, _ptr(out), ctypes.byref(bloom) if bloom is not None else None,
                               _ptr(ldr)))
    return out, ldr


def postprocess_device(ctx, width, height, d_hdr, bloom=None, d_ldr=None):
    """Output stage on device images (integer addresses), queued on the
    context's stream."""
    check(lib().rg_postprocess_device(ctx.h, width, height, d_hdr,
                                      ctypes.byref(bloom) if bloom is not None else None, d_ldr))


def postprocess_device_noalloc(ctx, width, height, d_hdr, bloom=None, d_ldr=None):
    """rg_postprocess_device_noalloc: as postprocess_device, with the glow
    buffers in the library's fixed per-GPU scratch array (no allocation)."""
    check(lib().rg_postprocess_device_noalloc(ctx.h, width, height, d_hdr,
                                              ctypes.byref(bloom) if bloom is not None else None, d_ldr))


def image_write(path, rgba, quality=0):
    """rg_image_write: the format follows the extension (.hdr RGBE of a
    linear image; .jpg/.png/.ppm/.bmp of an LDR image in [0, 1])."""
    a = np.ascontiguousarray(rgba, np.float32)
    h, w = a.shape[:2]
    if a.shape[2] != 4:
        raise ValueError("rgba image expected")
    check(lib().rg_image_write(str(path).encode(), w, h, _ptr(a), quality))

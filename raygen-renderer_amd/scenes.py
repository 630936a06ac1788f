"""Scene generators for the path-tracing workloads, in render (camera) space:
the camera sits at the origin looking down -Z, as RayRenderer::transformScene
leaves the world after its view transform (rayrenderer.cpp:157-363).

A scene is a dict of flat arrays (the shape the C ABI takes):
  v9 (n,9) vertices, n9 (n,9) vertex normals, obj (n,) object index per
  triangle, objects (structured OBJECT_DTYPE: one Material per SceneObject,
  material.h:20-70), media (MEDIUM_DTYPE), global_medium, env, camera.
"""
import numpy as np

OBJECT_DTYPE = np.dtype([("color", "<f4", (3,)), ("glossy", "<f4"), ("roughness", "<f4"),
                         ("transparency", "<f4"), ("refraction", "<f4"), ("refraction_ratio", "<f4"),
                         ("metallic", "<f4"), ("anisotropy", "<f4"), ("aniso_rotation", "<f4"),
                         ("emission", "<f4"), ("interior_medium", "<i4"), ("visible", "<u4"),
                         ("chroma_dispersion", "<f4"), ("pad", "<u4")])
# rg_medium (HomogeneousMedium's authoring fields, medium.h:37-214)
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


def medium(**kw):
    """One rg_medium with the reference's defaults (medium.h:37-214), fields
    overridden by keyword (path = a list of (p, radius, t))."""
    m = np.zeros(1, MEDIUM_DTYPE)[0]
    defaults = dict(density=1.0, cone_axis=(0.0, 0.0, -1.0), cone_length=1.0, cone_radius=0.3,
                    cone_inner=(0.6, 0.7, 1.0), cone_outer=(1.0, 0.4, 0.05), cone_intensity=50.0,
                    cone_peak_axial=0.15, cone_peak_sharpness=5.0, cone_samples=4, path_inner=(0.6, 0.7, 1.0),
                    path_outer=(0.3, 0.4, 0.6), path_intensity=1.0, path_falloff_power=2.0, path_samples=6,
                    noise_frequency=1.0, noise_octaves=4, noise_gain=0.5, noise_lacunarity=2.0,
                    noise_amplitude=1.0, ior_amplitude=0.005, ior_frequency=4.0, ior_octaves=3, ior_gain=0.5,
                    ior_lacunarity=2.0, ior_march_steps=16)
    defaults.update(kw)
    path = defaults.pop("path", None)
    m["path"]["radius"] = 0.1       # PathSample's default radius
    for k, v in defaults.items():
        m[k] = v
    if path is not None:
        m["n_path"] = len(path)
        for i, (p, r, t) in enumerate(path):
            m["path"][i]["p"] = p
            m["path"][i]["radius"] = r
            m["path"][i]["t"] = t
    return m
# rg_point_light: a point light source in render space (an emissive object
# without meshes, rayrenderer.cpp:333-353)
POINT_LIGHT_DTYPE = np.dtype([("pos", "<f4", (3,)), ("normal", "<f4", (3,)), ("color", "<f4", (3,)),
                              ("emission", "<f4"), ("spot_range", "<f4"), ("pad", "<u4")])


def point_light(pos, color=(1.0, 1.0, 1.0), emission=1.0, normal=(0.0, -1.0, 0.0), spot_range=0.0):
    """One point light record (spot_range 0 = omni)."""
    L = np.zeros(1, POINT_LIGHT_DTYPE)[0]
    L["pos"] = pos
    L["normal"] = normal
    L["color"] = color
    L["emission"] = emission
    L["spot_range"] = spot_range
    return L


def material(color=(0.8, 0.8, 0.8), glossy=0.0, roughness=0.5, transparency=0.0, refraction=0.0,
             refraction_ratio=1.45, metallic=0.0, anisotropy=0.0, aniso_rotation=0.0, emission=0.0,
             interior_medium=-1, visible=1, chroma_dispersion=0.0):
    """One object's material with the reference defaults (material.h:26-50)."""
    m = np.zeros(1, OBJECT_DTYPE)[0]
    m["color"] = color
    m["glossy"] = glossy
    m["roughness"] = roughness
    m["transparency"] = transparency
    m["refraction"] = refraction
    m["refraction_ratio"] = refraction_ratio
    m["metallic"] = metallic
    m["anisotropy"] = anisotropy
    m["aniso_rotation"] = aniso_rotation
    m["emission"] = emission
    m["interior_medium"] = interior_medium
    m["visible"] = visible
    m["chroma_dispersion"] = chroma_dispersion
    return m


class Builder:
    def __init__(self):
        self.v, self.n, self.o, self.uv, self.objects = [], [], [], [], []
        self.textures, self.tex_of_obj, self.tiling = [], [], []

    def add_object(self, mat, texture=-1, tiling=(1.0, 1.0)):
        """One SceneObject; texture = index into add_texture's table (-1
        none), tiling = its texTiling."""
        self.objects.append(mat)
        self.tex_of_obj.append(texture)
        self.tiling.append(tiling)
        return len(self.objects) - 1

    def add_texture(self, rgb, srgb=True):
        """An albedo texture: (H, W, 3) floats as stored (LDR byte / 255),
        row 0 at the top."""
        self.textures.append({"rgb": np.ascontiguousarray(rgb, np.float32), "srgb": srgb})
        return len(self.textures) - 1

    def quad(self, obj, p0, p1, p2, p3, normal=None, uv=((0, 0), (0, 0), (0, 0), (0, 0))):
        """Two triangles (p0,p1,p2), (p0,p2,p3); flat normal from the winding
        unless given; uv = the corners' texture coordinates."""
        p0, p1, p2, p3 = (np.asarray(p, np.float32) for p in (p0, p1, p2, p3))
        if normal is None:
            normal = np.cross(p1 - p0, p2 - p0)
            normal = normal / np.linalg.norm(normal)
        nn = np.tile(np.asarray(normal, np.float32), 3)
        t = [np.asarray(x, np.float32) for x in uv]
        for tri, tt in (((p0, p1, p2), (t[0], t[1], t[2])), ((p0, p2, p3), (t[0], t[2], t[3]))):
            self.v.append(np.concatenate(tri))
            self.n.append(nn)
            self.o.append(obj)
            self.uv.append(np.concatenate(tt))

    def box(self, obj, lo, hi):
        """Axis-aligned box, 12 triangles, outward flat normals (polygons.cpp:61-130 layout)."""
        x0, y0, z0 = lo
        x1, y1, z1 = hi
        self.quad(obj, (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1))  # +z
        self.quad(obj, (x1, y0, z0), (x0, y0, z0), (x0, y1, z0), (x1, y1, z0))  # -z
        self.quad(obj, (x1, y0, z1), (x1, y0, z0), (x1, y1, z0), (x1, y1, z1))  # +x
        self.quad(obj, (x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0))  # -x
        self.quad(obj, (x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (x0, y1, z0))  # +y
        self.quad(obj, (x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1))  # -y

    def sphere(self, obj, center, radius, stacks, slices):
        """SphereMesh (polygons.cpp:132-180): (stacks+1)(slices+1) vertices,
        phi = i/stacks * pi, theta = j/slices * 2pi, triangles (cur, next,
        cur+1) and (cur+1, next, next+1); smooth normals; 2*stacks*slices
        triangles (pole rows are degenerate, as in the reference)."""
        c = np.asarray(center, np.float32)
        i = np.arange(stacks + 1, dtype=np.float32)[:, None]
        j = np.arange(slices + 1, dtype=np.float32)[None, :]
        phi = ((i / np.float32(stacks)).astype(np.float64) * np.pi).astype(np.float32)
        theta = ((j / np.float32(slices) * np.float32(2.0)).astype(np.float64) * np.pi).astype(np.float32)
        x = np.sin(phi) * np.cos(theta)
        y = np.cos(phi) * np.ones_like(theta)
        z = np.sin(phi) * np.sin(theta)
        nrm = np.stack([x, y, z], axis=-1).reshape(-1, 3).astype(np.float32)
        pos = (nrm * np.float32(radius) + c).astype(np.float32)
        ii, jj = np.meshgrid(np.arange(stacks), np.arange(slices), indexing="ij")
        cur = (ii * (slices + 1) + jj).ravel()
        nxt = ((ii + 1) * (slices + 1) + jj).ravel()
        idx = np.stack([np.stack([cur, nxt, cur + 1], 1), np.stack([cur + 1, nxt, nxt + 1], 1)], 1).reshape(-1, 3)
        self.v.append(pos[idx].reshape(-1, 9))
        self.n.append(nrm[idx].reshape(-1, 9))
        self.o.append(np.full(len(idx), obj, np.uint32))
        self.uv.append(np.zeros((len(idx), 6), np.float32))  # uvCount 0: no UVs (mesh.h:92)

    def finish(self, camera, media=None, global_medium=-1, env=None, point_lights=None):
        v = [np.atleast_2d(x) for x in self.v]
        n = [np.atleast_2d(x) for x in self.n]
        o = [np.atleast_1d(np.asarray(x, np.uint32)) for x in self.o]
        scene = {
            "v9": np.ascontiguousarray(np.concatenate(v), np.float32),
            "n9": np.ascontiguousarray(np.concatenate(n), np.float32),
            "obj": np.ascontiguousarray(np.concatenate(o), np.uint32),
            "objects": np.array(self.objects, dtype=OBJECT_DTYPE),
            "media": np.array(media if media is not None else [], dtype=MEDIUM_DTYPE),
            "global_medium": global_medium,
            "camera": camera,
        }
        if env is not None:
            scene["env"] = env
        if point_lights is not None and len(point_lights):
            scene["point_lights"] = np.array(point_lights, dtype=POINT_LIGHT_DTYPE)
        if self.textures:
            scene["textures"] = self.textures
            scene["tex_of_obj"] = np.asarray(self.tex_of_obj, np.int32)
            scene["tiling"] = np.asarray(self.tiling, np.float32).reshape(-1, 2)
            scene["uv6"] = np.ascontiguousarray(np.concatenate([np.atleast_2d(x) for x in self.uv]), np.float32)
        return scene


def checker(n, cells, a, b):
    """n x n texture of cells x cells squares alternating colours a and b."""
    y, x = np.mgrid[0:n, 0:n]
    m = ((x * cells // n + y * cells // n) % 2)[..., None]
    return np.where(m == 0, np.asarray(a, np.float32), np.asarray(b, np.float32)).astype(np.float32)


def textured_room():
    """The cube room with albedo textures: a checkerboard floor (sRGB, tiled
    3 x 3), a colour-gradient back wall (sRGB, 64 x 32 texels) and a glossy
    metallic tall box whose wood-like stripes tint its Fresnel F0 (linear
    texture); the box faces carry no UVs, so they sample texel (0, 0), like a
    mesh without a UV set."""
    b = Builder()
    z0, z1 = -4.6, -2.6
    grey = material(color=(0.75, 0.75, 0.75))
    t_floor = b.add_texture(checker(64, 8, (0.9, 0.9, 0.85), (0.15, 0.2, 0.3)))
    yy, xx = np.mgrid[0:32, 0:64].astype(np.float32)
    grad = np.stack([xx / 63.0, yy / 31.0, 1.0 - xx / 63.0 * 0.5], -1)
    t_back = b.add_texture(np.floor(grad * 255.0) / 255.0)
    stripes = np.tile(np.stack([0.9 - 0.3 * (np.arange(16) % 2), 0.6 + 0.0 * np.arange(16),
                                0.3 + 0.2 * (np.arange(16) % 3 == 0)], -1)[None], (4, 1, 1))
    t_box = b.add_texture(stripes, srgb=False)
    floor = b.add_object(grey, texture=t_floor, tiling=(3.0, 3.0))
    b.quad(floor, (-1, -1, z1), (1, -1, z1), (1, -1, z0), (-1, -1, z0), normal=(0, 1, 0),
           uv=((0, 1), (1, 1), (1, 0), (0, 0)))
    ceil = b.add_object(grey)
    b.quad(ceil, (-1, 1, z0), (1, 1, z0), (1, 1, z1), (-1, 1, z1), normal=(0, -1, 0))
    back = b.add_object(grey, texture=t_back, tiling=(2.0, 1.0))
    b.quad(back, (-1, -1, z0), (1, -1, z0), (1, 1, z0), (-1, 1, z0), normal=(0, 0, 1),
           uv=((0, 1), (1, 1), (1, 0), (0, 0)))
    left = b.add_object(material(color=(0.75, 0.25, 0.25)))
    b.quad(left, (-1, -1, z1), (-1, -1, z0), (-1, 1, z0), (-1, 1, z1), normal=(1, 0, 0))
    right = b.add_object(material(color=(0.25, 0.75, 0.25)))
    b.quad(right, (1, -1, z0), (1, -1, z1), (1, 1, z1), (1, 1, z0), normal=(-1, 0, 0))
    tall = b.add_object(material(color=(0.95, 0.9, 0.8), glossy=0.8, roughness=0.3, metallic=1.0),
                        texture=t_box)
    b.box(tall, (-0.65, -1.0, -4.2), (-0.1, 0.2, -3.65))
    short = b.add_object(material(color=(0.7, 0.7, 0.7)))
    b.box(short, (0.1, -1.0, -3.5), (0.65, -0.4, -2.95))
    light = b.add_object(material(color=(1.0, 1.0, 1.0), emission=15.0))
    b.quad(light, (-0.25, 0.999, -3.85), (0.25, 0.999, -3.85), (0.25, 0.999, -3.35), (-0.25, 0.999, -3.35),
           normal=(0, -1, 0))
    return b.finish(camera={"fov_deg": 45.0, "exposure": 1.0, "aperture": 1.8, "depth_of_field": 0.0})


def cube_room():
    """Plain cube room: a 2x2x2 box open toward the camera (5 planes, 10
    triangles; floor/ceiling/back 0.75 grey, left red, right green), a tall
    and a short axis-aligned inner box (24 triangles, albedo 0.7) and one
    0.5x0.5 emissive ceiling quad (radiance 15) facing -Y: 36 triangles."""
    b = Builder()
    z0, z1 = -4.6, -2.6
    grey = material(color=(0.75, 0.75, 0.75))
    floor = b.add_object(grey)
    b.quad(floor, (-1, -1, z1), (1, -1, z1), (1, -1, z0), (-1, -1, z0), normal=(0, 1, 0))
    ceil = b.add_object(grey)
    b.quad(ceil, (-1, 1, z0), (1, 1, z0), (1, 1, z1), (-1, 1, z1), normal=(0, -1, 0))
    back = b.add_object(grey)
    b.quad(back, (-1, -1, z0), (1, -1, z0), (1, 1, z0), (-1, 1, z0), normal=(0, 0, 1))
    left = b.add_object(material(color=(0.75, 0.25, 0.25)))
    b.quad(left, (-1, -1, z1), (-1, -1, z0), (-1, 1, z0), (-1, 1, z1), normal=(1, 0, 0))
    right = b.add_object(material(color=(0.25, 0.75, 0.25)))
    b.quad(right, (1, -1, z0), (1, -1, z1), (1, 1, z1), (1, 1, z0), normal=(-1, 0, 0))
    tall = b.add_object(material(color=(0.7, 0.7, 0.7)))
    b.box(tall, (-0.65, -1.0, -4.2), (-0.1, 0.2, -3.65))
    short = b.add_object(material(color=(0.7, 0.7, 0.7)))
    b.box(short, (0.1, -1.0, -3.5), (0.65, -0.4, -2.95))
    light = b.add_object(material(color=(1.0, 1.0, 1.0), emission=15.0))
    b.quad(light, (-0.25, 0.999, -3.85), (0.25, 0.999, -3.85), (0.25, 0.999, -3.35), (-0.25, 0.999, -3.35),
           normal=(0, -1, 0))
    return b.finish(camera={"fov_deg": 45.0, "exposure": 1.0, "aperture": 1.8, "depth_of_field": 0.0})


def constant_env(value=0.3, w=64, h=32):
    """A uniform equirectangular environment (the "constant environment")."""
    return {"rgb": np.full((h, w, 3), value, np.float32), "intensity": 1.0, "rotation_deg": 0.0}


def bsdf_frame(grid=5, stacks=200, slices=200, seed=0xB5DF0002, env_value=0.3, dispersion=0.0):
    """Full BSDF frame: grid x grid tessellated spheres (2*stacks*slices
    triangles each; 5x5 at 200x200 = 2,000,000), seeded GGX materials with
    roughness U[0.05, 1], glossy U[0, 1] and refraction 0 or 0.9 at IOR 1.5
    (chromaDispersion `dispersion` on the refractive ones), one area-light
    quad, a constant environment and one homogeneous medium box (sigma_a 0.1,
    sigma_s 0.4, in a transparency-1 cube)."""
    rng = np.random.default_rng(seed)
    b = Builder()
    span = 4.0
    r = 0.42 * span / max(grid, 1)
    for gy in range(grid):
        for gx in range(grid):
            cx = (gx - (grid - 1) / 2) * span / max(grid - 1, 1) if grid > 1 else 0.0
            cy = (gy - (grid - 1) / 2) * span / max(grid - 1, 1) if grid > 1 else 0.0
            mat = material(color=tuple(rng.uniform(0.3, 0.9, 3)), roughness=float(rng.uniform(0.05, 1.0)),
                           glossy=float(rng.uniform(0.0, 1.0)), refraction=float(0.9 if rng.random() < 0.5 else 0.0),
                           refraction_ratio=1.5)
            if mat["refraction"] > 0:
                mat["chroma_dispersion"] = dispersion
            o = b.add_object(mat)
            b.sphere(o, (cx, cy, -6.0), r, stacks, slices)
    light = b.add_object(material(color=(1.0, 1.0, 1.0), emission=8.0))
    b.quad(light, (-1.0, 3.2, -7.0), (1.0, 3.2, -7.0), (1.0, 3.2, -5.0), (-1.0, 3.2, -5.0), normal=(0, -1, 0))
    box = b.add_object(material(color=(1.0, 1.0, 1.0), transparency=1.0, interior_medium=0))
    b.box(box, (-0.8, -1.0, -4.6), (0.8, 0.6, -3.0))
    med = np.zeros(1, MEDIUM_DTYPE)
    med[0]["sigma_a"] = (0.1, 0.1, 0.1)
    med[0]["sigma_s"] = (0.4, 0.4, 0.4)
    med[0]["density"] = 1.0
    return b.finish(camera={"fov_deg": 45.0, "exposure": 1.0, "aperture": 1.8, "depth_of_field": 0.0,
                            "view_far": 50.0},
                    media=med, env=constant_env(env_value))


def ico_ggx(level=8, amp=0.05, seed=0x1C05EED, env_value=0.3):
    """The traversal benchmark's displaced icosphere (20 * 4^level triangles,
    level 8 = 1,310,720) shaded with GGX (glossy 1, roughness 0.3, flat face
    normals), beside the BSDF frame's medium box (sigma_a 0.1, sigma_s 0.4 in
    a transparency-1 cube), lit by one area-light quad and a constant
    environment."""
    import rgapi

    v = rgapi.gen_icosphere(level, amp, seed).reshape(-1, 3, 3).astype(np.float32)
    v = (v * np.float32(1.3) + np.array([0.9, 0.0, -5.5], np.float32)).astype(np.float32)
    fn = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    fn = (fn / np.linalg.norm(fn, axis=1, keepdims=True)).astype(np.float32)
    b = Builder()
    ball = b.add_object(material(color=(0.9, 0.75, 0.35), glossy=1.0, roughness=0.3))
    b.v.append(v.reshape(-1, 9))
    b.n.append(np.repeat(fn, 3, axis=0).reshape(-1, 9))
    b.o.append(np.full(len(v), ball, np.uint32))
    floor = b.add_object(material(color=(0.6, 0.6, 0.6)))
    b.quad(floor, (-6, -1.5, -1), (6, -1.5, -1), (6, -1.5, -12), (-6, -1.5, -12), normal=(0, 1, 0))
    light = b.add_object(material(color=(1.0, 1.0, 1.0), emission=8.0))
    b.quad(light, (-1.0, 3.2, -6.5), (1.0, 3.2, -6.5), (1.0, 3.2, -4.5), (-1.0, 3.2, -4.5), normal=(0, -1, 0))
    box = b.add_object(material(color=(1.0, 1.0, 1.0), transparency=1.0, interior_medium=0))
    b.box(box, (-2.4, -1.49, -5.6), (-0.8, 0.1, -4.0))
    med = np.zeros(1, MEDIUM_DTYPE)
    med[0]["sigma_a"] = (0.1, 0.1, 0.1)
    med[0]["sigma_s"] = (0.4, 0.4, 0.4)
    med[0]["density"] = 1.0
    return b.finish(camera={"fov_deg": 45.0, "exposure": 1.0, "aperture": 1.8, "depth_of_field": 0.0,
                            "view_far": 50.0},
                    media=med, env=constant_env(env_value))


def sky_env(w=64, h=32, sun=(0.62, 0.22), sun_value=60.0, zenith=(0.35, 0.5, 0.9), horizon=(0.9, 0.85, 0.8),
            ground=(0.12, 0.1, 0.08), rotation_deg=0.0):
    """A non-constant equirectangular sky: a zenith-to-horizon gradient, a
    dark ground below the horizon and a small bright sun disc at
    (u, v) = sun, row 0 at the top."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    v = (y + 0.5) / h
    t = np.clip(v / 0.5, 0.0, 1.0)[..., None]
    up = (1 - t) * np.asarray(zenith, np.float32) + t * np.asarray(horizon, np.float32)
    img = np.where((v < 0.5)[..., None], up, np.asarray(ground, np.float32))
    du = (x + 0.5) / w - sun[0]
    dv = v - sun[1]
    img[(du * du + dv * dv) < (1.6 / w) ** 2] = sun_value
    return {"rgb": np.ascontiguousarray(img, np.float32), "intensity": 1.0, "rotation_deg": rotation_deg}


def materials_extra(stacks=16, slices=24, env=None):
    """The branches the BSDF frame leaves out: a delta mirror (roughness 0,
    metallic, and a dielectric one mixed with diffuse), anisotropic GGX of
    both signs with a brush rotation, a Henyey-Greenstein medium (g = 0.6) in
    a transparency-1 box, a level floor, one area light and a non-constant
    sky with a sun (rotated by 40 degrees)."""
    b = Builder()
    fl = b.add_object(material(color=(0.7, 0.7, 0.65)))
    b.quad(fl, (-8, -1.0, 2), (8, -1.0, 2), (8, -1.0, -14), (-8, -1.0, -14), normal=(0, 1, 0))
    mats = [
        material(color=(0.95, 0.8, 0.55), glossy=1.0, roughness=0.0, metallic=1.0),
        material(color=(0.3, 0.5, 0.8), glossy=0.6, roughness=0.0),
        material(color=(0.9, 0.9, 0.9), glossy=1.0, roughness=0.5, metallic=1.0, anisotropy=0.8,
                 aniso_rotation=30.0),
        material(color=(0.8, 0.4, 0.3), glossy=0.7, roughness=0.6, metallic=0.5, anisotropy=-0.6),
    ]
    for i, m in enumerate(mats):
        o = b.add_object(m)
        b.sphere(o, (-2.4 + 1.6 * i, -0.3, -6.0), 0.65, stacks, slices)
    box = b.add_object(material(color=(1.0, 1.0, 1.0), transparency=1.0, interior_medium=0))
    b.box(box, (-1.2, -0.99, -4.2), (0.4, 0.4, -3.0))
    light = b.add_object(material(color=(1.0, 0.95, 0.9), emission=10.0))
    b.quad(light, (-0.8, 3.0, -6.5), (0.8, 3.0, -6.5), (0.8, 3.0, -5.0), (-0.8, 3.0, -5.0), normal=(0, -1, 0))
    med = np.zeros(1, MEDIUM_DTYPE)
    med[0]["sigma_a"] = (0.05, 0.1, 0.2)
    med[0]["sigma_s"] = (0.6, 0.5, 0.4)
    med[0]["g"] = 0.6
    med[0]["density"] = 1.0
    return b.finish(camera={"fov_deg": 45.0, "exposure": 1.0, "aperture": 1.8, "depth_of_field": 0.0,
                            "view_far": 50.0},
                    media=med, env=env if env is not None else sky_env(rotation_deg=40.0))


def fog_room(scatter=0.15, absorb=0.05, g=0.3):
    """A scene inside a global medium without an environment: a floor, a
    box, an emissive quad and open space around them, so many eye rays cross
    the fog without meeting anything (the backColor case of traceEyeRay,
    rayrenderer.cpp:1263-1273)."""
    b = Builder()
    fl = b.add_object(material(color=(0.6, 0.6, 0.6)))
    b.quad(fl, (-2, -1.0, -2), (2, -1.0, -2), (2, -1.0, -6), (-2, -1.0, -6), normal=(0, 1, 0))
    bx = b.add_object(material(color=(0.7, 0.3, 0.2), glossy=0.3, roughness=0.4))
    b.box(bx, (-0.9, -1.0, -4.6), (-0.2, -0.2, -3.8))
    light = b.add_object(material(color=(1.0, 1.0, 1.0), emission=6.0))
    b.quad(light, (0.3, 0.2, -4.6), (1.1, 0.2, -4.6), (1.1, 1.0, -4.6), (0.3, 1.0, -4.6), normal=(0, 0, 1))
    med = np.zeros(1, MEDIUM_DTYPE)
    med[0]["sigma_a"] = absorb
    med[0]["sigma_s"] = scatter
    med[0]["g"] = g
    med[0]["density"] = 1.0
    return b.finish(camera={"fov_deg": 60.0, "exposure": 1.0, "aperture": 1.8, "depth_of_field": 0.0},
                    media=med, global_medium=0)


def banded_env(w=64, h=32, value=2.0):
    """A strongly coloured environment whose hue changes over short angles,
    so refraction at the three dispersion bands sees different colours."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    img = np.stack([np.cos(x / w * 12) * 0.5 + 0.5, np.sin(y / h * 9) * 0.5 + 0.5,
                    np.cos((x + y) / 7) * 0.5 + 0.5], -1) * value
    return {"rgb": np.ascontiguousarray(img, np.float32), "intensity": 1.0, "rotation_deg": 0.0}


def dispersive_window(stacks=24, slices=32):
    """The case where the reference's Mix branch hands its dispersion band to
    the transparency branch of an eye hit (rayrenderer.cpp:2889-2894,
    bsdf.cpp:402-407): a transparent, refractive, dispersive pane seen
    directly, a dispersive glass sphere behind it and a banded environment."""
    b = Builder()
    pane = b.add_object(material(color=(1, 1, 1), transparency=0.9, refraction=0.8, refraction_ratio=1.5,
                                 roughness=0.0, chroma_dispersion=0.15))
    b.quad(pane, (-3, -3, -2.5), (3, -3, -2.5), (3, 3, -2.5), (-3, 3, -2.5), normal=(0, 0, 1))
    glass = b.add_object(material(color=(1, 1, 1), refraction=1.0, refraction_ratio=1.5, roughness=0.0,
                                  chroma_dispersion=0.15))
    b.sphere(glass, (0, 0, -5), 1.8, stacks, slices)
    return b.finish(camera={"fov_deg": 40.0, "aperture": 1.8}, env=banded_env())


def point_lit_room(spot_range=14.0, env_value=0.0):
    """The cube room lit by its ceiling quad and two point lights of
    traceLight (rayrenderer.cpp:2205-2290): an omni light near the back wall
    and a spot above the short box aimed down (spot cosine read through
    RADIAN_TO_DEGREE, as the reference does); env_value > 0 adds a constant
    environment, so a diffuse event casts all three NEE rays (area, point,
    environment)."""
    sc = cube_room()
    sc["point_lights"] = np.array([
        point_light((-0.5, 0.5, -4.3), color=(1.0, 0.8, 0.6), emission=0.8),
        point_light((0.35, 0.6, -3.2), color=(0.6, 0.7, 1.0), emission=0.6, normal=(0.0, 1.0, 0.0),
                    spot_range=spot_range),
    ], dtype=POINT_LIGHT_DTYPE)
    if env_value > 0.0:
        sc["env"] = constant_env(env_value)
    return sc


def point_light_plane(height=1.0, emission=2.0, color=(1.0, 0.5, 0.25), albedo=(0.8, 0.6, 0.4)):
    """A diffuse plane y = -1 under one omni point light at (0, -1 + height,
    -3), nothing else: with no environment and no other surface, a pixel
    that sees the plane at p with normal n = +Y has, at depth 1 or more,
    the radiance albedo x clamp(color x emission x cos / r^2) (the point
    light's term of traceLight multiplied by the diffuse albedo)."""
    b = Builder()
    pl = b.add_object(material(color=albedo))
    b.quad(pl, (-4, -1.0, -1), (4, -1.0, -1), (4, -1.0, -7), (-4, -1.0, -7), normal=(0, 1, 0))
    return b.finish(camera={"fov_deg": 60.0, "exposure": 1.0, "aperture": 1.8, "depth_of_field": 0.0},
                    point_lights=[point_light((0.0, -1.0 + height, -3.0), color=color, emission=emission)])


def media_room(env_value=0.15):
    """The media of rank-4 "extended media": a floor and a back wall with an
    area light and a dim constant environment; inside transparent boxes an
    fBm cloud (scattering, delta tracking), a cone flame (emissive, also an
    NEE volume light), a path-tube glow and a heat-haze volume in front of a
    striped wall (medium.cpp:125-472; rayrenderer.cpp:1596-1864, 2066-2203)."""
    b = Builder()
    fl = b.add_object(material(color=(0.6, 0.6, 0.6)))
    b.quad(fl, (-4, -1.0, 0), (4, -1.0, 0), (4, -1.0, -8), (-4, -1.0, -8), normal=(0, 1, 0))
    back = b.add_object(material(color=(0.7, 0.7, 0.7)))
    b.quad(back, (-4, -1, -6), (4, -1, -6), (4, 3, -6), (-4, 3, -6), normal=(0, 0, 1))
    for i in range(6):     # stripes behind the haze volume
        st = b.add_object(material(color=(0.9, 0.2, 0.1) if i % 2 else (0.1, 0.3, 0.9)))
        x0 = 0.4 + 0.25 * i
        b.quad(st, (x0, -1, -5.98), (x0 + 0.25, -1, -5.98), (x0 + 0.25, 1.5, -5.98), (x0, 1.5, -5.98),
               normal=(0, 0, 1))
    light = b.add_object(material(color=(1.0, 1.0, 1.0), emission=6.0))
    b.quad(light, (-0.5, 2.5, -4.0), (0.5, 2.5, -4.0), (0.5, 2.5, -3.0), (-0.5, 2.5, -3.0), normal=(0, -1, 0))
    media = [
        medium(sigma_a=(0.05, 0.05, 0.05), sigma_s=(2.0, 2.0, 2.0), g=0.6, density_field=1, noise_frequency=1.6,
               noise_octaves=4, noise_amplitude=1.4, noise_bias=-0.25),
        medium(emission_mode=1, cone_origin=(-0.6, -0.2, -4.0), cone_axis=(-1.0, 0.3, 0.0), cone_length=1.2,
               cone_radius=0.3, cone_inner=(4.0, 6.0, 12.0), cone_outer=(10.0, 3.0, 0.4), cone_intensity=2.5,
               cone_peak_axial=0.18, cone_peak_sharpness=4.0, cone_samples=8, density_field=1,
               noise_frequency=2.0, noise_octaves=3, noise_amplitude=1.2, noise_bias=-0.1),
        medium(emission_mode=2, path=[((-0.4, 0.6, -3.0), 0.15, 0.0), ((0.0, 0.9, -3.2), 0.2, 0.5),
                                      ((0.4, 0.7, -3.1), 0.1, 1.0)], path_intensity=3.0, path_samples=6),
        medium(heat_haze=1, ior_amplitude=0.02, ior_frequency=4.0, ior_octaves=3, ior_march_steps=16),
    ]
    clear = dict(color=(1.0, 1.0, 1.0), transparency=1.0)
    cloud = b.add_object(material(interior_medium=0, **clear))
    b.box(cloud, (0.3, -0.6, -4.4), (1.7, 0.4, -3.4))
    flame = b.add_object(material(interior_medium=1, **clear))
    b.box(flame, (-2.0, -0.6, -4.4), (-0.5, 0.4, -3.6))
    tube = b.add_object(material(interior_medium=2, **clear))
    b.box(tube, (-0.7, 0.3, -3.5), (0.7, 1.2, -2.8))
    haze = b.add_object(material(interior_medium=3, **clear))
    b.box(haze, (0.3, -0.9, -2.8), (2.0, 1.2, -2.0))
    sc = b.finish(camera={"fov_deg": 60.0, "exposure": 1.0, "aperture": 1.8, "depth_of_field": 0.0},
                  media=np.array(media, MEDIUM_DTYPE), env=constant_env(env_value) if env_value > 0 else None)
    # the flame is registered as an NEE volume light (a cone of positive
    # intensity); the path tube is not (rayrenderer.cpp:193-203)
    sc["volume_lights"] = np.array([(1, (-1.25, -0.1, -4.0))], VOLUME_LIGHT_DTYPE)
    return sc

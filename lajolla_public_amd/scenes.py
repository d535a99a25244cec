"""Scene descriptions and seeded synthetic scene generators.

A :class:`SceneDesc` holds numpy arrays and plain Python records shaped like
lajolla's Scene (src/scene.h:42-82): a camera, materials, shapes, lights,
media, the environment-map light id and RenderOptions.  ``to_c()`` packs it
into the ``ljx_scene_desc`` C struct of include/lajolla_amd.h.

The generators build the workloads the benchmarks and tests use, each from a
recorded 64-bit seed:

* ``cornell_box``      -- 256x256, 16 spp, depth 6, 36 triangles, Lambertian
* ``terrain_spheres``  -- 1920x1080, 64 spp, depth 8, ~2.31M triangles,
                          Lambertian / RoughPlastic / DisneyBSDF, 1024
                          emissive triangles and a constant environment
* ``terrain_gap``      -- terrain_spheres under a ceiling with a 5% gap and a
                          small emitter above it carrying 90% of the light
                          power (the learned-guiding workload)
* ``nested_media``     -- 1024x1024, 128 spp, depth 16, homogeneous media
                          inside a 100k-triangle sphere nested in a
                          dielectric sphere, Henyey-Greenstein phase
"""
import ctypes as C
import math

import numpy as np

from . import _capi as capi


# ----------------------------------------------------------------- transforms
def normalize(v):
    v = np.asarray(v, dtype=np.float64)
    n = np.linalg.norm(v)
    return v / n if n > 0 else np.zeros(3)


def look_at(pos, look, up):
    """Camera-to-world matrix, same construction as src/transform.cpp:46-70."""
    pos = np.asarray(pos, np.float64)
    d = normalize(np.asarray(look, np.float64) - pos)
    left = normalize(np.cross(normalize(up), d))
    new_up = np.cross(d, left)
    m = np.eye(4)
    m[:3, 0] = left
    m[:3, 1] = new_up
    m[:3, 2] = d
    m[:3, 3] = pos
    return m


def constant_volume(value, scale=1.0):
    """ConstantVolume<Spectrum> (src/volume.h:13-16)."""
    return dict(type=capi.VOLUME_CONSTANT, value=tuple(float(v) for v in value), scale=float(scale))


def grid_volume(data, p_min, p_max, scale=1.0):
    """GridVolume<Spectrum> (src/volume.h:18-26). data: (rz, ry, rx) or
    (rz, ry, rx, 1|3) floats, x fastest in memory."""
    a = np.ascontiguousarray(data, np.float32)
    if a.ndim == 3:
        a = a[..., None]
    assert a.ndim == 4 and a.shape[3] in (1, 3)
    return dict(type=capi.VOLUME_GRID, data=a, p_min=tuple(map(float, p_min)), p_max=tuple(map(float, p_max)),
                scale=float(scale))


def write_vol(path, data, p_min, p_max):
    """A Mitsuba .vol grid (version 3, float32; the layout read by
    src/volume.cpp:6-114): data (rz, ry, rx) or (rz, ry, rx, 1|3)."""
    a = np.ascontiguousarray(data, np.float32)
    if a.ndim == 3:
        a = a[..., None]
    rz, ry, rx, ch = a.shape
    with open(path, "wb") as f:
        f.write(b"VOL" + bytes([3]))
        f.write(np.array([1, rx, ry, rz, ch], np.int32).tobytes())
        f.write(np.array(list(p_min) + list(p_max), np.float32).tobytes())
        f.write(a.tobytes())


def camera_desc(cam):
    """ljx_camera_desc from a camera dict (cam_to_world, fov, width, height,
    filter_type, filter_param, medium_id)."""
    c = capi.CameraDesc()
    c.cam_to_world = capi.D16(*np.asarray(cam["cam_to_world"], np.float64).reshape(16))
    c.fov = cam["fov"]
    c.width = cam["width"]
    c.height = cam["height"]
    c.filter_type = cam.get("filter_type", capi.FILTER_BOX)
    c.filter_param = cam.get("filter_param", 1.0)
    c.medium_id = cam.get("medium_id", -1)
    return c


# ------------------------------------------------------------------ the desc
class SceneDesc:
    def __init__(self):
        self.camera = dict(cam_to_world=np.eye(4), fov=45.0, width=64, height=64,
                           filter_type=capi.FILTER_BOX, filter_param=1.0, medium_id=-1)
        self.materials = []
        self.shapes = []
        self.lights = []
        self.media = []
        self.textures = []
        self.envmap_light_id = -1
        self.options = dict(integrator=capi.INTEGRATOR_PATH, samples_per_pixel=4,
                            max_depth=-1, rr_depth=5, vol_path_version=0,
                            max_null_collisions=1000, seed=capi.DEFAULT_SEED)
        self.name = "scene"
        self.meta = {}

    # -- builders
    def add_material(self, type, color_a=(0, 0, 0), color_b=(0, 0, 0), roughness=0.0, eta=1.5, textures=None,
                     **kw):
        """textures: {field name: texture id from add_image_texture /
        add_checkerboard} for the fields that vary over the surface."""
        m = dict(type=type, color_a=tuple(color_a), color_b=tuple(color_b),
                 roughness=roughness, eta=eta, textures=dict(textures or {}))
        for k in m["textures"]:
            assert k in capi.TEXTURE_SLOTS, k
        for k in ("specular_transmission", "metallic", "subsurface", "specular", "specular_tint",
                  "anisotropic", "sheen", "sheen_tint", "clearcoat", "clearcoat_gloss"):
            m[k] = kw.get(k, 0.0)
        self.materials.append(m)
        return len(self.materials) - 1

    def add_mesh(self, positions, indices, material_id, normals=None, uvs=None,
                 interior_medium_id=-1, exterior_medium_id=-1):
        s = dict(type=capi.SHAPE_TRIANGLE_MESH, material_id=material_id, area_light_id=-1,
                 interior_medium_id=interior_medium_id, exterior_medium_id=exterior_medium_id,
                 positions=np.ascontiguousarray(positions, np.float32).reshape(-1, 3),
                 indices=np.ascontiguousarray(indices, np.int32).reshape(-1, 3),
                 normals=None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3),
                 uvs=None if uvs is None else np.ascontiguousarray(uvs, np.float32).reshape(-1, 2),
                 center=(0.0, 0.0, 0.0), radius=0.0)
        self.shapes.append(s)
        return len(self.shapes) - 1

    def add_sphere(self, center, radius, material_id, interior_medium_id=-1, exterior_medium_id=-1):
        s = dict(type=capi.SHAPE_SPHERE, material_id=material_id, area_light_id=-1,
                 interior_medium_id=interior_medium_id, exterior_medium_id=exterior_medium_id,
                 positions=None, indices=None, normals=None, uvs=None,
                 center=tuple(float(c) for c in center), radius=float(radius))
        self.shapes.append(s)
        return len(self.shapes) - 1

    def add_area_light(self, shape_id, intensity):
        self.lights.append(dict(type=capi.LIGHT_DIFFUSE_AREA, shape_id=shape_id,
                                intensity=tuple(intensity)))
        self.shapes[shape_id]["area_light_id"] = len(self.lights) - 1
        return len(self.lights) - 1

    def add_image_texture(self, data, uscale=1.0, vscale=1.0, uoffset=0.0, voffset=0.0):
        """ImageTexture (src/texture.h:82-88): data (h, w) or (h, w, 1|3)
        floats, row 0 at the top."""
        a = np.ascontiguousarray(data, np.float32)
        if a.ndim == 2:
            a = a[..., None]
        assert a.ndim == 3 and a.shape[2] in (1, 3)
        self.textures.append(dict(type=capi.TEXTURE_IMAGE, data=a, color0=(0, 0, 0), color1=(0, 0, 0),
                                  uscale=uscale, vscale=vscale, uoffset=uoffset, voffset=voffset))
        return len(self.textures) - 1

    def add_checkerboard(self, color0=(0.4, 0.4, 0.4), color1=(0.2, 0.2, 0.2), uscale=1.0, vscale=1.0, uoffset=0.0,
                         voffset=0.0):
        """CheckerboardTexture (src/texture.h:90-97; defaults of parse_scene.cpp:349-350)."""
        self.textures.append(dict(type=capi.TEXTURE_CHECKERBOARD, data=None, color0=tuple(color0),
                                  color1=tuple(color1), uscale=uscale, vscale=vscale, uoffset=uoffset,
                                  voffset=voffset))
        return len(self.textures) - 1

    def add_envmap(self, image, scale=1.0, to_world=None):
        """An environment map of an (h, w, 3) lat-long image (src/lights/envmap.inl)."""
        self.lights.append(dict(type=capi.LIGHT_ENVMAP, shape_id=-1, intensity=(0, 0, 0),
                                env=np.ascontiguousarray(image, np.float32).reshape(image.shape[0], image.shape[1], 3),
                                to_world=np.eye(4) if to_world is None else np.asarray(to_world),
                                scale=scale))
        self.envmap_light_id = len(self.lights) - 1
        return self.envmap_light_id

    def add_constant_envmap(self, rgb, scale=1.0, to_world=None):
        self.lights.append(dict(type=capi.LIGHT_ENVMAP, shape_id=-1, intensity=(0, 0, 0),
                                env=np.asarray(rgb, np.float32).reshape(1, 1, 3),
                                to_world=np.eye(4) if to_world is None else np.asarray(to_world),
                                scale=scale))
        self.envmap_light_id = len(self.lights) - 1
        return self.envmap_light_id

    def add_medium(self, sigma_a, sigma_s, phase_type=capi.PHASE_ISOTROPIC, g=0.0):
        self.media.append(dict(type=capi.MEDIUM_HOMOGENEOUS, sigma_a=tuple(sigma_a),
                               sigma_s=tuple(sigma_s), phase_type=phase_type, g=g, albedo=None, density=None))
        return len(self.media) - 1

    def add_heterogeneous_medium(self, albedo, density, phase_type=capi.PHASE_ISOTROPIC, g=0.0):
        """albedo and density are volumes from constant_volume / grid_volume."""
        self.media.append(dict(type=capi.MEDIUM_HETEROGENEOUS, sigma_a=(0, 0, 0), sigma_s=(0, 0, 0),
                               phase_type=phase_type, g=g, albedo=albedo, density=density))
        return len(self.media) - 1

    @property
    def num_triangles(self):
        return int(sum(len(s["indices"]) for s in self.shapes if s["type"] == capi.SHAPE_TRIANGLE_MESH))

    # -- C packing
    def to_c(self):
        keep = []
        d = capi.SceneDescC()
        d.camera = camera_desc(self.camera)

        def ptr(a, ctype):
            if a is None:
                return C.POINTER(ctype)()
            a = np.ascontiguousarray(a)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))

        shapes = (capi.ShapeDesc * max(1, len(self.shapes)))()
        for i, s in enumerate(self.shapes):
            o = shapes[i]
            o.type = s["type"]
            o.material_id = s["material_id"]
            o.area_light_id = s["area_light_id"]
            o.interior_medium_id = s["interior_medium_id"]
            o.exterior_medium_id = s["exterior_medium_id"]
            if s["type"] == capi.SHAPE_TRIANGLE_MESH:
                o.positions = ptr(s["positions"], C.c_float)
                o.indices = ptr(s["indices"], C.c_int32)
                o.normals = ptr(s["normals"], C.c_float)
                o.uvs = ptr(s["uvs"], C.c_float)
                o.num_vertices = len(s["positions"])
                o.num_triangles = len(s["indices"])
            o.center = capi.D3(*s["center"])
            o.radius = s["radius"]
        mats = (capi.MaterialDesc * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            o = mats[i]
            o.type = m["type"]
            o.color_a = capi.D3(*m["color_a"])
            o.color_b = capi.D3(*m["color_b"])
            for k in ("roughness", "eta", "specular_transmission", "metallic", "subsurface", "specular",
                      "specular_tint", "anisotropic", "sheen", "sheen_tint", "clearcoat", "clearcoat_gloss"):
                setattr(o, k, float(m[k]))
            for k, tid in m.get("textures", {}).items():
                o.texture[capi.TEXTURE_SLOTS.index(k)] = tid + 1
        texs = (capi.TextureDesc * max(1, len(self.textures)))()
        for i, t in enumerate(self.textures):
            o = texs[i]
            o.type = t["type"]
            o.color0 = capi.D3(*t["color0"])
            o.color1 = capi.D3(*t["color1"])
            o.uscale, o.vscale, o.uoffset, o.voffset = t["uscale"], t["vscale"], t["uoffset"], t["voffset"]
            if t["type"] == capi.TEXTURE_IMAGE:
                o.height, o.width, o.channels = t["data"].shape
                o.data = ptr(t["data"], C.c_float)
        lights = (capi.LightDesc * max(1, len(self.lights)))()
        for i, l in enumerate(self.lights):
            o = lights[i]
            o.type = l["type"]
            o.shape_id = l["shape_id"]
            o.intensity = capi.D3(*l["intensity"])
            if l["type"] == capi.LIGHT_ENVMAP:
                o.env_rgb = ptr(l["env"], C.c_float)
                o.env_height, o.env_width = l["env"].shape[:2]
                o.to_world = capi.D16(*np.asarray(l["to_world"], np.float64).reshape(16))
                o.scale = l["scale"]
        media = (capi.MediumDesc * max(1, len(self.media)))()
        for i, m in enumerate(self.media):
            o = media[i]
            o.type = m["type"]
            o.sigma_a = capi.D3(*m["sigma_a"])
            o.sigma_s = capi.D3(*m["sigma_s"])
            o.phase_type = m["phase_type"]
            o.g = m["g"]
            for name in ("albedo", "density"):
                v = m.get(name)
                if v is None:
                    continue
                ov = getattr(o, name)
                ov.type = v["type"]
                ov.scale = v["scale"]
                if v["type"] == capi.VOLUME_CONSTANT:
                    ov.value = capi.D3(*v["value"])
                else:
                    data = v["data"]
                    ov.channels = data.shape[3]
                    ov.resolution[0], ov.resolution[1], ov.resolution[2] = data.shape[2], data.shape[1], data.shape[0]
                    ov.p_min = capi.D3(*v["p_min"])
                    ov.p_max = capi.D3(*v["p_max"])
                    ov.data = ptr(data, C.c_float)
        keep += [shapes, mats, lights, media, texs]
        d.textures = C.cast(texs, C.POINTER(capi.TextureDesc))
        d.num_textures = len(self.textures)
        d.shapes = C.cast(shapes, C.POINTER(capi.ShapeDesc))
        d.num_shapes = len(self.shapes)
        d.materials = C.cast(mats, C.POINTER(capi.MaterialDesc))
        d.num_materials = len(self.materials)
        d.lights = C.cast(lights, C.POINTER(capi.LightDesc))
        d.num_lights = len(self.lights)
        d.media = C.cast(media, C.POINTER(capi.MediumDesc))
        d.num_media = len(self.media)
        d.envmap_light_id = self.envmap_light_id
        d.options = render_options(**self.options)
        return d, keep


def render_options(integrator=capi.INTEGRATOR_PATH, samples_per_pixel=4, max_depth=-1, rr_depth=5,
                   vol_path_version=0, max_null_collisions=1000, seed=capi.DEFAULT_SEED):
    o = capi.RenderOptions()
    o.integrator = integrator
    o.samples_per_pixel = samples_per_pixel
    o.max_depth = max_depth
    o.rr_depth = rr_depth
    o.vol_path_version = vol_path_version
    o.max_null_collisions = max_null_collisions
    o.seed = seed
    return o


# --------------------------------------------------------------- geometry kit
def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); the normal is (p1-p0)x(p2-p0)."""
    P = np.array([p0, p1, p2, p3], np.float32)
    I = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    n = np.cross(P[1] - P[0], P[2] - P[0]).astype(np.float64)
    n /= np.linalg.norm(n)
    N = np.tile(n.astype(np.float32), (4, 1))
    return P, I, N


def box(center, size, angle_deg):
    """Axis-aligned box rotated about +y; 6 faces with outward flat normals."""
    cx, cy, cz = center
    sx, sy, sz = [s / 2 for s in size]
    a = math.radians(angle_deg)
    ca, sa = math.cos(a), math.sin(a)

    def xf(x, y, z):
        return (cx + ca * x + sa * z, cy + y, cz - sa * x + ca * z)

    c = [(-sx, -sy, -sz), (sx, -sy, -sz), (sx, sy, -sz), (-sx, sy, -sz),
         (-sx, -sy, sz), (sx, -sy, sz), (sx, sy, sz), (-sx, sy, sz)]
    faces = [(4, 5, 6, 7), (1, 0, 3, 2), (0, 4, 7, 3), (5, 1, 2, 6), (7, 6, 2, 3), (0, 1, 5, 4)]
    Ps, Is, Ns = [], [], []
    for k, f in enumerate(faces):
        P, I, N = quad(*[xf(*c[j]) for j in f])
        Ps.append(P)
        Is.append(I + 4 * k)
        Ns.append(N)
    return np.concatenate(Ps), np.concatenate(Is), np.concatenate(Ns)


def icosphere(subdiv):
    """Unit icosphere by midpoint subdivision; outward CCW winding."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [tuple(np.array(v) / np.linalg.norm(v)) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
         (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            k = (a, b) if a < b else (b, a)
            if k not in cache:
                m = np.add(V[a], V[b]) / 2
                V.append(tuple(m / np.linalg.norm(m)))
                cache[k] = len(V) - 1
            return cache[k]

        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = nf
    P = np.array(V, np.float64)
    I = np.array(F, np.int32)
    # orient outward
    n = np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]])
    flip = np.einsum("ij,ij->i", n, P[I].mean(axis=1)) < 0
    I[flip] = I[flip][:, [0, 2, 1]]
    return P, I


def uv_sphere(n_theta, n_phi):
    """UV sphere with n_theta*n_phi*2 triangles (poles use degenerate-free fans
    by clamping the ring count); outward winding and smooth normals."""
    th = np.linspace(0, np.pi, n_theta + 1)
    ph = np.linspace(0, 2 * np.pi, n_phi + 1)[:-1]
    T, Ph = np.meshgrid(th, ph, indexing="ij")
    P = np.stack([np.sin(T) * np.cos(Ph), np.cos(T), np.sin(T) * np.sin(Ph)], -1).reshape(-1, 3)
    idx = lambda i, j: i * n_phi + (j % n_phi)
    I = []
    for i in range(n_theta):
        for j in range(n_phi):
            a, b, c, d = idx(i, j), idx(i, j + 1), idx(i + 1, j + 1), idx(i + 1, j)
            I.append((a, b, c))
            I.append((a, c, d))
    I = np.array(I, np.int32)
    n = np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]])
    flip = np.einsum("ij,ij->i", n, P[I].mean(axis=1)) < 0
    I[flip] = I[flip][:, [0, 2, 1]]
    return P, I


def value_noise_fbm(x, z, rng, octaves=6, lacunarity=2.0, gain=0.5, base_freq=0.15):
    """Seeded fBm of value noise with smoothstep interpolation."""
    perm = rng.permutation(256)
    vals = rng.uniform(-1, 1, 256)

    def lattice(ix, iz):
        return vals[perm[(perm[ix & 255] + iz) & 255]]

    h = np.zeros_like(x)
    amp, freq = 1.0, base_freq
    for _ in range(octaves):
        fx, fz = x * freq, z * freq
        ix, iz = np.floor(fx).astype(np.int64), np.floor(fz).astype(np.int64)
        tx, tz = fx - ix, fz - iz
        sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)
        a = lattice(ix, iz)
        b = lattice(ix + 1, iz)
        c = lattice(ix, iz + 1)
        d = lattice(ix + 1, iz + 1)
        h += amp * ((a * (1 - sx) + b * sx) * (1 - sz) + (c * (1 - sx) + d * sx) * sz)
        amp *= gain
        freq *= lacunarity
    return h


# ---------------------------------------------------------------- generators
def cornell_box(seed=1, width=256, height=256, spp=16, max_depth=6):
    """Cornell-box-like room: 5 quads of a 2x2x2 room, two boxes rotated about
    y by seeded angles, a 0.5x0.5 ceiling light of radiance (17, 12, 4).
    Lambertian only; 36 triangles."""
    rng = np.random.default_rng(seed)
    s = SceneDesc()
    s.name = "cornell_box"
    refl = lambda: tuple(rng.uniform(0.1, 0.8, 3))
    m_floor, m_ceil, m_back = (s.add_material(capi.MAT_LAMBERTIAN, refl()) for _ in range(3))
    m_red = s.add_material(capi.MAT_LAMBERTIAN, (0.63, 0.065, 0.05))
    m_green = s.add_material(capi.MAT_LAMBERTIAN, (0.14, 0.45, 0.091))
    m_box1, m_box2 = s.add_material(capi.MAT_LAMBERTIAN, refl()), s.add_material(capi.MAT_LAMBERTIAN, refl())
    m_light = s.add_material(capi.MAT_LAMBERTIAN, (0.0, 0.0, 0.0))
    # room interior faces (normals point inward)
    walls = [
        (quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1)), m_floor),  # floor, +y
        (quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)), m_ceil),   # ceiling, -y
        (quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)), m_back),  # back, +z
        (quad((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1)), m_red),   # left, +x
        (quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)), m_green),     # right, -x
    ]
    for (P, I, N), m in walls:
        s.add_mesh(P, I, m, normals=N)
    a1, a2 = rng.uniform(0, 90, 2)
    P, I, N = box((-0.35, 0.6, -0.3), (0.55, 1.2, 0.55), a1)
    s.add_mesh(P, I, m_box1, normals=N)
    P, I, N = box((0.4, 0.3, 0.35), (0.55, 0.6, 0.55), a2)
    s.add_mesh(P, I, m_box2, normals=N)
    P, I, N = quad((-0.25, 1.999, -0.25), (0.25, 1.999, -0.25), (0.25, 1.999, 0.25), (-0.25, 1.999, 0.25))
    lid = s.add_mesh(P, I, m_light, normals=N)
    s.add_area_light(lid, (17.0, 12.0, 4.0))
    s.camera.update(cam_to_world=look_at((0, 1, 3.8), (0, 1, 0), (0, 1, 0)), fov=40.0,
                    width=width, height=height, filter_type=capi.FILTER_BOX, filter_param=1.0)
    s.options.update(samples_per_pixel=spp, max_depth=max_depth, rr_depth=5)
    s.meta = dict(seed=seed, box_angles=[float(a1), float(a2)])
    return s


def _random_material(s, rng, kind):
    if kind == 0:
        return s.add_material(capi.MAT_LAMBERTIAN, tuple(rng.uniform(0, 1, 3)))
    if kind == 1:
        return s.add_material(capi.MAT_ROUGH_PLASTIC, tuple(rng.uniform(0, 1, 3)),
                              tuple(rng.uniform(0, 1, 3)), roughness=float(rng.uniform(0.05, 0.6)), eta=1.5)
    p = rng.uniform(0, 1, 12)
    return s.add_material(capi.MAT_DISNEY_BSDF, tuple(rng.uniform(0, 1, 3)), eta=1.5,
                          specular_transmission=p[0], metallic=p[1], subsurface=p[2], specular=p[3],
                          roughness=p[4], specular_tint=p[5], anisotropic=p[6], sheen=p[7], sheen_tint=p[8],
                          clearcoat=p[9], clearcoat_gloss=p[10])


_ICO5 = None


def terrain_spheres(seed=7, width=1920, height=1080, spp=64, max_depth=8, grid=(1000, 500),
                    n_spheres=64, sphere_subdiv=5, n_lights=1024, env=(0.03, 0.04, 0.06)):
    """Large-mesh scene: a seeded fBm heightfield of grid[0] x grid[1] vertices,
    n_spheres icospheres (flattened instances), n_lights one-triangle
    emitters of log-uniform power, and a constant environment."""
    global _ICO5
    rng = np.random.default_rng(seed)
    s = SceneDesc()
    s.name = "terrain_spheres"
    nx, nz = grid
    ext_x, ext_z = 40.0, 20.0
    xs = np.linspace(-ext_x / 2, ext_x / 2, nx)
    zs = np.linspace(-ext_z / 2, ext_z / 2, nz)
    X, Z = np.meshgrid(xs, zs, indexing="xy")  # (nz, nx)
    H = 1.6 * value_noise_fbm(X, Z, rng)
    P = np.stack([X, H, Z], -1).reshape(-1, 3)
    # smooth normals from central differences
    dhdx = np.gradient(H, xs, axis=1)
    dhdz = np.gradient(H, zs, axis=0)
    N = np.stack([-dhdx, np.ones_like(H), -dhdz], -1).reshape(-1, 3)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    ii, jj = np.meshgrid(np.arange(nz - 1), np.arange(nx - 1), indexing="ij")
    v00 = (ii * nx + jj).ravel()
    v01 = v00 + 1
    v10 = v00 + nx
    v11 = v10 + 1
    # (v00, v10, v11) has normal +y for increasing z rows
    I = np.concatenate([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)]).astype(np.int32)
    m_terrain = _random_material(s, rng, int(rng.integers(0, 2)))
    s.add_mesh(P, I, m_terrain, normals=N)
    if _ICO5 is None or _ICO5[0] != sphere_subdiv:
        _ICO5 = (sphere_subdiv,) + icosphere(sphere_subdiv)
    _, SP, SI = _ICO5

    def height_at(x, z):
        j = np.clip(np.searchsorted(xs, x), 0, nx - 1)
        i = np.clip(np.searchsorted(zs, z), 0, nz - 1)
        return H[i, j]

    for k in range(n_spheres):
        r = float(rng.uniform(0.4, 1.3))
        cx, cz = float(rng.uniform(-17, 17)), float(rng.uniform(-8.5, 8.5))
        cy = float(height_at(cx, cz)) + r * float(rng.uniform(0.6, 1.4))
        m = _random_material(s, rng, k % 3)
        s.add_mesh(SP * r + np.array([cx, cy, cz]), SI, m, normals=SP)
    m_black = s.add_material(capi.MAT_LAMBERTIAN, (0.0, 0.0, 0.0))
    for k in range(n_lights):
        c = np.array([rng.uniform(-19, 19), rng.uniform(5.0, 9.0), rng.uniform(-9.5, 9.5)])
        size = rng.uniform(0.15, 0.45)
        a = rng.uniform(0, 2 * np.pi)
        p0 = c + size * np.array([np.cos(a), 0, np.sin(a)])
        p1 = c + size * np.array([np.cos(a + 2.1), 0, np.sin(a + 2.1)])
        p2 = c + size * np.array([np.cos(a + 4.2), 0, np.sin(a + 4.2)])
        tri = np.array([p0, p1, p2], np.float32)
        nrm = np.cross(tri[1] - tri[0], tri[2] - tri[0]).astype(np.float64)
        if nrm[1] > 0:  # face downwards
            tri = tri[[0, 2, 1]]
        nrm = np.cross(tri[1] - tri[0], tri[2] - tri[0]).astype(np.float64)
        area = 0.5 * np.linalg.norm(nrm)
        nrm /= np.linalg.norm(nrm)
        power = 10 ** rng.uniform(-1, 2)
        col = rng.uniform(0.3, 1.0, 3)
        lum = col @ np.array([0.212671, 0.715160, 0.072169])
        # light_power = luminance(I) * area * pi (diffuse_area_light.inl:1-3)
        intensity = col / lum * power / (area * np.pi)
        sid = s.add_mesh(tri, np.array([[0, 1, 2]], np.int32), m_black,
                         normals=np.tile(nrm.astype(np.float32), (3, 1)))
        s.add_area_light(sid, tuple(float(v) for v in intensity))
    s.add_constant_envmap(env)
    s.camera.update(cam_to_world=look_at((0, 9, 24), (0, 0, -1), (0, 1, 0)), fov=62.0,
                    width=width, height=height, filter_type=capi.FILTER_BOX, filter_param=1.0)
    s.options.update(samples_per_pixel=spp, max_depth=max_depth, rr_depth=5)
    s.meta = dict(seed=seed, grid=list(grid), n_spheres=n_spheres, sphere_subdiv=sphere_subdiv,
                  n_lights=n_lights, env=list(env))
    return s


def terrain_gap(seed=7, width=1920, height=1080, spp=64, max_depth=8, gap_frac=0.05, power_frac=0.9, **kw):
    """Guiding workload: the terrain_spheres scene under an occluding ceiling
    plane (y = 10.5, covering the terrain) with a square gap of gap_frac of
    its area, and a small emitter high above the gap that carries power_frac
    of the scene's total light power (lajolla's light power: pi * area *
    luminance for area lights, pi r^2 * luminance for the environment,
    src/lights/diffuse_area_light.inl:1-3, src/lights/envmap.inl:2-4).  The
    emitter lights a patch of the terrain through the gap; the rest of the
    scene sees it only by indirect light from that patch."""
    s = terrain_spheres(seed=seed, width=width, height=height, spp=spp, max_depth=max_depth, **kw)
    s.name = "terrain_gap"
    rng = np.random.default_rng(seed + 1000)
    y = 10.5
    x0, x1, z0, z1 = -22.0, 22.0, -12.0, 12.0
    area = (x1 - x0) * (z1 - z0)
    side = math.sqrt(gap_frac * area)
    gx, gz = float(rng.uniform(-8, 8)), float(rng.uniform(-4, 4))
    hx0, hx1, hz0, hz1 = gx - side / 2, gx + side / 2, gz - side / 2, gz + side / 2
    m_ceil = s.add_material(capi.MAT_LAMBERTIAN, (0.5, 0.5, 0.5))
    # four quads around the hole, normals facing down
    for (a0, a1, b0, b1) in ((x0, hx0, z0, z1), (hx1, x1, z0, z1), (hx0, hx1, z0, hz0), (hx0, hx1, hz1, z1)):
        P, I, N = quad((a0, y, b0), (a1, y, b0), (a1, y, b1), (a0, y, b1))
        if N[0][1] > 0:
            P, I, N = quad((a0, y, b0), (a0, y, b1), (a1, y, b1), (a1, y, b0))
        s.add_mesh(P, I, m_ceil, normals=N)
    # power of the lights already present
    lumw = np.array([0.212671, 0.715160, 0.072169])
    lo = np.min([sh["positions"].min(0) for sh in s.shapes if sh["positions"] is not None], axis=0)
    hi = np.max([sh["positions"].max(0) for sh in s.shapes if sh["positions"] is not None], axis=0)
    r = np.linalg.norm(hi - lo) / 2
    others = 0.0
    for l in s.lights:
        if l["type"] == capi.LIGHT_DIFFUSE_AREA:
            P = s.shapes[l["shape_id"]]["positions"].astype(np.float64)
            I = s.shapes[l["shape_id"]]["indices"]
            a = 0.5 * np.linalg.norm(np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]]), axis=1).sum()
            others += np.pi * a * float(np.asarray(l["intensity"]) @ lumw)
        else:
            others += np.pi * r * r * float(l["env"].reshape(3) @ lumw) * l["scale"]
    # the bright emitter: a 1 x 1 quad 25 above the gap, facing down
    ls = 1.0
    ly = y + 25.0
    P, I, N = quad((gx - ls / 2, ly, gz - ls / 2), (gx - ls / 2, ly, gz + ls / 2), (gx + ls / 2, ly, gz + ls / 2),
                   (gx + ls / 2, ly, gz - ls / 2))
    if N[0][1] > 0:
        P, I, N = quad((gx - ls / 2, ly, gz - ls / 2), (gx + ls / 2, ly, gz - ls / 2), (gx + ls / 2, ly, gz + ls / 2),
                       (gx - ls / 2, ly, gz + ls / 2))
    m_black = s.add_material(capi.MAT_LAMBERTIAN, (0.0, 0.0, 0.0))
    sid = s.add_mesh(P, I, m_black, normals=N)
    power = power_frac / (1 - power_frac) * others
    col = np.array([1.0, 0.85, 0.6])
    intensity = col / float(col @ lumw) * power / (ls * ls * np.pi)
    s.add_area_light(sid, tuple(float(v) for v in intensity))
    s.meta.update(gap_frac=gap_frac, power_frac=power_frac, gap_center=[gx, y, gz], gap_side=side,
                  emitter_power=float(power), other_power=float(others))
    return s


def orbit_camera(desc, rng, width, height, yaw_deg=20.0, pitch_deg=5.0, dist_range=(0.85, 1.15)):
    """A seeded camera pose around the scene camera's view: yaw and pitch
    about the point 10 units ahead, distance scaled; the fov is kept."""
    cam = dict(desc.camera)
    m = np.asarray(cam["cam_to_world"], np.float64)
    pos, fwd = m[:3, 3], m[:3, 2]
    target = pos + 10.0 * fwd
    off = pos - target
    yaw = math.radians(rng.uniform(-yaw_deg, yaw_deg))
    pitch = math.radians(rng.uniform(-pitch_deg, pitch_deg))
    cy, sy = math.cos(yaw), math.sin(yaw)
    off = np.array([cy * off[0] + sy * off[2], off[1], -sy * off[0] + cy * off[2]])
    r = np.linalg.norm(off)
    el = math.asin(np.clip(off[1] / r, -1, 1)) + pitch
    horiz = math.hypot(off[0], off[2])
    off = np.array([off[0] / horiz * r * math.cos(el), r * math.sin(el), off[2] / horiz * r * math.cos(el)])
    off *= rng.uniform(*dist_range)
    cam["cam_to_world"] = look_at(target + off, target, (0, 1, 0))
    cam["width"], cam["height"] = int(width), int(height)
    return cam


def batch_jobs(scenes_and_descs, n_jobs=96, seed=3, res_range=(64, 1024), spps=(4, 16, 64, 256), max_depth=8):
    """The batched-rendering workload: n_jobs jobs, each on one of the given
    (scene, desc) pairs in turn, with a seeded orbit pose, width and height
    log-uniform in res_range, spp drawn from spps."""
    rng = np.random.default_rng(seed)
    jobs = []
    lo, hi = math.log(res_range[0]), math.log(res_range[1])
    for k in range(n_jobs):
        sc, desc = scenes_and_descs[k % len(scenes_and_descs)]
        W = int(round(math.exp(rng.uniform(lo, hi))))
        H = int(round(math.exp(rng.uniform(lo, hi))))
        jobs.append(dict(scene=sc, camera=orbit_camera(desc, rng, W, H), spp=int(rng.choice(spps)),
                         max_depth=max_depth, seed=int(rng.integers(1, 1 << 62))))
    return jobs


def nested_media(seed=11, g=0.7, width=1024, height=1024, spp=128, max_depth=16, sphere_res=224):
    """Volumetric scene: a sphere_res^2*2-triangle index-matched sphere holding
    a homogeneous medium (sigma_a ~ U[0.01,0.5]^3, sigma_s ~ U[0.5,4]^3),
    nested in a RoughDielectric sphere holding a second medium, over a
    Lambertian floor, lit by two area lights; Henyey-Greenstein phase."""
    rng = np.random.default_rng(seed)
    s = SceneDesc()
    s.name = "nested_media"
    med_in = s.add_medium(tuple(rng.uniform(0.01, 0.5, 3)), tuple(rng.uniform(0.5, 4.0, 3)),
                          capi.PHASE_HENYEY_GREENSTEIN, g)
    med_out = s.add_medium(tuple(rng.uniform(0.01, 0.1, 3)), tuple(rng.uniform(0.05, 0.3, 3)),
                           capi.PHASE_HENYEY_GREENSTEIN, g)
    P, I = uv_sphere(sphere_res, sphere_res)
    s.add_mesh(P * 0.8 + np.array([0, 1.2, 0]), I, -1, normals=P,
               interior_medium_id=med_in, exterior_medium_id=med_out)
    m_glass = s.add_material(capi.MAT_ROUGH_DIELECTRIC, (1, 1, 1), (1, 1, 1), roughness=0.1, eta=1.5)
    SP, SI = icosphere(4)
    s.add_mesh(SP * 1.15 + np.array([0, 1.2, 0]), SI, m_glass, normals=SP,
               interior_medium_id=med_out, exterior_medium_id=-1)
    m_floor = s.add_material(capi.MAT_LAMBERTIAN, tuple(rng.uniform(0.3, 0.8, 3)))
    Pf, If, Nf = quad((-6, 0, -6), (-6, 0, 6), (6, 0, 6), (6, 0, -6))
    s.add_mesh(Pf, If, m_floor, normals=Nf)
    m_black = s.add_material(capi.MAT_LAMBERTIAN, (0, 0, 0))
    for c, col in (((-2.5, 4.0, 1.5), (18.0, 14.0, 10.0)), ((2.0, 3.0, -2.0), (4.0, 6.0, 12.0))):
        cx, cy, cz = c
        Pl, Il, Nl = quad((cx - 0.4, cy, cz - 0.4), (cx + 0.4, cy, cz - 0.4), (cx + 0.4, cy, cz + 0.4),
                          (cx - 0.4, cy, cz + 0.4))
        lid = s.add_mesh(Pl, Il, m_black, normals=Nl)
        s.add_area_light(lid, col)
    s.camera.update(cam_to_world=look_at((0, 2.2, 5.5), (0, 1.0, 0), (0, 1, 0)), fov=45.0,
                    width=width, height=height, filter_type=capi.FILTER_BOX, filter_param=1.0)
    s.options.update(integrator=capi.INTEGRATOR_VOLPATH, samples_per_pixel=spp, max_depth=max_depth,
                     rr_depth=5)
    s.meta = dict(seed=seed, g=g, sphere_res=sphere_res)
    return s


def smoke_density(res, seed=3, octaves=4):
    """A seeded, wispy smoke-like density on a res^3 grid (values in [0, ~1]):
    octaves of trilinearly upsampled value noise, thresholded, with a
    spherical falloff so the grid's border holds nothing."""
    rng = np.random.default_rng(seed)
    n = res
    g = np.linspace(0, 1, n)
    acc = np.zeros((n, n, n))
    amp, base = 1.0, 4
    for _ in range(octaves):
        coarse = rng.uniform(-1, 1, (base + 1,) * 3)
        idx = g * base
        i0 = np.minimum(idx.astype(int), base - 1)
        f = idx - i0
        # separable trilinear upsampling
        a = coarse[i0] * (1 - f)[:, None, None] + coarse[i0 + 1] * f[:, None, None]
        a = a[:, i0] * (1 - f)[None, :, None] + a[:, i0 + 1] * f[None, :, None]
        a = a[:, :, i0] * (1 - f)[None, None, :] + a[:, :, i0 + 1] * f[None, None, :]
        acc += amp * a
        amp *= 0.5
        base *= 2
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
    falloff = np.clip(1 - (r / 0.5) ** 2, 0, 1)
    return (np.clip(acc + 0.2, 0, None) * falloff).astype(np.float32)


def hetvol_smoke(seed=3, width=768, height=576, spp=64, max_depth=16, res=128, scale=40.0, colored=False):
    """Heterogeneous-medium scene in the manner of the reference's
    scenes/volpath_test/hetvol.xml: a res^3 smoke density grid (times scale)
    with a constant albedo of 0.9 (or a seeded RGB albedo grid when colored),
    inside an index-matched unit box above a diffuse plane, lit by a sphere
    emitter of radiance 16; Henyey-Greenstein g = 0."""
    s = SceneDesc()
    s.name = "hetvol_smoke"
    lo, hi = (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)
    dens = grid_volume(smoke_density(res, seed), lo, hi, scale)
    if colored:
        rng = np.random.default_rng(seed + 1)
        alb = grid_volume(rng.uniform(0.6, 0.98, (8, 8, 8, 3)), lo, hi)
    else:
        alb = constant_volume((0.9, 0.9, 0.9))
    med = s.add_heterogeneous_medium(alb, dens, capi.PHASE_HENYEY_GREENSTEIN, 0.0)
    P, I, _ = box((0, 0, 0), (1.0, 1.0, 1.0), 0.0)
    s.add_mesh(P, I, -1, interior_medium_id=med, exterior_medium_id=-1)
    m_floor = s.add_material(capi.MAT_LAMBERTIAN, (0.2, 0.2, 0.3))
    Pf, If, Nf = quad((-4, -0.52, -4), (-4, -0.52, 4), (4, -0.52, 4), (4, -0.52, -4))
    s.add_mesh(Pf, If, m_floor, normals=Nf)
    m_black = s.add_material(capi.MAT_LAMBERTIAN, (0, 0, 0))
    lid = s.add_sphere((-0.8, 2.0, -1.0), 0.6, m_black)
    s.add_area_light(lid, (16.0, 16.0, 16.0))
    s.camera.update(cam_to_world=look_at((0.4, 0.5, 2.2), (0, 0, 0), (0, 1, 0)), fov=45.0,
                    width=width, height=height, filter_type=capi.FILTER_BOX, filter_param=1.0)
    s.options.update(integrator=capi.INTEGRATOR_VOLPATH, samples_per_pixel=spp, max_depth=max_depth,
                     rr_depth=5)
    s.meta = dict(seed=seed, res=res, scale=scale, colored=colored)
    return s


def sky_image(w=64, h=32, sun=(0.3, 0.25), seed=0):
    """A lat-long sky: a blue-to-white gradient with a bright sun texel block
    at (u, v) = sun, and a dark lower hemisphere."""
    rng = np.random.default_rng(seed)
    v = (np.arange(h) + 0.5) / h
    img = np.zeros((h, w, 3), np.float32)
    up = np.clip(1 - 2 * v, 0, 1)[:, None]
    img[:] = (0.3 + 0.7 * up)[..., None] * np.array([0.4, 0.6, 1.0]) + (1 - up)[..., None] * 0.05
    img *= 1 + 0.05 * rng.standard_normal((h, w, 1))
    su, sv = int(sun[0] * w), int(sun[1] * h)
    img[max(sv - 1, 0):sv + 1, max(su - 1, 0):su + 1] = [60.0, 50.0, 40.0]
    return np.clip(img, 0, None).astype(np.float32)


def textured_showcase(seed=5, width=96, height=72, spp=16, max_depth=6, envmap=True):
    """Textured materials under an image environment map: a checkerboard
    floor, spheres with image-textured colour, a RoughPlastic with a
    roughness texture, a DisneyBSDF with metallic and base-colour textures,
    a RoughDielectric with a coloured transmittance image, and a small area
    light."""
    rng = np.random.default_rng(seed)
    s = SceneDesc()
    s.name = "textured_showcase"
    chk = s.add_checkerboard((0.8, 0.8, 0.8), (0.1, 0.12, 0.15), uscale=8, vscale=8)
    img = rng.uniform(0.05, 0.95, (32, 32, 3)).astype(np.float32)
    img_tex = s.add_image_texture(img, uscale=1 / (2 * np.pi), vscale=1 / np.pi)  # sphere uv is in radians
    rough = s.add_image_texture(np.tile(np.linspace(0.05, 0.9, 16, dtype=np.float32), (16, 1)),
                                uscale=1 / (2 * np.pi), vscale=1 / np.pi)
    metal = s.add_checkerboard((1, 1, 1), (0, 0, 0), uscale=2 / np.pi, vscale=2 / np.pi)
    trans = s.add_image_texture(rng.uniform(0.5, 1.0, (8, 8, 3)).astype(np.float32), uscale=1 / np.pi,
                                vscale=1 / np.pi)
    floor = s.add_material(capi.MAT_LAMBERTIAN, textures={"color_a": chk})
    Pf = np.array([[-6, 0, -6], [6, 0, -6], [6, 0, 6], [-6, 0, 6]], np.float64)
    s.add_mesh(Pf, np.array([[0, 2, 1], [0, 3, 2]]), floor, uvs=(Pf[:, [0, 2]] + 6) / 12,
               normals=np.tile([0, 1, 0], (4, 1)))
    m_img = s.add_material(capi.MAT_LAMBERTIAN, textures={"color_a": img_tex})
    m_plastic = s.add_material(capi.MAT_ROUGH_PLASTIC, (0.6, 0.2, 0.1), (1, 1, 1), roughness=0.3, eta=1.5,
                               textures={"roughness": rough})
    m_disney = s.add_material(capi.MAT_DISNEY_BSDF, (0.9, 0.6, 0.3), roughness=0.35, eta=1.5, specular=0.5,
                              textures={"metallic": metal, "color_a": img_tex})
    m_glass = s.add_material(capi.MAT_ROUGH_DIELECTRIC, (1, 1, 1), (1, 1, 1), roughness=0.05, eta=1.5,
                             textures={"color_b": trans})
    # the spheres float 5 cm above the floor: a tangent contact would leave which surface a ray
    # meets first to rounding
    for mid, c in ((m_img, (-1.8, 0.75, 0.0)), (m_plastic, (-0.6, 0.75, 0.8)), (m_disney, (0.6, 0.75, -0.6)),
                   (m_glass, (1.8, 0.75, 0.4))):
        s.add_sphere(c, 0.7, mid)
    black = s.add_material(capi.MAT_LAMBERTIAN, (0, 0, 0))
    Pl, Il, Nl = quad((-0.5, 4, -0.5), (0.5, 4, -0.5), (0.5, 4, 0.5), (-0.5, 4, 0.5))
    lid = s.add_mesh(Pl, Il, black, normals=Nl)
    s.add_area_light(lid, (8.0, 8.0, 8.0))
    if envmap:
        s.add_envmap(sky_image(seed=seed), scale=1.0)
    s.camera.update(cam_to_world=look_at((0.5, 2.2, 5.5), (0, 0.5, 0), (0, 1, 0)), fov=50.0, width=width,
                    height=height)
    s.options.update(integrator=capi.INTEGRATOR_PATH, samples_per_pixel=spp, max_depth=max_depth, rr_depth=5)
    return s

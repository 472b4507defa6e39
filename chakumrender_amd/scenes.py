"""Seeded synthetic scenes for the per-sample radiance path.

Scene layout follows the reference's scene front end:
  * triangle order = object triangles, then the dummy degenerate triangle at
    (1e6, 1e6, 1e6) (prepareScene.cpp:241), then mesh-light triangles
    (LightSource::addMSource appends them, lightSources.h:197-203);
  * a light's radiance is col.maxnorm() * intensity (lightSources.h:12-16);
  * light triangles carry an emitter material with colour 0 and
    reflectance 0 (lightSources.h:133);
  * the camera uses the reference's cx/cy convention (prepareScene.cpp:24-33).

Every array is plain numpy so the same scene feeds the HIP library, the
oracle and the tests.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

MATERIAL_DTYPE = np.dtype(
    [
        ("brdf", "<u4"),
        ("light", "<u4"),
        ("ks", "<f4"),
        ("alpha", "<f4"),
        ("phong_exp", "<f4"),
        ("int_ior", "<f4"),
        ("ext_ior", "<f4"),
        ("reflectance", "<f4"),
        ("color", "<f4", (3,)),
        ("radiance", "<f4", (3,)),
    ]
)
assert MATERIAL_DTYPE.itemsize == 56

LIGHT_DTYPE = np.dtype([("first_tri", "<u4"), ("n_tris", "<u4")])
# chk_sphere (Sphere with its material index) and chk_point_source
SPHERE_DTYPE = np.dtype([("center", "<f8", (3,)), ("radius", "<f8"), ("mat", "<u4"), ("pad", "<u4")])
POINT_DTYPE = np.dtype([("pos", "<f8", (3,)), ("power", "<f8", (3,))])

# chk_brdf_type / chk_light_type (materialTypes.h:6-7)
BRDF_NONE, CLASSIC_PHONG, GGX, BECKMANN, PHONG, GGXAPPROX = range(6)
LIGHT_NONE, LIGHT_POINT, LIGHT_VOLUME, LIGHT_PLANAR, LIGHT_MESH = range(5)


@dataclass
class Camera:
    pos: tuple
    dir: tuple
    up: tuple
    fovx_deg: float
    fovy_deg: float
    width: int
    height: int

    def basis(self):
        """(pos, dir, cx, cy) in float64, prepareScene.cpp:21-33."""
        pos = np.asarray(self.pos, np.float64)
        d = np.asarray(self.dir, np.float64)
        d = d / np.linalg.norm(d)
        cx = np.cross(d, np.asarray(self.up, np.float64))
        cx = cx / np.linalg.norm(cx)
        cx = cx * (self.width / self.height) * math.tan(self.fovx_deg * 3.14159265359 / 360.0)
        cy = np.cross(cx, d)
        cy = cy / np.linalg.norm(cy) * math.tan(self.fovy_deg * 3.14159265359 / 360.0)
        return pos, d, cx, cy


@dataclass
class EnvMap:
    """An environment light EnvSource(image, w, h, axis, col, intensity)
    (lightSources.h:33-109): image (h, w, 3) float64 with row 0 at theta = 0
    about the world axis `axis` (0 x, 1 y, 2 z).  The reference's front end
    would add EnvMap(readImage("env1.ppm"), 1, (1, 1, 1), 2)
    (prepareScene.cpp:150-154, commented out there)."""
    image: np.ndarray
    axis: int = 1
    col: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0

    @property
    def radiance(self):
        """BaseSource: col.maxnorm() * intensity (lightSources.h:12-16)."""
        c = np.asarray(self.col, np.float64)
        return tuple(float(v) for v in c / c.max() * self.intensity)


@dataclass
class Scene:
    tris: np.ndarray          # (n, 9) float32: A, B, C
    tri_mat: np.ndarray       # (n,) uint32 material index per triangle
    mats: np.ndarray          # (m,) MATERIAL_DTYPE
    lights: np.ndarray        # (L,) LIGHT_DTYPE
    camera: Camera
    name: str = ""
    meta: dict = field(default_factory=dict)
    # the other primitives and sources of chk_scene_desc
    spheres: np.ndarray = field(default_factory=lambda: np.zeros(0, SPHERE_DTYPE))
    tri_lights: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    sphere_sources: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    point_sources: np.ndarray = field(default_factory=lambda: np.zeros(0, POINT_DTYPE))
    env: EnvMap | None = None

    @property
    def n_tris(self):
        return int(self.tris.shape[0])

    def with_camera(self, **kw):
        cam = Camera(**{**self.camera.__dict__, **kw})
        return replace(self, camera=cam, meta=dict(self.meta))


def _material(brdf, ks=0.0, alpha=0.0, phong_exp=0.0, int_ior=1.5, ext_ior=1.0,
              reflectance=1.0, color=(1, 1, 1), light=LIGHT_NONE, radiance=(0, 0, 0)):
    m = np.zeros((), MATERIAL_DTYPE)
    m["brdf"] = brdf
    m["light"] = light
    m["ks"] = ks
    m["alpha"] = alpha
    m["phong_exp"] = phong_exp
    m["int_ior"] = int_ior
    m["ext_ior"] = ext_ior
    m["reflectance"] = reflectance
    m["color"] = color
    m["radiance"] = radiance
    return m


def emitter_material(col, intensity):
    """TriLight material: radiance = col.maxnorm() * intensity, colour 0,
    reflectance 0 (lightSources.h:12-16,133)."""
    col = np.asarray(col, np.float64)
    rad = col / col.max() * intensity
    return _material(BRDF_NONE, reflectance=0.0, color=(0, 0, 0), light=LIGHT_MESH, radiance=rad)


def lambert_material(albedo):
    """Lambert lobe as a CLASSIC_PHONG material with ks = 0: brdf = 1/pi,
    pdf = cos/pi (materialTypes.cpp:12-15,92)."""
    return _material(CLASSIC_PHONG, ks=0.0, phong_exp=1.0, color=albedo)


def quad(c, u, v):
    """Two triangles (c-u-v, c+u-v, c+u+v) and (c-u-v, c+u+v, c-u+v)."""
    c, u, v = (np.asarray(a, np.float64) for a in (c, u, v))
    p0, p1, p2, p3 = c - u - v, c + u - v, c + u + v, c - u + v
    return np.array([np.r_[p0, p1, p2], np.r_[p0, p2, p3]])


def uv_sphere(center, radius, slices, stacks):
    """UV sphere with triangle-fan poles: 2*slices*(stacks-1) triangles."""
    center = np.asarray(center, np.float64)
    th = np.linspace(0, math.pi, stacks + 1)
    ph = np.linspace(0, 2 * math.pi, slices + 1)
    P = np.stack([np.outer(np.sin(th), np.cos(ph)), np.outer(np.cos(th), np.ones_like(ph)),
                  np.outer(np.sin(th), np.sin(ph))], -1) * radius + center
    # per quad (i, j): triangle (a, b, c) unless i is the top stack, then
    # (a, c, d) unless i is the bottom stack; quads in row-major order
    a, b = P[:-1, :-1], P[:-1, 1:]
    c, d = P[1:, 1:], P[1:, :-1]
    t1 = np.concatenate([a, b, c], -1)  # (stacks, slices, 9)
    t2 = np.concatenate([a, c, d], -1)
    both = np.stack([t1, t2], 2)  # (stacks, slices, 2, 9)
    keep = np.ones((stacks, slices, 2), bool)
    keep[0, :, 0] = False
    keep[stacks - 1, :, 1] = False
    return both[keep]


def _dummy_triangle():
    # prepareScene.cpp:241: Triangle(1e6, 1e6, 1e6, ...) converts each 1e6
    # through Vec(double x) into the vertex (1e6, 0, 0)
    return np.array([[1e6, 0, 0] * 3], np.float64)


def _assemble(obj_tris, obj_mat, mats, light_quads, camera, name, meta):
    """Object triangles, dummy triangle, then light triangles (one mesh light
    per quad)."""
    n_obj = obj_tris.shape[0]
    mats = list(mats)
    dummy_mat = len(mats)
    mats.append(_material(CLASSIC_PHONG, ks=0.5, phong_exp=1.0, reflectance=0.9, color=(0, 0, 0.9)))
    tris = [obj_tris, _dummy_triangle()]
    tri_mat = [obj_mat.astype(np.uint32), np.array([dummy_mat], np.uint32)]
    lights = []
    first = n_obj + 1
    for (qt, col, inten) in light_quads:
        mi = len(mats)
        mats.append(emitter_material(col, inten))
        tris.append(qt)
        tri_mat.append(np.full(qt.shape[0], mi, np.uint32))
        lights.append((first, qt.shape[0]))
        first += qt.shape[0]
    tris = np.concatenate(tris).astype(np.float32)
    tri_mat = np.concatenate(tri_mat)
    mats_arr = np.zeros(len(mats), MATERIAL_DTYPE)
    for i, m in enumerate(mats):
        mats_arr[i] = m
    lights_arr = np.array(lights, LIGHT_DTYPE)
    return Scene(tris, tri_mat, mats_arr, lights_arr, camera, name, meta)


def _sphere_res(n_tris):
    """(slices, stacks) with 2*slices*(stacks-1) close to n_tris."""
    slices = max(6, int(round(math.sqrt(n_tris / 2 * 1.2))))
    stacks = max(3, int(round(n_tris / (2 * slices))) + 1)
    return slices, stacks


def cornell(seed=0x5EED0001, sphere_tris=960, n_lights=1, width=256, height=256,
            light_size=0.5, intensity=15.0, light_rng=False):
    """Cornell-box-style scene: 5 axis-aligned quad walls of the unit box (10
    triangles, Lambert, RGB albedo U[0.2,0.8] from the seed), two UV spheres of
    `sphere_tris` triangles each (GGX, alpha 0.3, IOR 1.5/1 -> F0 0.04, ks 1),
    and `n_lights` quad area lights just under the ceiling.  With the
    defaults: 10 + 2*960 + 1 + 2 = 1,933 triangles, one 0.5 x 0.5 light of
    radiance (15,15,15)."""
    rng = np.random.default_rng(seed)
    walls = [
        ((0.5, 0.0, 0.5), (0.5, 0, 0), (0, 0, 0.5)),      # floor
        ((0.5, 1.0, 0.5), (0.5, 0, 0), (0, 0, 0.5)),      # ceiling
        ((0.5, 0.5, 0.0), (0.5, 0, 0), (0, 0.5, 0)),      # back
        ((0.0, 0.5, 0.5), (0, 0.5, 0), (0, 0, 0.5)),      # left
        ((1.0, 0.5, 0.5), (0, 0.5, 0), (0, 0, 0.5)),      # right
    ]
    mats = []
    obj, omat = [], []
    for w in walls:
        mats.append(lambert_material(rng.uniform(0.2, 0.8, 3)))
        q = quad(*w)
        obj.append(q)
        omat.append(np.full(2, len(mats) - 1))
    sl, st = (24, 21) if sphere_tris == 960 else _sphere_res(sphere_tris)
    ggx = _material(GGX, ks=1.0, alpha=0.3, int_ior=1.5, ext_ior=1.0, color=(1, 1, 1))
    mats.append(ggx)
    for c, r in (((0.3, 0.2, 0.35), 0.2), ((0.72, 0.25, 0.62), 0.22)):
        s = uv_sphere(c, r, sl, st)
        obj.append(s)
        omat.append(np.full(s.shape[0], len(mats) - 1))
    obj = np.concatenate(obj)
    omat = np.concatenate(omat)
    light_quads = []
    if n_lights == 1 and not light_rng:
        h = light_size / 2
        light_quads.append((quad((0.5, 0.999, 0.5), (h, 0, 0), (0, 0, h)), (1, 1, 1), intensity))
    else:
        g = int(math.ceil(math.sqrt(n_lights)))
        for i in range(n_lights):
            gx, gz = (i % g + 0.5) / g, (i // g + 0.5) / g
            h = rng.uniform(0.2, 1.0) * 0.5 / g
            col = rng.uniform(0.5, 1.0, 3)
            inten = rng.uniform(5.0, 20.0)
            light_quads.append((quad((gx, 0.999 - 1e-4 * i, gz), (h, 0, 0), (0, 0, h)), col, inten))
    cam = Camera((0.5, 0.5, 2.5), (0, 0, -1), (0, 1, 0), 36.0, 36.0, width, height)
    return _assemble(obj, omat, mats, light_quads, cam, "cornell",
                     {"seed": seed, "sphere_tris": int(2 * sl * (st - 1))})


# ------------------------------------------------------------------ icosphere
def _icosahedron():
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    return v, f


def geodesic_sphere(freq):
    """Class-I geodesic icosphere: each icosahedron face split into freq^2
    triangles, vertices projected to the unit sphere; 20*freq^2 triangles,
    (n, 3, 3) float64."""
    V, F = _icosahedron()
    out = []
    ii, jj = np.meshgrid(np.arange(freq + 1), np.arange(freq + 1), indexing="ij")
    for a, b, c in F:
        A, B, C = V[a], V[b], V[c]

        def P(i, j):
            return (A * (freq - i - j)[..., None] + B * i[..., None] + C * j[..., None]) / freq

        i = ii[:-1, :-1].ravel()
        j = jj[:-1, :-1].ravel()
        m = i + j < freq
        i1, j1 = i[m], j[m]
        out.append(np.stack([P(i1, j1), P(i1 + 1, j1), P(i1, j1 + 1)], 1))
        m2 = i + j < freq - 1
        i2, j2 = i[m2], j[m2]
        out.append(np.stack([P(i2 + 1, j2), P(i2 + 1, j2 + 1), P(i2, j2 + 1)], 1))
    T = np.concatenate(out)
    T /= np.linalg.norm(T, axis=2, keepdims=True)
    return T


def value_noise(p, seed, octaves=3, base_freq=2.0):
    """3-octave value noise in [-1, 1]: lattice values from a hash of the
    integer lattice point, trilinear smoothstep interpolation, octave
    amplitudes 1, 1/2, 1/4 normalised by their sum."""
    p = np.asarray(p, np.float64)
    total = np.zeros(p.shape[:-1])
    amp_sum = 0.0
    for o in range(octaves):
        f = base_freq * (2 ** o)
        q = p * f
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        w = fr * fr * (3 - 2 * fr)
        acc = 0.0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    c = i0 + np.array([dx, dy, dz])
                    h = (c[..., 0] * 73856093) ^ (c[..., 1] * 19349663) ^ (c[..., 2] * 83492791) ^ (seed + o * 0x9E3779B1)
                    h = (h * 0x2545F4914F6CDD1D) & 0xFFFFFFFFFFFF
                    h ^= h >> 23
                    val = (h & 0xFFFF) / 32767.5 - 1.0
                    wx = w[..., 0] if dx else 1 - w[..., 0]
                    wy = w[..., 1] if dy else 1 - w[..., 1]
                    wz = w[..., 2] if dz else 1 - w[..., 2]
                    acc = acc + val * wx * wy * wz
        amp = 0.5 ** o
        total += amp * acc
        amp_sum += amp
    return total / amp_sum


def icosphere_scene(seed=0x5EED0002, freq=71, n_lights=8, width=1920, height=1080,
                    amplitude=0.1):
    """Seeded procedural mesh: a class-I geodesic icosphere (20*freq^2
    triangles; freq 71 -> 100,820) displaced radially by 3-octave value noise
    of amplitude 0.1, on a 2-triangle ground plane.  Per-triangle material
    uniform over {GGX, Beckmann, Phong} with roughness U[0.05, 0.8] (Phong
    exponent 2/r^2 - 2), RGB albedo U[0.1, 0.9], ks 0.5.  `n_lights` quad
    lights with random pose on a hemisphere of radius 5, size U[0.2, 1.0],
    intensity U[5, 20]."""
    rng = np.random.default_rng(seed)
    T = geodesic_sphere(freq)
    r = 1.0 + amplitude * value_noise(T, seed & 0xFFFFFF)
    center = np.array([0.0, 1.15, 0.0])
    T = T * r[..., None] + center
    n = T.shape[0]
    kinds = rng.integers(0, 3, n)
    rough = rng.uniform(0.05, 0.8, n)
    albedo = rng.uniform(0.1, 0.9, (n, 3))
    mats = np.zeros(n + 1, MATERIAL_DTYPE)
    mats["brdf"][:n] = np.array([GGX, BECKMANN, PHONG])[kinds]
    mats["ks"][:n] = 0.5
    mats["alpha"][:n] = np.where(kinds == 2, 2.0 / rough ** 2 - 2.0, rough)
    mats["int_ior"][:n] = 1.5
    mats["ext_ior"][:n] = 1.0
    mats["reflectance"][:n] = 1.0
    mats["color"][:n] = albedo
    ground = lambert_material((0.5, 0.5, 0.5))
    mats[n] = ground
    obj = np.concatenate([T.reshape(n, 9), quad((0, 0, 0), (6, 0, 0), (0, 0, 6))])
    omat = np.concatenate([np.arange(n), [n, n]])
    light_quads = []
    for _ in range(n_lights):
        # uniform direction on the upper hemisphere, elevation >= 15 degrees
        while True:
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            if d[1] > math.sin(math.radians(15)):
                break
        c = center + 5.0 * d
        w = -d
        a = np.cross(w, [0.0, 1.0, 0.0] if abs(w[1]) < 0.9 else [1.0, 0.0, 0.0])
        a /= np.linalg.norm(a)
        b = np.cross(w, a)
        rot = rng.uniform(0, 2 * math.pi)
        u = math.cos(rot) * a + math.sin(rot) * b
        v = np.cross(w, u)
        h = rng.uniform(0.2, 1.0) / 2
        light_quads.append((quad(c, u * h, v * h), (1, 1, 1), rng.uniform(5.0, 20.0)))
    mats_list = [m for m in mats]
    cam = Camera((0.0, 2.0, 4.6), (0.0, -0.22, -1.0), (0, 1, 0), 40.0, 40.0, width, height)
    return _assemble(obj, omat, mats_list, light_quads, cam, "icosphere",
                     {"seed": seed, "freq": freq})


def cornell_large(seed=0x5EED0005, total_tris=2_000_000, n_lights=32, width=1920, height=1080):
    """Config-1-style scene scaled to ~`total_tris` triangles (two spheres of
    ~total_tris/2 each) with `n_lights` quad lights of random size, colour and
    intensity in a grid under the ceiling."""
    return cornell(seed=seed, sphere_tris=total_tris // 2, n_lights=n_lights, width=width,
                   height=height, light_rng=True)


def batch_workload(seed=0x5EED0004, n_requests=64, n_scenes=16, min_tris=1_000,
                   max_tris=500_000, max_lights=16, sizes=(64, 1024), spp=(1, 4, 16, 64, 256),
                   max_depth=4):
    """Config-4 serving workload: a pool of `n_scenes` Cornell-generator
    scenes of log-uniform [min_tris, max_tris] sphere triangles and 1..max_lights
    lights, and `n_requests` requests drawing a scene from the pool, each
    side log-uniform in `sizes`, spp from `spp`.  Returns (scenes, requests),
    each request a dict(scene=index, width, height, spp, max_depth, seed)."""
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(n_scenes):
        tris = int(np.exp(rng.uniform(np.log(min_tris), np.log(max_tris))))
        nl = int(rng.integers(1, max_lights + 1))
        pool.append(cornell(seed=0x5EED0400 + k, sphere_tris=max(tris // 2, 8), n_lights=nl,
                            light_rng=nl > 1))
    reqs = []
    for i in range(n_requests):
        w = int(np.exp(rng.uniform(np.log(sizes[0]), np.log(sizes[1]))))
        h = int(np.exp(rng.uniform(np.log(sizes[0]), np.log(sizes[1]))))
        n = int(rng.choice(list(spp)))
        reqs.append(dict(scene=int(rng.integers(0, len(pool))), width=w, height=h, spp=n,
                         max_depth=max_depth, seed=0x5EED0400 + i))
    return pool, reqs


# ------------------------------------------------------------------ direct-lighting room
def direct_room(seed=0x5EED0006, width=64, height=64, reflectance=None, fan=True, panel=True,
                light_scale=1.0, light_intensity=None):
    """A unit room for the shadeDI estimators: Lambert floor, GGX back wall,
    Beckmann left wall, Phong right wall, a classic-Phong panel hanging under
    the ceiling (it shadows part of the floor), and two mesh lights under the
    ceiling: a 2-triangle quad and a 3-triangle fan of unequal areas (so the
    reference's 300-bin inverse CDF is not exact).  `reflectance` overrides
    every surface's reflectance; fan=False makes the second light a quad too,
    panel=False drops the panel (nothing then occludes a light);
    light_scale shrinks both lights about their centres; light_intensity
    replaces both intensities (1.0: radiance.maxnorm() is the radiance)."""
    rng = np.random.default_rng(seed)

    def R(r):
        return r if reflectance is None else reflectance

    mats = [
        lambert_material(rng.uniform(0.3, 0.8, 3)),
        _material(GGX, ks=0.5, alpha=0.3, reflectance=R(0.9), color=rng.uniform(0.4, 0.9, 3)),
        _material(BECKMANN, ks=0.5, alpha=0.2, reflectance=R(1.0), color=rng.uniform(0.4, 0.9, 3)),
        _material(PHONG, ks=0.333, alpha=2.0 / 0.3 ** 2 - 2.0, reflectance=R(0.8),
                  color=rng.uniform(0.4, 0.9, 3)),
        _material(CLASSIC_PHONG, ks=0.5, phong_exp=20.0, reflectance=R(0.9), color=(0.8, 0.8, 0.8)),
    ]
    if reflectance is not None:
        mats[0]["reflectance"] = reflectance
    obj = [quad((0.5, 0.0, 0.5), (0.5, 0, 0), (0, 0, 0.5)),   # floor
           quad((0.5, 0.5, 0.0), (0.5, 0, 0), (0, 0.5, 0)),   # back
           quad((0.0, 0.5, 0.5), (0, 0.5, 0), (0, 0, 0.5)),   # left
           quad((1.0, 0.5, 0.5), (0, 0.5, 0), (0, 0, 0.5)),   # right
           quad((0.45, 0.6, 0.5), (0.15, 0, 0), (0, 0, 0.12))]  # panel
    omat = np.repeat(np.arange(5), 2)
    if not panel:
        obj, omat = obj[:4], omat[:8]
    c = np.array([0.7, 0.999, 0.6])
    k = float(light_scale)
    if fan:
        second = np.array([np.r_[c, c + k * np.array((0.12, 0, -0.05)), c + k * np.array((0.02, 0, 0.1))],
                           np.r_[c, c + k * np.array((0.02, 0, 0.1)), c + k * np.array((-0.1, 0, 0.04))],
                           np.r_[c, c + k * np.array((-0.1, 0, 0.04)), c + k * np.array((-0.03, 0, -0.03))]])
    else:
        second = quad((0.7, 0.999, 0.6), (0.08 * k, 0, 0), (0, 0, 0.07 * k))
    i0, i1 = (12.0, 20.0) if light_intensity is None else (light_intensity, light_intensity)
    lights = [(quad((0.3, 0.999, 0.35), (0.12 * k, 0, 0), (0, 0, 0.1 * k)), (1.0, 0.9, 0.7), i0),
              (second, (0.6, 0.8, 1.0), i1)]
    cam = Camera((0.5, 0.5, 2.2), (0, 0, -1), (0, 1, 0), 40.0, 40.0, width, height)
    return _assemble(np.concatenate(obj), omat, mats, lights, cam, "direct_room", {"seed": seed})


def sources_room(seed=0x5EED0008, width=64, height=64, mesh=True, sphere_source=True,
                 tri_light=True, points=True, spheres=True):
    """direct_room with every other primitive and source kind of the
    reference: two spheres on the floor (GGX and Phong), a sphere source
    (LightSource::addSSource: the source is also a visible VOLUME sphere), a
    triangle light (TriLight, PLANAR) on the left wall and two point sources
    (PointSource).  mesh=False drops the two mesh lights."""
    sc = direct_room(seed=seed, width=width, height=height)
    rng = np.random.default_rng(seed)
    mats = list(sc.mats)
    cut = sc.n_tris if mesh else int(sc.lights["first_tri"].min())  # lights are the tail
    tris = [sc.tris[:cut]]
    tri_mat = [sc.tri_mat[:cut]]
    lights = sc.lights if mesh else np.zeros(0, LIGHT_DTYPE)
    tl = []
    if tri_light:
        mi = len(mats)
        m = emitter_material((0.9, 1.0, 0.8), 9.0)
        m["light"] = LIGHT_PLANAR
        mats.append(m)
        tl.append(sum(t.shape[0] for t in tris))
        tris.append(np.array([[0.001, 0.75, 0.3, 0.001, 0.85, 0.45, 0.001, 0.7, 0.5]], np.float32))
        tri_mat.append(np.array([mi], np.uint32))
    sph = []
    ss = []
    if spheres:
        g = len(mats)
        mats.append(_material(GGX, ks=0.6, alpha=0.25, reflectance=0.95,
                              color=rng.uniform(0.4, 0.9, 3)))
        mats.append(_material(PHONG, ks=0.4, alpha=30.0, reflectance=0.9,
                              color=rng.uniform(0.4, 0.9, 3)))
        sph.append(((0.3, 0.15, 0.35), 0.15, g))
        sph.append(((0.74, 0.12, 0.3), 0.12, g + 1))
    if sphere_source:
        mi = len(mats)
        col = np.array([1.0, 0.85, 0.6])
        m = _material(BRDF_NONE, reflectance=0.0, color=(0, 0, 0), light=LIGHT_VOLUME,
                      radiance=col / col.max() * 14.0)
        mats.append(m)
        ss.append(len(sph))
        sph.append(((0.78, 0.82, 0.62), 0.05, mi))
    pts = []
    if points:
        pts.append(((0.5, 0.9, 0.8), (0.9, 0.9, 0.9)))
        pts.append(((0.2, 0.7, 0.25), (0.6, 0.4, 0.2)))
    mats_arr = np.zeros(len(mats), MATERIAL_DTYPE)
    for i, m in enumerate(mats):
        mats_arr[i] = m
    sph_arr = np.zeros(len(sph), SPHERE_DTYPE)
    for i, (c, r, m) in enumerate(sph):
        sph_arr[i]["center"] = c
        sph_arr[i]["radius"] = r
        sph_arr[i]["mat"] = m
    pt_arr = np.zeros(len(pts), POINT_DTYPE)
    for i, (p, w) in enumerate(pts):
        pt_arr[i]["pos"] = p
        pt_arr[i]["power"] = w
    return Scene(np.concatenate(tris).astype(np.float32), np.concatenate(tri_mat).astype(np.uint32),
                 mats_arr, np.asarray(lights, LIGHT_DTYPE), sc.camera, "sources_room",
                 {"seed": seed}, spheres=sph_arr, tri_lights=np.array(tl, np.uint32),
                 sphere_sources=np.array(ss, np.uint32), point_sources=pt_arr)


def sky_image(seed=0x5EED000A, width=64, height=32, sun=(0.35, 0.6)):
    """A small 8-bit sky for an environment light: a blue gradient above the
    horizon, darker ground below and a saturated sun disc at (theta, phi) =
    sun * (pi, 2 pi); values k / 255 like a P6 file's (readImage,
    ppm.cpp:106-110).  (height, width, 3) float64."""
    rng = np.random.default_rng(seed)
    th = (np.arange(height) + 0.5) / height * math.pi
    ph = (np.arange(width) + 0.5) / width * 2 * math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    up = np.cos(T)
    sky = np.stack([0.35 + 0.25 * up, 0.5 + 0.25 * up, 0.9 + 0.1 * up], -1)
    ground = np.stack([0.25, 0.2, 0.15], -1) * (0.6 + 0.4 * rng.random((height, width, 1)))
    img = np.where(up[..., None] > 0, sky, ground)
    st, sp = sun[0] * math.pi, sun[1] * 2 * math.pi
    cosd = np.sin(T) * math.sin(st) * np.cos(P - sp) + np.cos(T) * math.cos(st)
    img = np.where((cosd > math.cos(0.25))[..., None], 1.0, img)
    return np.round(np.clip(img, 0, 1) * 255) / 255.0


def sphere_points(scene, k, seed=0x5EED0009, eye=(0.5, 0.55, 0.95), spheres=None):
    """k shading points on the scene's non-emitter spheres (or on `spheres`)
    seen from `eye`: (x, n, d, prim) like shading_points, prim = n_tris +
    sphere index and n the outward normal (the sphere hit's normal turned
    against d, shader.cpp:41-47)."""
    rng = np.random.default_rng(seed)
    S = scene.spheres
    if spheres is None:
        spheres = np.nonzero(scene.mats["light"][S["mat"]] == LIGHT_NONE)[0]
    sid = rng.choice(np.asarray(spheres), k)
    c, r = S["center"][sid], S["radius"][sid][:, None]
    e = np.asarray(eye, np.float64)
    x = np.empty((k, 3))
    for i in range(k):
        while True:
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            if np.dot(u, e - c[i]) > 0.2 * np.linalg.norm(e - c[i]):
                break
        x[i] = c[i] + r[i] * u
    n = x - c
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = x - e
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return x, n, d, (scene.n_tris + sid).astype(np.uint32)


def shading_points(scene, k, seed=0x5EED0007, eye=(0.5, 0.55, 0.95), tris=None):
    """k shading points on the scene's non-emitter triangles (or on `tris`)
    seen from `eye`: (x, n facing the eye, d = unit direction eye -> x, prim),
    FP64 from the FP32 triangles, the arguments the reference's estimators
    receive from shadeDirectOnly."""
    rng = np.random.default_rng(seed)
    T = scene.tris.astype(np.float64)
    if tris is None:
        tris = np.nonzero(scene.mats["light"][scene.tri_mat] == LIGHT_NONE)[0]
        tris = tris[np.abs(T[tris]).max(1) < 1e5]  # not the dummy triangle
    prim = rng.choice(np.asarray(tris), k).astype(np.uint32)
    A, B, C = T[prim, 0:3], T[prim, 3:6], T[prim, 6:9]
    u = rng.uniform(0.02, 0.98, (k, 2))
    flip = u.sum(1) > 1
    u[flip] = 1 - u[flip]
    x = A + (B - A) * u[:, :1] + (C - A) * u[:, 1:]
    d = x - np.asarray(eye, np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    n = np.cross(C - A, B - A)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n = np.where((n * d).sum(1, keepdims=True) > 0, -n, n)
    return x, n, d, prim


# ------------------------------------------------------------------ LTC points
def ltc_points(n, seed=0x5EED0003, n_quads=16, alpha_range=(0.02, 1.0)):
    """Shading points for the isolated LTC kernel: positions U[-1,1]^3,
    normals uniform on S^2, view directions uniform on the upper hemisphere
    of the normal, GGX roughness U[0.02, 1.0]; one shared set of `n_quads`
    quads with corners in a unit box centred at distance U[2,6] along a
    uniform direction.  Returns dict of float32 arrays."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v = np.where((v * nrm).sum(1, keepdims=True) < 0, -v, v)
    alpha = rng.uniform(*alpha_range, n)
    quads = []
    rad = []
    for _ in range(n_quads):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        c = d * rng.uniform(2, 6)
        # planar quad inside a unit box around c: random in-plane axes
        w = rng.normal(size=3)
        w /= np.linalg.norm(w)
        a = np.cross(w, rng.normal(size=3))
        a /= np.linalg.norm(a)
        b = np.cross(w, a)
        ha, hb = rng.uniform(0.15, 0.5, 2)
        quads.append(np.stack([c - a * ha - b * hb, c + a * ha - b * hb, c + a * ha + b * hb,
                               c - a * ha + b * hb]))
        rad.append(rng.uniform(1.0, 10.0, 3))
    return {
        "x": x.astype(np.float32), "n": nrm.astype(np.float32), "v": v.astype(np.float32),
        "alpha": alpha.astype(np.float32), "quads": np.array(quads, np.float32),
        "radiance": np.array(rad, np.float32),
    }


def quads_to_tris(quads, radiance):
    """Each quad (p0,p1,p2,p3) -> triangles (p0,p1,p2), (p0,p2,p3), the order
    in which a two-triangle MESHSOURCE quad is listed; radiance per triangle."""
    q = np.asarray(quads, np.float64)
    t1 = np.concatenate([q[:, 0], q[:, 1], q[:, 2]], 1)
    t2 = np.concatenate([q[:, 0], q[:, 2], q[:, 3]], 1)
    tris = np.stack([t1, t2], 1).reshape(-1, 9)
    rad = np.repeat(np.asarray(radiance, np.float64), 2, axis=0)
    return tris, rad

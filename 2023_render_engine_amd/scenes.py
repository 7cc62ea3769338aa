"""Scene descriptions of the benchmark workloads, as plain arrays.

A SceneDesc holds triangles (vertices, indices, per-triangle material),
analytic spheres, materials, point lights and the camera.  ``to_rt`` builds
the library scene (rt_scene_* through the C ABI); the tests hand the same
arrays to the CPU oracle.  Generators are deterministic (numpy PCG64 with
fixed seeds).

cornell():      the Cornell-box workload (256x256, 16 spp, depth 8).
procedural():   heightfield + icospheres + 8 point lights + a sphere light
                (1920x1080, 64 spp, depth 16), scalable for tests.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class SceneDesc:
    xyz: np.ndarray                 # (nv, 3) float32
    idx: np.ndarray                 # (nt, 3) uint32
    tri_mat: np.ndarray             # (nt,) uint32
    materials: list                 # dicts: type, kd, ior, roughness, eta, k, intensity
    spheres: list = field(default_factory=list)   # (center, radius, material, is_light)
    lights: list = field(default_factory=list)    # (pos, color, intensity)
    camera: dict = field(default_factory=dict)    # pos, lookat, up, cts, frame_y, width, height
    spp: int = 16
    max_depth: int = 8
    rr_depth: int = 3
    seed: int = 1
    planes: list = field(default_factory=list)    # (corner, axis_x, axis_y, normal, material): Plane shapes
    textures: list = field(default_factory=list)  # (height, width, 3) uint8 arrays, row y = pixelMatrix[y]
    tri_uv: object = None                         # (nt, 6) float32 corner texture coordinates, or None
    tri_nrm: object = None                        # (nt, 9) float32 corner normals (NaN row: none), or None

    @property
    def width(self):
        return int(self.camera["width"])

    @property
    def height(self):
        return int(self.camera["height"])

    def material_structs(self):
        from . import Material, f3  # noqa: F401  (package-relative)
        out = []
        for m in self.materials:
            out.append(Material(m["type"], f3(m.get("kd", (0, 0, 0))), m.get("ior", 1.5), m.get("roughness", 0.0),
                                f3(m.get("eta", (1, 1, 1))), f3(m.get("k", (0, 0, 0))), m.get("intensity", 0.0),
                                f3(m.get("ke", (0, 0, 0))), int(m.get("texture", 0))))
        return out

    def camera_struct(self):
        from . import Camera, f3
        c = self.camera
        return Camera(f3(c["pos"]), f3(c["lookat"]), f3(c["up"]), f3(c["cts"]), c["frame_y"], c["width"], c["height"])

    def to_rt(self, ctx):
        """Build and commit the library scene."""
        from . import Scene
        return Scene(ctx, self)


def _mat(type_, kd=(0, 0, 0), ior=1.5, roughness=0.0, eta=(1, 1, 1), k=(0, 0, 0), intensity=0.0, ke=(0, 0, 0),
         texture=0):
    return dict(type=type_, kd=tuple(float(x) for x in kd), ior=float(ior), roughness=float(roughness),
                eta=tuple(eta), k=tuple(k), intensity=float(intensity), ke=tuple(float(x) for x in ke),
                texture=int(texture))


def quad(p0, p1, p2, p3):
    """Two triangles (p0 p1 p2), (p0 p2 p3)."""
    v = np.array([p0, p1, p2, p3], np.float32)
    return v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def cornell(width=256, height=256, spp=16, max_depth=8, seed=1) -> SceneDesc:
    """Cornell box of 5 axis-aligned quads (10 triangles) in [-0.5, 0.5]^3,
    open towards -x where the camera of cornellBoxGuion.obejota sits
    (p = (-1.75, 0, 0), lookat origin, up +z, cts (0, 0, -0.8), frameY 0.6),
    a Lambertian sphere (0.73) and a glass sphere (IOR 1.5), radius 0.2,
    resting on the floor, and a point light of intensity 10 just below the
    ceiling centre."""
    LAMB, DIEL, LIGHT = 0, 1, 4
    mats = [_mat(LAMB, (0.73, 0.73, 0.73)), _mat(LAMB, (0.63, 0.06, 0.05)), _mat(LAMB, (0.14, 0.45, 0.09)),
            _mat(DIEL, (1, 1, 1), ior=1.5)]
    a = 0.5
    quads = [
        (quad((-a, -a, -a), (a, -a, -a), (a, a, -a), (-a, a, -a)), 0),   # floor z = -0.5
        (quad((-a, -a, a), (-a, a, a), (a, a, a), (a, -a, a)), 0),       # ceiling z = +0.5
        (quad((a, -a, -a), (a, -a, a), (a, a, a), (a, a, -a)), 0),       # back x = +0.5
        (quad((-a, -a, -a), (-a, -a, a), (a, -a, a), (a, -a, -a)), 1),   # left y = -0.5 (red)
        (quad((-a, a, -a), (a, a, -a), (a, a, a), (-a, a, a)), 2),       # right y = +0.5 (green)
    ]
    xyz, idx, tm = [], [], []
    for (v, f), m in quads:
        idx.append(f + len(xyz) * 4)
        xyz.append(v)
        tm += [m, m]
    xyz = np.concatenate(xyz).astype(np.float32)
    idx = np.concatenate(idx).astype(np.uint32)
    spheres = [((0.2, -0.2, -0.3), 0.2, 0, 0), ((-0.1, 0.2, -0.3), 0.2, 3, 0)]
    lights = [((0.0, 0.0, 0.5 - 1e-3), (1, 1, 1), 10.0)]
    cam = dict(pos=(-1.75, 0, 0), lookat=(0, 0, 0), up=(0, 0, 1), cts=(0, 0, -0.8), frame_y=0.6,
               width=width, height=height)
    return SceneDesc(xyz, idx, np.array(tm, np.uint32), mats, spheres, lights, cam, spp, max_depth, 3, seed)


def icosphere(subdiv: int):
    """Unit icosphere: 20 * 4**subdiv triangles."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts, np.float64), np.array(faces, np.uint32)


def value_noise(x, y, lattice):
    """Bilinear value noise with smoothstep weights on a periodic lattice."""
    n = lattice.shape[0]
    xi, yi = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64)
    fx, fy = x - xi, y - yi
    sx, sy = fx * fx * (3 - 2 * fx), fy * fy * (3 - 2 * fy)
    x0, y0 = xi % n, yi % n
    x1, y1 = (x0 + 1) % n, (y0 + 1) % n
    a = lattice[x0, y0] * (1 - sx) + lattice[x1, y0] * sx
    b = lattice[x0, y1] * (1 - sx) + lattice[x1, y1] * sx
    return a * (1 - sy) + b * sy


def procedural(grid=1024, n_spheres=64, subdiv=5, width=1920, height=1080, spp=64, max_depth=16,
               seed=2, n_point_lights=8) -> SceneDesc:
    """Heightfield of grid x grid vertices on [0,1]^2 displaced by 4 octaves
    of seeded value noise (amplitude 0.1) plus n_spheres icospheres
    (subdiv subdivisions) at uniform positions with radii U[0.02, 0.1];
    per-sphere materials uniform over {Lambertian albedo U[0.2,0.8]^3,
    Trowbridge-Reitz (F = 1) with alpha U[0.05, 0.5] (roughness = alpha^2),
    dielectric IOR 1.5}; a Lambertian ground; 8 point lights and one sphere
    area light."""
    LAMB, DIEL, GGXB, LIGHT = 0, 1, 3, 4
    rng = np.random.default_rng(seed)
    lattice = rng.uniform(-1, 1, (64, 64))
    g = np.linspace(0.0, 1.0, grid)
    X, Y = np.meshgrid(g, g, indexing="ij")
    f0 = 4.0

    def hfield(x, y):
        z = np.zeros_like(x)
        for o in range(4):
            z += 2.0 ** (-o) * value_noise(x * f0 * 2 ** o, y * f0 * 2 ** o, lattice)
        return 0.1 * z

    Z = hfield(X, Y)
    hv = np.stack([X, Y, Z], -1).reshape(-1, 3)
    ii = np.arange(grid - 1)
    a = (ii[:, None] * grid + ii[None, :]).reshape(-1)
    quads = np.stack([a, a + grid, a + grid + 1, a + 1], 1)
    hf = np.concatenate([quads[:, [0, 1, 2]], quads[:, [0, 2, 3]]]).astype(np.uint32)
    mats = [_mat(LAMB, rng.uniform(0.2, 0.8, 3))]
    tri_mat = [np.zeros(hf.shape[0], np.uint32)]
    xyz = [hv]
    idx = [hf]
    nv = hv.shape[0]
    iv, if_ = icosphere(subdiv)
    for s in range(n_spheres):
        c = np.array([rng.uniform(0.05, 0.95), rng.uniform(0.05, 0.95), 0.0])
        r = rng.uniform(0.02, 0.1)
        c[2] = float(hfield(np.array([c[0]]), np.array([c[1]]))[0]) + r
        kind = rng.integers(0, 3)
        if kind == 0:
            mats.append(_mat(LAMB, rng.uniform(0.2, 0.8, 3)))
        elif kind == 1:
            alpha = rng.uniform(0.05, 0.5)
            mats.append(_mat(GGXB, rng.uniform(0.2, 0.8, 3), roughness=alpha * alpha))
        else:
            mats.append(_mat(DIEL, (1, 1, 1), ior=1.5))
        xyz.append(iv * r + c)
        idx.append(if_ + nv)
        nv += iv.shape[0]
        tri_mat.append(np.full(if_.shape[0], len(mats) - 1, np.uint32))
    xyz = np.concatenate(xyz).astype(np.float32)
    idx = np.concatenate(idx).astype(np.uint32)
    tri_mat = np.concatenate(tri_mat)
    lights = []
    for _ in range(n_point_lights):
        p = (rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.5, 0.9))
        lights.append((p, tuple(rng.uniform(0.6, 1.0, 3)), 0.15))
    mats.append(_mat(LIGHT, (1.0, 0.9, 0.8), intensity=6.0))
    spheres = [((0.5, 0.5, 1.1), 0.12, len(mats) - 1, 1)]
    cam = dict(pos=(0.5, -0.55, 0.75), lookat=(0.5, 0.45, 0.0), up=(0, 0, 1), cts=(0, 0, -1.0), frame_y=0.9,
               width=width, height=height)
    return SceneDesc(xyz, idx, tri_mat, mats, spheres, lights, cam, spp, max_depth, 3, seed)


def procedural_sized(target_tris: int, size: int, seed: int, spp=32, max_depth=12) -> SceneDesc:
    """A scene from the procedural generator with about target_tris
    triangles at size x size pixels: 8 icospheres take ~40% of the budget
    (subdivision level rounded to the nearest power of 4), the heightfield
    the rest."""
    import math
    sph_budget = 0.4 * target_tris
    subdiv = int(min(5, max(0, round(math.log(max(sph_budget / (8 * 20), 1.0), 4)))))
    rest = max(2.0, target_tris - 8 * 20 * 4 ** subdiv)
    grid = int(max(8, round(math.sqrt(rest / 2.0)) + 1))
    return procedural(grid=grid, n_spheres=8, subdiv=subdiv, width=size, height=size, spp=spp,
                      max_depth=max_depth, seed=seed)


def batch_suite(n=32, seed=2024, spp=32, max_depth=12):
    """The batched-rendering workload: n procedural scenes with triangle
    counts log-uniform in [1e4, 5e5] and resolutions uniform over
    {256^2, 512^2, 1024^2}; scene i has generator seed 1000 + i (also its
    PCG seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        tris = int(round(math.exp(rng.uniform(math.log(1e4), math.log(5e5)))))
        size = int(rng.choice([256, 512, 1024]))
        out.append(procedural_sized(tris, size, 1000 + i, spp=spp, max_depth=max_depth))
    return out

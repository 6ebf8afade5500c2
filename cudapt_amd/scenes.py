"""Procedural scenes and a minimal OBJ reader, as flat numpy arrays.

A scene is a dict with
  pos   float32 [nv, 3]   vertex positions
  nrm   float32 [nv, 3]   vertex normals (RTVertex::norm, include/raytracer.h:58-64)
  idx   uint32  [nt, 3]   triangle vertex indices
  mat   uint32  [nt]      material index per triangle (RTTriangle::matInd)
  mats  numpy structured array of MATERIAL_DTYPE, RTMaterial field order
        (include/raytracer.h:82-100)
  camera dict(pos, dir, up, fov) with fov in radians

The reference builds scenes by assimp import (src/raytracer.cpp:238-282); the
workloads here are procedural and seeded, so every run sees the same bytes.
Scene-level randomness uses numpy's PCG64 generator; per-pixel randomness is the
renderer's own PCG32 stream.
"""
from __future__ import annotations

import math
import numpy as np

MATERIAL_DTYPE = np.dtype([
    ("diffuse", "<f4", 3), ("emissive", "<f4", 3),
    ("diffuseTexId", "<i4"), ("normalTexId", "<i4"), ("emissiveTexId", "<i4"), ("matType", "<i4"),
    ("transparency", "<f4"), ("subsurface", "<f4"), ("specular", "<f4"), ("metallic", "<f4"),
    ("roughness", "<f4"), ("anisotropic", "<f4"), ("sheen", "<f4"), ("sheenTint", "<f4"),
    ("clearcoat", "<f4"), ("clearcoatGloss", "<f4"),
])
assert MATERIAL_DTYPE.itemsize == 80


def material(diffuse=(0.8, 0.8, 0.8), emissive=(0.0, 0.0, 0.0), transparency=0.0, specular=0.0,
             metallic=0.0, roughness=1.0, sheen=0.0, clearcoat=0.0):
    m = np.zeros((), MATERIAL_DTYPE)
    m["diffuse"] = diffuse
    m["emissive"] = emissive
    m["diffuseTexId"] = m["normalTexId"] = m["emissiveTexId"] = -1
    m["transparency"] = transparency
    m["specular"] = specular
    m["metallic"] = metallic
    m["roughness"] = roughness
    m["sheen"] = sheen
    m["clearcoat"] = clearcoat
    return m


class _Builder:
    def __init__(self):
        self.pos, self.nrm, self.idx, self.mat, self.mats = [], [], [], [], []
        self.nv = 0

    def add_material(self, m) -> int:
        self.mats.append(m)
        return len(self.mats) - 1

    def add_quad(self, a, b, c, d, normal, m):
        """Quad a-b-c-d (counter-clockwise seen from `normal`) as two triangles."""
        base = self.nv
        self.pos += [a, b, c, d]
        self.nrm += [normal] * 4
        self.idx += [(base, base + 1, base + 2), (base, base + 2, base + 3)]
        self.mat += [m, m]
        self.nv += 4

    def add_mesh(self, pos, nrm, idx, m):
        base = self.nv
        self.pos += [tuple(p) for p in pos]
        self.nrm += [tuple(n) for n in nrm]
        self.idx += [tuple(int(i) + base for i in t) for t in idx]
        self.mat += [m] * len(idx)
        self.nv += len(pos)

    def add_box(self, center, half, yaw, m):
        """Axis-aligned box rotated by `yaw` about +y; flat normals, 12 triangles."""
        c, s = math.cos(yaw), math.sin(yaw)
        rot = lambda v: (c * v[0] + s * v[2], v[1], -s * v[0] + c * v[2])
        hx, hy, hz = half
        faces = [  # (normal, four corners in local space, ccw from outside)
            ((1, 0, 0), [(hx, -hy, hz), (hx, -hy, -hz), (hx, hy, -hz), (hx, hy, hz)]),
            ((-1, 0, 0), [(-hx, -hy, -hz), (-hx, -hy, hz), (-hx, hy, hz), (-hx, hy, -hz)]),
            ((0, 1, 0), [(-hx, hy, hz), (hx, hy, hz), (hx, hy, -hz), (-hx, hy, -hz)]),
            ((0, -1, 0), [(-hx, -hy, -hz), (hx, -hy, -hz), (hx, -hy, hz), (-hx, -hy, hz)]),
            ((0, 0, 1), [(-hx, -hy, hz), (hx, -hy, hz), (hx, hy, hz), (-hx, hy, hz)]),
            ((0, 0, -1), [(hx, -hy, -hz), (-hx, -hy, -hz), (-hx, hy, -hz), (hx, hy, -hz)]),
        ]
        for n, corners in faces:
            pts = [tuple(np.add(rot(p), center)) for p in corners]
            self.add_quad(*pts, rot(n), m)

    def build(self, camera):
        mats = np.zeros(len(self.mats), MATERIAL_DTYPE)
        for i, m in enumerate(self.mats):
            mats[i] = m
        return dict(pos=np.asarray(self.pos, np.float32).reshape(-1, 3),
                    nrm=np.asarray(self.nrm, np.float32).reshape(-1, 3),
                    idx=np.asarray(self.idx, np.uint32).reshape(-1, 3),
                    mat=np.asarray(self.mat, np.uint32), mats=mats, camera=camera)


def cornell(seed: int = 0xC1, light_radiance: float = 15.0):
    """Cornell box, 5 quads + 2 yaw-rotated cubes + 1 ceiling light = 36 triangles.

    Lambertian albedo U(0.2, 0.8) per RGB, expressed in the reference material
    model as specular = metallic = transparency = 0 (F = 0, reflProb = 0).
    """
    rng = np.random.default_rng(seed)
    b = _Builder()
    alb = lambda: material(diffuse=tuple(rng.uniform(0.2, 0.8, 3)))
    floor, ceil, back, left, right = (b.add_material(alb()) for _ in range(5))
    b.add_quad((-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1), (0, 1, 0), floor)
    b.add_quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), (0, -1, 0), ceil)
    b.add_quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1), (0, 0, 1), back)
    b.add_quad((-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (1, 0, 0), left)
    b.add_quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), (-1, 0, 0), right)
    tall, short = b.add_material(alb()), b.add_material(alb())
    yaw = rng.uniform(-math.pi / 4, math.pi / 4, 2)
    b.add_box((-0.35, 0.6, -0.3), (0.3, 0.6, 0.3), yaw[0], tall)
    b.add_box((0.35, 0.3, 0.3), (0.3, 0.3, 0.3), yaw[1], short)
    light = b.add_material(material(diffuse=(0.0, 0.0, 0.0), emissive=(light_radiance,) * 3))
    b.add_quad((-0.25, 1.98, -0.25), (0.25, 1.98, -0.25), (0.25, 1.98, 0.25), (-0.25, 1.98, 0.25), (0, -1, 0), light)
    cam = dict(pos=(0.0, 1.0, 3.4), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov=math.radians(40.0))
    return b.build(cam)


def textured_cornell(seed: int = 0xC1, glass: bool = False):
    """The Cornell box with textures on four materials (texture path test scene).

    Every quad's corners get uv (0,0), (s,0), (s,s), (0,s), s per material:
    - floor: 8 x 8 two-colour checker, s = 2.5 (wrap addressing);
    - back wall: 16 x 4 red-to-blue ramp (bilinear blending);
    - short box: 4 x 4 checker whose alpha alternates 1 and 0.3; with `glass`
      the material's transparency is 0.9 (transparency = alpha x transparency);
    - ceiling light: 2 x 2 emissive texture (emission from the texture).
    """
    sc = cornell(seed)
    nq = len(sc["pos"]) // 4
    quad_uv = np.asarray([(0, 0), (1, 0), (1, 1), (0, 1)], np.float32)
    mat_of_quad = sc["mat"][::2]                    # two triangles per quad, in order
    scale = np.ones(len(sc["mats"]), np.float32)
    scale[0] = 2.5
    uv = (quad_uv[None] * scale[mat_of_quad][:, None, None]).reshape(nq * 4, 2)
    c0, c1 = np.array([0.8, 0.8, 0.75, 1.0]), np.array([0.2, 0.3, 0.6, 1.0])
    ii, jj = np.mgrid[0:8, 0:8]
    checker = np.where(((ii + jj) % 2 == 0)[..., None], c0, c1).astype(np.float32)
    ramp = np.zeros((4, 16, 4), np.float32)
    t = np.linspace(0.0, 1.0, 16)
    ramp[..., 0] = 0.8 * (1 - t)
    ramp[..., 1] = 0.3
    ramp[..., 2] = 0.8 * t
    ramp[..., 3] = 1.0
    ii, jj = np.mgrid[0:4, 0:4]
    boxtex = np.where(((ii + jj) % 2 == 0)[..., None], np.array([0.9, 0.6, 0.3, 1.0]),
                     np.array([0.3, 0.9, 0.6, 0.3])).astype(np.float32)
    light = np.array([[[20, 20, 20, 1], [10, 14, 18, 1]], [[18, 14, 10, 1], [16, 16, 16, 1]]], np.float32)
    mats = sc["mats"].copy()
    mats[0]["diffuseTexId"] = 0
    mats[2]["diffuseTexId"] = 1
    mats[6]["diffuseTexId"] = 2
    mats[6]["transparency"] = 0.9 if glass else 0.0
    mats[7]["emissiveTexId"] = 3
    return dict(sc, uv=uv, mats=mats, textures=[checker, ramp, boxtex, light])


def uv_sphere(center, radius, n_lon=8, n_lat=5):
    """UV sphere with n_lat bands: 2*n_lon cap triangles + 2*n_lon*(n_lat-2) quads' triangles."""
    pos, nrm, idx = [], [], []
    for i in range(n_lat + 1):
        th = math.pi * i / n_lat
        for j in range(n_lon):
            ph = 2 * math.pi * j / n_lon
            n = (math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph))
            pos.append(tuple(center[k] + radius * n[k] for k in range(3)))
            nrm.append(n)
    vid = lambda i, j: i * n_lon + (j % n_lon)
    for i in range(n_lat):
        for j in range(n_lon):
            a, b_, c, d = vid(i, j), vid(i, j + 1), vid(i + 1, j + 1), vid(i + 1, j)
            if i == 0:
                idx.append((a, c, d))
            elif i == n_lat - 1:
                idx.append((a, b_, d))
            else:
                idx.append((a, b_, c))
                idx.append((a, c, d))
    return pos, nrm, idx


def icosphere(center, radius, subdiv=1):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [tuple(np.asarray(p) / np.linalg.norm(p)) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache, nf = {}, []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = np.add(v[a], v[b]) / 2.0
                v.append(tuple(m / np.linalg.norm(m)))
                cache[key] = len(v) - 1
            return cache[key]
        for a, b_, c in f:
            ab, bc, ca = mid(a, b_), mid(b_, c), mid(c, a)
            nf += [(a, ab, ca), (b_, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    pos = [tuple(center[k] + radius * p[k] for k in range(3)) for p in v]
    return pos, v, f


def sphere_grid(grid=64, seed=0x5EED0002, spacing=2.0, light_radiance=10.0):
    """Streaming-BDPT workload: grid x grid 64-triangle spheres on a ground plane + 2 sphere lights.

    64 x 64 spheres x 64 triangles = 262,144 sphere triangles (the stated total;
    a 320-triangle subdivision-2 icosphere would give 1,310,720).  Radius
    U(0.3, 1.0), jitter U(-0.5, 0.5), per-object BRDF parameters uniform random
    (metallic, roughness U(0.05, 1), specular, sheen, clearcoat).  The lights are
    emissive 80-triangle icospheres, because the reference estimator only knows
    emissive triangles.
    """
    rng = np.random.default_rng(seed)
    b = _Builder()
    half = grid * spacing / 2.0
    ground = b.add_material(material(diffuse=(0.5, 0.5, 0.5), specular=0.2, roughness=0.8))
    b.add_quad((-half - 2, 0, half + 2), (half + 2, 0, half + 2), (half + 2, 0, -half - 2), (-half - 2, 0, -half - 2),
               (0, 1, 0), ground)
    radius = rng.uniform(0.3, 1.0, grid * grid)
    jitter = rng.uniform(-0.5, 0.5, (grid * grid, 2))
    par = rng.uniform(0.0, 1.0, (grid * grid, 6))
    pos_all, nrm_all, idx_all, mat_all = [], [], [], []
    base = 0
    unit_pos, unit_nrm, unit_idx = uv_sphere((0.0, 0.0, 0.0), 1.0)
    unit_pos = np.asarray(unit_pos, np.float64)
    unit_nrm = np.asarray(unit_nrm, np.float32)
    unit_idx = np.asarray(unit_idx, np.uint32)
    for k in range(grid * grid):
        gx, gz = k % grid, k // grid
        c = np.array([-half + spacing * (gx + 0.5) + jitter[k, 0], radius[k], -half + spacing * (gz + 0.5) + jitter[k, 1]])
        m = b.add_material(material(diffuse=tuple(0.2 + 0.6 * par[k, :3]), metallic=par[k, 3],
                                    roughness=0.05 + 0.95 * par[k, 4], specular=par[k, 5],
                                    sheen=rng.uniform(), clearcoat=rng.uniform()))
        pos_all.append(c + radius[k] * unit_pos)
        nrm_all.append(unit_nrm)
        idx_all.append(unit_idx + (b.nv + base))
        mat_all.append(np.full(len(unit_idx), m, np.uint32))
        base += len(unit_pos)
    scene = b.build(None)
    lights = _Builder()
    lights.nv = scene["pos"].shape[0] + base
    lm = len(b.mats)
    lmat = material(diffuse=(0.0, 0.0, 0.0), emissive=(light_radiance,) * 3)
    for center in ((-half / 2, 24.0, -half / 3), (half / 2, 30.0, half / 3)):
        p, n, f = icosphere(center, 4.0, 1)
        lights.add_mesh(p, n, f, lm)
    mats = np.concatenate([scene["mats"], np.asarray([lmat], MATERIAL_DTYPE)])
    cam = dict(pos=(0.0, 38.0, half + 30.0), dir=tuple(np.array([0.0, -38.0, -(half + 30.0)]) /
                                                      np.linalg.norm([0.0, -38.0, -(half + 30.0)])),
               up=(0.0, 1.0, 0.0), fov=math.radians(45.0))
    return dict(
        pos=np.concatenate([scene["pos"], np.concatenate(pos_all).astype(np.float32),
                            np.asarray(lights.pos, np.float32)]),
        nrm=np.concatenate([scene["nrm"], np.concatenate(nrm_all), np.asarray(lights.nrm, np.float32)]),
        idx=np.concatenate([scene["idx"], np.concatenate(idx_all), np.asarray(lights.idx, np.uint32)]),
        mat=np.concatenate([scene["mat"], np.concatenate(mat_all), np.asarray(lights.mat, np.uint32)]),
        mats=mats, camera=cam)


def _value_noise(rng, shape, cells):
    """Bilinearly interpolated random lattice of `cells` cells over `shape`."""
    cy, cx = cells
    lat = rng.uniform(-1.0, 1.0, (cy + 2, cx + 2))
    h, w = shape
    y = np.linspace(0, cy, h, dtype=np.float64)
    x = np.linspace(0, cx, w, dtype=np.float64)
    y0, x0 = np.floor(y).astype(int), np.floor(x).astype(int)
    fy, fx = (y - y0)[:, None], (x - x0)[None, :]
    a = lat[y0][:, x0]
    b = lat[y0][:, x0 + 1]
    c = lat[y0 + 1][:, x0]
    d = lat[y0 + 1][:, x0 + 1]
    return (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy


# Adaptive-sampling settings of the heightfield workload, as renderer attributes:
# filter radius 7, colour / normal / relative-depth sigmas 0.2 / 0.1 / 0.05, a
# spatial sigma of radius / 2 (the reference would take the colour sigma, which
# at 0.2 leaves only the centre tap), the position term off (the depth term
# replaces it), and 4 to 256 eye paths per pixel.
ADAPTIVE_C3 = {"Filter Radius": 7, "Filter Color EuD": 0.2, "Filter Norm EuD": 0.1, "Filter Depth EuD": 0.05,
               "Filter Spatial EuD": 3.5, "Filter Pos EuD": 1e4, "Min Pixel Paths": 4, "Max Pixel Paths": 256}



def env_directions(width, height):
    """Unit directions of the texel centres of a width x height latitude-longitude
    map (u = atan2(z, x) / 2pi + 1/2 across, v = acos(y) / pi down)."""
    u = (np.arange(width) + 0.5) / width
    v = (np.arange(height) + 0.5) / height
    phi = (u - 0.5) * 2 * math.pi
    theta = v * math.pi
    st = np.sin(theta)[:, None]
    return np.stack([st * np.cos(phi)[None, :], np.cos(theta)[:, None] * np.ones((1, width)),
                     st * np.sin(phi)[None, :]], -1)


def env_lobes(width=512, height=256, seed=0xE7, lobes=8):
    """HDR environment of seeded Gaussian lobes: a sky gradient over the upper
    hemisphere, a dim ground below, a sun lobe and `lobes - 1` broader coloured
    lobes above the horizon, each I * exp(k (d . mu - 1)).  FP32 RGB, H x W x 3."""
    rng = np.random.default_rng(seed)
    d = env_directions(width, height)
    up = np.clip(d[..., 1], 0.0, 1.0)[..., None]
    sky = np.array([0.35, 0.45, 0.65]) * (0.4 + 0.6 * up) * (d[..., 1:2] >= 0)
    ground = np.array([0.08, 0.07, 0.06]) * (d[..., 1:2] < 0)
    img = sky + ground
    for k in range(lobes):
        el = rng.uniform(math.radians(10), math.radians(70))
        az = rng.uniform(0, 2 * math.pi)
        mu = np.array([math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az)])
        if k == 0:                                        # the sun: small and bright
            kappa, inten, col = 2000.0, 60.0, np.array([1.0, 0.92, 0.8])
        else:
            kappa, inten, col = rng.uniform(20, 300), rng.uniform(0.5, 4.0), rng.uniform(0.3, 1.0, 3)
        img = img + inten * col * np.exp(kappa * (d @ mu - 1.0))[..., None]
    return img.astype(np.float32)


def heightfield(nx=1024, nz=512, boxes=200, seed=0xC3, extent=(256.0, 128.0), amplitude=18.0, octaves=6,
                light_radiance=12.0, light="spheres"):
    """Adaptive-sampling workload: an nx x nz fBm heightfield (octaves of value
    noise, 2 triangles per cell: 1,048,576 triangles at 1024 x 512) with
    `boxes` random boxes (12 triangles each) standing on it.

    light = "env": lit by the stated 512 x 256 FP32 HDR environment of seeded
    Gaussian lobes (env_lobes), the environment-light extension; no emissive
    triangle, so there are no light subpaths and eye paths gather the
    environment when they leave the scene.
    light = "spheres": lit by four emissive icospheres above it instead, the
    reference estimator's own kind of light (it knows emissive triangles only,
    cudaRTBDPTStream.cu:57-77).
    """
    if light not in ("spheres", "env"):
        raise ValueError("light: 'spheres' or 'env'")
    rng = np.random.default_rng(seed)
    ex, ez = extent
    hgt = np.zeros((nz + 1, nx + 1))
    amp, cells = 1.0, (2, 4)
    for _ in range(octaves):
        hgt += amp * _value_noise(rng, hgt.shape, cells)
        amp *= 0.5
        cells = (cells[0] * 2, cells[1] * 2)
    hgt *= amplitude
    xs = np.linspace(-ex / 2, ex / 2, nx + 1)
    zs = np.linspace(-ez / 2, ez / 2, nz + 1)
    X, Z = np.meshgrid(xs, zs)
    pos = np.stack([X, hgt, Z], -1).reshape(-1, 3)
    gz, gx = np.gradient(hgt, zs, xs)
    nrm = np.stack([-gx, np.ones_like(gx), -gz], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    nrm = nrm.reshape(-1, 3)
    i = np.arange(nz)[:, None] * (nx + 1) + np.arange(nx)[None, :]
    a, b, c, d = i, i + 1, i + nx + 2, i + nx + 1
    idx = np.concatenate([np.stack([a, d, c], -1).reshape(-1, 3), np.stack([a, c, b], -1).reshape(-1, 3)])
    ground = material(diffuse=(0.45, 0.42, 0.38), specular=0.2, roughness=0.9)
    mats = [ground]
    mat = [np.zeros(len(idx), np.uint32)]
    bld = _Builder()
    bld.nv = len(pos)
    for k in range(boxes):
        bx, bz = rng.uniform(-ex / 2 * 0.9, ex / 2 * 0.9), rng.uniform(-ez / 2 * 0.9, ez / 2 * 0.9)
        ix = int(round((bx + ex / 2) / ex * nx))
        iz = int(round((bz + ez / 2) / ez * nz))
        half = rng.uniform(0.5, 3.0, 3)
        par = rng.uniform(0.0, 1.0, 6)
        m = len(mats)
        mats.append(material(diffuse=tuple(0.2 + 0.6 * par[:3]), metallic=par[3], roughness=0.05 + 0.95 * par[4],
                             specular=par[5], sheen=rng.uniform(), clearcoat=rng.uniform()))
        bld.add_box((bx, hgt[iz, ix] + half[1] * 0.8, bz), tuple(half), rng.uniform(0, 2 * math.pi), m)
    if light == "spheres":
        lm = len(mats)
        mats.append(material(diffuse=(0.0, 0.0, 0.0), emissive=(light_radiance,) * 3))
        for lx, lz in ((-ex / 4, -ez / 4), (ex / 4, -ez / 4), (-ex / 4, ez / 4), (ex / 4, ez / 4)):
            p, n, f = icosphere((lx, amplitude * 2.5 + 30.0, lz), 6.0, 1)
            bld.add_mesh(p, n, f, lm)
    mats_arr = np.zeros(len(mats), MATERIAL_DTYPE)
    for j, m in enumerate(mats):
        mats_arr[j] = m
    cam = dict(pos=(0.0, amplitude * 2.0 + 25.0, ez / 2 + 40.0), fov=math.radians(45.0), up=(0.0, 1.0, 0.0))
    d = np.array([0.0, -(amplitude + 25.0), -(ez / 2 + 40.0)])
    cam["dir"] = tuple(d / np.linalg.norm(d))
    extra_mat = np.asarray(bld.mat, np.uint32)
    return dict(pos=np.concatenate([pos.astype(np.float32), np.asarray(bld.pos, np.float32).reshape(-1, 3)]),
                nrm=np.concatenate([nrm.astype(np.float32), np.asarray(bld.nrm, np.float32).reshape(-1, 3)]),
                idx=np.concatenate([idx.astype(np.uint32), np.asarray(bld.idx, np.uint32).reshape(-1, 3)]),
                mat=np.concatenate(mat + [extra_mat]), mats=mats_arr, camera=cam,
                env=env_lobes(seed=seed ^ 0xE7) if light == "env" else None)


def heightfield_poses(n, seed0=0xCA3E0000, resolutions=((256, 256), (640, 480), (1280, 720), (1920, 1080)),
                      extent=(256.0, 128.0), amplitude=18.0):
    """n seeded camera views over the heightfield, resolutions drawn from `resolutions`."""
    ex, ez = extent
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed0 + i)
        w, h = resolutions[int(rng.integers(len(resolutions)))]
        pos = np.array([rng.uniform(-ex / 2, ex / 2), amplitude * rng.uniform(1.5, 3.0) + 10.0,
                        rng.uniform(-ez / 2, ez / 2)])
        tgt = np.array([rng.uniform(-ex / 3, ex / 3), 0.0, rng.uniform(-ez / 3, ez / 3)])
        d = tgt - pos
        d /= np.linalg.norm(d)
        out.append(dict(width=w, height=h, pos=tuple(pos), dir=tuple(d), up=(0.0, 1.0, 0.0),
                        fov=math.radians(50.0)))
    return out


def read_obj(text: str):
    """Positions and faces of an OBJ file; polygons are fan-triangulated
    (assimp's aiProcess_Triangulate order for convex faces)."""
    pos, idx = [], []
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v":
            pos.append(tuple(float(x) for x in parts[1:4]))
        elif parts[0] == "f":
            vs = [int(p.split("/")[0]) for p in parts[1:]]
            vs = [v - 1 if v > 0 else len(pos) + v for v in vs]
            for k in range(1, len(vs) - 1):
                idx.append((vs[0], vs[k], vs[k + 1]))
    return np.asarray(pos, np.float32).reshape(-1, 3), np.asarray(idx, np.uint32).reshape(-1, 3)


def mesh_scene(pos, idx, camera=None):
    """A scene around a bare mesh: flat normals, one grey material."""
    pos = np.asarray(pos, np.float32)
    idx = np.asarray(idx, np.uint32)
    p = pos[idx]  # nt x 3 x 3
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    # unshared vertices so flat normals survive
    mats = np.asarray([material()], MATERIAL_DTYPE)
    return dict(pos=p.reshape(-1, 3).astype(np.float32), nrm=np.repeat(n, 3, axis=0).astype(np.float32),
                idx=np.arange(len(idx) * 3, dtype=np.uint32).reshape(-1, 3),
                mat=np.zeros(len(idx), np.uint32), mats=mats, camera=camera)

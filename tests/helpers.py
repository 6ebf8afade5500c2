"""Shared inputs for the parity tests: seeded ray batches over the procedural scenes."""
from __future__ import annotations

import math

import numpy as np


def camera_rays(camera, W, H, jitter=False, seed=0):
    """Primary rays of the reference's pinhole (src/cudaRTBDPTStream.cu:629-637), evaluated in numpy float32."""
    d = np.asarray(camera["dir"], np.float32)
    up = np.asarray(camera["up"], np.float32)
    right = np.cross(d, up)
    right /= np.linalg.norm(right)
    up2 = np.cross(right, d)
    up2 /= np.linalg.norm(up2)
    tf = math.tan(camera["fov"] * 0.5)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float32)
    u = (2 * (xs + 0.5) / W - 1) * tf * W / H
    v = (2 * (ys + 0.5) / H - 1) * tf
    if jitter:
        rng = np.random.default_rng(seed)
        u = u + (rng.uniform(size=u.shape) - 0.5) / H * tf
        v = v + (rng.uniform(size=v.shape) - 0.5) / H * tf
    dirs = right[None, None] * u[..., None] + up2[None, None] * v[..., None] + d[None, None]
    dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
    o = np.broadcast_to(np.asarray(camera["pos"], np.float32), dirs.shape)
    return rays_array(o.reshape(-1, 3), dirs.reshape(-1, 3))


def rays_array(o, d, max_dist=1e20, eps=1e-4):
    r = np.zeros((len(o), 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = max_dist
    r[:, 4:7] = d
    r[:, 7] = eps
    return r


def random_rays(n, lo, hi, seed=0, axis_aligned_frac=0.1, eps=1e-4):
    """Origins uniform in the box [lo, hi], directions uniform on the sphere;
    a fraction get exactly axis-aligned directions (escapeZero path)."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    k = int(n * axis_aligned_frac)
    if k:
        ax = rng.integers(0, 3, k)
        d[:k] = 0
        d[np.arange(k), ax] = rng.choice([-1.0, 1.0], k)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return rays_array(o, d, eps=eps)


def vertex_rays(sc, n, seed=0):
    """Rays aimed exactly at mesh vertices and edge midpoints: hits on shared edges and ties."""
    rng = np.random.default_rng(seed)
    tri = sc["idx"][rng.integers(0, len(sc["idx"]), n)]
    p = sc["pos"][tri]                              # n x 3 x 3
    w = rng.integers(0, 3, n)
    target = np.where((rng.uniform(size=n) < 0.5)[:, None], p[np.arange(n), w],
                      0.5 * (p[np.arange(n), w] + p[np.arange(n), (w + 1) % 3]))
    o = target + rng.normal(size=(n, 3)).astype(np.float32) * 3.0
    d = (target - o).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return rays_array(o.astype(np.float32), d)


def hits_equal(a, b):
    """Bit equality of (tri, t, u, v) records."""
    a = np.ascontiguousarray(a).view(np.uint32).reshape(-1, 4)
    b = np.ascontiguousarray(b).view(np.uint32).reshape(-1, 4)
    return np.all(a == b, axis=1)


def rel_mse(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(((a - ref) ** 2).sum() / max((ref ** 2).sum(), 1e-300))


def oracle_adaptive(attrs):
    """Renderer adaptive attributes (e.g. scenes.ADAPTIVE_C3) as oracle render_params keywords."""
    keys = {"Filter Radius": "radius", "Filter Color EuD": "sc", "Filter Norm EuD": "sn", "Filter Depth EuD": "sz",
            "Filter Spatial EuD": "ss", "Filter Pos EuD": "sp", "Min Pixel Paths": "min_paths",
            "Max Pixel Paths": "max_paths"}
    return {keys[k]: v for k, v in attrs.items()}

"""Traversal-cost and BVH-depth heat maps: the counterpart of the reference's
BVH debug renderer (cudaRTBVHDebug: "TraceCost" mode, TraceCost in
src/cudaRTCommon.cu:232-273, over this library's traversal; "Normal" and
"BVHDepth" modes, TraceDepth / TraceDepthParent in src/cudaRTCommon.cu:129-231,
as written over the reference's binary tree).

One primary ray per pixel through the pixel centre (the debug renderer's
pinhole, src/cudaRTBVHDebug.cu:113-117: no jitter), traced closest-hit; the
map holds each ray's node visits and triangle tests in the chosen layout
(0: 128-byte 4-wide nodes, 1: the reference's binary tree, 2: 64-byte
quantized 4-wide nodes).  `to_rgb` renders it as the reference displays cost:
red for rays that reach no leaf, green for rays that do, value / 10.
"""
from __future__ import annotations

import math

import numpy as np

from . import MISS, Scene


def primary_rays(camera, width: int, height: int) -> np.ndarray:
    """(W*H, 8) rays, row y = 0 first: o, max distance 1e20, d, epsilon 1e-4."""
    d = np.asarray(camera["dir"], np.float64)
    up = np.asarray(camera["up"], np.float64)
    right = np.cross(d, up)
    right /= np.linalg.norm(right)
    up2 = np.cross(right, d)
    up2 /= np.linalg.norm(up2)
    tf = math.tan(camera["fov"] * 0.5)
    ys, xs = np.mgrid[0:height, 0:width].astype(np.float64)
    u = (2.0 * (xs + 0.5) / width - 1.0) * tf * width / height
    v = (2.0 * (ys + 0.5) / height - 1.0) * tf
    dirs = right * u[..., None] + up2 * v[..., None] + d
    dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
    r = np.zeros((width * height, 8), np.float32)
    r[:, 0:3] = np.asarray(camera["pos"], np.float32)
    r[:, 3] = 1e20
    r[:, 4:7] = dirs.reshape(-1, 3)
    r[:, 7] = 1e-4
    return r


def cost_map(scene: Scene, camera, width: int, height: int, layout: int = 2):
    """(H, W, 2) uint32 node visits and triangle tests, and (H, W) bool hit."""
    rays = primary_rays(camera, width, height)
    cost = scene.trace_cost(rays, layout=layout).reshape(height, width, 2)
    hit = (scene.trace(rays)["tri"] != MISS).reshape(height, width)
    return cost, hit


def depth_map(scene: Scene, camera, width: int, height: int, spec_depth=None):
    """The debug renderer's "Normal" mode (TraceDepth, spec_depth None): (H, W)
    depth, leaf and hit; its "BVHDepth" mode (TraceDepthParent): (H, W) depth
    (0xFFFFFFFF: no box at spec_depth), parent node and hit.  Over the reference
    tree, with the max distance M_INF as the debug renderer passes it."""
    rays = primary_rays(camera, width, height)
    out = scene.trace_depth(rays, spec_depth).reshape(height, width, 3)
    return out[..., 1], out[..., 2], out[..., 0] != 0


def depth_rgb(depth: np.ndarray, leaf: np.ndarray, scale: float = 0.1) -> np.ndarray:
    """"Normal" mode colouring: depth * scale, green where a leaf box was entered, red elsewhere."""
    v = depth.astype(np.float32) * np.float32(scale)
    img = np.zeros(depth.shape + (3,), np.float32)
    img[..., 1] = np.where(leaf, v, 0.0)
    img[..., 0] = np.where(leaf, 0.0, v)
    return img


def parent_rgb(depth: np.ndarray, parent: np.ndarray, hit: np.ndarray, spec_depth: int) -> np.ndarray:
    """"BVHDepth" mode colouring: green by a hash of the parent node where a box at
    spec_depth was entered, red 0.3 where boxes were entered but none at that
    depth.  The reference seeds cuRAND with the parent id for the green value;
    a splitmix64 hash stands in for it (the value only tells boxes apart)."""
    z = parent.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    col = (z >> np.uint64(40)).astype(np.float32) / np.float32(1 << 24)
    img = np.zeros(depth.shape + (3,), np.float32)
    at = hit & (depth == np.uint32(spec_depth))
    img[..., 1] = np.where(at, col, 0.0)
    img[..., 0] = np.where(hit & ~at, 0.3, 0.0)
    return img


def to_rgb(cost: np.ndarray, scale: float = 0.1) -> np.ndarray:
    """The debug renderer's colouring: (nodes + triangles) * scale in green for
    rays that tested a triangle, in red for rays that did not."""
    total = (cost[..., 0].astype(np.float32) + cost[..., 1]) * np.float32(scale)
    img = np.zeros(cost.shape[:2] + (3,), np.float32)
    leaf = cost[..., 1] > 0
    img[..., 1] = np.where(leaf, total, 0.0)
    img[..., 0] = np.where(leaf, 0.0, total)
    return img

"""ctypes front end of the CPU oracle (oracle/oracle.cpp).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        vp, c32, c64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        L.orc_scene_create.restype = vp
        L.orc_scene_create.argtypes = [f32p, f32p, c32, u32p, u32p, c32, vp, c32, c32]
        L.orc_scene_destroy.argtypes = [vp]
        L.orc_scene_set_uv.argtypes = [vp, f32p]
        L.orc_scene_set_env.argtypes = [vp, c32, c32, vp]
        L.orc_env_radiance.argtypes = [vp, f64p, ctypes.c_int, f64p]
        L.orc_scene_add_texture.restype = c32
        L.orc_scene_add_texture.argtypes = [vp, c32, c32, f32p]
        L.orc_tex_sample.argtypes = [vp, c32, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")]
        L.orc_scene_bvh.restype = c32
        L.orc_trace_depth.argtypes = [vp, f32p, c32, c32, u32p]
        L.orc_sample_bounce.argtypes = [f32p, f32p, ctypes.c_float, f32p, f32p, c64, c64, ctypes.c_int, f32p]
        L.orc_sample_bounce.restype = ctypes.c_int
        L.orc_scene_bvh.argtypes = [vp, vp, vp, vp, c32]
        L.orc_scene_light_tris.restype = c32
        L.orc_scene_light_tris.argtypes = [vp, vp, c32]
        L.orc_trace.argtypes = [vp, f32p, c32, c32, c32, c32, vp, vp]
        L.orc_renderer_create.restype = vp
        L.orc_renderer_create.argtypes = [vp, c32]
        L.orc_renderer_destroy.argtypes = [vp]
        L.orc_render_params_size.restype = c32
        L.orc_render.argtypes = [vp, vp, c32, c32, vp, vp, vp, vp]
        L.orc_render_result64.argtypes = [vp, f64p]
        L.orc_render_counts.argtypes = [vp, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")]
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_light_vertices.restype = c32
        L.orc_light_vertices.argtypes = [vp, f32p, c32]
        L.orc_pcg32_stream.restype = c32
        L.orc_pcg32_stream.argtypes = [c64, c64, c32, u32p]
        L.orc_stream_key.restype = c64
        L.orc_stream_key.argtypes = [c64, c32, c32, c32]
        L.orc_bilateral.argtypes = [f32p, f32p, f32p, vp, c32, c32, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float, c32, c32, f32p]
        L.orc_gaussian_table.argtypes = [ctypes.c_float, c32, f32p]
        L.orc_square_error.argtypes = [f32p, f32p, u32p, c32, c32, f32p]
        L.orc_empirical_error.argtypes = [f32p, f32p, u32p, c32, f32p]
        L.orc_alloc_proportional.argtypes = [f32p, c32, ctypes.c_float, c32, u32p]
        L.orc_det_eps.restype = ctypes.c_float
        L.orc_adaptive_state.argtypes = [vp, c32, vp, vp]
        L.orc_bilateral_defined.argtypes = [f32p, f32p, f32p, vp, c32, c32, ctypes.c_float, ctypes.c_float,
                                            ctypes.c_float, ctypes.c_float, c32, f32p]
        L.orc_bilateral_ext.argtypes = [f32p, f32p, f32p, vp, c32, c32] + [ctypes.c_float] * 6 + [f32p, c32, c32,
                                                                                                  f32p]
        L.orc_expf.restype = ctypes.c_float
        L.orc_expf.argtypes = [ctypes.c_float]
        d, i32p = ctypes.c_double, np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        ci = ctypes.c_int
        L.orc_d_ggx.restype = d
        L.orc_d_ggx.argtypes = [ci, d, d]
        L.orc_vis_smith.restype = d
        L.orc_vis_smith.argtypes = [ci, d, d, d]
        L.orc_f_schlick.argtypes = [ci, f64p, d, f64p]
        L.orc_sincos_def.argtypes = [f32p, c32, f32p, f32p]
        L.orc_atan2_def.argtypes = [f32p, f32p, c32, f32p]
        L.orc_acos_def.argtypes = [f32p, c32, f32p]
        L.orc_sample_ggx.argtypes = [ci, d, d, d, f64p, f64p]
        L.orc_basis.argtypes = [ci, f64p, f64p, f64p]
        L.orc_shading.argtypes = [ci, f64p, f64p]
        L.orc_sample_bounce_many.argtypes = [ci, f64p, f64p, d, f64p, f64p, c64, c32, ci, i32p, f64p, f64p, f64p]
        L.orc_disney_f.argtypes = [f64p] * 4 + [f64p]
        L.orc_disney_pdf.restype = d
        L.orc_disney_pdf.argtypes = [f64p] * 4
        L.orc_disney_sample_many.argtypes = [f64p, f64p, f64p, c64, c32, f64p, i32p]
        L.orc_tree_sum.restype = ctypes.c_double
        L.orc_tree_sum.argtypes = [f32p, c32]
        _lib = L
    return _lib


HIT_DTYPE = np.dtype([("tri", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])

RENDER_PARAMS_DTYPE = np.dtype([
    ("width", "<u4"), ("height", "<u4"), ("cam_pos", "<f4", 3), ("cam_dir", "<f4", 3), ("cam_up", "<f4", 3),
    ("fov", "<f4"), ("eye_depth", "<u4"), ("light_depth", "<u4"), ("light_paths", "<u4"), ("flags", "<u4"),
    ("seed", "<u8"), ("row0", "<u4"), ("rows", "<u4"),
    ("mode", "<u4"), ("nd", "<u4"), ("sc", "<f4"), ("sp", "<f4"), ("sn", "<f4"), ("sd", "<f4"),
    ("radius", "<u4"), ("filter_diff", "<u4"), ("warmup", "<u4"), ("pass_frames", "<u4"),
    ("min_prob", "<f4"), ("pad3", "<u4"), ("ss", "<f4"), ("sz", "<f4"), ("min_paths", "<u4"), ("max_paths", "<u4")])


def pcg32(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32_stream(initstate, initseq, n, out)
    return out


class Scene:
    """Oracle scene: literal bvhhelper build + BVH-leaf-order triangle data."""

    def __init__(self, sc: dict, max_depth: int = 32):
        L = lib()
        self.pos = np.ascontiguousarray(sc["pos"], np.float32)
        self.nrm = np.ascontiguousarray(sc["nrm"], np.float32)
        self.idx = np.ascontiguousarray(sc["idx"], np.uint32)
        self.mat = np.ascontiguousarray(sc["mat"], np.uint32)
        self.mats = np.ascontiguousarray(sc["mats"])
        self.h = L.orc_scene_create(self.pos.ravel(), self.nrm.ravel(), len(self.pos), self.idx.ravel(), self.mat,
                                    len(self.idx), self.mats.ctypes.data, len(self.mats), max_depth)
        if not self.h:
            raise RuntimeError("oracle: BVH build failed")
        self.nt = len(self.idx)
        if sc.get("uv") is not None:
            self.uv = np.ascontiguousarray(sc["uv"], np.float32).reshape(-1)
            L.orc_scene_set_uv(self.h, self.uv)
        self.textures = []
        for t in sc.get("textures", ()):
            self.add_texture(t)
        if sc.get("env") is not None:
            self.set_environment(sc["env"])

    def set_environment(self, rgb):
        """H x W x 3 latitude-longitude radiance map (None removes it)."""
        if rgb is None:
            lib().orc_scene_set_env(self.h, 0, 0, None)
            return
        self.env = np.ascontiguousarray(rgb, np.float32)
        lib().orc_scene_set_env(self.h, self.env.shape[1], self.env.shape[0], self.env.ctypes.data)

    def env_radiance(self, d, prec=64):
        out = np.zeros(3)
        lib().orc_env_radiance(self.h, np.ascontiguousarray(d, np.float64), prec, out)
        return out

    def add_texture(self, rgba) -> int:
        t = np.ascontiguousarray(rgba, np.float32)
        self.textures.append(t)
        return lib().orc_scene_add_texture(self.h, t.shape[1], t.shape[0], t.ravel())

    def tex_sample(self, tex_id: int, u: float, v: float, bits: int = 32) -> np.ndarray:
        """One texture lookup (wrap, bilinear) in FP32 or FP64."""
        out = np.zeros(4, np.float64)
        lib().orc_tex_sample(self.h, tex_id, float(u), float(v), bits, out)
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def bvh(self):
        L = lib()
        n = L.orc_scene_bvh(self.h, None, None, None, 0)
        bounds = np.zeros((n, 6), np.float32)
        off = np.zeros((n, 2), np.uint32)
        reorder = np.zeros(self.nt, np.uint32)
        L.orc_scene_bvh(self.h, bounds.ctypes.data, off.ctypes.data, reorder.ctypes.data, n)
        return bounds, off, reorder

    def light_tris(self):
        n = lib().orc_scene_light_tris(self.h, None, 0)
        out = np.zeros(max(n, 1), np.uint32)
        lib().orc_scene_light_tris(self.h, out.ctypes.data, n)
        return out[:n]

    def trace(self, rays: np.ndarray, mode: str = "brute", prec: int = 32, cull: bool = False):
        """rays: n x 8 float32 {ox,oy,oz,maxDist,dx,dy,dz,eps}. Returns (hits, stats)."""
        rays = np.ascontiguousarray(rays, np.float32)
        hits = np.zeros(len(rays), HIT_DTYPE)
        stats = np.zeros(2, np.uint32)
        m = {"brute": 0, "defined": 1, "reference": 2}[mode]
        lib().orc_trace(self.h, rays.ravel(), len(rays), int(cull), m, prec, hits.ctypes.data, stats.ctypes.data)
        return hits, stats

    def trace_depth(self, rays: np.ndarray, spec_depth=None):
        """(n, 3) uint32 of the BVH debug renderer's depth queries over the reference tree:
        TraceDepth {hit, depth, leaf} (spec_depth None) or TraceDepthParent(spec_depth)
        {hit, depth or 0xFFFFFFFF, parent}."""
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros((len(rays), 3), np.uint32)
        spec = 0xFFFFFFFF if spec_depth is None else int(spec_depth)
        lib().orc_trace_depth(self.h, rays.ravel(), len(rays), spec, out.ravel())
        return out


def max_threads() -> int:
    return int(lib().orc_max_threads())


def render_params(width, height, camera, eye_depth=5, light_depth=5, light_paths=128, seed=0x5EED0001, flags=0,
                  row0=0, rows=0, mode=0, nd=5, sc=50.0, sp=1.0, sn=0.25, sd=0.1, radius=5, filter_diff=0,
                  warmup=5, pass_frames=1, min_prob=0.0, ss=0.0, sz=0.0, min_paths=0, max_paths=0):
    """Reference-estimator parameters; mode 1 = adaptive proportional allocation with the
    reference's defaults (AdapBilatAtomic.cu:28-38)."""
    p = np.zeros((), RENDER_PARAMS_DTYPE)
    p["width"], p["height"] = width, height
    p["cam_pos"], p["cam_dir"], p["cam_up"] = camera["pos"], camera["dir"], camera["up"]
    p["fov"] = camera["fov"]
    p["eye_depth"], p["light_depth"], p["light_paths"] = eye_depth, light_depth, light_paths
    p["seed"] = seed
    p["flags"] = flags
    p["row0"], p["rows"] = row0, rows
    p["mode"], p["nd"], p["sc"], p["sp"], p["sn"], p["sd"] = mode, nd, sc, sp, sn, sd
    p["radius"], p["filter_diff"], p["warmup"] = radius, filter_diff, warmup
    p["pass_frames"] = pass_frames
    p["min_prob"] = min_prob
    p["ss"], p["sz"], p["min_paths"], p["max_paths"] = ss, sz, min_paths, max_paths
    return p


class Renderer:
    """Stateful reference-estimator run (running mean persists across calls)."""

    def __init__(self, params, prec: int = 32):
        assert lib().orc_render_params_size() == RENDER_PARAMS_DTYPE.itemsize
        self.p = np.ascontiguousarray(params)
        self.prec = prec
        self.h = lib().orc_renderer_create(self.p.ctypes.data, prec)
        self.W, self.H = int(params["width"]), int(params["height"])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_renderer_destroy(self.h)
            self.h = None

    def render(self, scene: Scene, frame0: int, nframes: int, gbuf: bool = False):
        n = self.W * self.H
        res = np.zeros((self.H, self.W, 3), np.float32)
        N = np.zeros((self.H, self.W), np.uint32)
        g = np.zeros((self.H, self.W, 9), np.float32) if gbuf else None
        lvn = np.zeros(1, np.uint32)
        lib().orc_render(scene.h, self.h, frame0, nframes, res.ctypes.data, N.ctypes.data,
                         g.ctypes.data if gbuf else None, lvn.ctypes.data)
        self.light_vertex_count = int(lvn[0])
        return (res, N, g) if gbuf else (res, N)

    def ray_counts(self):
        out = np.zeros(2, np.uint64)
        lib().orc_render_counts(self.h, out)
        return int(out[0]), int(out[1])

    def result64(self):
        out = np.zeros(self.W * self.H * 3, np.float64)
        lib().orc_render_result64(self.h, out)
        return out.reshape(self.H, self.W, 3)

    def adaptive_state(self):
        n = self.W * self.H
        out = {}
        for k, name, shape, dt in ((0, "gpos", (self.H, self.W, 3), np.float32), (1, "gnorm", (self.H, self.W, 3), np.float32),
                                   (2, "gdiff", (self.H, self.W, 3), np.float32), (3, "filtered", (self.H, self.W, 3), np.float32),
                                   (4, "err", (self.H, self.W), np.float32), (5, "counts", (self.H, self.W), np.uint32),
                                   (6, "sq", (self.H, self.W, 3), np.float32), (7, "rounds", (1,), np.uint32),
                                   (8, "paths", (self.H, self.W), np.uint32)):
            a = np.zeros(shape, dt)
            s = np.zeros(1, np.float32)
            lib().orc_adaptive_state(self.h, k, a.ctypes.data, s.ctypes.data)
            out[name] = a
            out["err_sum"] = float(s[0])
        return out

    def light_vertices(self):
        out = np.zeros((2048, 21), np.float32)
        n = lib().orc_light_vertices(self.h, out.ravel(), 2048)
        return out[:n]


def _v(x, n=3):
    a = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1))
    assert a.size == n
    return a


def d_ggx(rough, noh, prec=64):
    """D_GGX (cudaRTCommon.cu:309-315)."""
    return lib().orc_d_ggx(prec, float(rough), float(noh))


def vis_smith(rough, nov, nol, prec=64):
    """Vis_SmithJointApprox (cudaRTCommon.cu:299-307)."""
    return lib().orc_vis_smith(prec, float(rough), float(nov), float(nol))


def sincos_def(x):
    """The FP32 sin and cos of x >= 0 (Cephes' octant reduction and polynomials),
    the device's operation sequence."""
    x = np.ascontiguousarray(x, np.float32).ravel()
    sn, cs = np.zeros_like(x), np.zeros_like(x)
    lib().orc_sincos_def(x, x.size, sn, cs)
    return sn, cs


def atan2_def(y, x):
    """The FP32 atan2 of the environment lookup (Cephes' atanf core by octant)."""
    y = np.ascontiguousarray(y, np.float32).ravel()
    x = np.ascontiguousarray(x, np.float32).ravel()
    out = np.zeros_like(y)
    lib().orc_atan2_def(y, x, y.size, out)
    return out


def acos_def(y):
    """The FP32 acos of the environment lookup: atan2(sqrt((1 - y)(1 + y)), y)."""
    y = np.ascontiguousarray(y, np.float32).ravel()
    out = np.zeros_like(y)
    lib().orc_acos_def(y, y.size, out)
    return out


def f_schlick(f0, voh, prec=64):
    """F_Schlick (cudaRTCommon.cu:317-322)."""
    out = np.zeros(3)
    lib().orc_f_schlick(prec, _v(f0), float(voh), out)
    return out


def sample_ggx(xi0, xi1, rough, n, prec=64):
    """ImportanceSampleGGX (cudaRTCommon.cu:324-345): the half vector."""
    out = np.zeros(3)
    lib().orc_sample_ggx(prec, float(xi0), float(xi1), float(rough), _v(n), out)
    return out


def basis(w, prec=64):
    """The tangent frame of the samplers ((|w.x| > 0.1 ? y : x) x w)."""
    u, v = np.zeros(3), np.zeros(3)
    lib().orc_basis(prec, _v(w), u, v)
    return u, v


def shading(out_dir, in_dir, irrad, n, diff, metallic, rough, spec, dw, sw, prec=64):
    """GetShadingResult (cudaRTBDPTStream.cu:642-663)."""
    a = np.concatenate([_v(out_dir), _v(in_dir), _v(irrad), _v(n), _v(diff),
                        np.array([metallic, rough, spec, dw, sw], np.float64)])
    out = np.zeros(3)
    lib().orc_shading(prec, a, out)
    return out


def sample_bounce_many(ro, rd, t, norm, diff, trans, specular, metallic, roughness, nl, seed, n, prec=64,
                       refl_offset=False):
    """n eye-bounce BRDF samples at one hit (sample i from its own PCG32 stream):
    (type, orig, dir, mul) arrays; type 2 = specular branch, 1 = diffuse."""
    hs = np.concatenate([_v(norm), _v(diff), np.array([trans, specular, metallic, roughness], np.float64)])
    ty = np.zeros(n, np.int32)
    o, d, m = (np.zeros((n, 3)) for _ in range(3))
    lib().orc_sample_bounce_many(prec, _v(ro), _v(rd), float(t), hs, _v(nl), int(seed), int(n), int(refl_offset),
                                 ty, o.reshape(-1), d.reshape(-1), m.reshape(-1))
    return ty, o, d, m


def _disney_m(m):
    """Material vector of the Disney checks from a dict (base, metallic, subsurface,
    specular, roughness, sheen, sheenTint, clearcoat, clearcoatGloss)."""
    keys = ("metallic", "subsurface", "specular", "roughness", "sheen", "sheenTint", "clearcoat", "clearcoatGloss")
    return np.ascontiguousarray(list(m["base"]) + [float(m.get(k, 0.0)) for k in keys], np.float64)


def disney_f(m, n, wo, wi):
    """The Disney BRDF of the extension mode (FP64)."""
    out = np.zeros(3)
    lib().orc_disney_f(_disney_m(m), _v(n), _v(wo), _v(wi), out)
    return out


def disney_pdf(m, n, wo, wi):
    return lib().orc_disney_pdf(_disney_m(m), _v(n), _v(wo), _v(wi))


def disney_sample_many(m, n, wo, seed, count):
    """count BRDF samples: (directions count x 3, valid flags)."""
    wi = np.zeros((count, 3))
    ok = np.zeros(count, np.int32)
    lib().orc_disney_sample_many(_disney_m(m), _v(n), _v(wo), int(seed), int(count), wi.reshape(-1), ok)
    return wi, ok.astype(bool)


def sample_bounce(ro, rd, t, norm, diff, trans, specular, metallic, roughness, nl, state, seq, tir=False):
    """One FP32 BRDF sample at a hit: (origin, direction, throughput, ray type 1 diffuse / 2 specular)."""
    f = lambda v: np.ascontiguousarray(v, np.float32)
    hs = f([*norm, *diff, trans, specular, metallic, roughness])
    out = np.zeros(9, np.float32)
    ty = lib().orc_sample_bounce(f(ro), f(rd), float(t), hs, f(nl), int(state), int(seq), int(bool(tir)), out)
    return out[0:3], out[3:6], out[6:9], int(ty)


def bilateral(color, pos, norm, emit, sc, sp, sn, se, radius, prec=32):
    H, W = color.shape[:2]
    out = np.zeros((H, W, 3), np.float32)
    e = np.ascontiguousarray(emit, np.float32) if emit is not None else None
    lib().orc_bilateral(np.ascontiguousarray(color, np.float32).ravel(), np.ascontiguousarray(pos, np.float32).ravel(),
                        np.ascontiguousarray(norm, np.float32).ravel(), e.ctypes.data if e is not None else None,
                        W, H, sc, sp, sn, se, radius, prec, out.ravel())
    return out


def gaussian_table(sigma, radius):
    out = np.zeros(2 * radius + 1, np.float32)
    lib().orc_gaussian_table(sigma, radius, out)
    return out


def square_error(filtered, raw, N, nd):
    n = N.size
    out = np.zeros(n, np.float32)
    lib().orc_square_error(np.ascontiguousarray(filtered, np.float32).ravel(),
                           np.ascontiguousarray(raw, np.float32).ravel(),
                           np.ascontiguousarray(N, np.uint32).ravel(), nd, n, out)
    return out.reshape(N.shape)


def empirical_error(mean, sq, N):
    n = N.size
    out = np.zeros(n, np.float32)
    lib().orc_empirical_error(np.ascontiguousarray(mean, np.float32).ravel(), np.ascontiguousarray(sq, np.float32).ravel(),
                              np.ascontiguousarray(N, np.uint32).ravel(), n, out)
    return out.reshape(N.shape)


def alloc_rounds(err, gen, bisect, desired, min_prob=0.0, seed=0x5EED0001, frame=0):
    """Round-based allocation (AdapBilatBinary / AdapEmp) of an error map: (counts, rounds)."""
    e = np.ascontiguousarray(err, np.float32).ravel()
    out = np.zeros(e.size, np.uint32)
    L = lib()
    L.orc_alloc_rounds.restype = ctypes.c_uint32
    L.orc_alloc_rounds.argtypes = [f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_uint64, ctypes.c_uint32, u32p]
    r = L.orc_alloc_rounds(e, e.size, gen, int(bisect), desired, min_prob, seed, frame, out)
    return out.reshape(np.shape(err)), int(r)


def alloc_proportional(err, total, gen):
    out = np.zeros(err.size, np.uint32)
    lib().orc_alloc_proportional(np.ascontiguousarray(err, np.float32).ravel(), err.size, total, gen, out)
    return out.reshape(err.shape)


def bilateral_defined(color, pos, norm, emit, sc, sp, sn, se, radius):
    H, W = color.shape[:2]
    out = np.zeros((H, W, 3), np.float32)
    e = np.ascontiguousarray(emit, np.float32) if emit is not None else None
    lib().orc_bilateral_defined(np.ascontiguousarray(color, np.float32).ravel(),
                                np.ascontiguousarray(pos, np.float32).ravel(),
                                np.ascontiguousarray(norm, np.float32).ravel(), e.ctypes.data if e is not None else None,
                                W, H, sc, sp, sn, se, radius, out.ravel())
    return out


def bilateral_ext(color, pos, norm, emit, sc, sp, sn, se, radius, ss=0.0, sz=0.0, cam=(0.0, 0.0, 0.0), prec=32):
    """The filter with a spatial sigma of its own (ss > 0) and the relative-depth
    term ((z_q - z_p) / z_p)^2 / 2 sz^2 (sz > 0), z = |position - cam|; prec 32 the
    defined arithmetic of the device kernel, 64 the literal FP64 product."""
    H, W = color.shape[:2]
    out = np.zeros((H, W, 3), np.float32)
    f = lambda a: np.ascontiguousarray(a, np.float32).reshape(-1)
    e = f(emit) if emit is not None else None
    lib().orc_bilateral_ext(f(color), f(pos), f(norm), e.ctypes.data if e is not None else None, W, H, sc, sp, sn, se,
                            ss, sz, f(cam), radius, prec, out.reshape(-1))
    return out


def expf(x: float) -> float:
    return float(lib().orc_expf(x))


def tree_sum(e) -> float:
    e = np.ascontiguousarray(e, np.float32).ravel()
    return float(lib().orc_tree_sum(e, e.size))

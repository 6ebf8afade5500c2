"""cudapt_amd: MI355X (gfx950) streaming-BDPT hot path behind a C ABI.

Python front end over libcpt_hip.so (include/cpt.h) through ctypes.  There is
no CPU fallback: importing the library fails loudly when the HIP build is
missing, and every call raises CptError on a non-zero return code.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import scenes  # noqa: F401  (procedural inputs)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcpt_hip.so")
_lib = None

CPT_MODE_BDPT_STREAM = 0
CPT_MODE_BDPT_ADAPTIVE_PLL = 1
CPT_MODE_BDPT_ADAPTIVE_BINARY = 2
CPT_MODE_BDPT_ADAPTIVE_EMP = 3
CPT_MODE_PT_STREAM = 4
CPT_MODE_PT_STREAM_REGEN = 5
CPT_MODE_PT_DISNEY_MIS = 6
CPT_MODE_BDPT_REGEN = 7
(CPT_BUF_RESULT, CPT_BUF_SAMPLE_N, CPT_BUF_GBUF_POS, CPT_BUF_GBUF_NORM, CPT_BUF_GBUF_DIFF, CPT_BUF_FILTERED,
 CPT_BUF_ERROR, CPT_BUF_SAMPLE_COUNTS, CPT_BUF_ALLOC_ERROR, CPT_BUF_SECOND_MOMENT, CPT_BUF_PIXEL_PATHS,
 CPT_BUF_BOUNCE_RAYS) = range(12)
MLP_INPUTS, MLP_HIDDEN, MLP_PARAMS = 9, 64, 4865

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("max_dist", "<f4"), ("d", "<f4", 3), ("eps", "<f4")])
HIT_DTYPE = np.dtype([("tri", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
MISS = 0xFFFFFFFF


class CptError(RuntimeError):
    pass


_ATTR_ALIASES = {"eye_depth": "Eye Depth", "light_depth": "Light Depth", "light_paths": "Light Paths",
                 "seed": "Seed", "literal_shadow": "Literal Shadow Endpoint", "warmup": "Warm-up Frames",
                 "nd": "Desired Max Sampling", "sigma_color": "Filter Color EuD", "sigma_pos": "Filter Pos EuD",
                 "sigma_norm": "Filter Norm EuD", "sigma_diff": "Filter Diff EuD", "radius": "Filter Radius",
                 "filter_diff": "Filter Diffuse Flag", "in_flight": "Frames In Flight",
                 "pass_frames": "Adaptive Pass Frames", "megakernel": "Megakernel",
                 "sigma_spatial": "Filter Spatial EuD", "sigma_depth": "Filter Depth EuD",
                 "min_paths": "Min Pixel Paths", "max_paths": "Max Pixel Paths"}


class _Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("extension_rays", "shadow_rays", "eye_paths", "frames",
                                              "extend_launches", "extend_nodes", "extend_tris")] + \
               [(n, ctypes.c_double) for n in ("extend_ms", "shadow_ms", "shade_ms", "light_ms", "total_ms",
                                               "filter_ms", "error_ms", "gen_ms", "accum_ms", "mlp_ms")] + \
               [(n, ctypes.c_uint64) for n in ("shadow_nodes", "shadow_tris", "extend_spills", "shadow_spills",
                                              "tail_rays", "tail_nodes", "tail_tris")] + \
               [("tail_ms", ctypes.c_double)]


def lib():
    """The loaded HIP library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CptError(f"{LIB_PATH} is missing: run `python -m cudapt_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        vp, c32, i32, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_float
        L.cpt_last_error.restype = ctypes.c_char_p
        sigs = {
            "cpt_device_count": [vp],
            "cpt_device_alloc": [vp, ctypes.c_uint64],
            "cpt_device_free": [vp],
            "cpt_memcpy_htod": [vp, vp, ctypes.c_uint64],
            "cpt_memcpy_dtoh": [vp, vp, ctypes.c_uint64],
            "cpt_memcpy_dtod": [vp, vp, ctypes.c_uint64],
            "cpt_device_synchronize": [],
            "cpt_scene_create": [vp],
            "cpt_scene_set_mesh": [vp, vp, vp, vp, c32, vp, vp, c32],
            "cpt_scene_set_materials": [vp, vp, c32],
            "cpt_scene_update_materials": [vp],
            "cpt_bvh_build_sah": [vp, c32],
            "cpt_bvh_export": [vp, vp, vp, vp, vp],
            "cpt_bvh4_info": [vp, vp, vp, vp],
            "cpt_bvh4_set_leaf_size": [vp, c32],
            "cpt_scene_add_texture": [vp, c32, c32, vp, vp],
            "cpt_scene_clear_textures": [vp],
            "cpt_scene_set_environment": [vp, c32, c32, vp],
            "cpt_trace_cost": [vp, vp, c32, c32, c32, vp],
            "cpt_trace_depth": [vp, vp, c32, c32, vp],
            "cpt_scene_upload": [vp],
            "cpt_scene_destroy": [vp],
            "cpt_trace_closest": [vp, vp, c32, c32, vp],
            "cpt_trace_closest_device": [vp, vp, c32, c32, vp, vp, c32],
            "cpt_trace_occluded_device": [vp, vp, c32, c32, vp, vp],
            "cpt_trace_counters": [vp, vp, vp, vp, i32],
            "cpt_filter_cross_bilateral_device": [vp, vp, vp, vp, c32, c32, f32, f32, f32, f32, c32, vp, c32, vp,
                                                  vp, vp],
            "cpt_filter_cross_bilateral_device_noalloc": [vp, vp, vp, vp, c32, c32, f32, f32, f32, f32, c32, vp,
                                                          c32, vp, vp, vp],
            "cpt_renderer_create": [vp, i32, c32, c32],
            "cpt_render_iteration": [vp, vp, vp, vp, vp, f32, vp],
            "cpt_render_frames": [vp, vp, vp, vp, vp, f32, c32, vp],
            "cpt_renderer_attrib_count": [vp, vp],
            "cpt_renderer_get_attrib": [vp, c32, vp, c32, vp, vp],
            "cpt_renderer_set_attrib": [vp, c32, ctypes.c_double],
            "cpt_renderer_reset": [vp],
            "cpt_renderer_read_buffer": [vp, i32, vp],
            "cpt_renderer_stats": [vp, vp, i32],
            "cpt_renderer_set_profiling": [vp, i32],
            "cpt_renderer_destroy": [vp],
            "cpt_batch_create": [vp, vp, c32],
            "cpt_batch_get_config": [vp, vp],
            "cpt_batch_set_config": [vp, vp],
            "cpt_batch_render": [vp, vp, c32],
            "cpt_batch_read_result": [vp, c32, vp],
            "cpt_batch_sample_counts": [vp, c32, vp],
            "cpt_batch_stats": [vp, vp, i32],
            "cpt_batch_reset": [vp],
            "cpt_batch_autotune": [vp, vp, c32, vp, vp, c32],
            "cpt_batch_candidates": [vp, vp],
            "cpt_batch_candidate": [vp, c32, vp],
            "cpt_batch_destroy": [vp],
            "cpt_batch_read_buffer": [vp, c32, i32, vp],
            "cpt_batch_set_adaptive": [vp, i32, c32, vp],
            "cpt_batch_set_attrib": [vp, ctypes.c_char_p, ctypes.c_double],
            "cpt_mlp_create": [vp, ctypes.c_uint64],
            "cpt_mlp_destroy": [vp],
            "cpt_mlp_set_params": [vp, vp, c32],
            "cpt_mlp_get_params": [vp, vp, c32],
            "cpt_mlp_get_grad": [vp, vp, c32],
            "cpt_mlp_train_step": [vp, vp, vp, c32, f32, i32, vp, vp],
            "cpt_mlp_infer": [vp, vp, c32, vp, vp],
            "cpt_mlp_train_steps": [vp, vp, vp, c32, c32, c32, c32, f32, vp, vp],
            "cpt_renderer_mlp_features": [vp, vp],
            "cpt_renderer_mlp_target": [vp, vp, f32, vp],
            "cpt_renderer_set_mlp": [vp, vp],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise CptError(f"{what} failed ({rc}): {lib().cpt_last_error().decode()}")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(lib().cpt_device_count(ctypes.byref(n)), "cpt_device_count")
    return n.value


class DeviceBuffer:
    """HBM allocation owned through the library's HIP runtime."""

    def __init__(self, nbytes: int):
        p = ctypes.c_void_p()
        _check(lib().cpt_device_alloc(ctypes.byref(p), int(nbytes)), "cpt_device_alloc")
        self.ptr, self.nbytes = p.value or 0, int(nbytes)

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().cpt_memcpy_htod(ctypes.c_void_p(self.ptr), _ptr(a), a.nbytes), "cpt_memcpy_htod")

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.zeros(count, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().cpt_memcpy_dtoh(_ptr(out), ctypes.c_void_p(self.ptr), out.nbytes), "cpt_memcpy_dtoh")
        return out

    def __del__(self):
        if getattr(self, "ptr", 0) and _lib is not None:
            _lib.cpt_device_free(ctypes.c_void_p(self.ptr))
            self.ptr = 0


def synchronize():
    _check(lib().cpt_device_synchronize(), "cpt_device_synchronize")


class Scene:
    """RTScene counterpart: mesh + materials, SAH BVH, device upload."""

    def __init__(self, sc: dict | None = None, max_depth: int = 32, upload: bool = True):
        L = lib()
        h = ctypes.c_void_p()
        _check(L.cpt_scene_create(ctypes.byref(h)), "cpt_scene_create")
        self.h = h
        if sc is not None:
            self.set_mesh(sc["pos"], sc.get("nrm"), sc["idx"], sc["mat"], sc.get("uv"))
            self.set_materials(sc["mats"])
            for t in sc.get("textures", ()):
                self.add_texture(t)
            if sc.get("env") is not None:
                self.set_environment(sc["env"])
            self.build_bvh(max_depth)
            if upload:
                self.upload()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cpt_scene_destroy(self.h)
            self.h = None

    def set_mesh(self, pos, nrm, idx, mat, uv=None):
        self._pos = np.ascontiguousarray(pos, np.float32)
        self._nrm = None if nrm is None else np.ascontiguousarray(nrm, np.float32)
        self._uv = None if uv is None else np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        if self._uv is not None and len(self._uv) != len(self._pos):
            raise CptError("uv: one pair per vertex")
        self._idx = np.ascontiguousarray(idx, np.uint32)
        self._mat = np.ascontiguousarray(mat, np.uint32)
        self.nt = len(self._idx)
        _check(lib().cpt_scene_set_mesh(self.h, _ptr(self._pos), _ptr(self._nrm), _ptr(self._uv), len(self._pos),
                                        _ptr(self._idx), _ptr(self._mat), self.nt), "cpt_scene_set_mesh")

    def set_environment(self, rgb):
        """Environment light: an H x W x 3 float latitude-longitude radiance map (None removes it)."""
        if rgb is None:
            _check(lib().cpt_scene_set_environment(self.h, 0, 0, None), "cpt_scene_set_environment")
            return
        self._env = np.ascontiguousarray(rgb, np.float32)
        if self._env.ndim != 3 or self._env.shape[2] != 3:
            raise CptError("environment: H x W x 3 floats")
        _check(lib().cpt_scene_set_environment(self.h, self._env.shape[1], self._env.shape[0], _ptr(self._env)),
               "cpt_scene_set_environment")

    def add_texture(self, rgba) -> int:
        """Appends an H x W x 4 float texture (RGBA); returns its id."""
        t = np.ascontiguousarray(rgba, np.float32)
        if t.ndim != 3 or t.shape[2] != 4:
            raise CptError("texture: H x W x 4 floats")
        i = ctypes.c_uint32()
        _check(lib().cpt_scene_add_texture(self.h, t.shape[1], t.shape[0], _ptr(t), ctypes.byref(i)),
               "cpt_scene_add_texture")
        return i.value

    def set_materials(self, mats):
        self._mats = np.ascontiguousarray(mats)
        assert self._mats.dtype.itemsize == 80
        _check(lib().cpt_scene_set_materials(self.h, _ptr(self._mats), len(self._mats)), "cpt_scene_set_materials")

    def update_materials(self):
        _check(lib().cpt_scene_update_materials(self.h), "cpt_scene_update_materials")

    def build_bvh(self, max_depth: int = 32):
        _check(lib().cpt_bvh_build_sah(self.h, max_depth), "cpt_bvh_build_sah")

    def upload(self):
        _check(lib().cpt_scene_upload(self.h), "cpt_scene_upload")

    def bvh(self):
        n = ctypes.c_uint32(0)
        _check(lib().cpt_bvh_export(self.h, None, None, None, ctypes.byref(n)), "cpt_bvh_export")
        bounds = np.zeros((n.value, 6), np.float32)
        off = np.zeros((n.value, 2), np.uint32)
        reorder = np.zeros(self.nt, np.uint32)
        _check(lib().cpt_bvh_export(self.h, _ptr(bounds), _ptr(off), _ptr(reorder), ctypes.byref(n)),
               "cpt_bvh_export")
        return bounds, off, reorder

    def set_leaf_size(self, max_tris: int):
        """Most triangles per leaf of the 4-wide device BVH (1..64)."""
        _check(lib().cpt_bvh4_set_leaf_size(self.h, int(max_tris)), "cpt_bvh4_set_leaf_size")

    def bvh4_info(self):
        a, b, c = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().cpt_bvh4_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "cpt_bvh4_info")
        return dict(nodes=a.value, depth=b.value, leaves=c.value)

    def trace(self, rays: np.ndarray, cull: bool = False) -> np.ndarray:
        """rays: structured RAY_DTYPE or (n, 8) float32. Host arrays in and out."""
        r = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
        hits = np.zeros(len(r), HIT_DTYPE)
        _check(lib().cpt_trace_closest(self.h, _ptr(r), len(r), int(cull), _ptr(hits)), "cpt_trace_closest")
        return hits

    def trace_cost(self, rays: np.ndarray, layout: int = 2, cull: bool = False) -> np.ndarray:
        """(n, 2) uint32: node visits and triangle tests of each ray's closest-hit traversal."""
        r = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
        out = np.zeros((len(r), 2), np.uint32)
        _check(lib().cpt_trace_cost(self.h, _ptr(r), len(r), int(cull), layout, _ptr(out)), "cpt_trace_cost")
        return out

    def trace_depth(self, rays: np.ndarray, spec_depth=None) -> np.ndarray:
        """(n, 3) uint32 of the BVH debug renderer's depth queries over the reference tree:
        TraceDepth {hit, depth, leaf} (spec_depth None) or TraceDepthParent(spec_depth)
        {hit, depth or 0xFFFFFFFF, parent node}."""
        r = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
        out = np.zeros((len(r), 3), np.uint32)
        spec = 0xFFFFFFFF if spec_depth is None else int(spec_depth)
        _check(lib().cpt_trace_depth(self.h, _ptr(r), len(r), spec, _ptr(out)), "cpt_trace_depth")
        return out

    def trace_device(self, d_rays_ptr: int, n: int, d_hits_ptr: int, stream_ptr: int = 0, layout: int = 0,
                     cull: bool = False):
        _check(lib().cpt_trace_closest_device(self.h, ctypes.c_void_p(d_rays_ptr), n, int(cull),
                                              ctypes.c_void_p(d_hits_ptr),
                                              ctypes.c_void_p(stream_ptr) if stream_ptr else None, layout),
               "cpt_trace_closest_device")

    def occluded_device(self, d_rays_ptr: int, n: int, d_out_ptr: int, stream_ptr: int = 0, cull: bool = False):
        _check(lib().cpt_trace_occluded_device(self.h, ctypes.c_void_p(d_rays_ptr), n, int(cull),
                                               ctypes.c_void_p(d_out_ptr),
                                               ctypes.c_void_p(stream_ptr) if stream_ptr else None),
               "cpt_trace_occluded_device")

    def counters(self, reset: bool = False):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().cpt_trace_counters(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), int(reset)),
               "cpt_trace_counters")
        return dict(nodes=a.value, tris=b.value)


class Renderer:
    """RTRenderer + <ns>::Render counterpart (one call = one accumulation frame)."""

    def __init__(self, width: int, height: int, mode: int = CPT_MODE_BDPT_STREAM, **attribs):
        h = ctypes.c_void_p()
        _check(lib().cpt_renderer_create(ctypes.byref(h), mode, width, height), "cpt_renderer_create")
        self.h, self.W, self.H = h, width, height
        for k, v in attribs.items():
            self.set_attrib(k, v)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cpt_renderer_destroy(self.h)
            self.h = None

    def attribs(self):
        n = ctypes.c_uint32()
        _check(lib().cpt_renderer_attrib_count(self.h, ctypes.byref(n)), "cpt_renderer_attrib_count")
        out = {}
        for i in range(n.value):
            name = ctypes.create_string_buffer(64)
            t, v = ctypes.c_int(), ctypes.c_double()
            _check(lib().cpt_renderer_get_attrib(self.h, i, name, 64, ctypes.byref(t), ctypes.byref(v)),
                   "cpt_renderer_get_attrib")
            out[name.value.decode()] = (i, t.value, v.value)
        return out

    def set_attrib(self, name, value):
        key = _ATTR_ALIASES.get(name, name)
        idx = self.attribs()[key][0]
        _check(lib().cpt_renderer_set_attrib(self.h, idx, float(value)), "cpt_renderer_set_attrib")

    def render(self, scene: Scene, camera: dict, frames: int = 1, result: bool = True):
        out = np.zeros((self.H, self.W, 3), np.float32) if result else None
        args = (self.h, scene.h, _f3(camera["pos"]), _f3(camera["dir"]), _f3(camera["up"]),
                float(camera["fov"]))
        if frames == 1:
            _check(lib().cpt_render_iteration(*args, _ptr(out)), "cpt_render_iteration")
        else:
            _check(lib().cpt_render_frames(*args, frames, _ptr(out)), "cpt_render_frames")
        return out

    def sample_counts(self):
        out = np.zeros((self.H, self.W), np.uint32)
        _check(lib().cpt_renderer_read_buffer(self.h, CPT_BUF_SAMPLE_N, _ptr(out)), "cpt_renderer_read_buffer")
        return out

    def read(self, which: int):
        """A renderer buffer: CPT_BUF_* (result, counts, G-buffer, filtered, error, per-frame sample counts;
        CPT_BUF_BOUNCE_RAYS: the last wavefront's extension and shadow rays per bounce, 2 x 16)."""
        if which == CPT_BUF_BOUNCE_RAYS:
            out = np.zeros((2, 16), np.uint32)
        elif which in (CPT_BUF_SAMPLE_N, CPT_BUF_SAMPLE_COUNTS, CPT_BUF_PIXEL_PATHS):
            out = np.zeros((self.H, self.W), np.uint32)
        elif which in (CPT_BUF_ERROR, CPT_BUF_ALLOC_ERROR):
            out = np.zeros((self.H, self.W), np.float32)
        else:
            out = np.zeros((self.H, self.W, 3), np.float32)
        _check(lib().cpt_renderer_read_buffer(self.h, which, _ptr(out)), "cpt_renderer_read_buffer")
        return out

    def reset(self):
        _check(lib().cpt_renderer_reset(self.h), "cpt_renderer_reset")

    def mlp_features(self, d_x: "DeviceBuffer"):
        """Writes the W*H x 9 model inputs of the current adaptive state into d_x."""
        if d_x.nbytes < 36 * self.W * self.H:
            raise CptError("feature buffer smaller than W*H x 9 floats")
        _check(lib().cpt_renderer_mlp_features(self.h, ctypes.c_void_p(d_x.ptr)), "cpt_renderer_mlp_features")

    def mlp_target(self, d_ref: "DeviceBuffer", d_t: "DeviceBuffer", clamp: float = 0.0):
        """Writes the W*H training targets against the converged image in d_ref into d_t
        (radiance clamped to `clamp` on both sides when clamp > 0)."""
        if d_ref.nbytes < 12 * self.W * self.H or d_t.nbytes < 4 * self.W * self.H:
            raise CptError("reference or target buffer smaller than the image")
        _check(lib().cpt_renderer_mlp_target(self.h, ctypes.c_void_p(d_ref.ptr), float(clamp),
                                             ctypes.c_void_p(d_t.ptr)),
               "cpt_renderer_mlp_target")

    def set_mlp(self, model: "Mlp | None"):
        """Drives the adaptive allocation with `model` (None: the filter residual)."""
        _check(lib().cpt_renderer_set_mlp(self.h, model.h if model is not None else None), "cpt_renderer_set_mlp")
        self._mlp = model                 # the model must outlive the attachment

    def set_profiling(self, on: bool):
        _check(lib().cpt_renderer_set_profiling(self.h, int(on)), "cpt_renderer_set_profiling")

    def stats(self, reset: bool = False) -> dict:
        s = _Stats()
        _check(lib().cpt_renderer_stats(self.h, ctypes.byref(s), int(reset)), "cpt_renderer_stats")
        return {n: getattr(s, n) for n, _ in s._fields_}


class _View(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("cam_pos", ctypes.c_float * 3),
                ("cam_dir", ctypes.c_float * 3), ("cam_up", ctypes.c_float * 3), ("fov", ctypes.c_float)]


class BatchConfig(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("frames_in_flight", "extend_chunk", "extend_workgroups",
                                               "traversal_variant", "use_graph", "eye_depth", "light_depth",
                                               "light_paths")] + [("seed", ctypes.c_uint64),
                                                                  ("node_format", ctypes.c_uint32),
                                                                  ("refill", ctypes.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Batch:
    """Many camera views of uneven resolution rendered as one wavefront (plain
    streaming BDPT, or per-view adaptive sampling after set_adaptive)."""

    def __init__(self, views, **config):
        arr = (_View * len(views))()
        for a, v in zip(arr, views):
            a.width, a.height = int(v["width"]), int(v["height"])
            a.cam_pos[:] = [float(x) for x in v["pos"]]
            a.cam_dir[:] = [float(x) for x in v["dir"]]
            a.cam_up[:] = [float(x) for x in v["up"]]
            a.fov = float(v["fov"])
        h = ctypes.c_void_p()
        _check(lib().cpt_batch_create(ctypes.byref(h), arr, len(views)), "cpt_batch_create")
        self.h = h
        self.views = [(int(v["width"]), int(v["height"])) for v in views]
        if config:
            self.set_config(**config)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cpt_batch_destroy(self.h)
            self.h = None

    def config(self) -> dict:
        c = BatchConfig()
        _check(lib().cpt_batch_get_config(self.h, ctypes.byref(c)), "cpt_batch_get_config")
        return c.as_dict()

    def set_config(self, **kw):
        c = BatchConfig()
        _check(lib().cpt_batch_get_config(self.h, ctypes.byref(c)), "cpt_batch_get_config")
        for k, v in kw.items():
            setattr(c, k, int(v))
        _check(lib().cpt_batch_set_config(self.h, ctypes.byref(c)), "cpt_batch_set_config")

    def render(self, scene: Scene, frames: int):
        _check(lib().cpt_batch_render(self.h, scene.h, frames), "cpt_batch_render")

    def result(self, view: int) -> np.ndarray:
        w, h = self.views[view]
        out = np.zeros((h, w, 3), np.float32)
        _check(lib().cpt_batch_read_result(self.h, view, _ptr(out)), "cpt_batch_read_result")
        return out

    def sample_counts(self, view: int) -> np.ndarray:
        w, h = self.views[view]
        out = np.zeros((h, w), np.uint32)
        _check(lib().cpt_batch_sample_counts(self.h, view, _ptr(out)), "cpt_batch_sample_counts")
        return out

    def read(self, view: int, which: int) -> np.ndarray:
        """One view's CPT_BUF_* buffer."""
        w, h = self.views[view]
        if which in (CPT_BUF_SAMPLE_N, CPT_BUF_SAMPLE_COUNTS, CPT_BUF_PIXEL_PATHS):
            out = np.zeros((h, w), np.uint32)
        elif which in (CPT_BUF_ERROR, CPT_BUF_ALLOC_ERROR):
            out = np.zeros((h, w), np.float32)
        else:
            out = np.zeros((h, w, 3), np.float32)
        _check(lib().cpt_batch_read_buffer(self.h, view, which, _ptr(out)), "cpt_batch_read_buffer")
        return out

    def set_adaptive(self, mode: int = 1, pass_frames: int = 1, model: "Mlp | None" = None, **attribs):
        """Per-view adaptive sampling (mode CPT_MODE_BDPT_ADAPTIVE_PLL, optionally
        driven by a learned error model), or plain frames (CPT_MODE_BDPT_STREAM);
        then the given renderer attributes (names or their aliases)."""
        _check(lib().cpt_batch_set_adaptive(self.h, mode, pass_frames, model.h if model is not None else None),
               "cpt_batch_set_adaptive")
        self._model = model                  # the batch holds a pointer to it
        for k, v in attribs.items():
            self.set_attrib(k, v)

    def set_attrib(self, name: str, value):
        key = _ATTR_ALIASES.get(name, name)
        _check(lib().cpt_batch_set_attrib(self.h, key.encode(), float(value)), "cpt_batch_set_attrib")

    def stats(self, reset: bool = False) -> dict:
        s = _Stats()
        _check(lib().cpt_batch_stats(self.h, ctypes.byref(s), int(reset)), "cpt_batch_stats")
        return {n: getattr(s, n) for n, _ in s._fields_}

    def reset(self):
        _check(lib().cpt_batch_reset(self.h), "cpt_batch_reset")

    def candidates(self) -> list:
        n = ctypes.c_uint32()
        _check(lib().cpt_batch_candidates(self.h, ctypes.byref(n)), "cpt_batch_candidates")
        out = []
        for i in range(n.value):
            c = BatchConfig()
            _check(lib().cpt_batch_candidate(self.h, i, ctypes.byref(c)), "cpt_batch_candidate")
            out.append(c.as_dict())
        return out

    def autotune(self, scene: Scene, iterations: int = 2):
        """Times every candidate; returns (chosen config, per-candidate ms per iteration)."""
        n = len(self.candidates())
        ms = (ctypes.c_double * n)()
        best = BatchConfig()
        _check(lib().cpt_batch_autotune(self.h, scene.h, iterations, ctypes.byref(best), ms, n), "cpt_batch_autotune")
        return best.as_dict(), list(ms)


class Mlp:
    """Per-pixel error model 9 -> 64 -> 64 -> 1 (BF16 MFMA, FP32 master weights, Adam)."""

    def __init__(self, seed: int = 0xA22):
        h = ctypes.c_void_p()
        _check(lib().cpt_mlp_create(ctypes.byref(h), seed), "cpt_mlp_create")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cpt_mlp_destroy(self.h)
            self.h = None

    def params(self) -> np.ndarray:
        out = np.zeros(MLP_PARAMS, np.float32)
        _check(lib().cpt_mlp_get_params(self.h, _ptr(out), MLP_PARAMS), "cpt_mlp_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().cpt_mlp_set_params(self.h, _ptr(p), p.size), "cpt_mlp_set_params")

    def grad(self) -> np.ndarray:
        out = np.zeros(MLP_PARAMS, np.float32)
        _check(lib().cpt_mlp_get_grad(self.h, _ptr(out), MLP_PARAMS), "cpt_mlp_get_grad")
        return out

    def train_step_device(self, d_x: "DeviceBuffer", d_t: "DeviceBuffer", n: int, lr: float = 1e-3,
                          apply: bool = True, want_loss: bool = True, x_offset: int = 0):
        """One step on n samples starting at sample x_offset of device arrays d_x (n x 9) and d_t."""
        if n < 1 or x_offset < 0 or 36 * (x_offset + n) > d_x.nbytes or 4 * (x_offset + n) > d_t.nbytes:
            raise CptError(f"train batch [{x_offset}, {x_offset + n}) exceeds the device arrays")
        loss = ctypes.c_float()
        vp = ctypes.c_void_p
        _check(lib().cpt_mlp_train_step(self.h, vp(d_x.ptr + 36 * x_offset), vp(d_t.ptr + 4 * x_offset), n, lr,
                                        int(apply), ctypes.byref(loss) if want_loss else None, None),
               "cpt_mlp_train_step")
        return loss.value if want_loss else None

    def train_steps_device(self, d_x: "DeviceBuffer", d_t: "DeviceBuffer", n_samples: int, batch: int, steps: int,
                           lr: float = 1e-3, first_batch: int = 0, want_loss: bool = False):
        """`steps` Adam steps over consecutive batches of the device data set in one call
        (cpt_mlp_train_steps), starting at batch first_batch; returns the last step's loss
        when want_loss."""
        if batch < 1 or n_samples < batch or 36 * n_samples > d_x.nbytes or 4 * n_samples > d_t.nbytes:
            raise CptError(f"{n_samples} samples / batch {batch} do not fit the device arrays")
        loss = ctypes.c_float()
        _check(lib().cpt_mlp_train_steps(self.h, ctypes.c_void_p(d_x.ptr), ctypes.c_void_p(d_t.ptr), n_samples,
                                         batch, first_batch, steps, lr, ctypes.byref(loss) if want_loss else None,
                                         None), "cpt_mlp_train_steps")
        return loss.value if want_loss else None

    def train_step(self, x, t, lr: float = 1e-3, apply: bool = True) -> float:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        dx, dt = DeviceBuffer.from_array(x), DeviceBuffer.from_array(t)
        return self.train_step_device(dx, dt, len(t), lr, apply)

    def infer_device(self, d_x: "DeviceBuffer", n: int, d_y: "DeviceBuffer"):
        if 36 * n > d_x.nbytes or 4 * n > d_y.nbytes:
            raise CptError(f"{n} samples exceed the device arrays")
        _check(lib().cpt_mlp_infer(self.h, ctypes.c_void_p(d_x.ptr), n, ctypes.c_void_p(d_y.ptr), None),
               "cpt_mlp_infer")

    def infer(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        n = len(x)
        dx, dy = DeviceBuffer.from_array(x), DeviceBuffer(max(n, 1) * 4)
        self.infer_device(dx, n, dy)
        return dy.download(np.float32, n)


def make_rays(o, d, max_dist=1e20, eps=1e-4) -> np.ndarray:
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    r = np.zeros(len(o), RAY_DTYPE)
    r["o"], r["d"] = o, d
    r["max_dist"], r["eps"] = max_dist, eps
    return r


def filter_cross_bilateral(color, pos, norm, diff, sample_n, sigma_color, sigma_pos, sigma_norm, sigma_diff, radius,
                           nd=5, noalloc=False):
    """Device cross-bilateral + weighted squared residual on host images (copies in and out).

    noalloc: call cpt_filter_cross_bilateral_device_noalloc, which stages its
    table in the library's static device scratch instead of allocating."""
    H, W = color.shape[:2]
    bufs = [DeviceBuffer.from_array(np.ascontiguousarray(a, np.float32)) for a in (color, pos, norm)]
    db = DeviceBuffer.from_array(np.ascontiguousarray(diff, np.float32)) if diff is not None else None
    dn = DeviceBuffer.from_array(np.ascontiguousarray(sample_n, np.uint32))
    df = DeviceBuffer(W * H * 12)
    de = DeviceBuffer(W * H * 4)
    vp = ctypes.c_void_p
    name = "cpt_filter_cross_bilateral_device_noalloc" if noalloc else "cpt_filter_cross_bilateral_device"
    _check(getattr(lib(), name)(vp(bufs[0].ptr), vp(bufs[1].ptr), vp(bufs[2].ptr), vp(db.ptr) if db else None, W, H,
                                sigma_color, sigma_pos, sigma_norm, sigma_diff, radius, vp(dn.ptr), nd, vp(df.ptr),
                                vp(de.ptr), None), name)
    return df.download(np.float32, W * H * 3).reshape(H, W, 3), de.download(np.float32, W * H).reshape(H, W)

"""Scene ingest: OBJ + MTL models and the reference's .NPRTENV presets.

OBJ import follows what RTScene::AddModel (src/raytracer.cpp:238-282) takes
from its importer (aiProcess_Triangulate | FlipUVs | GenNormals):

- one vertex per face corner (no vertex joining), position, normal (the
  file's `vn`, else the face's flat normal) and texture coordinate with v
  flipped (1 - v);
- polygons fan-triangulated (0, k, k + 1);
- material table: a default material first (the importer's "DefaultMaterial",
  diffuse 0.6), then the .mtl's `newmtl` entries in file order; a face takes
  the material of the last `usemtl` before it (the default material before any);
- RTMaterial per entry (AssimpProcessSceneMaterial, src/raytracer.cpp:214-235):
  diffuse = Kd, emissive = Ka + Ke, transparency = 1 - d (or Tr), the other
  fields RTMaterial's defaults (include/raytracer.h:110-124).

Triangles keep file order.  The importer may regroup faces into one mesh per
(group, material) before the reference numbers them; that order only decides
ties between equal hit distances, so it is not reproduced.

.NPRTENV (CUDAPTWindow::BrowseEnvSetting, src/cudapt.cpp:623-737; tokenizer
NPConfFileHelper::txtConfFile, src/conffilehelper.cpp:32-110): a whitespace
token stream with // and /* */ comments; variables `campos` (3 floats),
`campitchyaw` (pitch, yaw), `rttech` (RTRenderer::RENDERER_MODE index),
`models` (count, paths), `materials` (count, 19 values each, applied only when
the scene already holds that many materials) and `convergedresult` (W*H*3
floats).  Other tokens are skipped, as the reference's read loop does.
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

from . import (CPT_MODE_BDPT_ADAPTIVE_BINARY, CPT_MODE_BDPT_ADAPTIVE_PLL, CPT_MODE_BDPT_STREAM, CPT_MODE_PT_STREAM)
from .scenes import MATERIAL_DTYPE

# RTRenderer::RENDERER_MODE with FULLSPECTRAL off (include/raytracer.h:1-25,295-304)
REFERENCE_MODES = ["CPU_DEBUG", "rastBasic", "cudaRTPT", "cudaRTPTStream", "cudaRTBDPTRegen", "cudaRTBDPTStream",
                   "cudaRTBDPTStreamBilat", "cudaRTBDPTStreamAdapBilatBinary", "cudaRTBDPTStreamAdapBilatAtomic",
                   "cudaRTBDPTAtomicAdapBilatAtomic", "cudaRTDebug", "cudaRTBVHDebug"]
# the ones this library renders
MODE_OF_REFERENCE = {"cudaRTBDPTStream": CPT_MODE_BDPT_STREAM, "cudaRTPTStream": CPT_MODE_PT_STREAM,
                     "cudaRTBDPTStreamAdapBilatBinary": CPT_MODE_BDPT_ADAPTIVE_BINARY,
                     "cudaRTBDPTStreamAdapBilatAtomic": CPT_MODE_BDPT_ADAPTIVE_PLL}

# the reference's camera field of view (src/cudapt.cpp:389)
FOV = math.pi / 4


def rt_material(diffuse=(1.0, 1.0, 1.0), emissive=(1.0, 1.0, 1.0), transparency=0.0):
    """RTMaterial with the reference constructor's defaults (include/raytracer.h:110-124)."""
    m = np.zeros((), MATERIAL_DTYPE)
    m["diffuse"] = diffuse
    m["emissive"] = emissive
    m["diffuseTexId"] = m["normalTexId"] = m["emissiveTexId"] = -1
    m["transparency"] = transparency
    m["specular"] = 0.5
    m["roughness"] = 0.5
    return m


def read_mtl(text: str):
    """[(name, {key: [floats]})] in `newmtl` order."""
    out = []
    for line in text.splitlines():
        parts = line.split()
        if not parts or parts[0].startswith("#"):
            continue
        if parts[0] == "newmtl":
            out.append((" ".join(parts[1:]), {}))
        elif out:
            try:
                out[-1][1][parts[0]] = [float(x) for x in parts[1:]]
            except ValueError:
                out[-1][1][parts[0]] = parts[1:]
    return out


def _material_from_mtl(props: dict):
    kd = props.get("Kd", [0.6, 0.6, 0.6])[:3]
    ka = props.get("Ka", [0.0, 0.0, 0.0])[:3]
    ke = props.get("Ke", [0.0, 0.0, 0.0])[:3]
    if "d" in props:
        opacity = props["d"][0]
    elif "Tr" in props:
        opacity = 1.0 - props["Tr"][0]
    else:
        opacity = 1.0
    em = tuple(float(np.float32(a) + np.float32(e)) for a, e in zip(ka, ke))
    return rt_material(diffuse=tuple(kd), emissive=em, transparency=float(np.float32(1.0) - np.float32(opacity)))


def default_material():
    """The importer's material for faces without `usemtl`: diffuse 0.6, no emission."""
    return rt_material(diffuse=(0.6, 0.6, 0.6), emissive=(0.0, 0.0, 0.0))


def read_obj_model(text: str, mtl_text: str | None = None):
    """A scene dict (pos, nrm, uv, idx, mat, mats) of one OBJ model."""
    mtl = read_mtl(mtl_text) if mtl_text else []
    names = {n: i + 1 for i, (n, _) in enumerate(mtl)}
    mats = [default_material()] + [_material_from_mtl(p) for _, p in mtl]
    v, vn, vt = [], [], []
    pos, nrm, uv, mat = [], [], [], []
    cur = 0
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        tag = parts[0]
        if tag == "v":
            v.append([float(x) for x in parts[1:4]])
        elif tag == "vn":
            vn.append([float(x) for x in parts[1:4]])
        elif tag == "vt":
            vt.append([float(x) for x in parts[1:3]] + [0.0] * (2 - len(parts[1:3])))
        elif tag == "usemtl":
            cur = names.get(" ".join(parts[1:]), 0)
        elif tag == "f":
            corners = []
            for p in parts[1:]:
                f = p.split("/")
                iv = int(f[0])
                iv = iv - 1 if iv > 0 else len(v) + iv
                it = int(f[1]) if len(f) > 1 and f[1] else None
                if it is not None:
                    it = it - 1 if it > 0 else len(vt) + it
                ino = int(f[2]) if len(f) > 2 and f[2] else None
                if ino is not None:
                    ino = ino - 1 if ino > 0 else len(vn) + ino
                corners.append((iv, it, ino))
            for k in range(1, len(corners) - 1):
                tri = (corners[0], corners[k], corners[k + 1])
                p3 = np.asarray([v[c[0]] for c in tri], np.float32)
                flat = np.cross(p3[1] - p3[0], p3[2] - p3[0]).astype(np.float64)
                ln = np.linalg.norm(flat)
                flat = flat / ln if ln > 0 else np.zeros(3)
                for c in tri:
                    pos.append(v[c[0]])
                    nrm.append(vn[c[2]] if c[2] is not None else list(flat))
                    t = vt[c[1]] if c[1] is not None else [0.0, 0.0]
                    uv.append([t[0], 1.0 - t[1]])
                mat.append(cur)
    nt = len(mat)
    return dict(pos=np.asarray(pos, np.float32).reshape(-1, 3), nrm=np.asarray(nrm, np.float32).reshape(-1, 3),
                uv=np.asarray(uv, np.float32).reshape(-1, 2),
                idx=np.arange(3 * nt, dtype=np.uint32).reshape(-1, 3), mat=np.asarray(mat, np.uint32),
                mats=np.asarray(mats, MATERIAL_DTYPE))


def merge_models(models):
    """RTScene::AddModel called once per model: vertex indices and material
    indices offset by what the earlier models added (src/raytracer.cpp:108,140)."""
    pos, nrm, uv, idx, mat, mats = [], [], [], [], [], []
    nv = nm = 0
    for m in models:
        pos.append(m["pos"]); nrm.append(m["nrm"]); uv.append(m["uv"])
        idx.append(m["idx"] + np.uint32(nv))
        mat.append(m["mat"] + np.uint32(nm))
        mats.append(m["mats"])
        nv += len(m["pos"])
        nm += len(m["mats"])
    cat = lambda xs, shape, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(shape, dt)  # noqa: E731
    return dict(pos=cat(pos, (0, 3), np.float32), nrm=cat(nrm, (0, 3), np.float32), uv=cat(uv, (0, 2), np.float32),
                idx=cat(idx, (0, 3), np.uint32), mat=cat(mat, (0,), np.uint32),
                mats=np.concatenate(mats) if mats else np.zeros(0, MATERIAL_DTYPE))


def fly_camera(pos, pitch, yaw, fov=FOV):
    """NPCamHelper::FlyCamera (include/camhelper.h:142-227): dir = normalize(
    cos(yaw) sin(pitch), sin(yaw), cos(yaw) cos(pitch)), up = +y."""
    d = np.array([math.cos(yaw) * math.sin(pitch), math.sin(yaw), math.cos(yaw) * math.cos(pitch)], np.float32)
    d = d / np.float32(np.linalg.norm(d))
    return dict(pos=tuple(float(x) for x in pos), dir=tuple(float(x) for x in d), up=(0.0, 1.0, 0.0), fov=fov)


def _tokens(text: str):
    """The tokenizer's stream: // comments to the end of the line, /* */ blocks."""
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    out = []
    for line in text.splitlines():
        c = line.find("//")
        if c >= 0:
            line = line[:c]
        out.extend(line.split())
    return out


# RTMaterial fields in the order a `materials` record lists them
# (src/cudapt.cpp:678-693): scalars except diffuse / emissive (3 floats each)
PRESET_MATERIAL_FIELDS = [("anisotropic", 1), ("clearcoat", 1), ("clearcoatGloss", 1), ("diffuse", 3),
                          ("diffuseTexId", 1), ("emissive", 3), ("emissiveTexId", 1), ("metallic", 1),
                          ("normalTexId", 1), ("roughness", 1), ("sheen", 1), ("sheenTint", 1), ("specular", 1),
                          ("subsurface", 1), ("transparency", 1)]
PRESET_MATERIAL_VALUES = sum(n for _, n in PRESET_MATERIAL_FIELDS)   # 19


def _resolve(model: str, base_dir: str, model_dirs):
    rel = model.replace("\\", "/")
    cands = [os.path.normpath(os.path.join(base_dir, rel))]
    cands += [os.path.join(d, os.path.basename(rel)) for d in model_dirs]
    for c in cands:
        if os.path.exists(c):
            return c
    raise FileNotFoundError(f"model {model!r} not found (tried {cands})")


def load_model(path: str):
    """An OBJ file and the .mtl its `mtllib` names (looked up beside it)."""
    with open(path) as f:
        text = f.read()
    mtl_text = None
    m = re.search(r"^mtllib\s+(.+?)\s*$", text, flags=re.M)
    if m:
        mp = os.path.join(os.path.dirname(path), m.group(1).replace("\\", "/"))
        if os.path.exists(mp):
            with open(mp) as f:
                mtl_text = f.read()
    return read_obj_model(text, mtl_text)


def parse_preset(text: str):
    """The variables of a .NPRTENV text, in file order: campos, campitchyaw,
    rttech, models (list of paths), materials (list of 19-value records),
    convergedresult (floats)."""
    tok = _tokens(text)
    out = {}
    i = 0
    while i < len(tok):
        t = tok[i]
        i += 1
        if t == "campos":
            out["campos"] = tuple(float(x) for x in tok[i:i + 3]); i += 3
        elif t == "campitchyaw":
            out["campitchyaw"] = (float(tok[i]), float(tok[i + 1])); i += 2
        elif t == "rttech":
            out["rttech"] = int(tok[i]); i += 1
        elif t == "models":
            n = int(tok[i]); i += 1
            out.setdefault("models", []).extend(tok[i:i + n]); i += n
        elif t == "materials":
            n = int(tok[i]); i += 1
            vals = [float(x) for x in tok[i:i + n * PRESET_MATERIAL_VALUES]]
            i += n * PRESET_MATERIAL_VALUES
            out["materials"] = [vals[k * PRESET_MATERIAL_VALUES:(k + 1) * PRESET_MATERIAL_VALUES] for k in range(n)]
            out["materials_after_models"] = len(out.get("models", []))
        elif t == "convergedresult":
            rest = []
            while i < len(tok):
                try:
                    rest.append(float(tok[i]))
                except ValueError:
                    break
                i += 1
            out["convergedresult"] = np.asarray(rest, np.float32)
    return out


def apply_preset_materials(mats, records):
    """The `materials` records over the first entries of the scene's table,
    only when the table holds at least that many (src/cudapt.cpp:674-695)."""
    mats = np.array(mats, MATERIAL_DTYPE, copy=True)
    if len(mats) < len(records):
        return mats
    for k, rec in enumerate(records):
        j = 0
        for name, n in PRESET_MATERIAL_FIELDS:
            val = rec[j:j + n]
            j += n
            mats[k][name] = val if n > 1 else val[0]
    return mats


def load_preset(path: str, model_dirs=(), width: int | None = None, height: int | None = None):
    """Scene, camera and renderer mode of a .NPRTENV preset.

    Model paths are resolved against the preset's directory, then by file name
    in `model_dirs`.  Returns dict(scene=..., camera=..., rttech=int,
    reference_mode=name, mode=library mode or None when this library does not
    implement that renderer, converged=H x W x 3 array or None)."""
    with open(path) as f:
        pre = parse_preset(f.read())
    base = os.path.dirname(os.path.abspath(path))
    models = [load_model(_resolve(m, base, model_dirs)) for m in pre.get("models", [])]
    scene = merge_models(models)
    if "materials" in pre:
        scene["mats"] = apply_preset_materials(scene["mats"], pre["materials"])
    pitch, yaw = pre.get("campitchyaw", (0.0, 0.0))
    camera = fly_camera(pre.get("campos", (0.0, 0.0, 0.0)), pitch, yaw)
    scene["camera"] = camera
    rt = pre.get("rttech")
    name = REFERENCE_MODES[rt] if rt is not None and 0 <= rt < len(REFERENCE_MODES) else None
    conv = pre.get("convergedresult")
    if conv is not None and width and height and conv.size == width * height * 3:
        conv = conv.reshape(height, width, 3)
    return dict(scene=scene, camera=camera, rttech=rt, reference_mode=name, mode=MODE_OF_REFERENCE.get(name),
                converged=conv)

"""The reference's RMSE-versus-time experiment on a preset scene.

CUDAPTWindow's protocol (src/cudapt.cpp:268-337,951-965): load a .NPRTENV
preset, capture a converged image, then render from a fresh accumulation and
record "sec, frames, rmse" rows every `interval` seconds of rendering time
(asset/experimentResult*/ hold the reference's rows for its scenes, 800 x 600,
interval 5 s).  Here the converged image is plain streaming BDPT at
`converged_spp` with a seed of its own, and each renderer mode in `modes` is
run from scratch with RmseHarness.
"""
from __future__ import annotations

import time

from . import CPT_MODE_BDPT_ADAPTIVE_PLL, CPT_MODE_BDPT_STREAM, Renderer, Scene, synchronize
from . import preset as _preset
from .rmse import RmseHarness

CONVERGED_SEED = 0x0EF5EED5
MODE_NAMES = {CPT_MODE_BDPT_STREAM: "sc", CPT_MODE_BDPT_ADAPTIVE_PLL: "atomic", 2: "binary", 3: "emp"}


def converged_image(scene: Scene, camera, width, height, spp, in_flight=64, seed=CONVERGED_SEED, **attribs):
    R = Renderer(width, height, mode=CPT_MODE_BDPT_STREAM, seed=seed, **attribs)
    R.set_attrib("Frames In Flight", min(in_flight, spp))
    img = R.render(scene, camera, frames=spp)
    synchronize()
    return img


def run_preset(preset_path, model_dirs=(), width=800, height=600, modes=(CPT_MODE_BDPT_STREAM,
                                                                         CPT_MODE_BDPT_ADAPTIVE_PLL),
               converged_spp=4096, interval=0.05, count=20, frames_per_step=1, max_frames=1 << 20, **attribs):
    """{mode name: [Sample(sec, frames, rmse)]} plus the converged image and its
    render time.  `attribs` go to every renderer (eye_depth=..., light_depth=...)."""
    p = _preset.load_preset(preset_path, model_dirs=model_dirs)
    S = Scene(p["scene"])
    cam = p["camera"]
    t0 = time.perf_counter()
    conv = converged_image(S, cam, width, height, converged_spp, **attribs)
    conv_s = time.perf_counter() - t0
    out = {"converged": conv, "converged_s": conv_s, "camera": cam, "reference_mode": p["reference_mode"],
           "triangles": int(len(p["scene"]["idx"])), "curves": {}}
    for m in modes:
        R = Renderer(width, height, mode=m, **attribs)
        # first-call allocations outside the timing, on the same camera; the
        # reset below restarts the accumulation the harness measures
        R.render(S, cam, frames=1, result=False)
        R.reset()
        synchronize()
        h = RmseHarness(R, S, cam, conv, frames_per_step=frames_per_step)
        out["curves"][MODE_NAMES.get(m, str(m))] = h.collect(interval, count, max_frames=max_frames)
    return out


def _main():
    import argparse
    import json
    import os
    ap = argparse.ArgumentParser(description="RMSE-versus-time rows of the reference's protocol on a preset scene")
    ap.add_argument("preset")
    ap.add_argument("--model-dir", action="append", default=[])
    ap.add_argument("--width", type=int, default=800)
    ap.add_argument("--height", type=int, default=600)
    ap.add_argument("--converged-spp", type=int, default=4096)
    ap.add_argument("--interval", type=float, default=0.05)
    ap.add_argument("--count", type=int, default=20)
    ap.add_argument("--frames-per-step", type=int, default=1)
    ap.add_argument("--eye-depth", type=int, default=5)
    ap.add_argument("--out", default=None, help="directory for <mode>_<preset>.csv")
    a = ap.parse_args()
    r = run_preset(a.preset, a.model_dir, a.width, a.height, converged_spp=a.converged_spp, interval=a.interval,
                   count=a.count, frames_per_step=a.frames_per_step, eye_depth=a.eye_depth, light_depth=a.eye_depth)
    name = os.path.splitext(os.path.basename(a.preset))[0]
    summary = {"preset": name, "width": a.width, "height": a.height, "triangles": r["triangles"],
               "converged_spp": a.converged_spp, "converged_s": round(r["converged_s"], 3), "curves": {}}
    for mode, rows in r["curves"].items():
        summary["curves"][mode] = [[round(s.sec, 5), s.frames, round(s.rmse, 6)] for s in rows]
        if a.out:
            os.makedirs(a.out, exist_ok=True)
            with open(os.path.join(a.out, f"{mode}_{name}.csv"), "w") as f:
                f.write("".join(f"{s.sec}, {s.frames}, {s.rmse}\n" for s in rows))
    print(json.dumps(summary))


if __name__ == "__main__":
    _main()

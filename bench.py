"""Benchmark of the streaming-BDPT hot path on one MI355X.

Workload (one step): the 262,144-triangle sphere-grid scene (64 x 64
64-triangle spheres on a ground plane, 2 emissive sphere lights), 1280 x 720,
eye depth 8, 64 frames = 64 samples per pixel through the wavefront BDPT
(light-vertex cache every third frame).  Inputs are built and uploaded to HBM
before timing; the timed region is K renders of 64 frames each.

metric/value: Mrays/s = (extension + shadow rays traced) / wall time.
roofline: the closest-hit extension kernel (by default k_trace_dyn: persistent
  waves that take new rays from a global queue once 64 (first bounce) or 48
  (later bounces) of their lanes are idle), algorithmic bytes per launch =
  32 B per ray read from memory (first-bounce eye rays are computed in the
  kernel and read nothing) + 16 B hit out per ray + 128 B (FP32 layout) or
  64 B (quantized layout) per 4-wide node visited + 48 B per triangle tested,
  over its HIP-event duration on the renderer stream.  Beside it, from
  rocprofv3 counter passes: the HBM bytes actually moved (FETCH_SIZE +
  WRITE_SIZE, calibrated on a copy), the VALUs' busy share and lane use
  (VALUBusy, VALUUtilization), and the limiter they name.
cpu_baseline: the oracle's FP32 reference estimator on the host cores (OpenMP),
  a bounded sample of the same frame (a band of rows at 1 spp).
time_to_equal_mse: the learned-vs-plain adaptive protocol of cudapt_amd/learned.py
  on the same scene and resolution: converged images by plain BDPT (4096 spp for
  the held-out poses, 1024 for the training poses), per-pixel error model trained
  on plain adaptive runs of the training poses (2000 Adam steps, batch 65,536),
  then on each held-out pose the time the learned renderer needs to reach the
  MSE (radiance clamped to [0, 1]) plain adaptive BDPT reaches after 64 frames.
mlp: MFMA throughput of the model's training step (forward + backward + Adam)
  at batch 65,536: 28,416 FLOP per sample (3x the 9,472 FLOP forward).
multiview: 32 seeded views of uneven resolution (256x256 .. 1920x1080) over the
  1,051,296-triangle fBm heightfield, one wavefront per iteration, autotuned
  extend queue chunk / workgroups / traversal variant, hipGraph-replayed
  iterations, 32 spp per view.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WIDTH, HEIGHT, SPP, EYE_DEPTH, LIGHT_DEPTH = 1280, 720, 64, 8, 8



def extend_kernel_name(args, demangled=False) -> str:
    """Template instance the renderer launches for extension rays (demangled:
    the prefix of the name the profiler reports; with refilling waves it covers
    both instances, the first bounce's, which computes its eye rays, and the
    later bounces')."""
    stk, minw = {806: (8, 6), 807: (8, 7), 808: (8, 8), 1606: (16, 6), 8: (8, 0), 24: (24, 0)}.get(args.lds_stack,
                                                                                                 (16, 0))
    q = "true" if args.node_format else "false"
    layout = 2 if args.node_format else 0
    if demangled:
        return (f"k_trace_dyn<1, {stk}, {minw}, {q}," if args.refill
                else f"k_trace<{layout}, false, 1, {stk}, {minw}>")
    if args.refill:
        return (f"k_trace_dyn<1,{stk},{minw},{q},EYE> (closest-hit extension, refill at {args.refill} idle lanes; "
                f"EYE at the first bounce: eye rays computed in the kernel)")
    return f"k_trace<{layout},false,1,{stk},{minw}> (closest-hit extension, one thread per ray)"

def queue_bytes(bounce_rays, eye_in_kernel):
    """Algorithmic SoA queue bytes of one wavefront per bounce (path buffers,
    hits and the shadow queue; not the scene): the extension reads each ray
    (32 B) and writes its hit (16 B); shading reads the hit and the path state
    (100 B) and writes each survivor's state (100 B) and each shadow ray
    (48 B); the shadow kernel reads the shadow ray and updates the path's
    accumulator (16 B read + 16 B written).  First-bounce eye rays computed in
    the kernels read no ray and no state."""
    ext, sh = [int(x) for x in bounce_rays[0]], [int(x) for x in bounce_rays[1]]
    out = []
    for b in range(EYE_DEPTH):
        fresh = b == 0 and eye_in_kernel
        nxt = ext[b + 1] if b + 1 < len(ext) else 0
        out.append(ext[b] * (16 + 16 + (0 if fresh else 32 + 100)) + nxt * 100 + sh[b] * (48 + 48 + 32))
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_peak_gbs(ca, mib=1024, reps=8):
    """Measured device-to-device copy rate (bytes read + written per second) of
    a `mib` MiB buffer: the practical HBM ceiling beside the nominal 8 TB/s."""
    nbytes = mib << 20
    src, dst = ca.DeviceBuffer(nbytes), ca.DeviceBuffer(nbytes)
    copy = lambda: ca._check(ca.lib().cpt_memcpy_dtod(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), nbytes),
                             "cpt_memcpy_dtod")
    copy()
    ca.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        copy()
    ca.synchronize()
    dt = time.perf_counter() - t0
    del src, dst
    return 2.0 * nbytes * reps / dt / 1e9


def cpu_baseline(sc, seconds_target=12.0, max_frames=64):
    """Oracle FP32 reference estimator over whole 1280x720 frames at 1 spp each
    (frames 0, 1, ... of the same stream) until ~seconds_target of wall time."""
    import oracle
    O = oracle.Scene(sc)
    threads = oracle.max_threads()
    p = oracle.render_params(WIDTH, HEIGHT, sc["camera"], eye_depth=EYE_DEPTH, light_depth=LIGHT_DEPTH)
    R = oracle.Renderer(p, 32)
    t_used, frames = 0.0, 0
    while t_used < seconds_target and frames < max_frames:
        t0 = time.perf_counter()
        R.render(O, frames, 1)
        t_used += time.perf_counter() - t0
        frames += 1
    e, s = R.ray_counts()
    return {"value": round((e + s) / t_used / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP32 reference estimator, {frames} full {WIDTH}x{HEIGHT} frame(s) at 1 spp "
                      f"({e + s} rays in {t_used:.1f} s, OpenMP {threads} threads, "
                      f"{platform.processor() or platform.machine()})"}


MLP_FLOP_PER_SAMPLE = 3 * 2 * (9 * 64 + 64 * 64 + 64)
BF16_PEAK_TFLOPS = 2500.0


def mlp_step_throughput(ca, m, X, T, batch=65536, reps=200):
    """Wall time and MFMA throughput of one training step at `batch` (forward,
    backward, gradient sum and Adam), steps issued back to back."""
    dx = ca.DeviceBuffer.from_array(np.ascontiguousarray(X[:batch], np.float32))
    dt = ca.DeviceBuffer.from_array(np.ascontiguousarray(T[:batch], np.float32))
    m.train_steps_device(dx, dt, batch, batch, 20, lr=0.0)
    ca.synchronize()
    ts = time.perf_counter()
    m.train_steps_device(dx, dt, batch, batch, reps, lr=0.0)          # one call, steps back to back
    ca.synchronize()
    step_s = (time.perf_counter() - ts) / reps
    tflops = MLP_FLOP_PER_SAMPLE * batch / step_s / 1e12
    return {"train_step_us": round(step_s * 1e6, 2), "flop_per_step": MLP_FLOP_PER_SAMPLE * batch,
            "achieved": round(tflops, 2), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / BF16_PEAK_TFLOPS, 5), "bound": "mfma",
            "note": f"{reps} steps (k_mlp_train + k_mlp_adam) in one cpt_mlp_train_steps call, wall time per step"}


def learned_leg(ca, args):
    """Time-to-equal-MSE of learned vs plain adaptive BDPT on the adaptive
    heightfield workload (cudapt_amd/learned.py): the 1,050,976-triangle
    heightfield under its HDR environment, 1920x1080, 16 training poses of
    low-spp / 4096-spp pairs, 4 held-out poses."""
    from cudapt_amd import learned, scenes
    W, H = args.learn_width, args.learn_height
    t0 = time.perf_counter()
    sc = scenes.heightfield(light="env")
    S = ca.Scene(sc)
    n_tr, n_te = args.learn_train, args.learn_test
    cams = learned.heightfield_poses(n_tr + n_te, width=W, height=H)
    attrs = dict(scenes.ADAPTIVE_C3)
    log(f"learned leg: scene in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    refs = [learned.converged(S, c, W, H, args.learn_ref_spp, depth=EYE_DEPTH) for c in cams]
    ref_s = time.perf_counter() - t0
    log(f"learned leg: {len(refs)} converged {args.learn_ref_spp}-spp images in {ref_s:.1f} s")
    t0 = time.perf_counter()
    X, T = learned.collect(S, cams[:n_tr], refs[:n_tr], W, H, checkpoints=(6, 8, 12, 16, 24, 32),
                           keep=args.learn_keep, depth=EYE_DEPTH, pass_frames=args.pass_frames, attribs=attrs)
    log(f"learned leg: {len(T)} training samples in {time.perf_counter() - t0:.1f} s")
    m = ca.Mlp(0xA22)
    losses, train_s = learned.train(m, X, T, steps=args.learn_steps, batch=65536, lr=1e-3)
    log(f"learned leg: trained in {train_s:.2f} s, loss {losses[0][1]:.3f} -> {losses[-1][1]:.3f}")
    rows = learned.time_to_equal_mse(S, m, cams[n_tr:], refs[n_tr:], W, H, frames=args.learn_frames,
                                     depth=EYE_DEPTH, pass_frames=args.pass_frames, signal=args.learn_signal,
                                     attribs=attrs)
    ratios = [r["learned_s"] / r["plain_adaptive_s"] for r in rows if r["learned_s"]]
    bdpt = [r["plain_bdpt_s"] / r["plain_adaptive_s"] for r in rows if r["plain_bdpt_s"]]
    gm = float(np.exp(np.mean(np.log(ratios)))) if len(ratios) == len(rows) else None
    for r in rows:
        r["learned_over_plain_adaptive"] = r["learned_s"] / r["plain_adaptive_s"] if r["learned_s"] else None
    mlp = mlp_step_throughput(ca, m, X, T)
    del S
    return m, {
        "learned_over_plain_adaptive": round(gm, 4) if gm is not None else None,
        "plain_bdpt_over_plain_adaptive": round(float(np.exp(np.mean(np.log(bdpt)))), 4) if bdpt else None,
        "target": "<= 0.5",
        "poses": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()} for r in rows],
        "protocol": {"scene": "heightfield 1,050,976 tris, 512x256 HDR environment of Gaussian lobes",
                     "width": W, "height": H, "depth": EYE_DEPTH, "train_poses": n_tr, "heldout_poses": n_te,
                     "ref_spp": args.learn_ref_spp, "training_pairs": "plain adaptive at 6, 8, 12, 16, 24 and "
                     f"32 spp against the {args.learn_ref_spp}-spp image", "plain_frames": args.learn_frames,
                     "pass_frames": args.pass_frames, "adaptive": attrs, "train_steps": args.learn_steps,
                     "batch": 65536, "lr": 1e-3, "train_samples": int(len(T)),
                     "mse": "radiance clamped to [0, 1]", "signal": args.learn_signal,
                     "timing": "synchronised wall time per render call, MSE readback excluded",
                     "reference_images_s": round(ref_s, 2)},
        "train_s": round(train_s, 3),
        "final_loss": round(float(losses[-1][1]), 4),
    }, mlp


def adaptive_leg(ca, args):
    """Adaptive sampling + cross-bilateral filter on the heightfield under its
    HDR environment, 1920x1080, 32 spp on average (5 warm-up frames, then
    per-frame proportional allocation), filter radius 7, colour / normal /
    relative-depth sigmas 0.2 / 0.1 / 0.05, 4 to 256 eye paths per pixel."""
    from cudapt_amd import scenes
    W, H, F = 1920, 1080, args.adaptive_spp
    t0 = time.perf_counter()
    sc = scenes.heightfield(light="env")
    S = ca.Scene(sc)
    log(f"adaptive leg: scene in {time.perf_counter() - t0:.1f} s")
    R = ca.Renderer(W, H, mode=ca.CPT_MODE_BDPT_ADAPTIVE_PLL, eye_depth=EYE_DEPTH, light_depth=LIGHT_DEPTH,
                    **scenes.ADAPTIVE_C3)
    R.set_attrib("Tail Bounce", args.adaptive_tail)
    cam = sc["camera"]
    R.render(S, dict(cam, pos=tuple(np.asarray(cam["pos"]) + 1e-3)), frames=8, result=False)   # allocations
    R.stats(reset=True)
    R.set_profiling(True)
    ca.synchronize()
    t0 = time.perf_counter()
    R.render(S, cam, frames=F, result=False)
    ca.synchronize()
    wall = time.perf_counter() - t0
    st = R.stats(reset=True)
    paths = R.read(ca.CPT_BUF_PIXEL_PATHS)
    filt_frames = F - 5
    res = {"workload": f"heightfield 1,050,976 tris + HDR environment, {W}x{H}, {F} spp average, proportional "
                       "allocation per frame, filter radius 7, sigmas colour/normal/depth 0.2/0.1/0.05, "
                       "4..256 paths per pixel",
           "seconds": round(wall, 4), "frames": F, "tail_bounce": args.adaptive_tail,
           "mrays_per_s": round((st["extension_rays"] + st["shadow_rays"]) / wall / 1e6, 2),
           "host_readback": "none in the timed region (the running mean stays in HBM)",
           "samples_per_s": round(st["eye_paths"] / wall, 1),
           "paths_per_pixel": {"mean": round(float(paths.mean()), 3), "min": int(paths.min()),
                               "max": int(paths.max())},
           "filter_ms_per_frame": round(st["filter_ms"] / max(filt_frames, 1), 4),
           "filter_gpix_per_s": round(W * H * filt_frames / (st["filter_ms"] / 1e3) / 1e9, 3)
           if st["filter_ms"] > 0 else None,
           "phase_ms": {k: round(st[k], 3) for k in ("extend_ms", "shade_ms", "shadow_ms", "tail_ms", "filter_ms",
                                                   "error_ms", "gen_ms", "accum_ms", "total_ms")}}
    del R, S
    return res


def counter_per_launch(path, ext_name, calib, counter=None):
    """(mean counter value per extension dispatch, scale) from one `rocprofv3
    --pmc <counters> --output-format csv` file, over the rows of `counter`
    (None: every row).  Dispatches whose name contains ext_name are averaged.
    For the byte counters (FETCH_SIZE / WRITE_SIZE, KiB) the scale is calib /
    what the counter read for the largest copy, which must be the calibration
    copy of calib bytes (the counter runs make no larger one); calib None: raw
    values, scale 1.  None if missing."""
    import csv
    ext, copy = [], []
    for row in csv.DictReader(open(path)):
        if counter is not None and row["Counter_Name"] != counter:
            continue
        v = float(row["Counter_Value"]) * (1024.0 if calib else 1.0)
        if ext_name in row["Kernel_Name"]:
            ext.append(v)
        elif "copyBuffer" in row["Kernel_Name"]:
            copy.append(v)
    if not ext:
        return None
    if calib is None:
        return sum(ext) / len(ext), 1.0
    if not copy or max(copy) <= 0:
        return None
    return sum(ext) / len(ext), calib / max(copy)


# One rocprofv3 --pmc pass each: the HBM byte counters, then the VALU
# counters with the derived VALUBusy (SQ_ACTIVE_INST_VALU over GUI-active
# cycles per CU: the share of time the VALUs work, dual issue counted once)
# and VALUUtilization (active lanes per VALU instruction).
COUNTER_PASSES = (("FETCH_SIZE",), ("WRITE_SIZE",), ("SQ_INSTS_VALU", "VALUBusy", "VALUUtilization"))


def counter_passes(args):
    """Per-extension-launch counters from rocprofv3 --pmc passes over one step
    of the same workload: FETCH_SIZE and WRITE_SIZE (HBM bytes, each scaled by
    what the counter reads for a 256 MiB device-to-device copy), and the VALU
    counters of COUNTER_PASSES."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    calib = 256 << 20
    ext_name = extend_kernel_name(args, demangled=True)
    out = {}
    for counters in COUNTER_PASSES:
        d = tempfile.mkdtemp(prefix="cpt_pmc_")
        cmd = ["timeout", "-s", "KILL", "120", prof, "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "t",
               "--", sys.executable, os.path.abspath(__file__), "--steps", "1", "--warmup", "0", "--no-cpu",
               "--no-traffic", "--no-learned", "--no-multiview", "--baseline-steps", "0", "--adaptive-spp", "0",
               "--frames", str(args.frames), "--in-flight", str(args.in_flight),
               "--lds-stack", str(args.lds_stack), "--lds-stack-shadow", str(args.lds_stack_shadow),
               "--node-format", str(args.node_format),
               "--refill", str(args.refill), "--refill-secondary", str(args.refill_secondary),
               "--refill-groups", str(args.refill_groups),
               "--refill-shadow", str(args.refill_shadow), "--tail-bounce", str(args.tail_bounce),
               "--copy-calib", str(calib >> 20)]
        log(f"counter pass: {' '.join(counters)}")
        r = subprocess.run(cmd, capture_output=True, text=True)
        files = [os.path.join(root, f) for root, _, fs in os.walk(d) for f in fs if f.endswith("counter_collection.csv")]
        if r.returncode != 0 or not files:
            log(f"counter pass {counters} failed (rc {r.returncode}); reported as null")
            return None
        for c in counters:
            got = counter_per_launch(files[0], ext_name, calib if c.endswith("_SIZE") else None, c)
            if got is None:
                return None
            out[c] = got
        shutil.rmtree(d, ignore_errors=True)
    fetch, ffac = out["FETCH_SIZE"]
    write, wfac = out["WRITE_SIZE"]
    return {"bytes_per_launch": int(fetch * ffac + write * wfac), "fetch_bytes": int(fetch * ffac),
            "write_bytes": int(write * wfac), "fetch_scale": round(ffac, 3), "write_scale": round(wfac, 3),
            "valu_insts_per_launch": int(out["SQ_INSTS_VALU"][0]), "valu_busy_pct": round(out["VALUBusy"][0], 2),
            "valu_lane_util_pct": round(out["VALUUtilization"][0], 2)}


def baselines_leg(ca, S, sc, args):
    """The reference's comparison set on the same scene, resolution and spp:
    the megakernel schedule of the same estimator (one thread per sample through
    every bounce; one step compared bit for bit with the wavefront's), and the
    unidirectional path tracers (cudaRTPTStream, cudaRTPTStreamRegen, eye depth
    10) and the "Path Regen" BDPT (cudaRTBDPTRegen: one thread per pixel, camera
    paths restarted in place, its own 640-vertex light cache), each timed over
    `--baseline-steps` steps."""
    def renderer(mode=0, mega=False):
        R = ca.Renderer(WIDTH, HEIGHT, mode=mode)
        if mode == 0:
            R.set_attrib("Eye Depth", EYE_DEPTH)
            R.set_attrib("Light Depth", LIGHT_DEPTH)
            R.set_attrib("Megakernel", 1 if mega else 0)
        R.set_attrib("Frames In Flight", args.in_flight)
        R.set_attrib("Traversal LDS Stack", args.lds_stack)
        R.set_attrib("Shadow LDS Stack", args.lds_stack_shadow or args.lds_stack)
        R.set_attrib("Node Format", args.node_format)
        R.set_attrib("Traversal Refill", args.refill)
        R.set_attrib("Secondary Refill", args.refill_secondary)
        R.set_attrib("Refill Workgroups", args.refill_groups)
        R.set_attrib("Shadow Refill", args.refill_shadow)
        return R

    def timed(R, label):
        R.render(S, cam, frames=args.frames, result=False)          # first-call allocations
        R.stats(reset=True)
        R.set_profiling(True)
        ca.synchronize()
        t0 = time.perf_counter()
        for k in range(args.baseline_steps):
            R.render(S, cam, frames=args.frames, result=False)
            ca.synchronize()
            log(f"{label} step {k + 1}/{args.baseline_steps}")
        wall = time.perf_counter() - t0
        st = R.stats(reset=True)
        rays = st["extension_rays"] + st["shadow_rays"]
        return {"mrays_per_s": round(rays / wall / 1e6, 2), "samples_per_s": round(st["eye_paths"] / wall, 1),
                "ms_per_step": round(wall / args.baseline_steps * 1e3, 3), "steps": args.baseline_steps}

    cam = sc["camera"]
    out = {}
    Rw, Rm = renderer(), renderer(mega=True)
    img_w = Rw.render(S, cam, frames=args.frames)
    img_m = Rm.render(S, cam, frames=args.frames)
    same = bool(np.array_equal(img_w.view(np.uint32), img_m.view(np.uint32)))
    del Rw, img_w, img_m
    log(f"megakernel: image identical to the wavefront's: {same}")
    Rm.reset()
    out["megakernel"] = timed(Rm, "megakernel")
    out["megakernel"]["same_image"] = same
    del Rm
    for mode, name in ((ca.CPT_MODE_PT_STREAM, "pt_stream"), (ca.CPT_MODE_PT_STREAM_REGEN, "pt_stream_regen")):
        R = renderer(mode)
        out[name] = timed(R, name)
        out[name]["eye_depth"] = 10
        del R
    R = ca.Renderer(WIDTH, HEIGHT, mode=ca.CPT_MODE_BDPT_REGEN)
    R.set_attrib("Eye Depth", EYE_DEPTH)
    R.set_attrib("Traversal LDS Stack", args.lds_stack)
    R.set_attrib("Node Format", args.node_format)
    out["bdpt_path_regen"] = timed(R, "bdpt_path_regen")
    out["bdpt_path_regen"]["eye_depth"] = EYE_DEPTH
    del R
    return out


def multiview_leg(ca, args, model=None):
    """Batched multi-view rendering: autotune, then 32 spp of every view, plain
    and with per-view (learned) adaptive sampling."""
    from cudapt_amd import scenes
    t0 = time.perf_counter()
    sc = scenes.heightfield(light="env")
    S = ca.Scene(sc)
    views = scenes.heightfield_poses(args.views)
    log(f"multiview: scene {len(sc['idx'])} triangles, {len(views)} views built in {time.perf_counter() - t0:.1f} s")
    B = ca.Batch(views, frames_in_flight=args.mv_in_flight, eye_depth=EYE_DEPTH, light_depth=LIGHT_DEPTH)
    t0 = time.perf_counter()
    best, cand_ms = B.autotune(S, iterations=1)
    log(f"multiview: autotuned in {time.perf_counter() - t0:.1f} s: {best}")
    B.reset()
    B.stats(reset=True)
    ca.synchronize()
    t0 = time.perf_counter()
    B.render(S, args.mv_spp)
    wall = time.perf_counter() - t0
    st = B.stats()
    px = sum(v["width"] * v["height"] for v in views)
    res = {"views": len(views), "pixels": px, "spp": args.mv_spp, "triangles": int(len(sc["idx"])),
           "seconds": round(wall, 4), "gpu_ms": round(st["total_ms"], 3),
           "mrays_per_s": round((st["extension_rays"] + st["shadow_rays"]) / wall / 1e6, 2),
           "host_readback": "none in the timed region (the running mean stays in HBM)",
           "samples_per_s": round(st["eye_paths"] / wall, 1),
           "resolutions": sorted({(v["width"], v["height"]) for v in views}),
           "autotuned": best, "candidates_ms_per_iteration": [round(m, 3) for m in cand_ms]}
    # the same views with per-view adaptive sampling (learned when a model is
    # given): 32 paths per pixel on average per view (5 warm-up frames, then
    # one-frame passes), the adaptive workload's filter and path bounds
    B.set_adaptive(ca.CPT_MODE_BDPT_ADAPTIVE_PLL, 1, model, **scenes.ADAPTIVE_C3)
    if model is not None:
        B.set_attrib("Learned Signal", args.learn_signal)
    B.render(S, 6)                                   # allocations, outside the timing
    B.reset()
    B.stats(reset=True)
    ca.synchronize()
    t0 = time.perf_counter()
    B.render(S, args.mv_spp)
    wall = time.perf_counter() - t0
    st = B.stats()
    paths = np.concatenate([B.read(v, ca.CPT_BUF_PIXEL_PATHS).ravel() for v in range(len(views))])
    res["adaptive"] = {"learned": model is not None, "seconds": round(wall, 4),
                       "mrays_per_s": round((st["extension_rays"] + st["shadow_rays"]) / wall / 1e6, 2),
                       "samples_per_s": round(st["eye_paths"] / wall, 1),
                       "paths_per_pixel": {"mean": round(float(paths.mean()), 3), "min": int(paths.min()),
                                           "max": int(paths.max())},
                       "phase_ms": {k: round(st[k], 3) for k in ("extend_ms", "shade_ms", "shadow_ms", "filter_ms",
                                                               "error_ms", "gen_ms", "accum_ms", "mlp_ms")}}
    del B, S
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--frames", type=int, default=SPP)
    ap.add_argument("--in-flight", type=int, default=64, help="frames traced as one wavefront")
    ap.add_argument("--lds-stack", type=int, default=806, help="traversal variant: LDS stack entries (8/16/24; "
                    "806/807/808/1606 = 8 or 16 entries with >= 6, 7 or 8 waves per SIMD)")
    ap.add_argument("--lds-stack-shadow", type=int, default=806, help="the same for shadow rays (0: as --lds-stack)")
    ap.add_argument("--node-format", type=int, default=1, help="BVH nodes: 0 = 128-byte FP32, 1 = 64-byte quantized")
    ap.add_argument("--refill", type=int, default=64,
                    help="traversal lanes refill from a ray queue once this many of a wave's are idle (0: off)")
    ap.add_argument("--refill-secondary", type=int, default=48,
                    help="the same for extension rays after the first bounce")
    ap.add_argument("--refill-shadow", type=int, default=16, help="the same for shadow rays")
    ap.add_argument("--refill-groups", type=int, default=8, help="workgroups per CU of the refilling grid")
    ap.add_argument("--tail-bounce", type=int, default=0,
                    help="bounces from this one on run as one k_tail launch (0: every bounce as wavefront passes)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM-traffic passes")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--no-learned", action="store_true", help="skip the time-to-equal-MSE leg")
    ap.add_argument("--no-multiview", action="store_true", help="skip the batched multi-view leg")
    ap.add_argument("--baseline-steps", type=int, default=2,
                    help="timed steps of each comparison baseline: megakernel, path tracers (0: skip)")
    ap.add_argument("--views", type=int, default=32)
    ap.add_argument("--mv-spp", type=int, default=32)
    ap.add_argument("--mv-in-flight", type=int, default=3)
    ap.add_argument("--learn-train", type=int, default=16)
    ap.add_argument("--learn-test", type=int, default=4)
    ap.add_argument("--learn-width", type=int, default=1920)
    ap.add_argument("--learn-height", type=int, default=1080)
    ap.add_argument("--learn-ref-spp", type=int, default=4096)
    ap.add_argument("--learn-keep", type=float, default=1.0 / 32)
    ap.add_argument("--learn-steps", type=int, default=2000)
    ap.add_argument("--learn-frames", type=int, default=32, help="plain adaptive frames whose MSE is the target")
    ap.add_argument("--learn-signal", type=int, default=3, help="allocation signal of the learned renderer")
    ap.add_argument("--pass-frames", type=int, default=1, help="adaptive pass size in frames (learned leg)")
    ap.add_argument("--adaptive-spp", type=int, default=32, help="average spp of the adaptive leg (0: skip)")
    ap.add_argument("--mlp-only", action="store_true",
                    help="only the model's training-step throughput on synthetic features (for profiling)")
    ap.add_argument("--multiview-only", action="store_true",
                    help="only the multi-view leg, plain and adaptive without a model (for profiling)")
    ap.add_argument("--adaptive-only", action="store_true", help="only the adaptive leg (for profiling)")
    ap.add_argument("--adaptive-tail", type=int, default=2,
                    help="Tail Bounce of the adaptive leg: bounces t.. of its adaptive frames in one launch (0: off)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the library renders on one "
              "device; several GPUs would each take their own views or frames)", file=sys.stderr)
        sys.exit(2)

    import cudapt_amd as ca
    from cudapt_amd import scenes

    if args.copy_calib:
        nbytes = args.copy_calib << 20
        src, dst = ca.DeviceBuffer(nbytes), ca.DeviceBuffer(nbytes)
        ca._check(ca.lib().cpt_memcpy_dtod(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), nbytes),
                  "cpt_memcpy_dtod")
        ca.synchronize()
        log(f"calibration copy: {nbytes} bytes device to device")
        del src, dst
    if args.mlp_only:
        rng = np.random.default_rng(7)
        X = rng.normal(size=(65536, 9)).astype(np.float32)
        T = (X[:, :3].sum(axis=1) + 0.1 * rng.normal(size=65536)).astype(np.float32)
        m = ca.Mlp(0xA22)
        print(json.dumps({"mlp": mlp_step_throughput(ca, m, X, T)}), flush=True)
        return
    if args.multiview_only:
        print(json.dumps({"multiview": multiview_leg(ca, args)}), flush=True)
        return
    if args.adaptive_only:
        print(json.dumps({"adaptive": adaptive_leg(ca, args)}), flush=True)
        return
    t0 = time.perf_counter()
    sc = scenes.sphere_grid()
    S = ca.Scene(sc)
    info = S.bvh4_info()
    log(f"scene: {len(sc['idx'])} triangles, bvh4 {info}, built+uploaded in {time.perf_counter() - t0:.1f} s")
    R = ca.Renderer(WIDTH, HEIGHT, eye_depth=EYE_DEPTH, light_depth=LIGHT_DEPTH)
    R.set_attrib("Frames In Flight", args.in_flight)
    R.set_attrib("Traversal LDS Stack", args.lds_stack)
    R.set_attrib("Shadow LDS Stack", args.lds_stack_shadow or args.lds_stack)
    R.set_attrib("Node Format", args.node_format)
    R.set_attrib("Traversal Refill", args.refill)
    R.set_attrib("Secondary Refill", args.refill_secondary)
    R.set_attrib("Refill Workgroups", args.refill_groups)
    R.set_attrib("Shadow Refill", args.refill_shadow)
    R.set_attrib("Tail Bounce", args.tail_bounce)
    cam = sc["camera"]
    for w in range(args.warmup):
        R.render(S, cam, frames=args.frames, result=False)
        log(f"warmup {w + 1}/{args.warmup} done")
    R.stats(reset=True)
    R.set_profiling(True)
    ca.synchronize()
    times = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        R.render(S, cam, frames=args.frames, result=False)   # ends with a stream synchronisation
        ca.synchronize()
        times.append(time.perf_counter() - ts)
        log(f"step {k + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms")
    wall = time.perf_counter() - t_start
    st = R.stats(reset=True)
    R.set_profiling(False)
    bounce_rays = R.read(ca.CPT_BUF_BOUNCE_RAYS)      # the last step's per-bounce queue lengths
    rays = st["extension_rays"] + st["shadow_rays"]
    mrays = rays / wall / 1e6
    launches = max(st["extend_launches"], 1)
    node_bytes = 64 if args.node_format else 128
    # first-bounce eye rays are computed inside the extension kernel (no 32 B
    # ray read) when the refilling grid runs the first bounce over the
    # quantized nodes, as the renderer's schedule does above Refill Min Paths
    eye_in_kernel = args.refill >= 1 and args.node_format != 0 and WIDTH * HEIGHT * args.in_flight >= (1 << 20)
    ext_rays = st["extension_rays"] - st["tail_rays"]     # the tail kernel's rays are not the extension kernel's
    rays_read = ext_rays - (st["eye_paths"] if eye_in_kernel else 0)
    bytes_alg = 32 * rays_read + 16 * ext_rays + node_bytes * st["extend_nodes"] + 48 * st["extend_tris"]
    ext_s = st["extend_ms"] / 1e3
    achieved = bytes_alg / ext_s / 1e9 if ext_s > 0 else 0.0
    launch_s = ext_s / launches
    res = {
        "metric": "Mrays/s + time-to-equal-MSE (learned vs plain adaptive BDPT), 1 MI355X",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "streaming BDPT, sphere grid 262,144 tris + 2 sphere lights, 1280x720, "
                               "depth 8, 64 spp per step, light-vertex cache every 3rd frame",
                   "triangles": int(len(sc["idx"])), "width": WIDTH, "height": HEIGHT, "spp": args.frames,
                   "eye_depth": EYE_DEPTH, "light_depth": LIGHT_DEPTH, "frames_in_flight": args.in_flight,
                   "traversal_variant": args.lds_stack,
                   "shadow_variant": args.lds_stack_shadow or args.lds_stack, "node_format": args.node_format,
                   "traversal_refill": args.refill, "secondary_refill": args.refill_secondary,
                   "refill_groups": args.refill_groups,
                   "shadow_refill": args.refill_shadow, "tail_bounce": args.tail_bounce},
        "host_readback": "none in the timed region: the running mean stays in HBM (the reference copies it "
                         "to the host after every frame, src/cudaRTBDPTStream.cu:949, 11 MB per 1280x720 image)",
        "samples_per_s": round(st["eye_paths"] / wall, 1),
        "rays": {"extension": st["extension_rays"], "shadow": st["shadow_rays"],
                 "extension_per_bounce": [int(x) for x in bounce_rays[0][:EYE_DEPTH]],
                 "shadow_per_bounce": [int(x) for x in bounce_rays[1][:EYE_DEPTH]],
                 "queue_bytes_per_bounce": queue_bytes(bounce_rays, eye_in_kernel)},
        "kernel_ms": {"extend": round(st["extend_ms"], 3), "shade": round(st["shade_ms"], 3),
                      "shadow": round(st["shadow_ms"], 3), "light_pass": round(st["light_ms"], 3),
                      "tail": round(st["tail_ms"], 3), "all": round(st["total_ms"], 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "traffic_detail": None,
                     "kernel": extend_kernel_name(args),
                     "launches": launches, "avg_launch_us": round(st["extend_ms"] * 1e3 / launches, 2),
                     "bytes_per_launch": int(bytes_alg / launches),
                     "nodes_per_ray": round(st["extend_nodes"] / max(ext_rays, 1), 2),
                     "tris_per_ray": round(st["extend_tris"] / max(ext_rays, 1), 2),
                     "shadow_nodes_per_ray": round(st["shadow_nodes"] / max(st["shadow_rays"], 1), 2),
                     "shadow_tris_per_ray": round(st["shadow_tris"] / max(st["shadow_rays"], 1), 2),
                     "eye_rays_computed_in_kernel": bool(eye_in_kernel),
                     "stack_spills_per_ray": round(st["extend_spills"] / max(ext_rays, 1), 5),
                     "shadow_stack_spills_per_ray": round(st["shadow_spills"] / max(st["shadow_rays"], 1), 5),
                     "box_tests_per_s": round(4 * st["extend_nodes"] / ext_s, 1) if ext_s > 0 else None,
                     "tri_tests_per_s": round(st["extend_tris"] / ext_s, 1) if ext_s > 0 else None,
                     "hbm_measured": None, "valu": None, "limiter": None},
        "time_to_equal_mse": None,
    }
    del R
    if not args.copy_calib:      # a counter pass scales by its one calibration copy: no other large copies
        peak = copy_peak_gbs(ca)
        log(f"device-to-device copy: {peak:.0f} GB/s")
        res["roofline"]["copy_peak_gbs"] = round(peak, 1)
        res["roofline"]["frac_of_copy_peak"] = round(achieved / peak, 4)
    res["baselines"] = None
    if args.baseline_steps > 0:
        res["baselines"] = baselines_leg(ca, S, sc, args)
        res["baselines"]["wavefront_over_megakernel"] = round(
            mrays / max(res["baselines"]["megakernel"]["mrays_per_s"], 1e-9), 3)
        res["baselines"]["wavefront_over_path_regen"] = round(
            mrays / max(res["baselines"]["bdpt_path_regen"]["mrays_per_s"], 1e-9), 3)
    del S
    res["adaptive"] = adaptive_leg(ca, args) if args.adaptive_spp > 0 else None
    res["mlp"] = None
    model = None
    if not args.no_learned:
        model, res["time_to_equal_mse"], res["mlp"] = learned_leg(ca, args)
    res["multiview"] = None
    if not args.no_multiview:
        res["multiview"] = multiview_leg(ca, args, model)
    if not args.no_traffic:
        tr = counter_passes(args)
        rl = res["roofline"]
        rl["traffic_detail"] = tr
        rl["traffic"] = tr["bytes_per_launch"] if tr else None
        if tr and launch_s > 0:
            # what HBM actually delivered (FETCH_SIZE + WRITE_SIZE) against its
            # peak, beside the algorithmic-byte fraction above
            gbs = tr["bytes_per_launch"] / launch_s / 1e9
            rl["hbm_measured"] = {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(gbs / HBM_PEAK_GBS, 4)}
            # VALU: VALUBusy, the share of the kernel's time the CUs' VALUs
            # work (gfx950 can issue two VALU instructions in one cycle, so the
            # instruction rate alone does not bound it)
            valu_rate = tr["valu_insts_per_launch"] / launch_s
            rl["valu"] = {"busy_pct": tr["valu_busy_pct"], "frac": round(tr["valu_busy_pct"] / 100.0, 4),
                          "lane_util_pct": tr["valu_lane_util_pct"],
                          "wave_insts_per_launch": tr["valu_insts_per_launch"],
                          "wave_insts_per_cu_cycle": round(valu_rate / (256 * 2.4e9), 3),
                          "per_ray": round(tr["valu_insts_per_launch"] * 64 / max(ext_rays / launches, 1),
                                           1)}
            hf, vf = rl["hbm_measured"]["frac"], rl["valu"]["frac"]
            rl["limiter"] = ("HBM bandwidth" if hf >= 0.6 else
                             f"VALU (busy {vf:.0%} of the kernel's time, {tr['valu_lane_util_pct']:.0f} % of lanes "
                             f"active; HBM {hf:.0%} of peak)" if vf >= 0.6 else
                             f"latency of dependent node / triangle loads (HBM {hf:.0%} of peak, VALU busy "
                             f"{vf:.0%}; bytes served mostly by L2 / MALL)")
    if not args.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(sc)
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

"""Learned adaptive sampling: training data, training loop and the
time-to-equal-MSE protocol (learned vs plain adaptive BDPT).

Protocol (all on the device, one scene, one resolution):
  1. Converged images: plain streaming BDPT at `ref_spp` samples per pixel for
     every pose, traced with a seed of its own so that its noise is independent
     of the runs it scores.
  2. Training data: plain adaptive runs (filter-residual allocation) on the
     training poses; at each checkpoint frame the 9 features of every pixel
     and the target log(N |r - ref|^2 / 3) are taken (a random subset of pixels
     per checkpoint).
  3. Training: Adam on mean squared error over shuffled mini-batches.
  4. Evaluation on held-out poses: the plain adaptive renderer runs `frames`
     frames; its final MSE against the converged image (radiance clamped to
     CLAMP per channel on both sides, so that isolated fireflies of the
     heavy-tailed estimator do not decide the comparison) is the target.  The
     learned renderer (same settings, model attached) runs until its MSE is at
     or below that target; the metric is the ratio of the two rendering times
     (per-frame wall time, synchronised; MSE readback excluded).  Plain
     non-adaptive BDPT is timed to the same MSE as a second baseline.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import (CPT_MODE_BDPT_ADAPTIVE_PLL, CPT_MODE_BDPT_STREAM, DeviceBuffer, Mlp, Renderer, Scene,
               synchronize)

POSE_SEED = 0xCA3E0000
REF_SEED = 0x0EF5EED5
# the error is measured on radiance clamped to [0, 1] (the displayed image), the
# form left commented out in CUDAPTWindow::CalculateRMSE (src/cudapt.cpp:960)
CLAMP = 1.0


def orbit_poses(center, radius_range, height_range, target_extent, n, seed0=POSE_SEED, fov=math.radians(45.0)):
    """n camera poses on an orbit around `center`, pose i drawn from seed seed0 + i."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed0 + i)
        a = rng.uniform(0, 2 * math.pi)
        r = rng.uniform(*radius_range)
        h = rng.uniform(*height_range)
        pos = np.array([center[0] + r * math.cos(a), center[1] + h, center[2] + r * math.sin(a)])
        tgt = np.array([center[0] + rng.uniform(-target_extent, target_extent), center[1],
                        center[2] + rng.uniform(-target_extent, target_extent)])
        d = tgt - pos
        d /= np.linalg.norm(d)
        out.append(dict(pos=tuple(pos), dir=tuple(d), up=(0.0, 1.0, 0.0), fov=fov))
    return out


def sphere_grid_poses(n, seed0=POSE_SEED, grid=64, spacing=2.0):
    half = grid * spacing / 2.0
    return orbit_poses((0.0, 0.0, 0.0), (half * 0.6, half + 30.0), (15.0, 45.0), half / 2.0, n, seed0)


def heightfield_poses(n, seed0=POSE_SEED, width=1920, height=1080):
    """n seeded views over the heightfield (scenes.heightfield_poses), all at
    width x height: pose i from seed seed0 + i."""
    from .scenes import heightfield_poses as hp
    return [dict(p, width=width, height=height) for p in hp(n, seed0=seed0, resolutions=((width, height),))]


def converged(scene: Scene, cam, W, H, spp, depth=8, in_flight=64, seed=REF_SEED):
    R = Renderer(W, H, mode=CPT_MODE_BDPT_STREAM, eye_depth=depth, light_depth=depth, seed=seed)
    R.set_attrib("Frames In Flight", min(in_flight, spp))
    return R.render(scene, cam, frames=spp)


def _adaptive(W, H, depth, pass_frames=1, signal=0, mode=CPT_MODE_BDPT_ADAPTIVE_PLL, attribs=None):
    R = Renderer(W, H, mode=mode, eye_depth=depth, light_depth=depth, **(attribs or {}))
    R.set_attrib("Adaptive Pass Frames", pass_frames)
    R.set_attrib("Learned Signal", signal)
    return R


def collect(scene: Scene, cams, refs, W, H, checkpoints=(6, 8, 12, 16, 24, 32, 48, 64), keep=0.125, depth=8,
            seed=0xDA7A, pass_frames=1, clamp=CLAMP, attribs=None):
    """Features and targets of plain adaptive runs at the checkpoint frames
    (attribs: adaptive renderer attributes, e.g. scenes.ADAPTIVE_C3)."""
    rng = np.random.default_rng(seed)
    n = W * H
    dx, dt = DeviceBuffer(n * 36), DeviceBuffer(n * 4)
    X, T = [], []
    for cam, ref in zip(cams, refs):
        d_ref = DeviceBuffer.from_array(np.ascontiguousarray(ref, np.float32))
        R = _adaptive(W, H, depth, pass_frames, attribs=attribs)
        done = 0
        for c in checkpoints:
            R.render(scene, cam, frames=c - done, result=False)
            done = c
            R.mlp_features(dx)
            R.mlp_target(d_ref, dt, clamp)
            sel = rng.random(n) < keep
            X.append(dx.download(np.float32, n * 9).reshape(n, 9)[sel])
            T.append(dt.download(np.float32, n)[sel])
    return np.concatenate(X), np.concatenate(T)


def train(model: Mlp, X, T, steps=2000, batch=65536, lr=1e-3, seed=0x7EA1, log_every=250):
    """Adam over shuffled mini-batches of (X, T); returns (losses, seconds)."""
    rng = np.random.default_rng(seed)
    m = len(T)
    batch = min(batch, m)
    perm = rng.permutation(m)
    dx = DeviceBuffer.from_array(np.ascontiguousarray(X[perm], np.float32))
    dt = DeviceBuffer.from_array(np.ascontiguousarray(T[perm], np.float32))
    losses = []
    synchronize()
    t0 = time.perf_counter()
    nb = m // batch
    first = model.train_step_device(dx, dt, batch, lr=lr, want_loss=True)      # step 1 reads batch 0
    losses.append((1, first))
    done = 1
    while done < steps:                       # the rest in device-side runs between loss readings
        k = min(log_every - done % log_every, steps - done)
        loss = model.train_steps_device(dx, dt, nb * batch, batch, k, lr=lr, first_batch=done % nb, want_loss=True)
        done += k
        losses.append((done, loss))
    synchronize()
    return losses, time.perf_counter() - t0


def mse(img, ref, clamp=CLAMP) -> float:
    c = clamp if clamp and clamp > 0 else np.inf
    d = np.minimum(np.asarray(img, np.float64), c) - np.minimum(np.asarray(ref, np.float64), c)
    return float(np.mean(d * d))


def mse_curve(scene: Scene, cam, ref, W, H, frames, model: Mlp | None = None, mode=CPT_MODE_BDPT_ADAPTIVE_PLL,
              depth=8, stop_below=None, pass_frames=1, clamp=CLAMP, signal=0, attribs=None):
    """(frames, cumulative render seconds, MSE, eye paths per pixel) after each
    render call of pass_frames frames.  With stop_below the run ends at the
    first call whose MSE is at or below it."""
    if mode != CPT_MODE_BDPT_STREAM:
        R = _adaptive(W, H, depth, pass_frames, signal, mode=mode, attribs=attribs)
        if model is not None:
            R.set_mlp(model)
    else:
        R = Renderer(W, H, mode=mode, eye_depth=depth, light_depth=depth)
        R.set_attrib("Frames In Flight", pass_frames)
    # allocations happen on the first render: done on a nearby camera (whose
    # accumulation the timed run's camera change discards), outside the timing
    warm = dict(cam, pos=tuple(np.asarray(cam["pos"]) + 1e-3))
    R.render(scene, warm, frames=min(frames, 6 + 2 * pass_frames), result=False)
    synchronize()
    R.stats(reset=True)
    fs, ts, ms, ps = [], [], [], []
    total = 0.0
    done = 0
    while done < frames:
        k = min(pass_frames, frames - done)
        synchronize()
        t0 = time.perf_counter()
        R.render(scene, cam, frames=k, result=False)
        synchronize()
        total += time.perf_counter() - t0
        done += k
        e = mse(R.read(0), ref, clamp)
        fs.append(done)
        ts.append(total)
        ms.append(e)
        ps.append(R.stats()["eye_paths"] / (W * H))
        if stop_below is not None and e <= stop_below:
            break
    return np.array(fs), np.array(ts), np.array(ms), np.array(ps)


def time_to_reach(ts, ms, target):
    """Time at which a curve first reaches `target` MSE, interpolating log MSE
    linearly between frames; None when it never does."""
    for i, e in enumerate(ms):
        if e <= target:
            if i == 0:
                return float(ts[0])
            t0, t1 = ts[i - 1], ts[i]
            l0, l1 = math.log(ms[i - 1]), math.log(e)
            w = 0.0 if l0 == l1 else (l0 - math.log(target)) / (l0 - l1)
            return float(t0 + w * (t1 - t0))
    return None


def time_to_equal_mse(scene: Scene, model: Mlp, cams, refs, W, H, frames=64, depth=8, max_frames=None,
                      pass_frames=1, clamp=CLAMP, signal=0, attribs=None):
    """Per held-out pose: plain adaptive time, learned time and plain BDPT time
    to plain adaptive's MSE after `frames` frames; eye paths per pixel each
    spent; the learned renderer's MSE after `frames` frames (a run of its own,
    not stopped at the target); and plain BDPT's MSE after as many frames as
    the learned renderer spent paths per pixel (rounded up): uniform sampling
    at the learned renderer's sample count."""
    max_frames = max_frames or 4 * frames
    # warm the code paths once (module load, first allocations)
    mse_curve(scene, cams[0], refs[0], W, H, 2 * pass_frames + 5, model, depth=depth, pass_frames=pass_frames,
              attribs=attribs)
    rows = []
    for cam, ref in zip(cams, refs):
        fp, tp, mp, pp = mse_curve(scene, cam, ref, W, H, frames, None, depth=depth, pass_frames=pass_frames,
                               clamp=clamp, attribs=attribs)
        target = float(mp[-1])
        fl, tl, ml, pl = mse_curve(scene, cam, ref, W, H, max_frames, model, depth=depth, stop_below=target,
                               pass_frames=pass_frames, clamp=clamp, signal=signal, attribs=attribs)
        fu, tu, mu, pu = mse_curve(scene, cam, ref, W, H, max_frames, None, mode=CPT_MODE_BDPT_STREAM, depth=depth,
                               stop_below=target, pass_frames=pass_frames, clamp=clamp)
        _, _, me, _ = mse_curve(scene, cam, ref, W, H, frames, model, depth=depth, pass_frames=pass_frames,
                                clamp=clamp, signal=signal, attribs=attribs)
        k = int(math.ceil(pl[-1]))                     # uniform frames = paths per pixel
        at = np.nonzero(fu >= k)[0]
        rows.append(dict(target_mse=target, plain_adaptive_s=float(tp[-1]),
                         learned_s=time_to_reach(tl, ml, target), plain_bdpt_s=time_to_reach(tu, mu, target),
                         learned_frames=int(fl[-1]), plain_bdpt_frames=int(fu[-1]),
                         plain_adaptive_paths_per_pixel=float(pp[-1]), learned_paths_per_pixel=float(pl[-1]),
                         plain_bdpt_paths_per_pixel=float(pu[-1]),
                         learned_mse_at_equal_frames=float(me[-1]),
                         plain_bdpt_mse_at_learned_paths=float(mu[at[0]]) if at.size and fu[at[0]] == k else None))
    return rows

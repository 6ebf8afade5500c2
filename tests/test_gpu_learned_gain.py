"""What the trained error model gains over the method without it: learned
against plain adaptive sampling on held-out poses of the environment-lit
heightfield, a small instance of the time-to-equal-MSE protocol
(cudapt_amd/learned.py):

  - 320 x 180, depth 8, the adaptive workload's filter and path bounds;
  - converged images: plain BDPT at 512 spp on a seed of their own;
  - training: features and targets of plain adaptive runs on 4 poses,
    300 Adam steps at batch 16,384;
  - evaluation on 2 held-out poses: plain adaptive's MSE after 24 frames
    (radiance clamped to [0, 1]) is the target.

Rendering, training and allocation are deterministic (fixed seeds and
fixed-order sums), so the outcome is too.  Gates, on every held-out pose:
the learned renderer reaches the target in fewer frames and with fewer eye
paths per pixel than plain adaptive, and its MSE after those 24 frames is
lower.  The same protocol with the untrained model (its random
initialisation) is printed beside it as the control.
"""
import numpy as np
import pytest

from cudapt_amd import learned, scenes

pytestmark = pytest.mark.gpu

W, H, FRAMES = 320, 180, 24


@pytest.fixture(scope="module")
def protocol(cpt):
    print("learned gain: heightfield 320x180, 4 training + 2 held-out poses", flush=True)
    sc = scenes.heightfield(light="env")
    S = cpt.Scene(sc)
    cams = learned.heightfield_poses(6, width=W, height=H)
    refs = [learned.converged(S, c, W, H, 512) for c in cams]
    attrs = dict(scenes.ADAPTIVE_C3)
    X, T = learned.collect(S, cams[:4], refs[:4], W, H, checkpoints=(6, 8, 12, 16, 24), keep=0.5, attribs=attrs)
    untrained = cpt.Mlp(0xA22)
    m = cpt.Mlp(0xA22)
    losses, _ = learned.train(m, X, T, steps=300, batch=16384, lr=1e-3, log_every=100)
    print(f"learned gain: {len(T)} samples, loss {losses[0][1]:.3f} -> {losses[-1][1]:.3f}", flush=True)
    out = {"losses": losses}
    for name, model in (("trained", m), ("untrained", untrained)):
        rows = []
        for cam, ref in zip(cams[4:], refs[4:]):
            fp, _, mp, pp = learned.mse_curve(S, cam, ref, W, H, FRAMES, None, attribs=attrs)
            fl, _, ml, pl = learned.mse_curve(S, cam, ref, W, H, FRAMES, model, signal=3, attribs=attrs)
            hit = np.nonzero(ml <= mp[-1])[0]
            first = int(fl[hit[0]]) if hit.size else None
            rows.append(dict(target=float(mp[-1]), plain_paths=float(pp[-1]), learned_mse=float(ml[-1]),
                             first=first, learned_paths=float(pl[hit[0]]) if hit.size else None))
            print(f"learned gain [{name}]: target {mp[-1]:.5f} at {FRAMES} frames ({pp[-1]:.2f} paths/px); learned "
                  f"reaches it at frame {first} ({rows[-1]['learned_paths']} paths/px), MSE at {FRAMES} frames "
                  f"{ml[-1]:.5f}", flush=True)
        out[name] = rows
    return out


def test_training_lowers_the_loss(protocol):
    losses = protocol["losses"]
    assert losses[-1][1] < 0.5 * losses[0][1]


def test_trained_model_reaches_plain_adaptive_mse_sooner(protocol):
    for r in protocol["trained"]:
        assert r["first"] is not None and r["first"] < FRAMES, r
        assert r["learned_paths"] < r["plain_paths"], r
        assert r["learned_mse"] < r["target"], r

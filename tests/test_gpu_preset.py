"""The reference's own experiment scene (cornellwtranball.NPRTENV: the Cornell
box with a glass sphere, one area light) loaded through the preset reader and
rendered by the GPU streaming BDPT, against the oracle on the same scene.

The FP32 host run is the device operation sequence, transcendentals included,
so the FP32 GPU image and counts equal it exactly, refraction through the
sphere and all.  Against FP64 the gate is 1e-3 (measured 1.7e-4: the FP32 host
run itself differs from FP64 by 1.8e-4, a flipped Russian-roulette or branch
decision through the glass changing a whole path)."""
import os

import numpy as np
import pytest

import oracle
from cudapt_amd import preset
from tests.helpers import rel_mse

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "presets")
W, H, SPP = 96, 72, 8


@pytest.fixture(scope="module")
def runs(cpt):
    p = preset.load_preset(os.path.join(GOLD, "cornellwtranball.NPRTENV"), model_dirs=[GOLD])
    sc = p["scene"]
    S = cpt.Scene(sc)
    R = cpt.Renderer(W, H)
    gpu = R.render(S, p["camera"], frames=SPP)
    O = oracle.Scene(sc)
    prm = oracle.render_params(W, H, p["camera"])
    r32, n32 = oracle.Renderer(prm, 32).render(O, 0, SPP)
    o64 = oracle.Renderer(prm, 64)
    o64.render(O, 0, SPP)
    return dict(gpu=gpu, n=R.sample_counts(), r32=r32, n32=n32, r64=o64.result64(), R=R, S=S, p=p)


def test_preset_scene_matches_fp32_oracle(runs):
    g, r = runs["gpu"], runs["r32"]
    e = rel_mse(g, r)
    close = (np.abs(g - r) <= 1e-4 * np.maximum(np.abs(r), 1e-3)).all(axis=-1).mean()
    print(f"preset scene: relative MSE vs FP32 oracle {e:.3e}, pixels within 1e-4: {close:.4f}")
    assert e <= 1e-4
    assert close >= 0.97
    assert (runs["n"] == runs["n32"]).mean() > 0.99
    assert np.array_equal(g.view(np.uint32), r.view(np.uint32)) and np.array_equal(runs["n"], runs["n32"])


def test_preset_scene_matches_fp64_oracle(runs):
    e = rel_mse(runs["gpu"], runs["r64"])
    print(f"preset scene: relative MSE vs FP64 oracle {e:.3e}")
    assert e <= 1e-3


def test_preset_scene_is_lit_and_finite(runs):
    g = runs["gpu"]
    assert np.isfinite(g).all() and g.mean() > 0.05 and (g > 0).mean() > 0.9


def test_rmse_experiment_rows(cpt):
    from cudapt_amd import experiment
    out = experiment.run_preset(os.path.join(GOLD, "cornellwtranball.NPRTENV"), [GOLD], width=64, height=48,
                                converged_spp=256, interval=0.004, count=4)
    assert set(out["curves"]) == {"sc", "atomic"} and out["converged"].shape == (48, 64, 3)
    for mode, rows in out["curves"].items():
        assert len(rows) == 4, mode
        secs = [r.sec for r in rows]
        frames = [r.frames for r in rows]
        assert all(b > a for a, b in zip(secs, secs[1:])) and all(b > a for a, b in zip(frames, frames[1:]))
        assert all(s >= 0.004 * (i + 1) for i, s in enumerate(secs))
        assert rows[-1].rmse < rows[0].rmse, mode

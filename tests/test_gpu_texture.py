"""Texture path on the GPU against the oracle (textured Cornell box: checker
floor with wrap, ramp wall, textured box, emissive-textured light).

The FP32 GPU and FP32 host runs (the device operation sequence, transcendentals
included) are identical, with and without refraction; against FP64 the gates
are the Cornell box's (1e-4).  With the box made of textured glass
(transparency = texture alpha x 0.9) light subpaths refract through it, and a
last-bit difference between FP32 and FP64 changes cached light vertices that
every pixel connects to: the host's FP32 and FP64 runs differ by relative MSE
8.6e-4.  That case is gated at the noise level against FP64, 3e-3 and mean
radiance within 1 %, which still fails if alpha is ignored (8.2e-3 and 3 %)."""
import numpy as np
import pytest

import oracle
from cudapt_amd import scenes
from tests.helpers import rel_mse

pytestmark = pytest.mark.gpu

W, H, SPP = 128, 128, 16


def _runs(cpt, sc):
    S = cpt.Scene(sc)
    R = cpt.Renderer(W, H)
    gpu = R.render(S, sc["camera"], frames=SPP)
    O = oracle.Scene(sc)
    p = oracle.render_params(W, H, sc["camera"])
    r32, n32 = oracle.Renderer(p, 32).render(O, 0, SPP)
    o64 = oracle.Renderer(p, 64)
    o64.render(O, 0, SPP)
    return dict(sc=sc, S=S, gpu=gpu, n=R.sample_counts(), r32=r32, n32=n32, r64=o64.result64())


@pytest.fixture(scope="module")
def runs(cpt):
    return _runs(cpt, scenes.textured_cornell())


def test_textured_scene_matches_fp32_oracle(runs):
    e = rel_mse(runs["gpu"], runs["r32"])
    print(f"textured scene: relative MSE vs FP32 oracle {e:.3e}")
    assert e <= 1e-6
    assert (runs["n"] == runs["n32"]).mean() > 0.995
    assert np.array_equal(runs["gpu"].view(np.uint32), runs["r32"].view(np.uint32))
    assert np.array_equal(runs["n"], runs["n32"])


def test_textured_scene_matches_fp64_oracle(runs):
    e = rel_mse(runs["gpu"], runs["r64"])
    print(f"textured scene: relative MSE vs FP64 oracle {e:.3e}")
    assert e <= 1e-4


def test_textured_glass_within_noise(cpt):
    r = _runs(cpt, scenes.textured_cornell(glass=True))
    e = rel_mse(r["gpu"], r["r64"])
    dm = abs(r["gpu"].mean() - r["r64"].mean()) / r["r64"].mean()
    print(f"textured glass: relative MSE vs FP64 oracle {e:.3e}, mean radiance off by {dm:.4f}")
    assert e <= 3e-3 and dm <= 0.01
    assert np.array_equal(r["gpu"].view(np.uint32), r["r32"].view(np.uint32)) and np.array_equal(r["n"], r["n32"])


def test_textures_change_the_image(cpt, runs):
    plain = cpt.Renderer(W, H).render(cpt.Scene(scenes.cornell()), runs["sc"]["camera"], frames=SPP)
    assert rel_mse(plain, runs["gpu"]) > 1e-2


def test_out_of_range_ids_mean_no_texture(cpt, runs):
    sc = runs["sc"]
    mats = sc["mats"].copy()
    mats[0]["diffuseTexId"] = 99                    # no texture 99: the material colour
    a = cpt.Renderer(32, 32).render(cpt.Scene(dict(sc, mats=mats)), sc["camera"], frames=2)
    mats[0]["diffuseTexId"] = -1
    b = cpt.Renderer(32, 32).render(cpt.Scene(dict(sc, mats=mats)), sc["camera"], frames=2)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_texture_argument_errors(cpt):
    S = cpt.Scene(scenes.cornell())
    with pytest.raises(cpt.CptError):
        S.add_texture(np.zeros((0, 4, 4), np.float32))
    with pytest.raises(cpt.CptError):
        S.add_texture(np.zeros((4, 4, 3), np.float32))
    assert S.add_texture(np.ones((2, 3, 4), np.float32)) == 0
    assert S.add_texture(np.ones((1, 1, 4), np.float32)) == 1

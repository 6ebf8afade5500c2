"""The Disney-BRDF path tracer (CPT_MODE_PT_DISNEY_MIS; an extension with no
reference counterpart) on the GPU against the oracle's FP64 restatement
(mode 7), on the same counter-based RNG stream: the Cornell box with random
Disney parameters per surface (metallic, roughness, specular, sheen, clearcoat),
next-event estimation to the ceiling light with power-heuristic MIS.

Gates: FP32 GPU radiance within relative MSE 1e-4 of the FP64 oracle and 1e-5
of the FP32 oracle; one count per path (the light samples' shadow rays add no
count); a grey floor under a constant sky renders the floor's albedo times the
sky (the oracle pins that value by quadrature).
"""
import numpy as np
import pytest

import oracle
from cudapt_amd import scenes
from tests.helpers import rel_mse
from tests.test_oracle_disney import _disney_cornell

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def disney_runs(cpt):
    print("Disney mode: Cornell 96x96, 16 spp, depth 6", flush=True)
    sc = _disney_cornell()
    W, H, F = 96, 96, 16
    S = cpt.Scene(sc)
    R = cpt.Renderer(W, H, mode=cpt.CPT_MODE_PT_DISNEY_MIS, eye_depth=6)
    gpu = R.render(S, sc["camera"], frames=F)
    out = dict(gpu=gpu, N=R.sample_counts(), stats=R.stats(), F=F)
    O = oracle.Scene(sc)
    p = oracle.render_params(W, H, sc["camera"], mode=7, eye_depth=6)
    o32 = oracle.Renderer(p, 32)
    out["r32"], _ = o32.render(O, 0, F)
    o64 = oracle.Renderer(p, 64)
    o64.render(O, 0, F)
    out["r64"] = o64.result64()
    return out


def test_disney_against_fp64_oracle(disney_runs):
    r = disney_runs
    e64, e32 = rel_mse(r["gpu"], r["r64"]), rel_mse(r["gpu"], r["r32"])
    print(f"Disney rel MSE vs FP64 {e64:.3e}, vs FP32 {e32:.3e}", flush=True)
    assert e64 <= 1e-4 and e32 <= 1e-5
    assert np.isfinite(r["gpu"]).all() and r["gpu"].mean() > 0.01


def test_disney_one_count_per_path(disney_runs):
    r = disney_runs
    assert (r["N"] == r["F"]).all()
    assert r["stats"]["shadow_rays"] > 0


def test_disney_frames_in_flight_bit_identical(cpt):
    sc = _disney_cornell()
    S = cpt.Scene(sc)
    imgs = []
    for k in (1, 5):
        R = cpt.Renderer(40, 32, mode=cpt.CPT_MODE_PT_DISNEY_MIS, eye_depth=6)
        R.set_attrib("Frames In Flight", k)
        imgs.append(R.render(S, sc["camera"], frames=7))
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))


def test_disney_floor_under_constant_sky(cpt):
    L = np.array([1.0, 1.0, 1.0], np.float32)
    b = scenes._Builder()
    mat = scenes.material(diffuse=(0.6, 0.6, 0.6), metallic=0.2, specular=0.5, roughness=0.6, sheen=0.3,
                          clearcoat=0.5)
    mat["clearcoatGloss"] = 0.5
    k = b.add_material(mat)
    b.add_quad((-1e3, 0, 1e3), (1e3, 0, 1e3), (1e3, 0, -1e3), (-1e3, 0, -1e3), (0, 1, 0), k)
    cam = dict(pos=(0.0, 2.0, 0.0), dir=(0.0, -1.0, 0.0), up=(0.0, 0.0, -1.0), fov=0.02)
    sc = b.build(cam)
    sc["env"] = np.broadcast_to(L, (8, 16, 3)).copy()
    img = cpt.Renderer(64, 64, mode=cpt.CPT_MODE_PT_DISNEY_MIS, eye_depth=3).render(cpt.Scene(sc), cam, frames=32)
    o = oracle.Renderer(oracle.render_params(64, 64, cam, mode=7, eye_depth=3), 64)
    o.render(oracle.Scene(sc), 0, 32)
    np.testing.assert_allclose(img.mean(), o.result64().mean(), rtol=2e-3)

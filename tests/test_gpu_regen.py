"""The "Path Regen" BDPT renderer (CPT_MODE_BDPT_REGEN; the reference's
megakernel estimator, src/cudaRTBDPTRegen.cu) on the GPU against the oracle's
restatement of it (mode 8).

Gates, with the same counter-based RNG streams:
  - FP32 GPU radiance within relative MSE 1e-4 of the FP64 host run, on the
    Cornell box, the textured glass box and a small sphere grid, or within
    1.25 x the FP32 host run's own distance from FP64 where that is larger
    (glass: a refraction branch taken in one precision and not the other
    moves a whole camera path, and the FP32 restatement is itself 1.4e-4
    from FP64 there);
  - against the FP32 host run (the device operation sequence, transcendentals
    included) identical: relative MSE 0 and every sample count equal;
  - an emitter filling the view gives L E / (L + 1) and 3 (L + 1) samples per
    pixel over three frames, for L levels (the closed form the oracle is
    pinned to);
  - frames split over calls equal one call; a camera change restarts the mean
    and the light cache; more than 8 levels is an error.
"""
import numpy as np
import pytest

import oracle
from cudapt_amd import scenes
from tests.helpers import rel_mse

pytestmark = pytest.mark.gpu

ORACLE_MODE = 8

SCENES = {
    "cornell": (lambda: scenes.cornell(), 96, 96, 9, 5),
    "textured_glass": (lambda: scenes.textured_cornell(glass=True), 80, 64, 6, 5),
    "sphere_grid": (lambda: scenes.sphere_grid(grid=12), 96, 64, 6, 8),
}


def _renderer(cpt, W, H, depth):
    R = cpt.Renderer(W, H, mode=cpt.CPT_MODE_BDPT_REGEN)
    R.set_attrib("Eye Depth", depth)
    return R


@pytest.fixture(scope="module", params=list(SCENES))
def runs(cpt, request):
    make, W, H, frames, depth = SCENES[request.param]
    print(f"path regen vs oracle: {request.param} {W}x{H}, {frames} frames, {depth} levels", flush=True)
    sc = make()
    S = cpt.Scene(sc)
    R = _renderer(cpt, W, H, depth)
    gpu = R.render(S, sc["camera"], frames=frames)
    gpu_n = R.sample_counts()
    O = oracle.Scene(sc)
    p = oracle.render_params(W, H, sc["camera"], mode=ORACLE_MODE, eye_depth=depth)
    o32 = oracle.Renderer(p, 32)
    r32, n32 = o32.render(O, 0, frames)
    o64 = oracle.Renderer(p, 64)
    o64.render(O, 0, frames)
    return dict(name=request.param, sc=sc, S=S, W=W, H=H, frames=frames, depth=depth, gpu=gpu, gpu_n=gpu_n,
                stats=R.stats(), r32=r32, n32=n32, r64=o64.result64())


def test_fp32_gpu_vs_fp64_oracle(runs):
    e = rel_mse(runs["gpu"], runs["r64"])
    own = rel_mse(runs["r32"], runs["r64"])
    print(f"path regen {runs['name']}: rel MSE vs FP64 oracle {e:.3e} (FP32 oracle vs FP64: {own:.3e})")
    bound = max(1e-4, 1.25 * own)
    assert e <= bound, f"{runs['name']}: relative MSE {e:.3e} > {bound:.3e}"


def test_gpu_vs_fp32_oracle_and_counts(runs):
    e = rel_mse(runs["gpu"], runs["r32"])
    same = (runs["gpu_n"] == runs["n32"]).mean()
    print(f"path regen {runs['name']}: rel MSE vs FP32 oracle {e:.3e}, counts equal on {same:.4%} of pixels")
    assert e <= 1e-5, f"{runs['name']}: relative MSE {e:.3e}"
    assert same > 0.99
    assert np.array_equal(runs["gpu"].view(np.uint32), runs["r32"].view(np.uint32))
    assert np.array_equal(runs["gpu_n"], runs["n32"])


def test_image_lit_and_counted(runs):
    g = runs["gpu"]
    assert np.isfinite(g).all() and g.mean() > 0.01
    assert (runs["gpu_n"] >= runs["frames"]).all()       # at least one camera path per pixel and frame
    st = runs["stats"]
    assert st["frames"] == runs["frames"]
    assert st["extension_rays"] >= runs["W"] * runs["H"] * runs["frames"]
    assert st["shadow_rays"] > 0


@pytest.mark.parametrize("levels", [5, 3, 1, 8])
def test_emitter_view_closed_form(cpt, levels):
    b = scenes._Builder()
    m = b.add_material(scenes.material(diffuse=(0, 0, 0), emissive=(2.5, 1.5, 0.5)))
    b.add_quad((-50, -50, -1), (50, -50, -1), (50, 50, -1), (-50, 50, -1), (0, 0, 1), m)
    cam = dict(pos=(0.0, 0.0, 0.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov=0.5)
    S = cpt.Scene(b.build(cam))
    R = _renderer(cpt, 32, 24, levels)
    img = R.render(S, cam, frames=3)
    assert (R.sample_counts() == 3 * (levels + 1)).all()
    E = np.array([2.5, 1.5, 0.5], np.float32)
    np.testing.assert_allclose(img, np.broadcast_to(E * levels / (levels + 1), img.shape), rtol=1e-6)


def test_frames_split_over_calls_equal_one_call(cpt):
    sc = scenes.cornell()
    S = cpt.Scene(sc)
    a = _renderer(cpt, 48, 40, 5)
    a.render(S, sc["camera"], frames=4)
    ra = a.render(S, sc["camera"], frames=3)
    b = _renderer(cpt, 48, 40, 5)
    rb = b.render(S, sc["camera"], frames=7)
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    assert np.array_equal(a.sample_counts(), b.sample_counts())


def test_camera_change_restarts(cpt):
    sc = scenes.cornell()
    S = cpt.Scene(sc)
    cam2 = dict(sc["camera"], pos=(0.0, 1.0, 3.3))
    a = _renderer(cpt, 48, 40, 5)
    a.render(S, sc["camera"], frames=2)       # frame 1 ends, the next call starts again at frame 0
    ra = a.render(S, cam2, frames=2)
    b = _renderer(cpt, 48, 40, 5)
    rb = b.render(S, cam2, frames=2)
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))


def test_more_than_eight_levels_is_an_error(cpt):
    sc = scenes.cornell()
    S = cpt.Scene(sc)
    R = _renderer(cpt, 16, 16, 9)
    with pytest.raises(cpt.CptError):
        R.render(S, sc["camera"], frames=1)

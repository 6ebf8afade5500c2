"""Streaming BDPT on the GPU against the oracle's reference estimator.

Gate: with the same counter-based RNG stream, the FP32 GPU radiance lies within
relative MSE 1e-4 of the FP64 host run at equal spp (Cornell box, 256x256,
16 spp, depth 5).  The FP32 host run uses the device operation sequence,
transcendentals included (the shared trigonometry and Schlick power), so the
FP32 GPU image and sample counts equal it exactly.
"""
import numpy as np
import pytest

import oracle
from cudapt_amd import scenes
from tests.helpers import rel_mse

pytestmark = pytest.mark.gpu

W, H, SPP = 256, 256, 16


@pytest.fixture(scope="module")
def cornell_runs(cpt):
    sc = scenes.cornell()
    S = cpt.Scene(sc)
    R = cpt.Renderer(W, H)
    gpu = R.render(S, sc["camera"], frames=SPP)
    gpu_n = R.sample_counts()
    O = oracle.Scene(sc)
    p = oracle.render_params(W, H, sc["camera"])
    o32 = oracle.Renderer(p, 32)
    r32, n32 = o32.render(O, 0, SPP)
    o64 = oracle.Renderer(p, 64)
    o64.render(O, 0, SPP)
    r64 = o64.result64()
    return dict(sc=sc, S=S, gpu=gpu, gpu_n=gpu_n, r32=r32, n32=n32, r64=r64)


def test_fp32_gpu_vs_fp64_oracle_rel_mse(cornell_runs):
    r = cornell_runs
    e = rel_mse(r["gpu"], r["r64"])
    assert e <= 1e-4, f"relative MSE {e:.3e} > 1e-4"


def test_gpu_vs_fp32_oracle_tight(cornell_runs):
    r = cornell_runs
    e = rel_mse(r["gpu"], r["r32"])
    print(f"plain BDPT Cornell: rel MSE vs FP32 oracle {e:.3e}, counts equal on "
          f"{(r['gpu_n'] == r['n32']).mean():.4%} of pixels")
    assert e <= 1e-6, f"relative MSE {e:.3e}"
    assert (r["gpu_n"] == r["n32"]).mean() > 0.995
    assert np.array_equal(r["gpu"].view(np.uint32), r["r32"].view(np.uint32))
    assert np.array_equal(r["gpu_n"], r["n32"])


def test_image_is_lit(cornell_runs):
    g = cornell_runs["gpu"]
    assert np.isfinite(g).all()
    assert g.mean() > 0.05
    assert (cornell_runs["gpu_n"] >= SPP).all()


def test_batched_frames_equal_sequential(cpt, cornell_runs):
    sc, S = cornell_runs["sc"], cornell_runs["S"]
    R = cpt.Renderer(64, 48)
    for _ in range(7):
        seq = R.render(S, sc["camera"], frames=1)
    R2 = cpt.Renderer(64, 48)
    bat = R2.render(S, sc["camera"], frames=7)
    assert np.array_equal(seq.view(np.uint32), bat.view(np.uint32))
    assert np.array_equal(R.sample_counts(), R2.sample_counts())


@pytest.mark.parametrize("in_flight", [1, 3, 6])
def test_frames_in_flight_bit_identical(cpt, cornell_runs, in_flight):
    """Frames traced as one wavefront equal frames traced one at a time."""
    sc, S = cornell_runs["sc"], cornell_runs["S"]
    ref = cpt.Renderer(64, 48)
    ref.set_attrib("Frames In Flight", 1)
    for _ in range(8):
        a = ref.render(S, sc["camera"], frames=1)
    R = cpt.Renderer(64, 48)
    R.set_attrib("Frames In Flight", in_flight)
    b = R.render(S, sc["camera"], frames=5)
    b = R.render(S, sc["camera"], frames=3)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(ref.sample_counts(), R.sample_counts())


def test_camera_change_restarts_accumulation(cpt, cornell_runs):
    sc, S = cornell_runs["sc"], cornell_runs["S"]
    cam2 = dict(sc["camera"], pos=(0.0, 1.0, 3.3))
    R = cpt.Renderer(64, 48)
    R.render(S, sc["camera"], frames=3)
    a = R.render(S, cam2, frames=1)
    R2 = cpt.Renderer(64, 48)
    b = R2.render(S, cam2, frames=1)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_emitter_only_scene(cpt):
    """Every primary ray hits the light: the image is exactly the radiance."""
    b = scenes._Builder()
    m = b.add_material(scenes.material(diffuse=(0, 0, 0), emissive=(2.5, 1.5, 0.5)))
    b.add_quad((-50, -50, -1), (50, -50, -1), (50, 50, -1), (-50, 50, -1), (0, 0, 1), m)
    cam = dict(pos=(0.0, 0.0, 0.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov=0.5)
    sc = b.build(cam)
    S = cpt.Scene(sc)
    R = cpt.Renderer(32, 32)
    img = R.render(S, cam, frames=4)
    assert np.allclose(img, np.array([2.5, 1.5, 0.5], np.float32), rtol=0, atol=1e-6)


def test_black_scene(cpt):
    sc = scenes.cornell(light_radiance=0.0)
    S = cpt.Scene(sc)
    R = cpt.Renderer(32, 32)
    img = R.render(S, sc["camera"], frames=3)
    assert (img == 0).all()


def test_rmse_harness_on_device(cpt, cornell_runs):
    """RMSE against the 16-spp FP64 image falls as frames accumulate."""
    from cudapt_amd import rmse
    sc, S = cornell_runs["sc"], cornell_runs["S"]
    R = cpt.Renderer(W, H)
    h = rmse.RmseHarness(R, S, sc["camera"], cornell_runs["r64"].astype(np.float32), frames_per_step=2)
    rows = h.collect(interval=1e-9, count=8)
    assert [r.frames for r in rows] == [2, 4, 6, 8, 10, 12, 14, 16]
    # the same streams as the 16-frame FP64 run: the last row differs by precision only
    assert rows[-1].rmse < 1e-2 * float(np.sqrt((cornell_runs["r64"] ** 2).mean()))
    assert rows[0].rmse > 10 * rows[-1].rmse


@pytest.mark.parametrize("variant", [806, 807, 1606, 16])
def test_quantized_node_format_same_image(cpt, cornell_runs, variant):
    """Traversal over 64-byte quantized nodes returns the same hits, so the
    rendered frames are bit-identical to the 128-byte layout's."""
    sc, S = cornell_runs["sc"], cornell_runs["S"]
    a = cpt.Renderer(64, 48)
    a.set_attrib("Node Format", 0)
    ra = a.render(S, sc["camera"], frames=6)
    b = cpt.Renderer(64, 48)
    b.set_attrib("Node Format", 1)
    b.set_attrib("Traversal LDS Stack", variant)
    b.set_attrib("Shadow LDS Stack", variant)
    rb = b.render(S, sc["camera"], frames=6)
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))


@pytest.mark.parametrize("variant,node_format,refill", [(806, 0, 8), (807, 1, 1), (1606, 0, 64), (16, 1, 16),
                                                         (808, 1, 64), (24, 0, 32)])
def test_refill_traversal_same_image(cpt, cornell_runs, variant, node_format, refill):
    """Lanes that refill themselves from the ray queue trace the same rays with
    the same per-node operations: frames bit-identical to one ray per thread.
    With quantized nodes the refilling side's first bounce also computes each
    slot's eye ray and path state in the extension and shading kernels, where
    the other side reads what k_gen_eye stored."""
    sc, S = cornell_runs["sc"], cornell_runs["S"]
    a, b = cpt.Renderer(64, 48), cpt.Renderer(64, 48)
    for R in (a, b):
        R.set_attrib("Node Format", node_format)
        R.set_attrib("Traversal LDS Stack", variant)
        R.set_attrib("Shadow LDS Stack", variant)
    a.set_attrib("Traversal Refill", 0)
    a.set_attrib("Secondary Refill", 0)
    a.set_attrib("Shadow Refill", 0)
    b.set_attrib("Traversal Refill", refill)
    b.set_attrib("Secondary Refill", max(1, refill // 2))
    b.set_attrib("Shadow Refill", 65 - refill)
    b.set_attrib("Refill Min Paths", 0)
    ra = a.render(S, sc["camera"], frames=6)
    rb = b.render(S, sc["camera"], frames=6)
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    assert np.array_equal(a.sample_counts(), b.sample_counts())
    sa, sb = a.stats(), b.stats()
    assert sa["extension_rays"] == sb["extension_rays"] and sa["extend_nodes"] == sb["extend_nodes"]
    assert sa["shadow_nodes"] == sb["shadow_nodes"] > 0 and sa["shadow_tris"] == sb["shadow_tris"]


@pytest.mark.parametrize("mode", [0, 4, 5])
def test_computed_eye_rays_same_image(cpt, cornell_runs, mode):
    """In every stream mode (BDPT, path tracing, path tracing with restarts),
    a first bounce on refilling waves, which computes each slot's eye ray and
    path state from the slot index, gives the frames of a first bounce at one
    ray per thread, which reads what k_gen_eye stored.  Three frames per
    wavefront over six frames: the slot's frame offset and the batch's first
    frame both enter the eye path.  Both sides call the same eye_slot; the
    oracle comparisons above pin it through k_gen_eye."""
    sc, S = cornell_runs["sc"], cornell_runs["S"]
    a, b = cpt.Renderer(64, 48, mode=mode), cpt.Renderer(64, 48, mode=mode)
    for R in (a, b):
        R.set_attrib("Node Format", 1)
        R.set_attrib("Refill Min Paths", 0)
        R.set_attrib("Frames In Flight", 3)
    a.set_attrib("Traversal Refill", 0)
    b.set_attrib("Traversal Refill", 64)
    ra = a.render(S, sc["camera"], frames=6)
    rb = b.render(S, sc["camera"], frames=6)
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    assert np.array_equal(a.sample_counts(), b.sample_counts())
    sa, sb = a.stats(), b.stats()
    assert sa["extension_rays"] == sb["extension_rays"] > 64 * 48 * 6
    assert sa["extend_nodes"] == sb["extend_nodes"]
    assert ra.any()



@pytest.mark.parametrize("attr", ["Traversal LDS Stack", "Shadow LDS Stack"])
def test_unknown_traversal_variant_is_an_error(cpt, cornell_runs, attr):
    sc, S = cornell_runs["sc"], cornell_runs["S"]
    R = cpt.Renderer(16, 12)
    R.set_attrib(attr, 805)
    with pytest.raises(cpt.CptError, match="traversal variant"):
        R.render(S, sc["camera"], frames=1)


def test_bounce_ray_counts_sum_to_the_ray_counters(cpt):
    """CPT_BUF_BOUNCE_RAYS of a one-wavefront render: extension rays per
    bounce start at every path, never grow, and sum to the extension-ray
    counter; shadow rays per bounce sum to the shadow-ray counter."""
    sc = scenes.cornell()
    S = cpt.Scene(sc)
    R = cpt.Renderer(64, 48)
    R.set_attrib("Frames In Flight", 4)
    R.render(S, sc["camera"], frames=4, result=False)
    b = R.read(cpt.CPT_BUF_BOUNCE_RAYS)
    st = R.stats()
    assert b[0][0] == 64 * 48 * 4
    assert all(b[0][i + 1] <= b[0][i] for i in range(15))
    assert int(b[0].sum()) == st["extension_rays"]
    assert int(b[1].sum()) == st["shadow_rays"] > 0

"""The megakernel schedule of the streaming estimator (one thread carries its
sample through every bounce, as the reference's Path Regen renderer,
src/cudaRTBDPTRegen.cu, schedules its own estimator) against the wavefront
passes.

Both run the same per-path functions in the same order, so the gate is
equality: the same frames bit for bit, the same sample counts, and the same
ray and traversal counters."""
import numpy as np
import pytest

from cudapt_amd import scenes

pytestmark = pytest.mark.gpu

SCENES = {
    "cornell": lambda: scenes.cornell(),
    "textured_glass": lambda: scenes.textured_cornell(glass=True),
    "sphere_grid": lambda: scenes.sphere_grid(grid=12),
}


@pytest.fixture(scope="module")
def scene_objs(cpt):
    out = {}
    for k, f in SCENES.items():
        sc = f()
        out[k] = (sc, cpt.Scene(sc))
    return out


def _pair(cpt, W, H, **attribs):
    a, b = cpt.Renderer(W, H), cpt.Renderer(W, H)
    for R in (a, b):
        for k, v in attribs.items():
            R.set_attrib(k, v)
    b.set_attrib("megakernel", 1)
    return a, b


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("variant,node_format", [(806, 1), (16, 0)])
def test_megakernel_same_image(cpt, scene_objs, name, variant, node_format):
    sc, S = scene_objs[name]
    a, b = _pair(cpt, 96, 64, **{"Node Format": node_format, "Traversal LDS Stack": variant,
                                 "Shadow LDS Stack": variant, "in_flight": 3, "eye_depth": 8})
    ra = a.render(S, sc["camera"], frames=7)
    rb = b.render(S, sc["camera"], frames=7)
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    assert np.array_equal(a.sample_counts(), b.sample_counts())
    sa, sb = a.stats(), b.stats()
    for k in ("extension_rays", "shadow_rays", "eye_paths", "frames", "extend_nodes", "extend_tris",
              "shadow_nodes", "shadow_tris"):
        assert sa[k] == sb[k], k
    assert sb["shadow_rays"] > 0 and sb["extension_rays"] > sb["eye_paths"]


def test_megakernel_large_batch_same_image(cpt, scene_objs):
    """A batch above the refill threshold: the wavefront side runs the
    refilling persistent traversal, the megakernel one thread per sample."""
    sc, S = scene_objs["sphere_grid"]
    a, b = _pair(cpt, 320, 240, **{"in_flight": 16})
    ra = a.render(S, sc["camera"], frames=16)
    rb = b.render(S, sc["camera"], frames=16)
    assert 320 * 240 * 16 >= a.attribs()["Refill Min Paths"][2]
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    assert a.stats()["extension_rays"] == b.stats()["extension_rays"]


def test_megakernel_not_an_adaptive_attribute(cpt):
    R = cpt.Renderer(16, 12, mode=cpt.CPT_MODE_BDPT_ADAPTIVE_PLL)
    with pytest.raises(KeyError):
        R.set_attrib("megakernel", 1)


def test_megakernel_unknown_variant_is_an_error(cpt, scene_objs):
    sc, S = scene_objs["cornell"]
    R = cpt.Renderer(16, 12)
    R.set_attrib("megakernel", 1)
    R.set_attrib("Traversal LDS Stack", 805)
    with pytest.raises(cpt.CptError):
        R.render(S, sc["camera"], frames=1)


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("tail", [1, 4, 7])
def test_tail_bounces_in_one_launch_same_image(cpt, scene_objs, name, tail):
    """"Tail Bounce": bounces tail.. of every path run in one k_tail launch
    (one thread carries each live slot through the remaining bounces).  The
    frames, counts and ray counters equal the wavefront's; the tail's own
    node counters take the rest of the traversal work."""
    sc, S = scene_objs[name]
    a, b = _pair(cpt, 96, 64, **{"in_flight": 5, "eye_depth": 8})
    b.set_attrib("megakernel", 0)
    b.set_attrib("Tail Bounce", tail)
    ra = a.render(S, sc["camera"], frames=10)
    rb = b.render(S, sc["camera"], frames=10)
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    assert np.array_equal(a.sample_counts(), b.sample_counts())
    sa, sb = a.stats(), b.stats()
    for k in ("extension_rays", "shadow_rays", "eye_paths", "frames"):
        assert sa[k] == sb[k], k
    assert sb["tail_rays"] > 0
    assert sb["extend_nodes"] + sb["shadow_nodes"] + sb["tail_nodes"] == sa["extend_nodes"] + sa["shadow_nodes"]
    assert sb["extend_tris"] + sb["shadow_tris"] + sb["tail_tris"] == sa["extend_tris"] + sa["shadow_tris"]


def test_tail_bounce_on_a_refilling_batch(cpt, scene_objs):
    """Above the refill threshold (refilling waves, computed eye rays, the side
    stream) and in a multi-view batch (graph replays) the tail keeps the image."""
    sc, S = scene_objs["sphere_grid"]
    a, b = _pair(cpt, 320, 240, **{"in_flight": 16})
    b.set_attrib("megakernel", 0)
    b.set_attrib("Tail Bounce", 3)
    ra = a.render(S, sc["camera"], frames=16)
    rb = b.render(S, sc["camera"], frames=16)
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    views = [dict(width=64, height=48, **sc["camera"]), dict(width=40, height=30, **sc["camera"])]
    B1, B2 = cpt.Batch(views, frames_in_flight=4), cpt.Batch(views, frames_in_flight=4)
    B2.set_attrib("Tail Bounce", 2)
    B1.render(S, 8)
    B2.render(S, 8)
    for v in range(2):
        assert np.array_equal(B1.result(v).view(np.uint32), B2.result(v).view(np.uint32))


@pytest.mark.parametrize("mode", ["pll", "binary", "emp"])
@pytest.mark.parametrize("tail", [1, 3])
def test_tail_bounce_in_adaptive_frames(cpt, scene_objs, mode, tail):
    """The adaptive frames after the warm-up (sample records, first-hit
    G-buffer at bounce 0) with their last bounces in one k_tail launch: image,
    sample counts, per-pixel paths and error map bit-identical."""
    m = {"pll": cpt.CPT_MODE_BDPT_ADAPTIVE_PLL, "binary": cpt.CPT_MODE_BDPT_ADAPTIVE_BINARY,
         "emp": cpt.CPT_MODE_BDPT_ADAPTIVE_EMP}[mode]
    sc, S = scene_objs["textured_glass"]
    out = []
    for t in (0, tail):
        R = cpt.Renderer(64, 48, mode=m, eye_depth=8)
        R.set_attrib("Tail Bounce", t)
        R.render(S, sc["camera"], frames=9, result=False)
        out.append((R.read(cpt.CPT_BUF_RESULT), R.sample_counts(), R.read(cpt.CPT_BUF_PIXEL_PATHS),
                    R.read(cpt.CPT_BUF_ERROR), R.stats()))
    (ia, na, pa, ea, sa), (ib, nb, pb, eb, sb) = out
    assert np.array_equal(ia.view(np.uint32), ib.view(np.uint32))
    assert np.array_equal(na, nb) and np.array_equal(pa, pb)
    assert np.array_equal(ea.view(np.uint32), eb.view(np.uint32))
    assert sa["extension_rays"] == sb["extension_rays"] and sb["tail_rays"] > 0


def test_tail_bounce_in_adaptive_environment_frames(cpt):
    """The adaptive workload's scene (heightfield under its environment map) at
    a small size: the tail keeps every adaptive frame bit for bit."""
    from cudapt_amd import scenes
    sc = scenes.heightfield(light="env")
    S = cpt.Scene(sc)
    cam = scenes.heightfield_poses(1, resolutions=((96, 64),))[0]
    imgs = []
    for t in (0, 2):
        R = cpt.Renderer(96, 64, mode=cpt.CPT_MODE_BDPT_ADAPTIVE_PLL, eye_depth=8, **scenes.ADAPTIVE_C3)
        R.set_attrib("Tail Bounce", t)
        imgs.append(R.render(S, cam, frames=10))
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))

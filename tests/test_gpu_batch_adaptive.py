"""Per-view adaptive sampling over a multi-view batch (cpt_batch_set_adaptive).

Every view runs the adaptive renderer's frames over its own pixels (warm-up
frames, its own cross-bilateral filter, error sum or learned error estimate,
and a proportional allocation of pass_frames x its pixel count paths) and all
views' paths are traced as one wavefront.  Each view's running mean, sample
counts, per-pixel path counts and error map must equal those of a one-view
adaptive renderer of that camera with the same attributes, seed and model,
bit for bit, with and without the model and across pass sizes and calls.
"""
import math

import numpy as np
import pytest

from cudapt_amd import scenes

pytestmark = pytest.mark.gpu

ATTRS = {"Filter Radius": 7, "Filter Color EuD": 0.2, "Filter Norm EuD": 0.1, "Filter Depth EuD": 0.05,
         "Filter Spatial EuD": 3.5, "Filter Pos EuD": 1e4, "Min Pixel Paths": 4, "Max Pixel Paths": 24}


def _views(sc):
    c = sc["camera"]
    v0 = dict(width=64, height=48, **c)
    v1 = dict(width=40, height=30, pos=(0.4, 1.2, 3.0), dir=(-0.1, -0.05, -1.0), up=(0.0, 1.0, 0.0), fov=0.8)
    d = np.array([0.3, -0.2, -1.0]); d /= np.linalg.norm(d)
    v2 = dict(width=23, height=17, pos=(-0.3, 0.8, 3.2), dir=tuple(d), up=(0.0, 1.0, 0.0), fov=math.radians(50))
    return [v0, v1, v2]


def _one_view(cpt, S, v, frames, pass_frames, model, signal, calls):
    R = cpt.Renderer(v["width"], v["height"], mode=cpt.CPT_MODE_BDPT_ADAPTIVE_PLL, **ATTRS)
    R.set_attrib("Adaptive Pass Frames", pass_frames)
    R.set_attrib("Learned Signal", signal)
    if model is not None:
        R.set_mlp(model)
    for f in calls:
        R.render(S, v, frames=f, result=False)
    out = dict(img=R.read(cpt.CPT_BUF_RESULT), n=R.sample_counts(), paths=R.read(cpt.CPT_BUF_PIXEL_PATHS),
               err=R.read(cpt.CPT_BUF_ERROR))
    return out


@pytest.fixture(scope="module")
def scene(cpt):
    sc = scenes.cornell()
    return dict(S=cpt.Scene(sc), views=_views(sc))


@pytest.mark.parametrize("pass_frames,learned,calls", [(1, False, (9,)), (2, False, (6, 7)), (1, True, (8, 3)),
                                                        (3, True, (12,))],
                         ids=["pass1", "pass2-two-calls", "learned-pass1", "learned-pass3"])
def test_views_equal_one_view_adaptive_renderers(cpt, scene, pass_frames, learned, calls):
    print(f"batch adaptive: pass {pass_frames}, learned {learned}, calls {calls}", flush=True)
    S, views = scene["S"], scene["views"]
    model = cpt.Mlp(0xBA7C) if learned else None
    signal = 3 if learned else 0
    B = cpt.Batch(views, frames_in_flight=4)
    B.set_adaptive(cpt.CPT_MODE_BDPT_ADAPTIVE_PLL, pass_frames, model, **ATTRS)
    B.set_attrib("Learned Signal", signal)
    for f in calls:
        B.render(S, f)
    for v, view in enumerate(views):
        ref = _one_view(cpt, S, view, sum(calls), pass_frames, model, signal, calls)
        assert np.array_equal(B.result(v).view(np.uint32), ref["img"].view(np.uint32)), f"view {v}"
        assert np.array_equal(B.sample_counts(v), ref["n"]), f"view {v}"
        assert np.array_equal(B.read(v, cpt.CPT_BUF_PIXEL_PATHS), ref["paths"]), f"view {v}"
        assert np.array_equal(B.read(v, cpt.CPT_BUF_ERROR).view(np.uint32), ref["err"].view(np.uint32)), f"view {v}"
    paths = [B.read(v, cpt.CPT_BUF_PIXEL_PATHS) for v in range(len(views))]
    assert all(p.min() >= 4 and p.max() <= ATTRS["Max Pixel Paths"] for p in paths)
    st = B.stats()
    assert st["frames"] == sum(calls) and st["extension_rays"] > 0


def test_adaptive_batch_budget_per_view(cpt, scene):
    """Each view spends at most pass_frames x its pixel count paths per pass
    (its own budget; the proportional counts round down, AdapBilatAtomic.cu:830-831,
    so a pixel loses less than one path), whatever the other views' errors."""
    S, views = scene["S"], scene["views"]
    B = cpt.Batch(views)
    B.set_adaptive(cpt.CPT_MODE_BDPT_ADAPTIVE_PLL, 2, None, **{"Filter Radius": 5, "Filter Color EuD": 0.2})
    B.render(S, 5)                                   # warm-up: one path per pixel and frame
    before = [B.read(v, cpt.CPT_BUF_PIXEL_PATHS).sum() for v in range(len(views))]
    B.render(S, 2)                                   # one pass of two frames
    for v, (w, h) in enumerate(B.views):
        got = int(B.read(v, cpt.CPT_BUF_PIXEL_PATHS).sum() - before[v])
        counts = B.read(v, cpt.CPT_BUF_SAMPLE_COUNTS)
        assert got == int(counts.sum())
        assert 2 * w * h - (counts.size - 1) <= got <= 2 * w * h, (v, got, 2 * w * h)


def test_adaptive_batch_rejections(cpt, scene):
    S, views = scene["S"], scene["views"]
    B = cpt.Batch(views)
    with pytest.raises(cpt.CptError):
        B.set_adaptive(cpt.CPT_MODE_BDPT_ADAPTIVE_BINARY, 1)     # only the proportional mode
    with pytest.raises(cpt.CptError):
        B.set_adaptive(cpt.CPT_MODE_BDPT_STREAM, 1, cpt.Mlp(1))   # a model needs the adaptive mode
    B.set_adaptive(cpt.CPT_MODE_BDPT_ADAPTIVE_PLL, 1)
    with pytest.raises(cpt.CptError):
        B.set_attrib("Eye Depth", 3)                             # owned by the batch configuration
    with pytest.raises(cpt.CptError):
        B.set_attrib("No Such Attribute", 1)
    with pytest.raises(cpt.CptError):
        B.autotune(S, 1)
    B.render(S, 6)
    B.set_adaptive(cpt.CPT_MODE_BDPT_STREAM, 1)                  # back to plain frames, from frame 0
    B.render(S, 4)
    P = cpt.Batch(views)
    P.render(S, 4)
    for v in range(len(views)):
        assert np.array_equal(B.result(v).view(np.uint32), P.result(v).view(np.uint32)), f"view {v}"

"""Multi-view batch: views of uneven resolution traced as one wavefront.

Every view's samples are the ones a one-view renderer of that camera draws
(eye streams keyed by the view-local pixel, light caches by frame), so each
view's image must equal that renderer's bit for bit, whatever the batch's
frames in flight, persistent-queue extend variant or graph replay."""
import math

import numpy as np
import pytest

from cudapt_amd import scenes

pytestmark = pytest.mark.gpu


def _views(sc):
    c = sc["camera"]
    v0 = dict(width=64, height=48, **c)
    v1 = dict(width=40, height=30, pos=(0.4, 1.2, 3.0), dir=(-0.1, -0.05, -1.0), up=(0.0, 1.0, 0.0), fov=0.8)
    d = np.array([0.3, -0.2, -1.0]); d /= np.linalg.norm(d)
    v2 = dict(width=23, height=17, pos=(-0.3, 0.8, 3.2), dir=tuple(d), up=(0.0, 1.0, 0.0), fov=math.radians(50))
    return [v0, v1, v2]


@pytest.fixture(scope="module")
def setup(cpt):
    sc = scenes.cornell()
    S = cpt.Scene(sc)
    views = _views(sc)
    refs = []
    for v in views:
        R = cpt.Renderer(v["width"], v["height"])
        img = R.render(S, v, frames=11)
        refs.append((img, R.sample_counts()))
    return dict(S=S, views=views, refs=refs)


@pytest.mark.parametrize("cfg", [dict(frames_in_flight=4), dict(frames_in_flight=6, use_graph=0),
                                 dict(frames_in_flight=3, extend_chunk=256, extend_workgroups=4, refill=0),
                                 dict(frames_in_flight=11, extend_chunk=1024, extend_workgroups=8,
                                      traversal_variant=1606, refill=0),
                                 dict(frames_in_flight=5, refill=16, extend_workgroups=2),
                                 dict(frames_in_flight=4, refill=0, node_format=0, use_graph=0)],
                         ids=["graph-k4", "eager-k6", "persistent-k3", "persistent-k11", "refill16-k5",
                              "per-thread-fp32-k4"])
def test_views_equal_one_view_renderers(cpt, setup, cfg):
    B = cpt.Batch(setup["views"], **cfg)
    B.render(setup["S"], 7)
    B.render(setup["S"], 4)                      # 11 frames, uneven iteration split
    for v, (img, n) in enumerate(setup["refs"]):
        assert np.array_equal(B.result(v).view(np.uint32), img.view(np.uint32)), f"view {v}"
        assert np.array_equal(B.sample_counts(v), n)
    st = B.stats()
    total = sum(w * h for w, h in B.views)
    assert st["frames"] == 11 and st["eye_paths"] == 11 * total


def test_reset_and_config_restart_accumulation(cpt, setup):
    B = cpt.Batch(setup["views"], frames_in_flight=4)
    B.render(setup["S"], 5)
    B.reset()
    B.render(setup["S"], 11)
    img, _ = setup["refs"][1]
    assert np.array_equal(B.result(1).view(np.uint32), img.view(np.uint32))


def test_autotuner_picks_a_candidate_and_stays_exact(cpt, setup):
    B = cpt.Batch(setup["views"], frames_in_flight=3)
    cands = B.candidates()
    assert len(cands) == 20 and len({tuple(sorted(c.items())) for c in cands}) == 20
    best, ms = B.autotune(setup["S"], iterations=1)
    assert best in cands and all(m > 0 for m in ms)
    assert best == B.config()
    B.render(setup["S"], 11)
    for v, (img, _) in enumerate(setup["refs"]):
        assert np.array_equal(B.result(v).view(np.uint32), img.view(np.uint32))


def test_batch_argument_errors(cpt, setup):
    with pytest.raises(cpt.CptError):
        cpt.Batch([])
    with pytest.raises(cpt.CptError):
        cpt.Batch([dict(setup["views"][0], width=0)])
    B = cpt.Batch(setup["views"])
    with pytest.raises(cpt.CptError):
        B.set_config(frames_in_flight=0)
    with pytest.raises(cpt.CptError):
        B.set_config(refill=65)
    with pytest.raises(cpt.CptError):
        B.set_config(traversal_variant=805)
    with pytest.raises(cpt.CptError):
        B.result(0)                              # nothing rendered yet

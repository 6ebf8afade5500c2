"""Pins of the oracle's "Path Regen" BDPT (mode 8; the reference's megakernel
renderer, src/cudaRTBDPTRegen.cu).

Each test checks something the restatement does not compute itself:
  - an emitter filling the view: every level hits the emitter, whose weight 0
    fails Russian roulette and restarts a camera path (count + 1, :562-571);
    every level starts a camera path and banks its emission (:598-606), and
    the level bound adds 1 (:627-631): the pixel is exactly L E / (L + 1) for
    L levels.  A count without the restarts, a bank only at the first level or
    a missing bound count each give another value;
  - a black floor: weight 0 at every level, L restarts, no connection: image 0,
    count L + 1 per frame;
  - the sky: a miss at the first level counts 1 and adds nothing;
  - the light cache (bdpt_light_kernel, :80-330): 128 paths of 5 slots; a path
    starts on the emitter with irrad = emission and no direction; the next
    vertex carries (1, 1, 1), not the emission (:179, :307); a vertex after a
    diffuse vertex carries that vertex's weight times what it carried (:302); a
    miss leaves an empty vertex and the next slot starts a new path;
  - FP32 and FP64 restatements agree on a Cornell box.
"""
import numpy as np
import pytest

import oracle
from cudapt_amd import scenes
from tests.helpers import rel_mse

MODE = 8


def _emitter_view():
    b = scenes._Builder()
    m = b.add_material(scenes.material(diffuse=(0, 0, 0), emissive=(2.5, 1.5, 0.5)))
    b.add_quad((-50, -50, -1), (50, -50, -1), (50, 50, -1), (-50, 50, -1), (0, 0, 1), m)
    cam = dict(pos=(0.0, 0.0, 0.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov=0.5)
    return b.build(cam), cam


@pytest.mark.parametrize("levels", [5, 3, 1])
def test_emitter_view_closed_form(levels):
    sc, cam = _emitter_view()
    O = oracle.Scene(sc)
    E = np.array([2.5, 1.5, 0.5])
    for prec in (32, 64):
        R = oracle.Renderer(oracle.render_params(8, 8, cam, mode=MODE, eye_depth=levels), prec)
        img, N = R.render(O, 0, 3)
        assert (N == 3 * (levels + 1)).all()
        np.testing.assert_allclose(img, np.broadcast_to(E * levels / (levels + 1), img.shape), rtol=1e-6)


def test_black_floor_restarts_every_level():
    b = scenes._Builder()
    floor = b.add_material(scenes.material(diffuse=(0, 0, 0)))
    lmat = b.add_material(scenes.material(diffuse=(0, 0, 0), emissive=(4.0, 4.0, 4.0)))
    b.add_quad((-10, 0, 10), (10, 0, 10), (10, 0, -10), (-10, 0, -10), (0, 1, 0), floor)
    b.add_quad((-0.1, 3, -0.1), (0.1, 3, -0.1), (0.1, 3, 0.1), (-0.1, 3, 0.1), (0, -1, 0), lmat)
    cam = dict(pos=(0.0, 1.0, 0.0), dir=(0.0, -1.0, 0.0), up=(0.0, 0.0, -1.0), fov=0.5)
    sc = b.build(cam)
    R = oracle.Renderer(oracle.render_params(8, 8, cam, mode=MODE, eye_depth=5), 64)
    img, N = R.render(oracle.Scene(sc), 0, 2)
    assert (img == 0).all() and (N == 2 * 6).all()


def test_sky_counts_one_per_frame():
    b = scenes._Builder()
    m = b.add_material(scenes.material(diffuse=(0.5, 0.5, 0.5)))
    b.add_quad((-1, -1, 5), (1, -1, 5), (1, 1, 5), (-1, 1, 5), (0, 0, -1), m)     # behind the camera
    cam = dict(pos=(0.0, 0.0, 0.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov=0.5)
    sc = b.build(cam)
    R = oracle.Renderer(oracle.render_params(8, 8, cam, mode=MODE), 64)
    img, N = R.render(oracle.Scene(sc), 0, 4)
    assert (img == 0).all() and (N == 4).all()


def test_light_cache_structure():
    """An emitter facing down onto a grey Lambert floor, both grey Lambert
    (specular 0, so the Fresnel term and the reflection probability are 0 and
    every bounce is diffuse with weight = albedo)."""
    a, e = 0.4, 0.5
    b = scenes._Builder()
    floor = b.add_material(scenes.material(diffuse=(a, a, a)))
    lmat = b.add_material(scenes.material(diffuse=(e, e, e), emissive=(3.0, 2.0, 1.0)))
    b.add_quad((-50, 0, 50), (50, 0, 50), (50, 0, -50), (-50, 0, -50), (0, 1, 0), floor)
    b.add_quad((-0.5, 2, -0.5), (0.5, 2, -0.5), (0.5, 2, 0.5), (-0.5, 2, 0.5), (0, -1, 0), lmat)
    cam = dict(pos=(0.0, 1.0, 3.0), dir=(0.0, -0.3, -1.0), up=(0.0, 1.0, 0.0), fov=0.8)
    sc = b.build(cam)
    R = oracle.Renderer(oracle.render_params(4, 4, cam, mode=MODE), 64)
    R.render(oracle.Scene(sc), 0, 1)
    lv = R.light_vertices()
    assert lv.shape == (640, 21)
    pos, irrad, idir = lv[:, 0:3], lv[:, 6:9], lv[:, 9:12]
    on_light = np.abs(pos[:, 1] - 2.0) < 1e-5
    on_floor = (np.abs(pos[:, 1]) < 1e-5) & (np.abs(irrad).sum(1) > 0)
    empty = np.abs(lv).sum(1) == 0
    assert (on_light | on_floor | empty).all()
    start = on_light & (np.abs(idir).sum(1) == 0)           # a path starting on the emitter
    paths = np.arange(640).reshape(128, 5)
    assert start[paths[:, 0]].all()
    np.testing.assert_allclose(irrad[start], np.broadcast_to([3.0, 2.0, 1.0], irrad[start].shape))
    seen = {"start": 0, "floor": 0, "hit": 0, "empty": 0}
    for p in paths:
        for k in range(1, 5):
            prev, cur = p[k - 1], p[k]
            if start[prev]:
                # the emitter faces the floor: its ray lands there carrying (1, 1, 1)
                # (or, at a grazing angle, leaves the floor's edge)
                assert on_floor[cur] or empty[cur]
                if on_floor[cur]:
                    np.testing.assert_allclose(irrad[cur], [1.0, 1.0, 1.0], rtol=1e-6)
                seen["start"] += 1
            elif on_floor[prev]:
                # a diffuse bounce upward: the sky, or the emitter's face, carrying the
                # floor's weight times what the floor vertex had (vecMul(lightMulTerm, rayIrrad))
                assert empty[cur] or (on_light[cur] and not start[cur])
                if on_light[cur]:
                    np.testing.assert_allclose(irrad[cur], a * irrad[prev], rtol=1e-5)
                seen["floor"] += 1
            elif on_light[prev]:
                # the emitter's face bounces down onto the floor: its weight times what it had
                assert on_floor[cur] or empty[cur]
                if on_floor[cur]:
                    np.testing.assert_allclose(irrad[cur], e * irrad[prev], rtol=1e-5)
                    seen["hit"] += 1
            else:
                assert empty[prev] and start[cur]       # a new path after a miss
                seen["empty"] += 1
    assert seen["start"] > 100 and seen["floor"] > 50 and seen["empty"] > 20 and seen["hit"] > 5


def test_fp32_matches_fp64_cornell():
    sc = scenes.cornell()
    O = oracle.Scene(sc)
    p = oracle.render_params(48, 48, sc["camera"], mode=MODE)
    r32 = oracle.Renderer(p, 32)
    img32, n32 = r32.render(O, 0, 9)
    r64 = oracle.Renderer(p, 64)
    r64.render(O, 0, 9)
    assert rel_mse(img32, r64.result64()) < 1e-6
    assert img32.mean() > 0.05

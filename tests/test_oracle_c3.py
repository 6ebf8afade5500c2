"""Pins of the oracle's extensions for the adaptive heightfield workload (no
reference counterpart): the environment light, the filter's spatial sigma and
relative-depth term, and the per-pixel path bounds.

  - environment lookup: texel centres return their texel, a constant map is
    constant, u wraps, v clamps, bilinear halfway values;
  - environment estimator: a grey diffuse floor under a constant sky renders
    exactly albedo * sky (the bounce always survives at 3a >= 1 and always
    escapes), and the sky itself exactly sky;
  - depth term: a hand-worked three-pixel case, scale invariance, a huge
    sigma switching it off, the FP32 defined arithmetic against the literal
    FP64 product;
  - path bounds: every pixel ends with at least min and (after warm-up) at
    most max eye paths, and the cap moves samples to other pixels.
"""
import math

import numpy as np
import pytest

import oracle
from cudapt_amd import scenes


# ---------------------------------------------------------------- environment lookup
def _env_scene(env):
    b = scenes._Builder()
    m = b.add_material(scenes.material(diffuse=(0.5, 0.5, 0.5)))
    b.add_quad((-1, -5, 1), (1, -5, 1), (1, -5, -1), (-1, -5, -1), (0, 1, 0), m)
    sc = b.build(dict(pos=(0.0, 0.0, 0.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov=1.0))
    sc["env"] = env
    return sc


def test_env_texel_centres_and_constant():
    W, H = 16, 8
    rng = np.random.default_rng(2)
    env = rng.uniform(0, 5, (H, W, 3)).astype(np.float32)
    O = oracle.Scene(_env_scene(env))
    dirs = scenes.env_directions(W, H)
    for j in range(H):
        for i in range(W):
            np.testing.assert_allclose(O.env_radiance(dirs[j, i]), env[j, i], rtol=1e-9, atol=1e-9)
    O.set_environment(np.full((H, W, 3), 2.5, np.float32))
    for d in rng.normal(size=(50, 3)):
        np.testing.assert_allclose(O.env_radiance(d / np.linalg.norm(d)), 2.5, rtol=1e-12)


def test_env_bilinear_wrap_and_clamp():
    W, H = 8, 4
    env = np.zeros((H, W, 3), np.float32)
    env[:, :, 0] = np.arange(W)[None, :]            # ramp across u
    env[:, :, 1] = np.arange(H)[:, None]            # ramp down v
    O = oracle.Scene(_env_scene(env))

    def d_of(u, v):
        phi, th = (u - 0.5) * 2 * math.pi, v * math.pi
        return np.array([math.sin(th) * math.cos(phi), math.cos(th), math.sin(th) * math.sin(phi)])

    # halfway between texel centres 2 and 3 across, 1 and 2 down
    r = O.env_radiance(d_of(3.0 / W, 2.0 / H))
    np.testing.assert_allclose(r[:2], [2.5, 1.5], atol=1e-9)
    # u wraps: halfway between the last column (7) and the first (0)
    np.testing.assert_allclose(O.env_radiance(d_of(1.0 - 1e-12, 0.6))[0], 3.5, atol=1e-6)
    # v clamps: straight up and straight down see the first and last rows
    assert O.env_radiance(np.array([0.0, 1.0, 0.0]))[1] == pytest.approx(0.0, abs=1e-12)
    assert O.env_radiance(np.array([0.0, -1.0, 0.0]))[1] == pytest.approx(H - 1, abs=1e-12)
    # FP32 lookups agree with FP64 to float rounding
    rng = np.random.default_rng(4)
    for d in rng.normal(size=(40, 3)):
        d /= np.linalg.norm(d)
        np.testing.assert_allclose(O.env_radiance(d, prec=32), O.env_radiance(d), rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("prec", [32, 64])
def test_env_furnace_floor(prec):
    """Camera looking down at a large grey Lambertian floor (albedo 0.5) under a
    constant sky L: a floor sample survives Russian roulette surely
    (|T| |f| = 3 * 0.5 >= 1), its diffuse bounce leaves the plane and sees L,
    so every floor pixel is exactly 0.5 L; with no emissive triangle there is no
    light subpath and no connection."""
    L = np.array([1.5, 1.0, 0.5], np.float32)
    b = scenes._Builder()
    m = b.add_material(scenes.material(diffuse=(0.5, 0.5, 0.5)))
    b.add_quad((-100, 0, 100), (100, 0, 100), (100, 0, -100), (-100, 0, -100), (0, 1, 0), m)
    cam = dict(pos=(0.0, 2.0, 0.0), dir=(0.0, -1.0, 0.0), up=(0.0, 0.0, -1.0), fov=1.0)
    sc = b.build(cam)
    sc["env"] = np.broadcast_to(L, (8, 16, 3)).copy()
    O = oracle.Scene(sc)
    R = oracle.Renderer(oracle.render_params(16, 16, cam), prec)
    img, N = R.render(O, 0, 4)
    np.testing.assert_allclose(img, np.broadcast_to(0.5 * L, img.shape), rtol=1e-6 if prec == 32 else 1e-13)
    assert (N == 4).all()
    ext, shadow = R.ray_counts()
    assert ext == 16 * 16 * 4 * 2 and shadow == 0
    # looking up at the sky: the sky itself
    cam_up = dict(pos=(0.0, 2.0, 0.0), dir=(0.0, 1.0, 0.0), up=(0.0, 0.0, -1.0), fov=1.0)
    img, _ = oracle.Renderer(oracle.render_params(8, 8, cam_up), prec).render(O, 0, 2)
    np.testing.assert_allclose(img, np.broadcast_to(L, img.shape), rtol=1e-7)


def test_env_is_off_without_a_map():
    sc = _env_scene(None)
    O = oracle.Scene(sc)
    img, _ = oracle.Renderer(oracle.render_params(8, 8, sc["camera"]), 64).render(O, 0, 2)
    assert (img == 0).all()


# ---------------------------------------------------------------- depth term
def test_depth_term_hand_case():
    """Three pixels in a row, equal colours' distance and normals, the position
    term off (sigma huge): the weights are G(j) * exp(-dz^2 / 2 sz^2) with
    dz = (z_q - z_p) / z_p.  Depths 1, 2, 1 about the camera at the origin."""
    color = np.array([[[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]]], np.float32)
    pos = np.array([[[1.0, 0, 0], [0, 2.0, 0], [0, 0, 1.0]]], np.float32)
    nrm = np.zeros((1, 3, 3), np.float32)
    sc_, sz, ss = 1e6, 0.5, 1.0
    out = oracle.bilateral_ext(color, pos, nrm, None, sc_, 1e6, 1.0, 1.0, 1, ss=ss, sz=sz, cam=(0, 0, 0), prec=64)
    g1 = math.exp(-1.0 / (2 * ss * ss))
    w = math.exp(-(0.5 ** 2) / (2 * sz * sz))          # from the middle: dz = (1 - 2) / 2
    # rows clamp onto row 0 (height 1) and contribute equally; the centre
    # column's left tap of pixel 0 wraps to the high edge (the reference clamp)
    mid = np.array([g1 * w, 1.0, g1 * w]) / (2 * g1 * w + 1.0)
    np.testing.assert_allclose(out[0, 1], mid, rtol=1e-6)
    # from pixel 0 (z = 1): the middle pixel has dz = (2 - 1) / 1 = 1
    w0 = math.exp(-1.0 / (2 * sz * sz))
    # pixel 0's taps: j = -1 wraps to the last pixel (z = 1, dz = 0), j = 0 itself, j = +1 the middle
    p0 = np.array([1.0, g1 * w0, g1]) / (1.0 + g1 * w0 + g1)
    np.testing.assert_allclose(out[0, 0], p0, rtol=1e-6)


def test_depth_term_scale_invariant_and_off_when_huge():
    rng = np.random.default_rng(8)
    H, W = 14, 19
    col = rng.uniform(0, 1, (H, W, 3)).astype(np.float32)
    pos = rng.uniform(-3, 3, (H, W, 3)).astype(np.float32) + np.float32([0, 0, -10])
    nrm = rng.normal(size=(H, W, 3)).astype(np.float32)
    args = (col, pos, nrm, None, 0.5, 1e6, 0.3, 0.1, 3)
    base = oracle.bilateral_ext(*args, ss=2.0, prec=64)
    with_z = oracle.bilateral_ext(*args, ss=2.0, sz=0.05, cam=(0.0, 0.0, 0.0), prec=64)
    assert np.abs(with_z - base).max() > 1e-3
    scaled = oracle.bilateral_ext(col, pos * 4, nrm, None, 0.5, 1e6, 0.3, 0.1, 3, ss=2.0, sz=0.05,
                                  cam=(0.0, 0.0, 0.0), prec=64)
    np.testing.assert_allclose(scaled, with_z, rtol=1e-5, atol=1e-6)
    off = oracle.bilateral_ext(*args, ss=2.0, sz=1e7, cam=(0.0, 0.0, 0.0), prec=64)
    np.testing.assert_allclose(off, base, rtol=1e-6, atol=1e-7)
    f32 = oracle.bilateral_ext(*args, ss=2.0, sz=0.05, cam=(0.0, 0.0, 0.0), prec=32)
    np.testing.assert_allclose(f32, with_z, rtol=2e-5, atol=1e-6)
    # a spatial sigma of its own: 0 falls back to the colour sigma
    np.testing.assert_array_equal(oracle.bilateral_ext(*args, ss=0.0, prec=32),
                                  oracle.bilateral_defined(col, pos, nrm, None, 0.5, 1e6, 0.3, 0.1, 3))


# ---------------------------------------------------------------- per-pixel path bounds
@pytest.mark.parametrize("mode", [1, 2])
def test_path_bounds(mode):
    sc = scenes.cornell()
    O = oracle.Scene(sc)
    W, H, F = 32, 32, 14
    p = oracle.render_params(W, H, sc["camera"], mode=mode, min_paths=6, max_paths=12)
    R = oracle.Renderer(p, 32)
    R.render(O, 0, F)
    st = R.adaptive_state()
    paths = st["paths"]
    assert paths.min() >= 6                          # six warm-up frames instead of five
    assert paths.max() <= 12
    assert paths.sum() <= W * H * F
    # without the cap some pixel goes beyond it
    free = oracle.Renderer(oracle.render_params(W, H, sc["camera"], mode=mode, min_paths=6), 32)
    free.render(O, 0, F)
    assert free.adaptive_state()["paths"].max() > 12

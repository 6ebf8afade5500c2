"""The oracle's texture lookup (wrap addressing, bilinear filtering,
src/cudaRTCommon.cu:416-441) pinned by closed forms: exact texel centres,
periodicity, linearity in the texels, exact reproduction of linear ramps."""
import numpy as np
import pytest

import oracle
from cudapt_amd import scenes


@pytest.fixture(scope="module")
def scene():
    return oracle.Scene(scenes.cornell())


@pytest.mark.parametrize("bits", [32, 64])
def test_texel_centres_are_exact(scene, bits):
    rng = np.random.default_rng(1)
    t = rng.uniform(0, 4, (8, 16, 4)).astype(np.float32)
    i = scene.add_texture(t)
    for y in range(8):
        for x in range(16):
            got = scene.tex_sample(i, (x + 0.5) / 16, (y + 0.5) / 8, bits)
            assert np.array_equal(got, t[y, x].astype(np.float64)), (x, y)


@pytest.mark.parametrize("bits", [32, 64])
def test_wrap_is_periodic(scene, bits):
    rng = np.random.default_rng(2)
    i = scene.add_texture(rng.uniform(0, 1, (4, 8, 4)).astype(np.float32))
    for u, v in [(0.1875, 0.3125), (0.0, 0.0), (0.96875, 0.03125), (0.5, 0.75)]:
        ref = scene.tex_sample(i, u, v, bits)
        for k, m in [(1, 0), (-1, 2), (3, -3), (-5, -1)]:
            assert np.array_equal(scene.tex_sample(i, u + k, v + m, bits), ref)
    # the seam: u = 0 blends the last and the first column equally
    t = np.zeros((1, 4, 4), np.float32)
    t[0, 0] = 1.0
    t[0, 3] = 3.0
    j = scene.add_texture(t)
    assert np.allclose(scene.tex_sample(j, 0.0, 0.5, bits), 2.0)


def test_linear_in_texels(scene):
    rng = np.random.default_rng(3)
    a, b = rng.uniform(0, 1, (2, 5, 7, 4)).astype(np.float64)
    ia, ib, iab = scene.add_texture(a), scene.add_texture(b), scene.add_texture(0.25 * a + 2.0 * b)
    for u, v in rng.uniform(-3, 3, (50, 2)):
        lhs = scene.tex_sample(iab, u, v, 64)
        rhs = 0.25 * scene.tex_sample(ia, u, v, 64) + 2.0 * scene.tex_sample(ib, u, v, 64)
        np.testing.assert_allclose(lhs, rhs, rtol=1e-6, atol=1e-7)   # texels stored in FP32


def test_reproduces_linear_ramp_inside(scene):
    W, H = 16, 8
    jj, ii = np.mgrid[0:H, 0:W].astype(np.float64)
    t = np.zeros((H, W, 4))
    t[..., 0] = 1.0 + 0.5 * ii
    t[..., 1] = 2.0 - 0.25 * jj
    t[..., 2] = 0.125 * ii + 0.0625 * jj
    t[..., 3] = 1.0
    k = scene.add_texture(t.astype(np.float32))
    rng = np.random.default_rng(4)
    for _ in range(200):
        x, y = rng.uniform(0.5, W - 0.5), rng.uniform(0.5, H - 0.5)   # texel-space, away from the seam
        got = scene.tex_sample(k, x / W, y / H, 64)
        fx, fy = x - 0.5, y - 0.5
        np.testing.assert_allclose(got[:3], [1.0 + 0.5 * fx, 2.0 - 0.25 * fy, 0.125 * fx + 0.0625 * fy],
                                   rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("bits", [32, 64])
def test_non_finite_coordinates_sample_origin(scene, bits):
    rng = np.random.default_rng(5)
    i = scene.add_texture(rng.uniform(0, 1, (4, 4, 4)).astype(np.float32))
    ref = scene.tex_sample(i, 0.0, 0.0, bits)
    for bad in (np.nan, np.inf, -np.inf, 1e31):
        assert np.array_equal(scene.tex_sample(i, bad, 0.0, bits), ref)
        assert np.array_equal(scene.tex_sample(i, 0.0, bad, bits), ref)


def test_constant_textures_equal_material_colours():
    """Textures of one value c render like materials of colour c (FP64; the blend
    weights sum to 1 up to rounding, so the images agree to rounding noise)."""
    from tests.helpers import rel_mse
    sc = scenes.textured_cornell()
    flat = dict(sc)
    mats = sc["mats"].copy()
    texs = []
    for m in range(len(mats)):
        mats[m]["diffuseTexId"] = -1
        mats[m]["emissiveTexId"] = -1
    for k, (m, field) in enumerate([(0, "diffuseTexId"), (2, "diffuseTexId")]):
        c = np.array([0.5, 0.25, 0.75, 1.0], np.float32)
        texs.append(np.broadcast_to(c, (3, 5, 4)).copy())
        mats[m][field] = k
    flat_mats = mats.copy()
    for m in (0, 2):
        flat_mats[m]["diffuseTexId"] = -1
        flat_mats[m]["diffuse"] = (0.5, 0.25, 0.75)
    a = dict(sc, mats=mats, textures=texs)
    b = dict(sc, mats=flat_mats, textures=[])
    p = oracle.render_params(48, 48, sc["camera"])
    ra = oracle.Renderer(p, 64)
    ra.render(oracle.Scene(a), 0, 4)
    rb = oracle.Renderer(p, 64)
    rb.render(oracle.Scene(b), 0, 4)
    assert rel_mse(ra.result64(), rb.result64()) < 1e-8

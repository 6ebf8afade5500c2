"""Pins of the oracle's Disney-BRDF path tracer (mode 7; an extension with no
reference counterpart: CudaPT reads the Disney parameters and never uses them).

  - the BRDF: reciprocity, a hand-worked value at normal incidence, the
    directional albedo of a metal at most 1 (quadrature), the cosine lobe alone
    at zero specular / roughness 1 reaching Burley's retro-reflective form;
  - the sampler: E[f cos / pdf] over its samples equals the quadrature of
    f cos (the weight the path tracer multiplies its throughput by), which fails
    for any sampler whose pdf is not its density; every sample has pdf > 0;
  - the estimator: next-event estimation with power-heuristic MIS, BRDF
    sampling alone and light sampling alone have one expectation (three
    renders of a Cornell box with Disney materials agree within their noise),
    and a floor under a constant sky renders albedo * sky.
"""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle
from cudapt_amd import scenes

MATS = [
    dict(base=(0.8, 0.3, 0.2), metallic=0.0, specular=0.5, roughness=0.7, sheen=0.4, sheenTint=0.5,
         clearcoat=0.0, clearcoatGloss=0.0, subsurface=0.3),
    dict(base=(0.9, 0.7, 0.4), metallic=1.0, specular=0.5, roughness=0.25),
    dict(base=(0.2, 0.5, 0.8), metallic=0.3, specular=0.8, roughness=0.45, clearcoat=1.0, clearcoatGloss=0.7),
]


def _dir(theta, phi):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


N = np.array([0.0, 0.0, 1.0])


@pytest.mark.parametrize("m", MATS)
def test_reciprocity(m):
    rng = np.random.default_rng(5)
    for _ in range(200):
        a = _dir(math.acos(rng.uniform(0.02, 1)), rng.uniform(0, 2 * math.pi))
        b = _dir(math.acos(rng.uniform(0.02, 1)), rng.uniform(0, 2 * math.pi))
        np.testing.assert_allclose(oracle.disney_f(m, N, a, b), oracle.disney_f(m, N, b, a), rtol=1e-12, atol=1e-15)


def test_hand_value_normal_incidence():
    # wo = wi = n: FL = FV = FH = 0, Fd = 1, ss = 1.25 * (1/2 * 1 + 1/2) ... with
    # subsurface 0 only Fd counts; Ds = 1 / (pi a^2), a = r^2 = 0.25; Gs = (1/2)^2;
    # Fs = Cspec0 = 0.08 * 0.5 = 0.04
    m = dict(base=(0.5, 0.5, 0.5), metallic=0.0, specular=0.5, roughness=0.5)
    want = 0.5 / math.pi + 0.04 * 0.25 / (math.pi * 0.25 ** 2)
    np.testing.assert_allclose(oracle.disney_f(m, N, N, N), want, rtol=1e-13)
    # metallic: Cspec0 = base, no diffuse
    m2 = dict(base=(0.9, 0.6, 0.3), metallic=1.0, specular=0.5, roughness=0.5)
    np.testing.assert_allclose(oracle.disney_f(m2, N, N, N), np.array([0.9, 0.6, 0.3]) * 0.25 / (math.pi * 0.0625),
                               rtol=1e-13)
    # below the surface: nothing
    assert (oracle.disney_f(m, N, N, -N) == 0).all() and oracle.disney_pdf(m, N, N, -N) == 0


def _albedo(m, wo):
    def g(th, ph):
        wi = _dir(th, ph)
        return oracle.disney_f(m, N, wo, wi)[0] * math.cos(th) * math.sin(th)
    v, _ = integrate.dblquad(lambda th, ph: g(th, ph), 0, 2 * math.pi, 0, math.pi / 2, epsabs=1e-6, epsrel=1e-5)
    return v


@pytest.mark.parametrize("rough", [0.2, 0.5, 1.0])
def test_metal_albedo_at_most_one(rough):
    m = dict(base=(1.0, 1.0, 1.0), metallic=1.0, specular=0.5, roughness=rough)
    for th in (0.0, 0.6, 1.2):
        a = _albedo(m, _dir(th, 0.3))
        assert 0.0 < a <= 1.0 + 1e-4, (rough, th, a)
    if rough == 1.0:
        # a = 1: D = 1 / pi, Gs = 1 / ((NL + 1)(NV + 1)), F = 1 for a white metal; at
        # normal view the albedo is int_0^1 mu / (mu + 1) dmu = 1 - ln 2
        assert _albedo(m, N) == pytest.approx(1.0 - math.log(2.0), rel=1e-4)
    if rough == 0.2:
        assert _albedo(m, N) > 0.9                        # a smooth metal loses little


@pytest.mark.parametrize("m", MATS)
def test_sampler_weight_matches_quadrature(m):
    """E over the sampler of f cos / pdf (the path tracer's throughput factor)
    equals the directional albedo by quadrature."""
    wo = _dir(0.7, 1.1)
    wi, ok = oracle.disney_sample_many(m, N, wo, 0xD15, 40000)
    w = np.zeros(len(wi))
    for i in np.nonzero(ok)[0]:
        p = oracle.disney_pdf(m, N, wo, wi[i])
        assert p > 0
        w[i] = oracle.disney_f(m, N, wo, wi[i])[0] * wi[i, 2] / p
    est, se = w.mean(), w.std() / math.sqrt(len(w))
    ref = _albedo(m, wo)
    assert abs(est - ref) < 5 * se + 2e-4, (est, ref, se)


def _disney_cornell():
    sc = scenes.cornell()
    mats = sc["mats"].copy()
    rng = np.random.default_rng(3)
    for i in range(len(mats)):
        if mats[i]["emissive"].max() > 0:
            continue
        mats[i]["metallic"] = rng.uniform(0, 1)
        mats[i]["roughness"] = rng.uniform(0.2, 1)
        mats[i]["specular"] = rng.uniform(0, 1)
        mats[i]["sheen"] = rng.uniform(0, 1)
        mats[i]["clearcoat"] = rng.uniform(0, 1)
        mats[i]["clearcoatGloss"] = rng.uniform(0, 1)
    sc["mats"] = mats
    return sc


def _frame_means(O, sc, flags, W, H, F, depth=6):
    """Per-frame image means of mode 7 (frame images from the running mean,
    which weights every frame's one sample per pixel equally)."""
    R = oracle.Renderer(oracle.render_params(W, H, sc["camera"], mode=7, eye_depth=depth, flags=flags), 64)
    out, prev = [], None
    for f in range(F):
        R.render(O, f, 1)
        cur = R.result64()
        frame = cur if prev is None else (f + 1) * cur - f * prev
        out.append(frame.mean())
        prev = cur
    return np.array(out), cur


def test_mis_and_single_strategies_agree():
    """Next-event estimation with MIS, BRDF sampling alone and light sampling
    alone estimate the same image: their frame means agree within 4 standard
    errors, and MIS has the smallest per-pixel noise."""
    sc = _disney_cornell()
    O = oracle.Scene(sc)
    W, H, F = 16, 16, 160
    res = {name: _frame_means(O, sc, flags, W, H, F) for name, flags in (("mis", 0), ("bsdf", 2), ("light", 4))}
    mis = res["mis"][0]
    for name in ("bsdf", "light"):
        other = res[name][0]
        se = math.sqrt(mis.var(ddof=1) / F + other.var(ddof=1) / F)
        assert abs(mis.mean() - other.mean()) < 4 * se, (name, mis.mean(), other.mean(), se)
    assert mis.var() <= 1.05 * min(res["bsdf"][0].var(), res["light"][0].var())
    assert mis.mean() > 0.01


def test_direct_lighting_strategies_agree():
    """Depth 2 (direct lighting only): the three strategies differ only in how
    the light is reached, with little noise, so a weight pair that does not sum
    to one shows."""
    sc = _disney_cornell()
    O = oracle.Scene(sc)
    W, H, F = 16, 16, 400
    res = {name: _frame_means(O, sc, flags, W, H, F, depth=2)[0] for name, flags in (("mis", 0), ("bsdf", 2),
                                                                                      ("light", 4))}
    for name in ("bsdf", "light"):
        se = math.sqrt(res["mis"].var(ddof=1) / F + res[name].var(ddof=1) / F)
        assert abs(res["mis"].mean() - res[name].mean()) < 4 * se, (name, res["mis"].mean(), res[name].mean(), se)


def test_russian_roulette_is_unbiased():
    """Deep paths in a bright closed box: with and without Russian roulette the
    frame means agree (survivors are reweighted by 1 / q)."""
    sc = _disney_cornell()
    mats = sc["mats"].copy()
    for i in range(len(mats)):
        if mats[i]["emissive"].max() == 0:
            mats[i]["diffuse"] = (0.9, 0.9, 0.9)
    sc["mats"] = mats
    O = oracle.Scene(sc)
    W, H, F = 12, 12, 300
    rr = _frame_means(O, sc, 0, W, H, F, depth=12)[0]
    full = _frame_means(O, sc, 8, W, H, F, depth=12)[0]
    se = math.sqrt(rr.var(ddof=1) / F + full.var(ddof=1) / F)
    assert abs(rr.mean() - full.mean()) < 4 * se, (rr.mean(), full.mean(), se)


def test_floor_under_constant_sky():
    """One bounce off a grey Disney floor, then the constant sky: the image
    mean equals sky * the floor's directional albedo (quadrature), averaged
    over the view directions (small field of view, straight down)."""
    L = np.array([1.0, 1.0, 1.0], np.float32)
    b = scenes._Builder()
    m = dict(base=(0.6, 0.6, 0.6), metallic=0.2, specular=0.5, roughness=0.6, sheen=0.3, clearcoat=0.5,
             clearcoatGloss=0.5)
    mat = scenes.material(diffuse=m["base"], metallic=m["metallic"], specular=m["specular"], roughness=m["roughness"],
                          sheen=m["sheen"], clearcoat=m["clearcoat"])
    mat["clearcoatGloss"] = m["clearcoatGloss"]
    k = b.add_material(mat)
    b.add_quad((-1e3, 0, 1e3), (1e3, 0, 1e3), (1e3, 0, -1e3), (-1e3, 0, -1e3), (0, 1, 0), k)
    cam = dict(pos=(0.0, 2.0, 0.0), dir=(0.0, -1.0, 0.0), up=(0.0, 0.0, -1.0), fov=0.02)
    sc = b.build(cam)
    sc["env"] = np.broadcast_to(L, (8, 16, 3)).copy()
    O = oracle.Scene(sc)
    R = oracle.Renderer(oracle.render_params(16, 16, cam, mode=7, eye_depth=3), 64)
    img, _ = R.render(O, 0, 256)
    want = _albedo(m, N)
    # 65536 one-bounce samples: a few 1e-3 of noise
    assert abs(img[..., 0].mean() - want) < 0.01 * want, (img[..., 0].mean(), want)

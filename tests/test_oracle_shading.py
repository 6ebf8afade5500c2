"""Pins of the oracle's BRDF evaluation, BRDF sampling, Russian roulette and
light-subpath bookkeeping (src/cudaRTCommon.cu:284-365,
src/cudaRTBDPTStream.cu:180-363,395-543,642-663).

Each test checks the oracle against something it does not compute itself:
quadrature, a hand-worked value, the closed-form distribution of a sampler, a
statistic of a Monte Carlo count, or an exact identity.  Each is chosen so that
a plausible one-line mistake fails it:
  - D_GGX with alpha = r instead of r^2, or a lost pi        -> normal-incidence value, quadrature
  - a sampler drawing from another lobe than D_GGX (n.h)    -> KS against the quadrature CDF
  - Schlick with another exponent or without the grazing term -> hand values at VoH = 0, 1/2, 1
  - Smith with k = (r+1)^2/8 or k = r^2                      -> hand values
  - a 'corrected' sqrt(1 - r2cos^2) hemisphere sampler        -> KS against the r2sin = 1 - r2cos^2 CDF
  - the specular lobe divided by 4 NoL NoV                    -> hand value of GetShadingResult
  - a Lambert term without 1/pi or without (1 - F)            -> white furnace
  - a wrong branch probability or weight                      -> branch frequencies and exact weights
  - Russian roulette on max channel, or survivors reweighted  -> mean path length in a closed box
  - irrad kept after a specular light bounce                  -> two-bounce light subpath by hand
"""
import math

import numpy as np
import pytest
from scipy import integrate, stats

import oracle
from cudapt_amd import scenes

PI = math.pi


# ---------------------------------------------------------------- D_GGX, ImportanceSampleGGX
@pytest.mark.parametrize("rough", [0.15, 0.3, 0.5, 0.8, 1.0])
def test_ggx_normalised_over_the_hemisphere(rough):
    # integral of D(h) (n.h) over the hemisphere of h = 2 pi int_0^1 D(c) c dc = 1
    a2 = rough ** 4
    pts = [max(0.0, 1 - 50 * a2), 1 - 5 * a2, 1 - a2] if a2 < 0.02 else None
    val, _ = integrate.quad(lambda c: oracle.d_ggx(rough, c) * c, 0.0, 1.0, points=pts, limit=400,
                            epsabs=1e-12, epsrel=1e-10)
    assert abs(2 * PI * val - 1.0) < 1e-7


@pytest.mark.parametrize("rough", [0.2, 0.5, 1.0])
def test_ggx_peak_value_uses_alpha_r_squared(rough):
    # D(n.h = 1) = 1 / (pi alpha^2) with alpha = roughness^2 (UE4 convention)
    assert oracle.d_ggx(rough, 1.0) == pytest.approx(1.0 / (PI * rough ** 4), rel=1e-12)
    # and D at n.h = 0 is alpha^2 / pi
    assert oracle.d_ggx(rough, 0.0) == pytest.approx(rough ** 4 / PI, rel=1e-12)


@pytest.mark.parametrize("rough", [0.25, 0.6, 1.0])
def test_ggx_sampler_draws_the_d_ggx_lobe(rough):
    """The half vectors of ImportanceSampleGGX follow D(h) (n.h) dw: the CDF of
    cos(theta_h) from quadrature of the oracle's own d_ggx."""
    rng = np.random.default_rng(7)
    n_s = 6000
    xi = rng.random((n_s, 2))
    normal = np.array([0.3, -0.5, 0.81])
    normal /= np.linalg.norm(normal)
    h = np.array([oracle.sample_ggx(x0, x1, rough, normal) for x0, x1 in xi])
    np.testing.assert_allclose(np.linalg.norm(h, axis=1), 1.0, atol=1e-12)
    cos_h = np.clip(h @ normal, -1, 1)

    def cdf(c):  # P(cos <= c) = 1 - 2 pi int_c^1 D(x) x dx
        c = np.atleast_1d(c)
        out = np.empty_like(c, dtype=float)
        for i, ci in enumerate(c):
            tail, _ = integrate.quad(lambda x: oracle.d_ggx(rough, x) * x, ci, 1.0, limit=200, epsabs=1e-12)
            out[i] = 1.0 - 2 * PI * tail
        return out

    grid = np.unique(np.concatenate([[0.0], 1.0 - np.geomspace(1e-10, 1.0, 1500), [1.0]]))
    table = cdf(grid)
    ks = stats.kstest(cos_h, lambda c: np.interp(c, grid, table))
    assert ks.pvalue > 1e-3, ks
    # azimuth about the normal is uniform
    u, v = oracle.basis(normal)
    phi = np.arctan2(h @ v, h @ u)
    assert stats.kstest((phi + PI) / (2 * PI), "uniform").pvalue > 1e-3


def test_sampler_frame_is_orthonormal():
    rng = np.random.default_rng(3)
    for w in rng.normal(size=(200, 3)):
        w /= np.linalg.norm(w)
        u, v = oracle.basis(w)
        m = np.stack([u, v, w])
        np.testing.assert_allclose(m @ m.T, np.eye(3), atol=1e-12)


# ---------------------------------------------------------------- F_Schlick, Vis_SmithJointApprox
def _rn32(q):
    """The float32 nearest the rational q (ties to even)."""
    from fractions import Fraction
    f = np.float32(float(q))
    best = None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        e = abs(Fraction(float(c)) - q)
        if best is None or e < best[0] or (e == best[0] and int(c.view(np.uint32)) % 2 == 0):
            best = (e, c)
    return best[1]


def test_schlick_fifth_power_fp32_is_correctly_rounded():
    """The FP32 restatement's (1 - VoH)^5, the device's double-evaluated power:
    equal to the correctly rounded x^5 on (nearly) every argument and never
    more than an ulp away (F = Fc exactly in channel x when f0 = (0, 0, 1))."""
    from fractions import Fraction
    rng = np.random.default_rng(5)
    voh = rng.uniform(0.0, 1.0, 4000).astype(np.float32)
    exact = same = 0
    for v in voh:
        x = np.float32(1.0) - v
        fc = np.float32(oracle.f_schlick(np.array([0.0, 0.0, 1.0]), float(v), prec=32)[0])
        want = _rn32(Fraction(float(x)) ** 5)
        ulp = abs(int(fc.view(np.int32)) - int(want.view(np.int32)))
        assert ulp <= 1, (v, fc, want)
        same += ulp == 0
        exact += 1
    assert same >= 0.999 * exact
    # powers of two are exact: x = 1/2 -> 1/32
    assert np.float32(oracle.f_schlick(np.array([0.0, 0.0, 1.0]), 0.5, prec=32)[0]) == np.float32(1 / 32)


def test_schlick_hand_values():
    f0 = np.array([0.04, 0.04, 0.04])
    # VoH = 1: Fc = 0, F = f0 exactly
    assert (oracle.f_schlick(f0, 1.0) == f0).all()
    coloured = np.array([0.9, 0.6, 0.3])
    assert (oracle.f_schlick(coloured, 1.0) == coloured).all()
    # VoH = 0: Fc = 1, every channel = saturate(50 f0.z)
    np.testing.assert_array_equal(oracle.f_schlick(np.array([0.5, 0.2, 0.01]), 0.0), [0.5, 0.5, 0.5])
    np.testing.assert_array_equal(oracle.f_schlick(f0, 0.0), [1.0, 1.0, 1.0])      # 50 * 0.04 = 2 -> 1
    np.testing.assert_array_equal(oracle.f_schlick(np.zeros(3), 0.0), [0.0, 0.0, 0.0])
    # VoH = 1/2: Fc = 1/32, F = 1/32 + 0.04 * 31/32 = 0.07
    np.testing.assert_allclose(oracle.f_schlick(f0, 0.5), 0.07, rtol=1e-14)
    # z drives the grazing term of every channel: (0.9, 0.6, 0.01) -> saturate(0.5) = 0.5
    np.testing.assert_allclose(oracle.f_schlick(np.array([0.9, 0.6, 0.01]), 0.5),
                               0.5 / 32 + np.array([0.9, 0.6, 0.01]) * 31 / 32, rtol=1e-14)


def test_smith_hand_values():
    # k = r^2 / 2; G = NoV / (NoV (1 - k) + k) * NoL / (NoL (1 - k) + k)
    assert oracle.vis_smith(0.5, 0.5, 1.0) == pytest.approx(8.0 / 9.0, rel=1e-14)       # k = 1/8
    assert oracle.vis_smith(1.0, 0.25, 0.5) == pytest.approx(0.4 * (2.0 / 3.0), rel=1e-14)  # k = 1/2
    assert oracle.vis_smith(0.8, 0.6, 0.3) == pytest.approx(
        (0.6 / (0.6 * 0.68 + 0.32)) * (0.3 / (0.3 * 0.68 + 0.32)), rel=1e-14)              # k = 0.32
    for r in (0.1, 0.5, 1.0):
        assert oracle.vis_smith(r, 1.0, 1.0) == pytest.approx(1.0, rel=1e-15)


# ---------------------------------------------------------------- GetShadingResult
def test_lambert_white_furnace():
    """A white Lambertian surface (specular 0, metallic 0: f0 = 0, F = 0) under
    unit irradiance from every direction reflects exactly 1."""
    n = np.array([0.0, 0.0, 1.0])
    out_dir = np.array([0.3, 0.0, math.sqrt(1 - 0.09)])

    def f(theta, phi):
        w = np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])
        return oracle.shading(out_dir, w, np.ones(3), n, np.ones(3), 0.0, 0.7, 0.0, 1.0, 1.0)[0] * math.sin(theta)

    val, _ = integrate.dblquad(lambda th, ph: f(th, ph), 0, 2 * PI, 0, PI / 2, epsabs=1e-9, epsrel=1e-9)
    assert val == pytest.approx(1.0, abs=1e-7)
    # light from below the surface gives nothing
    assert (oracle.shading(out_dir, -out_dir, np.ones(3), n, np.ones(3), 0.0, 0.7, 0.0, 1.0, 1.0) == 0).all()


def test_diffuse_weight_and_fresnel_split():
    # diffuse term carries (1 - F) and diffW; F at VoH from the half vector
    n = np.array([0.0, 0.0, 1.0])
    o = np.array([0.0, 0.0, 1.0])
    i = np.array([0.6, 0.0, 0.8])
    h = (o + i) / np.linalg.norm(o + i)
    voh = float(o @ h)
    F = 0.04 + (1 - 0.04) * 0 + (1.0 - voh) ** 5 * (1.0 - 0.04)      # saturate(50 * 0.04) = 1
    diff = np.array([0.5, 0.25, 0.125])
    got = oracle.shading(o, i, np.ones(3), n, diff, 0.0, 1.0, 0.5, 0.3, 0.0)  # specular weight 0
    np.testing.assert_allclose(got, 0.8 * (1 - F) * diff / PI * 0.3, rtol=1e-12)


def test_specular_lobe_is_d_times_g_without_quarter_denominator():
    """The reference's GetShadingResult returns D * Vis * F * NoL with Vis the
    Schlick-Smith G itself, not divided by 4 NoL NoV (cudaRTBDPTStream.cu:659).
    White metal: f0 = diff = 1, F = 1, no diffuse term."""
    n = np.array([0.0, 0.0, 1.0])
    s = 1 / math.sqrt(2)
    o, i = np.array([s, 0.0, s]), np.array([-s, 0.0, s])       # h = n
    r = 0.5
    D = 1.0 / (PI * r ** 4)
    k = r * r / 2
    G = (s / (s * (1 - k) + k)) ** 2
    got = oracle.shading(o, i, np.ones(3), n, np.ones(3), 1.0, r, 0.0, 1.0, 1.0)
    np.testing.assert_allclose(got, s * D * G, rtol=1e-12)      # 3.2556..., not 3.2556 / (4 * 0.5)
    assert got[0] == pytest.approx(3.25543, rel=1e-5)


# ---------------------------------------------------------------- the eye bounce's BRDF sampling
def _diffuse_cos_cdf(c):
    """CDF of n.d for the reference's hemisphere sampler r2cos = sqrt(u),
    r2sin = 1 - r2cos^2 (cudaRTBDPTStream.cu:482-490): c^2 = u / (1 - u + u^2),
    increasing in u, so P(n.d <= c) is the root u of c^2 u^2 - (c^2 + 1) u + c^2 = 0
    in [0, 1]."""
    c2 = np.asarray(c, float) ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        u = ((c2 + 1) - np.sqrt((c2 + 1) ** 2 - 4 * c2 * c2)) / (2 * c2)
    return np.where(c2 > 0, u, 0.0)


@pytest.mark.parametrize("prec", [64, 32])
def test_diffuse_sampler_distribution_and_weight(prec):
    nl = np.array([0.2, 0.9, -0.1])
    nl /= np.linalg.norm(nl)
    rd = -nl
    diff = np.array([0.7, 0.5, 0.3])
    n_s = 20000
    ty, orig, d, mul = oracle.sample_bounce_many(np.zeros(3), rd, 2.0, nl, diff, 0.0, 0.0, 0.0, 1.0, nl, 11, n_s,
                                                 prec=prec)
    assert (ty == 1).all()                                   # specular 0, metallic 0: F = 0, reflProb 0
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    cos = np.clip(d @ nl, 0.0, 1.0)
    ks = stats.kstest(cos, _diffuse_cos_cdf)
    assert ks.pvalue > 1e-3, ks
    # a cosine-weighted sampler (r2sin = sqrt(1 - u)) has CDF c^2: clearly rejected
    assert stats.kstest(cos, lambda c: np.asarray(c) ** 2).pvalue < 1e-20
    u, v = oracle.basis(nl)
    phi = np.arctan2(d @ v, d @ u)
    assert stats.kstest((phi + PI) / (2 * PI), "uniform").pvalue > 1e-3
    # weight: pi * diff / pi / ((1 - refr) (1 - refl)) = diff, no cosine or pdf
    np.testing.assert_allclose(mul, np.broadcast_to(diff, mul.shape), rtol=1e-6 if prec == 32 else 1e-14)
    # origin: the hit point stepped 1e-4 along the new direction
    np.testing.assert_allclose(orig, 2.0 * rd + 1e-4 * d, atol=1e-6 if prec == 32 else 1e-14)


@pytest.mark.parametrize("trans", [0.0, 0.3])
def test_branch_frequencies_and_weights(trans):
    """Near-mirror microfacets (roughness 1e-3) at normal incidence: h = n, so
    F = f0 = 0.08 * specular (metallic 0), reflProb = |F| = 0.04 sqrt(3),
    refrProb = transparency.  Branch frequencies reflProb, (1 - reflProb) trans,
    (1 - reflProb)(1 - trans); weights F G VoH / (NoV NoH reflProb) = F / |F|,
    cos * diff / (trans (1 - reflProb)) with cos = 1, and
    diff / ((1 - trans)(1 - reflProb))."""
    nl = np.array([0.0, 1.0, 0.0])
    rd = -nl
    diff = np.array([0.6, 0.4, 0.2])
    n_s = 40000
    ty, orig, d, mul = oracle.sample_bounce_many(np.array([0, 3.0, 0]), rd, 3.0, nl, diff, trans, 0.5, 0.0, 1e-3,
                                                 nl, 5, n_s)
    p_refl = 0.04 * math.sqrt(3)
    spec = ty == 2
    refl = spec & (d @ nl > 0)
    refr = spec & (d @ nl < 0)
    dif = ty == 1
    for got, p in ((refl.mean(), p_refl), (refr.mean(), (1 - p_refl) * trans),
                   (dif.mean(), (1 - p_refl) * (1 - trans))):
        sd = math.sqrt(max(p * (1 - p), 1e-12) / n_s)
        assert abs(got - p) <= 5 * sd + 1e-12, (got, p)
    np.testing.assert_allclose(mul[refl], 1 / math.sqrt(3), rtol=2e-5)
    np.testing.assert_allclose(d[refl], np.broadcast_to(nl, d[refl].shape), atol=2e-3)
    # reflection origin: the hit point stepped back 1e-4 along the incoming ray
    np.testing.assert_allclose(orig[refl], np.broadcast_to([0, 1e-4, 0], orig[refl].shape), atol=1e-9)
    np.testing.assert_allclose(mul[dif], np.broadcast_to(diff / ((1 - trans) * (1 - p_refl)), mul[dif].shape),
                               rtol=2e-5)
    if trans:
        np.testing.assert_allclose(mul[refr], np.broadcast_to(diff / (trans * (1 - p_refl)), mul[refr].shape),
                                   rtol=2e-5)
        np.testing.assert_allclose(d[refr], np.broadcast_to(-nl, d[refr].shape), atol=2e-3)
        # transmission origin: past the hit point by 1e-4, then 1e-4 along the refraction
        np.testing.assert_allclose(orig[refr], np.broadcast_to([0, -2e-4, 0], orig[refr].shape), atol=1e-8)


def test_metal_always_reflects_with_weight_f():
    nl = np.array([0.0, 0.0, 1.0])
    diff = np.array([0.9, 0.6, 0.3])
    ty, _, d, mul = oracle.sample_bounce_many(np.zeros(3), -nl, 1.0, nl, diff, 0.0, 0.5, 1.0, 1e-3, nl, 9, 2000)
    assert (ty == 2).all()                                   # reflProb = lerp(|F|, 1, 1) = 1
    np.testing.assert_allclose(mul, np.broadcast_to(diff, mul.shape), rtol=2e-5)   # F(VoH = 1) = f0 = diff


# ---------------------------------------------------------------- Russian roulette
def _closed_box(albedo):
    b = scenes._Builder()
    m = b.add_material(scenes.material(diffuse=(albedo,) * 3))
    b.add_box((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.3, m)
    cam = dict(pos=(0.0, 0.0, 0.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov=1.2)
    return b.build(cam), cam


@pytest.mark.parametrize("albedo,depth", [(0.3, 5), (0.5, 5), (0.5, 8)])
def test_russian_roulette_kill_rate(albedo, depth):
    """A closed grey Lambertian box with no emitter: every ray hits a wall, no
    light vertex exists, and a path survives bounce j with probability
    min(1, |T_j| |f|) = min(1, 3 a^(j+1)) (T_j = a^j (1, 1, 1), f = (a, a, a),
    survivors not reweighted), up to the depth cap.  The mean number of
    extension rays per path is sum_k prod_{j<k} min(1, 3 a^(j+1))."""
    sc, cam = _closed_box(albedo)
    O = oracle.Scene(sc)
    W = H = 64
    F = 4
    R = oracle.Renderer(oracle.render_params(W, H, cam, eye_depth=depth), 64)
    img, N = R.render(O, 0, F)
    ext, shadow = R.ray_counts()
    assert shadow == 0 and (img == 0).all() and (N == F).all()
    s = [min(1.0, 3 * albedo ** (j + 1)) for j in range(depth)]
    mean = sum(math.prod(s[:k]) for k in range(depth))
    second = sum((2 * k + 1) * math.prod(s[:k]) for k in range(depth))   # E[L^2] = sum (2k+1) P(L > k)
    sd = math.sqrt((second - mean ** 2) / (W * H * F))
    assert abs(ext / (W * H * F) - mean) < 5 * sd, (ext / (W * H * F), mean, sd)


# ---------------------------------------------------------------- light subpaths
def _light_and_ceiling(ceiling):
    b = scenes._Builder()
    lm = b.add_material(scenes.material(diffuse=(0.5, 0.5, 0.5), emissive=(3.0, 2.0, 1.0)))
    cm = b.add_material(ceiling)
    Ex = 200.0
    b.add_quad((-Ex, 0, Ex), (Ex, 0, Ex), (Ex, 0, -Ex), (-Ex, 0, -Ex), (0, 1, 0), lm)        # light, facing up
    b.add_quad((-Ex, 1, -Ex), (Ex, 1, -Ex), (Ex, 1, Ex), (-Ex, 1, Ex), (0, -1, 0), cm)       # ceiling, facing down
    cam = dict(pos=(0.0, 0.5, 0.0), dir=(0.0, 1.0, 0.0), up=(0.0, 0.0, -1.0), fov=0.5)
    return b.build(cam), cam


@pytest.mark.parametrize("kind", ["diffuse", "metal"])
def test_two_bounce_light_subpath_by_hand(kind):
    """Light paths leave an emissive floor (E = (3, 2, 1)), bounce off a ceiling
    and come back to the floor, which is an emitter and ends them
    (cudaRTBDPTStream.cu:180-363,545-617).  The cache holds, in order:
      slots [0, L): emission points, irrad = E, irradDir = 0;
      slots [L, 2L): the ceiling hits, irrad = the path's sample so far = E,
        zeroed when the next bounce is not diffuse (:331-332), irradDir = -dir;
      then the survivors' floor hits, irrad = E * f_ceiling, the diffuse weight
        being the albedo itself (no cosine, no pdf)."""
    a = 0.2
    ceiling = (scenes.material(diffuse=(a, a, a)) if kind == "diffuse"
               else scenes.material(diffuse=(0.9, 0.9, 0.9), metallic=1.0, roughness=1e-3))
    sc, cam = _light_and_ceiling(ceiling)
    O = oracle.Scene(sc)
    L = 64
    R = oracle.Renderer(oracle.render_params(4, 4, cam, light_paths=L, light_depth=5), 64)
    R.render(O, 0, 1)
    lv = R.light_vertices()                    # pos 0:3 norm 3:6 irrad 6:9 irradDir 9:12 diff emissive spec metal rough
    E = np.array([3.0, 2.0, 1.0], np.float32)
    org, b1 = lv[:L], lv[L:2 * L]
    assert (org[:, 1] == 0).all() and (org[:, 6:9] == E).all() and (org[:, 9:12] == 0).all()
    assert (org[:, 3:6] == [0, 1, 0]).all()
    hit = b1[:, 1] == 1.0                      # a grazing ray can leave past the edge: then the origin repeats
    assert hit.mean() > 0.95
    np.testing.assert_array_equal(b1[~hit], org[~hit])
    b1 = b1[hit]
    np.testing.assert_array_equal(b1[:, 3:6], np.broadcast_to([0, -1, 0], b1[:, 3:6].shape))
    assert (b1[:, 10] < 0).all()               # irradDir = -dir points back down to the light
    np.testing.assert_allclose(np.linalg.norm(b1[:, 9:12], axis=1), 1.0, atol=1e-6)
    rest = lv[2 * L:]
    if kind == "diffuse":
        np.testing.assert_array_equal(b1[:, 6:9], np.broadcast_to(E, b1[:, 6:9].shape))
        # survivors: Russian roulette at min(1, |(1,1,1)| |(a,a,a)|) = 3a = 0.6
        assert 0.3 < len(rest) / hit.sum() < 0.9
        floor = rest[(rest[:, 1] == 0) & (np.abs(rest[:, 9:12]).sum(axis=1) > 0)]   # not a repeated origin
        np.testing.assert_allclose(floor[:, 6:9], np.broadcast_to(E * a, floor[:, 6:9].shape), rtol=1e-7)
        np.testing.assert_array_equal(floor[:, 3:6], np.broadcast_to([0, 1, 0], floor[:, 3:6].shape))
    else:
        # specular next bounce: the cached irradiance is zeroed
        assert (b1[:, 6:9] == 0).all()
        floor = rest[(rest[:, 1] == 0) & (np.abs(rest[:, 9:12]).sum(axis=1) > 0)]
        assert len(floor) > 0 and (floor[:, 6:9] > 0).all()
    # nothing goes past the floor hit: emitters end light paths
    assert len(rest) <= hit.sum()

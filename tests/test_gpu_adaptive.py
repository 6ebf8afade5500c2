"""Adaptive BDPT on the GPU (cudaRTBDPTStreamAdapBilatAtomic) against the oracle.

The cross-bilateral kernel uses the defined one-exponential-per-tap arithmetic,
so its output and the error map equal the oracle's FP32 restatement bit for
bit, and the sample allocation derived from them is identical.  The literal
FP64 filter (product of the reference's per-term exponentials) bounds the
defined arithmetic's deviation.
"""
import numpy as np
import pytest

import oracle
from cudapt_amd import scenes
from tests.helpers import rel_mse

pytestmark = pytest.mark.gpu


def _images(seed, H, W):
    rng = np.random.default_rng(seed)
    col = rng.uniform(0, 2, (H, W, 3)).astype(np.float32)
    pos = rng.uniform(-1, 1, (H, W, 3)).astype(np.float32)
    nrm = rng.normal(size=(H, W, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    dif = rng.uniform(0, 1, (H, W, 3)).astype(np.float32)
    N = rng.integers(1, 60, (H, W)).astype(np.uint32)
    return col, pos, nrm, dif, N


@pytest.mark.parametrize("radius,use_diff,W,H", [(0, False, 19, 13), (3, True, 40, 23), (5, False, 64, 48),
                                                 (7, True, 37, 50), (12, False, 33, 29)])
def test_filter_bit_exact_against_oracle(cpt, radius, use_diff, W, H):
    col, pos, nrm, dif, N = _images(radius + 1, H, W)
    sig = (0.5, 0.4, 0.25, 0.1)
    f, e = cpt.filter_cross_bilateral(col, pos, nrm, dif if use_diff else None, N, *sig, radius, nd=5)
    ref = oracle.bilateral_defined(col, pos, nrm, dif if use_diff else None, *sig, radius)
    assert np.array_equal(f.view(np.uint32), ref.view(np.uint32))
    ref_e = oracle.square_error(ref, col, N, 5)
    assert np.array_equal(e.view(np.uint32), ref_e.view(np.uint32))
    lit = oracle.bilateral(col, pos, nrm, dif if use_diff else None, *sig, radius, prec=64)
    np.testing.assert_allclose(f, lit, rtol=2e-5, atol=1e-6)


@pytest.fixture(scope="module")
def adaptive_runs(cpt):
    sc = scenes.cornell()
    W, H, F = 128, 128, 12
    S = cpt.Scene(sc)
    R = cpt.Renderer(W, H, mode=cpt.CPT_MODE_BDPT_ADAPTIVE_PLL)
    gpu = R.render(S, sc["camera"], frames=F)
    out = dict(sc=sc, S=S, gpu=gpu, N=R.sample_counts(), counts=R.read(cpt.CPT_BUF_SAMPLE_COUNTS),
               err=R.read(cpt.CPT_BUF_ERROR), filtered=R.read(cpt.CPT_BUF_FILTERED),
               gpos=R.read(cpt.CPT_BUF_GBUF_POS), stats=R.stats())
    O = oracle.Scene(sc)
    p = oracle.render_params(W, H, sc["camera"], mode=1)
    o32 = oracle.Renderer(p, 32)
    out["r32"], out["n32"] = o32.render(O, 0, F)
    out["st32"] = o32.adaptive_state()
    o64 = oracle.Renderer(p, 64)
    o64.render(O, 0, F)
    out["r64"] = o64.result64()
    out["W"], out["H"] = W, H
    return out


def test_adaptive_allocation_matches_oracle(adaptive_runs):
    r = adaptive_runs
    c, c32 = r["counts"], r["st32"]["counts"]
    assert c.sum() <= r["W"] * r["H"]
    assert c.sum() > 0.5 * r["W"] * r["H"]          # floor(W*H*e_i/sum) drops the fractional parts
    assert (c == c32).mean() > 0.999
    # the filter is bit-exact on equal inputs (test above); here its inputs, the
    # accumulated images, differ in low bits through device vs libm transcendentals
    close = np.isclose(r["err"], r["st32"]["err"], rtol=1e-3, atol=1e-9)
    assert close.mean() > 0.99


def test_adaptive_fp32_vs_fp64_rel_mse(adaptive_runs):
    r = adaptive_runs
    e = rel_mse(r["gpu"], r["r64"])
    assert e <= 1e-4, f"relative MSE {e:.3e}"


def test_adaptive_gpu_vs_fp32_oracle_tight(adaptive_runs):
    r = adaptive_runs
    assert rel_mse(r["gpu"], r["r32"]) <= 1e-6
    assert (r["N"] == r["n32"]).mean() > 0.99
    np.testing.assert_allclose(r["gpos"], r["st32"]["gpos"], rtol=1e-4, atol=1e-4)


def test_adaptive_is_deterministic(cpt, adaptive_runs):
    r = adaptive_runs
    R = cpt.Renderer(r["W"], r["H"], mode=cpt.CPT_MODE_BDPT_ADAPTIVE_PLL)
    again = R.render(r["S"], r["sc"]["camera"], frames=12)
    assert np.array_equal(again.view(np.uint32), r["gpu"].view(np.uint32))


def test_adaptive_spends_samples_where_error_is(adaptive_runs):
    r = adaptive_runs
    err, counts = r["err"].ravel(), r["counts"].ravel()
    hi = err > np.quantile(err, 0.9)
    lo = err < np.quantile(err, 0.5)
    assert counts[hi].mean() > 3 * max(counts[lo].mean(), 1e-3)


@pytest.fixture(scope="module")
def pass_runs(cpt):
    """Adaptive passes of 4 frames' worth of samples (Adaptive Pass Frames = 4)."""
    sc = scenes.cornell()
    W, H, F = 96, 80, 17
    S = cpt.Scene(sc)
    R = cpt.Renderer(W, H, mode=cpt.CPT_MODE_BDPT_ADAPTIVE_PLL)
    R.set_attrib("Adaptive Pass Frames", 4)
    gpu = R.render(S, sc["camera"], frames=F)
    out = dict(gpu=gpu, N=R.sample_counts(), counts=R.read(cpt.CPT_BUF_SAMPLE_COUNTS), stats=R.stats())
    O = oracle.Scene(sc)
    p = oracle.render_params(W, H, sc["camera"], mode=1, pass_frames=4)
    o32 = oracle.Renderer(p, 32)
    out["r32"], out["n32"] = o32.render(O, 0, F)
    out["st32"] = o32.adaptive_state()
    o64 = oracle.Renderer(p, 64)
    o64.render(O, 0, F)
    out["r64"] = o64.result64()
    out["W"], out["H"], out["F"] = W, H, F
    return out


def test_pass_frames_match_oracle(pass_runs):
    r = pass_runs
    n = r["W"] * r["H"]
    # last pass: frames 13..16, 4 * W * H samples from one error map
    assert r["counts"].sum() <= 4 * n and r["counts"].sum() > 2 * n
    assert (r["counts"] == r["st32"]["counts"]).mean() > 0.999
    assert rel_mse(r["gpu"], r["r32"]) <= 1e-6
    assert rel_mse(r["gpu"], r["r64"]) <= 1e-4
    assert (r["N"] == r["n32"]).mean() > 0.99
    assert r["stats"]["frames"] == r["F"]


def test_heavily_sampled_pixels_take_the_workgroup_path(pass_runs):
    """Pixels with more than 32 records in a pass are summed by one workgroup each."""
    r = pass_runs
    big = r["counts"] > 32
    assert big.any()
    np.testing.assert_allclose(r["gpu"][big], r["r32"][big], rtol=1e-4, atol=1e-6)

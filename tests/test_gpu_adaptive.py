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
    """Pixels with more than 32 records in a pass are summed by one wave each."""
    r = pass_runs
    big = r["counts"] > 32
    assert big.any()
    np.testing.assert_allclose(r["gpu"][big], r["r32"][big], rtol=1e-4, atol=1e-6)


@pytest.fixture(scope="module", params=[(2, 1), (2, 3), (3, 1), (3, 3)],
                ids=["binary", "binary-pass3", "empirical", "empirical-pass3"])
def round_runs(cpt, request):
    """Round-based allocation modes (AdapBilatBinary, AdapEmp) against the oracle.
    The state before the last pass is read back so that the pass's error map and
    allocation can be recomputed from the device's own inputs."""
    mode, pf = request.param
    sc = scenes.cornell()
    W, H, F = 80, 64, 12
    last = (F - 5) % pf or pf                      # frames in the last pass (5 warm-up frames)
    S = cpt.Scene(sc)
    R = cpt.Renderer(W, H, mode=mode)
    R.set_attrib("Adaptive Pass Frames", pf)
    pre = R.render(S, sc["camera"], frames=F - last)
    out = dict(mode=mode, pre=pre, pre_N=R.sample_counts(), pre_sq=R.read(cpt.CPT_BUF_SECOND_MOMENT),
               pre_g=[R.read(b) for b in (cpt.CPT_BUF_GBUF_POS, cpt.CPT_BUF_GBUF_NORM)])
    gpu = R.render(S, sc["camera"], frames=last)
    out.update(gpu=gpu, N=R.sample_counts(), counts=R.read(cpt.CPT_BUF_SAMPLE_COUNTS),
               err=R.read(cpt.CPT_BUF_ERROR), sq=R.read(cpt.CPT_BUF_SECOND_MOMENT), stats=R.stats())
    O = oracle.Scene(sc)
    p = oracle.render_params(W, H, sc["camera"], mode=mode, pass_frames=pf)
    o32 = oracle.Renderer(p, 32)
    out["r32"], out["n32"] = o32.render(O, 0, F)
    out["st32"] = o32.adaptive_state()
    o64 = oracle.Renderer(p, 64)
    o64.render(O, 0, F)
    out["r64"] = o64.result64()
    out["W"], out["H"], out["F"], out["pf"], out["last"] = W, H, F, pf, last
    return out


def test_round_allocation_bit_exact_on_device_error_map(round_runs):
    """The pass's counts equal the oracle's rounds on the device's error map."""
    r = round_runs
    n = r["W"] * r["H"]
    counts, rounds = oracle.alloc_rounds(r["err"], r["last"] * n, bisect=r["mode"] == 2, desired=n / 5,
                                         frame=r["F"] - r["last"])
    assert rounds >= 1
    assert np.array_equal(r["counts"], counts)
    assert r["counts"].sum() == r["last"] * n


def test_empirical_error_bit_exact_on_device_state(round_runs):
    r = round_runs
    if r["mode"] != 3:
        pytest.skip("empirical error only")
    a = r["pre"].astype(np.float32)
    b = r["pre_sq"].astype(np.float32)
    N = r["pre_N"].astype(np.uint32)
    num = ((b[..., 0] - a[..., 0] * a[..., 0]) + (b[..., 1] - a[..., 1] * a[..., 1])) + (b[..., 2] - a[..., 2] * a[..., 2])
    with np.errstate(divide="ignore", invalid="ignore"):
        e = num * N.astype(np.float32) / ((N - np.uint32(1)).astype(np.float32) * np.float32(3))
    e = np.fmin(e, np.float32(1)).astype(np.float32)
    assert np.array_equal(e.view(np.uint32), r["err"].view(np.uint32))


def test_filter_residual_bit_exact_on_device_state(round_runs):
    r = round_runs
    if r["mode"] != 2:
        pytest.skip("filter residual only")
    f = oracle.bilateral_defined(r["pre"], r["pre_g"][0], r["pre_g"][1], None, 50.0, 1.0, 0.25, 0.1, 5)
    e = oracle.square_error(f, r["pre"], r["pre_N"], 5)
    assert np.array_equal(e.view(np.uint32), r["err"].view(np.uint32))


def test_round_modes_match_oracle(round_runs):
    """End to end against the oracle.  One pixel that crosses a keep threshold
    changes the slots filled after its round and so the streams of every later
    round, so any last-bit difference of the images would make the allocations
    diverge; the gate is the noise of 12 frames against FP64 (the allocation
    itself is bit-exact on equal error maps, test above)."""
    r = round_runs
    assert rel_mse(r["gpu"], r["r64"]) <= 3e-2
    assert abs(r["gpu"].mean() / r["r64"].mean() - 1) < 0.02
    assert (r["N"] > 0).all()
    if r["mode"] == 3:
        assert r["err"].max() <= 1.0


def test_round_modes_are_deterministic(cpt, round_runs):
    r = round_runs
    sc = scenes.cornell()
    S = cpt.Scene(sc)
    R = cpt.Renderer(r["W"], r["H"], mode=r["mode"])
    R.set_attrib("Adaptive Pass Frames", r["pf"])
    again = R.render(S, sc["camera"], frames=r["F"])
    assert np.array_equal(again.view(np.uint32), r["gpu"].view(np.uint32))


def test_empirical_mode_has_no_filter(cpt):
    sc = scenes.cornell()
    S = cpt.Scene(sc)
    R = cpt.Renderer(16, 16, mode=cpt.CPT_MODE_BDPT_ADAPTIVE_EMP)
    assert list(R.attribs())[:4] == ["Adaptive Mode", "Desired Max Sampling", "Min Trace Probability", "Debug Mode"]
    R.render(S, sc["camera"], frames=6)
    with pytest.raises(cpt.CptError):
        R.read(cpt.CPT_BUF_FILTERED)
    with pytest.raises(cpt.CptError):
        R.set_mlp(cpt.Mlp(1))


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_debug_mode_selects_the_returned_image(cpt, mode):
    """Debug Mode (AdapBilatAtomic.cu:1539-1552) picks the image a Render call
    returns and leaves the estimator alone: the running mean is bit-identical
    to a renderer with Debug Mode 0, and switching the mode back restarts
    nothing."""
    sc = scenes.cornell()
    W, H, F = 48, 40, 8
    S = cpt.Scene(sc)
    kw = {"Min Trace Probability": 0.05}
    base = cpt.Renderer(W, H, mode=mode, **kw).render(S, sc["camera"], frames=F)
    outs = {}
    for dm in (1, 2, 3, 4):
        if dm == 4 and mode == cpt.CPT_MODE_BDPT_ADAPTIVE_EMP:
            continue
        R = cpt.Renderer(W, H, mode=mode, **kw)
        R.set_attrib("Debug Mode", dm)
        out = R.render(S, sc["camera"], frames=F)
        assert np.array_equal(R.read(cpt.CPT_BUF_RESULT).view(np.uint32), base.view(np.uint32))
        err, counts = R.read(cpt.CPT_BUF_ERROR), R.read(cpt.CPT_BUF_SAMPLE_COUNTS)
        want = {1: counts.astype(np.float32), 2: np.maximum(err, 0.0), 3: np.maximum(err, np.float32(0.05))}
        if dm == 4:
            assert np.array_equal(out, R.read(cpt.CPT_BUF_FILTERED))
        else:
            assert np.array_equal(out, np.repeat(want[dm][..., None], 3, axis=2))
        outs[dm] = out
        R.set_attrib("Debug Mode", 0)               # back to the running mean, one more frame, no restart
        nxt = R.render(S, sc["camera"], frames=1)
        assert R.sample_counts().sum() > 0 and np.isfinite(nxt).all()
        ref = cpt.Renderer(W, H, mode=mode, **kw).render(S, sc["camera"], frames=F + 1)
        assert np.array_equal(nxt.view(np.uint32), ref.view(np.uint32))
    assert not np.array_equal(outs[2], base)


def test_debug_mode_checks(cpt):
    sc = scenes.cornell()
    S = cpt.Scene(sc)
    R = cpt.Renderer(16, 16, mode=cpt.CPT_MODE_BDPT_ADAPTIVE_EMP)
    R.set_attrib("Debug Mode", 4)                    # the empirical mode has no filtered image
    with pytest.raises(cpt.CptError, match="Debug Mode 4"):
        R.render(S, sc["camera"], frames=1)
    R.set_attrib("Debug Mode", 5)
    with pytest.raises(cpt.CptError, match="not in 0..4"):
        R.render(S, sc["camera"], frames=1)


def test_proportional_allocation_bit_exact_across_sum_chunks(cpt):
    """480 x 270 pixels: the error sum spans two chunks of 256 block sums
    (k_err_final); counts and offsets from the device's own error map equal
    the oracle's sum and proportional rule bit for bit."""
    print("PLL allocation at 480x270 (two error-sum chunks)", flush=True)
    sc = scenes.cornell()
    S = cpt.Scene(sc)
    W, H = 480, 270
    R = cpt.Renderer(W, H, mode=cpt.CPT_MODE_BDPT_ADAPTIVE_PLL)
    R.render(S, sc["camera"], frames=6, result=False)
    err = R.read(cpt.CPT_BUF_ERROR)
    counts = R.read(cpt.CPT_BUF_SAMPLE_COUNTS)
    assert (W * H + 255) // 256 > 256
    total = np.float32(oracle.tree_sum(err))
    want = oracle.alloc_proportional(err, total, W * H)
    assert np.array_equal(counts.ravel(), np.asarray(want).ravel())

"""Per-pixel error MLP on gfx950 MFMA against the FP64 oracle (oracle/mlp.py).

Tolerances: the device rounds X, W1, W2, H1, dZ2 and dZ1 to bfloat16 (the
oracle's emulated mode rounds the same operands) but accumulates in FP32 and
uses the device tanh; an operand whose value lands within an ulp of a bfloat16
rounding boundary may round the other way.  Outputs agree to 2e-3 absolute
(relative to max(1, |y|)), gradients to 1e-2 in relative L2 norm; a transposed
operand or a wrong MFMA lane mapping gives errors of order 1.
"""
import numpy as np
import pytest

import oracle
from cudapt_amd import scenes
from oracle import mlp as M

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    p = M.glorot_init(seed, oracle.pcg32).astype(np.float64)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(0, 0.2, 64)
    p[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.2, 64)
    p[M.OFF_B3] = 0.3
    return p.astype(np.float32)


def _data(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, M.N_IN)).astype(np.float32)
    t = (np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] - 0.5 * X[:, 8]).astype(np.float32)
    return X, t


def test_init_matches_oracle(cpt):
    m = cpt.Mlp(0xA22)
    assert np.array_equal(m.params(), M.glorot_init(0xA22, oracle.pcg32))


@pytest.mark.parametrize("n", [1, 64, 1000, 65536])
def test_forward_matches_oracle(cpt, n):
    m = cpt.Mlp(1)
    p = _params(7)
    m.set_params(p)
    X, _ = _data(n, n)
    y = m.infer(X)
    yb = M.forward(p, X, exact=False)
    ye = M.forward(p, X, exact=True)
    err = np.abs(y - yb) / np.maximum(1.0, np.abs(yb))
    assert err.max() <= 2e-3, f"max deviation {err.max():.2e} from the bf16-emulated oracle"
    # the emulated oracle is the closer one: bf16 rounding is the dominant error
    assert np.abs(y - yb).mean() < np.abs(y - ye).mean()


@pytest.mark.parametrize("n", [64, 777, 8192, 65536, 100000])
def test_gradient_matches_oracle(cpt, n):
    m = cpt.Mlp(1)
    p = _params(11)
    m.set_params(p)
    X, t = _data(100 + n, n)
    loss = m.train_step(X, t, apply=False)
    g = m.grad()
    l_ref, g_ref = M.loss_grad(p, X, t, exact=False)
    assert abs(loss - l_ref) <= 2e-3 * max(1.0, l_ref)
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    assert rel <= 1e-2, f"gradient relative L2 error {rel:.2e}"
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.N_PARAMS)):
        rb = np.linalg.norm(g[a:b] - g_ref[a:b]) / np.linalg.norm(g_ref[a:b])
        assert rb <= 3e-2, f"block [{a},{b}) relative error {rb:.2e}"
    # exact FP64 gradient: within bf16 distance
    _, g_ex = M.loss_grad(p, X, t, exact=True)
    assert np.linalg.norm(g - g_ex) / np.linalg.norm(g_ex) <= 0.1


def test_adam_steps_match_recurrence(cpt):
    m = cpt.Mlp(1)
    p0 = _params(13)
    m.set_params(p0)
    X, t = _data(5, 2048)
    pm, mm, vm = p0.astype(np.float64), np.zeros(M.N_PARAMS), np.zeros(M.N_PARAMS)
    for step in (1, 2, 3):
        m.train_step(X, t, lr=1e-2, apply=True)
        g = m.grad().astype(np.float64)           # the gradient the step used
        pm, mm, vm = M.adam(pm, mm, vm, g, step, lr=1e-2)
        np.testing.assert_allclose(m.params(), pm, rtol=1e-5, atol=1e-7)


def test_multi_step_call_equals_single_steps(cpt):
    """cpt_mlp_train_steps runs the same kernels in the same order as one
    cpt_mlp_train_step per batch: identical parameters, bit for bit."""
    p0 = _params(17)
    X, t = _data(21, 4 * 4096)
    dx, dt = cpt.DeviceBuffer.from_array(X), cpt.DeviceBuffer.from_array(t)
    a, b = cpt.Mlp(1), cpt.Mlp(1)
    a.set_params(p0)
    b.set_params(p0)
    for k in range(7):                                   # batches 2, 3, 0, 1, 2, 3, 0
        a.train_step_device(dx, dt, 4096, lr=5e-3, want_loss=False, x_offset=((2 + k) % 4) * 4096)
    loss = b.train_steps_device(dx, dt, 4 * 4096, 4096, 7, lr=5e-3, first_batch=2, want_loss=True)
    assert np.array_equal(a.params().view(np.uint32), b.params().view(np.uint32))
    assert np.isfinite(loss)


def test_training_reduces_loss(cpt):
    m = cpt.Mlp(3)
    X, t = _data(9, 8192)
    first = m.train_step(X, t, lr=3e-3)
    last = first
    for _ in range(400):
        last = m.train_step(X, t, lr=3e-3)
    assert np.isfinite(last) and last < 0.2 * first, (first, last)


def test_argument_errors(cpt):
    m = cpt.Mlp(1)
    with pytest.raises(cpt.CptError):
        m.set_params(np.zeros(10, np.float32))
    d = cpt.DeviceBuffer(36 * 64)
    with pytest.raises(cpt.CptError):
        m.train_step_device(d, d, 0)
    with pytest.raises(cpt.CptError):
        m.train_step_device(d, d, 65)
    assert m.infer(np.zeros((0, 9), np.float32)).size == 0


@pytest.fixture(scope="module")
def adaptive_state(cpt):
    sc = scenes.cornell()
    W, H = 64, 48
    S = cpt.Scene(sc)
    R = cpt.Renderer(W, H, mode=cpt.CPT_MODE_BDPT_ADAPTIVE_PLL)
    R.render(S, sc["camera"], frames=8)
    dx = cpt.DeviceBuffer(W * H * 36)
    R.mlp_features(dx)
    x = dx.download(np.float32, W * H * 9).reshape(-1, 9)
    st = {k: R.read(b) for k, b in (("result", cpt.CPT_BUF_RESULT), ("filtered", cpt.CPT_BUF_FILTERED),
                                    ("err", cpt.CPT_BUF_ERROR), ("gpos", cpt.CPT_BUF_GBUF_POS),
                                    ("gnrm", cpt.CPT_BUF_GBUF_NORM), ("gdif", cpt.CPT_BUF_GBUF_DIFF),
                                    ("sq", cpt.CPT_BUF_SECOND_MOMENT), ("counts", cpt.CPT_BUF_SAMPLE_COUNTS))}
    st["N"] = R.sample_counts()
    return dict(sc=sc, S=S, R=R, x=x, st=st, W=W, H=H)


def test_features_match_formulas(adaptive_state):
    a = adaptive_state
    st = a["st"]
    sqf = oracle.bilateral_defined(st["sq"], st["gpos"], st["gnrm"], None, 50.0, 1.0, 0.25, 0.1, 5)
    ref = M.features(st["result"], st["filtered"], st["err"], st["sq"], sqf, st["gpos"], st["gnrm"], st["gdif"],
                     st["N"], a["sc"]["camera"]["pos"])
    # the variance proxies cancel (s - r^2): compared where they are well above rounding
    ok = np.ones(ref.shape, bool)
    for col, (s, r) in ((3, (st["sq"], st["result"])), (5, (sqf, st["filtered"]))):
        s, r = s.reshape(-1, 3).astype(np.float64), r.reshape(-1, 3).astype(np.float64)
        ok[:, col] = (s - r * r).mean(1) > 1e-4 * np.maximum(s.mean(1), 1e-30)
        assert ok[:, col].mean() > 0.5
    np.testing.assert_allclose(a["x"][ok], ref[ok], rtol=1e-5, atol=2e-6)
    # the filter state is the one the next frame allocates from
    np.testing.assert_array_equal(st["filtered"].view(np.uint32),
                                  oracle.bilateral_defined(st["result"], st["gpos"], st["gnrm"], None, 50.0, 1.0,
                                                           0.25, 0.1, 5).view(np.uint32))


@pytest.mark.parametrize("clamp", [0.0, 0.5])
def test_target_matches_formula(cpt, adaptive_state, clamp):
    a = adaptive_state
    rng = np.random.default_rng(4)
    ref = (a["st"]["result"] + rng.normal(0, 0.05, a["st"]["result"].shape)).astype(np.float32)
    d_ref = cpt.DeviceBuffer.from_array(ref)
    d_t = cpt.DeviceBuffer(a["W"] * a["H"] * 4)
    a["R"].mlp_target(d_ref, d_t, clamp)
    t = d_t.download(np.float32, a["W"] * a["H"])
    exp = M.target(a["st"]["result"], ref, a["st"]["N"], clamp)
    np.testing.assert_allclose(t, exp, rtol=1e-5, atol=1e-5)
    if clamp:
        assert (np.abs(t - M.target(a["st"]["result"], ref, a["st"]["N"])) > 1e-3).any()


def test_learned_allocation_follows_model(cpt, adaptive_state):
    a = adaptive_state
    m = cpt.Mlp(5)
    m.set_params(_params(17))
    y = m.infer(a["x"])
    expect = M.learned_error(y, a["st"]["N"], 5).astype(np.float32)
    R = a["R"]
    assert R.attribs()["Learned Signal"][2] == 0
    R.set_mlp(m)
    R.render(a["S"], a["sc"]["camera"], frames=1)
    sig = R.read(cpt.CPT_BUF_ALLOC_ERROR).ravel()
    np.testing.assert_allclose(sig, expect, rtol=1e-5, atol=0)
    counts = R.read(cpt.CPT_BUF_SAMPLE_COUNTS).ravel()
    ref_counts = oracle.alloc_proportional(sig, oracle.tree_sum(sig), a["W"] * a["H"]).ravel()
    assert np.array_equal(counts, ref_counts)
    R.set_mlp(None)


@pytest.mark.parametrize("signal", [1, 2, 3])
def test_expected_error_signal(cpt, signal):
    sc = scenes.cornell()
    S = cpt.Scene(sc)
    R = cpt.Renderer(40, 30, mode=cpt.CPT_MODE_BDPT_ADAPTIVE_PLL)
    R.set_attrib("Learned Signal", signal)
    R.render(S, sc["camera"], frames=6)
    dx = cpt.DeviceBuffer(40 * 30 * 36)
    R.mlp_features(dx)
    m = cpt.Mlp(6)
    y = m.infer(dx.download(np.float32, 40 * 30 * 9).reshape(-1, 9))
    N = R.sample_counts()
    R.set_mlp(m)
    R.render(S, sc["camera"], frames=1)
    sig = R.read(cpt.CPT_BUF_ALLOC_ERROR).ravel()
    plain = R.read(cpt.CPT_BUF_ERROR).ravel()          # the pass's filter residual
    np.testing.assert_allclose(sig, M.learned_error(y, N, 5, signal=signal, plain=plain), rtol=1e-5, atol=0)

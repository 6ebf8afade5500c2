"""Pins of the FP64 MLP restatement (oracle/mlp.py): finite differences,
hand-expanded Adam steps, bfloat16 rounding bit patterns, initialisation from
the PCG32 stream, and the feature/target/allocation-signal formulas."""
import numpy as np
import pytest

import oracle
from oracle import mlp as M


def _problem(seed=3, n=24):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, M.N_PARAMS)
    X = rng.normal(0, 1, (n, M.N_IN))
    t = rng.normal(0, 1, n)
    return p, X, t


def test_layout():
    assert (M.OFF_B1, M.OFF_W2, M.OFF_B2, M.OFF_W3, M.OFF_B3, M.N_PARAMS) == (576, 640, 4736, 4800, 4864, 4865)


def test_gradient_matches_central_differences_everywhere():
    p, X, t = _problem()
    _, g = M.loss_grad(p, X, t, exact=True)
    h = 1e-6
    fd = np.empty_like(g)
    for i in range(M.N_PARAMS):
        q = p.copy()
        q[i] += h
        lp = M.loss(q, X, t)
        q[i] -= 2 * h
        lm = M.loss(q, X, t)
        fd[i] = (lp - lm) / (2 * h)
    scale = np.abs(g).max()
    assert np.abs(fd - g).max() <= 1e-6 * scale
    # every block carries gradient (a dropped term would leave zeros)
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3)):
        assert np.abs(g[a:b]).max() > 1e-4 * scale


def test_gelu_closed_forms():
    z = np.array([0.0, 1.0, -1.0, 30.0, -30.0])
    g, dg = M.gelu(z)
    assert g[0] == 0.0 and dg[0] == 0.5
    assert abs(g[3] - 30.0) < 1e-12 and abs(g[4]) < 1e-12
    assert abs(dg[3] - 1.0) < 1e-12 and abs(dg[4]) < 1e-12
    th = np.tanh(0.7978845608028654 * 1.044715)
    assert abs(g[1] - 0.5 * (1 + th)) < 1e-15
    # derivative by central differences
    zz = np.linspace(-4, 4, 41)
    fd = (M.gelu(zz + 1e-6)[0] - M.gelu(zz - 1e-6)[0]) / 2e-6
    assert np.abs(fd - M.gelu(zz)[1]).max() < 1e-8


def test_adam_first_two_steps_hand_expanded():
    g1 = np.array([0.5, -2.0, 1e-3, 0.0])
    g2 = np.array([-0.25, 1.0, 1e-3, 4.0])
    p0 = np.array([1.0, -1.0, 0.5, 0.0])
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    p1, m, v = M.adam(p0, np.zeros(4), np.zeros(4), g1, 1, lr)
    # step 1: m_hat = g, v_hat = g^2, so the step is lr * g / (|g| + eps)
    np.testing.assert_allclose(p1, p0 - lr * g1 / (np.abs(g1) + eps), rtol=0, atol=1e-15)
    p2, m, v = M.adam(p1, m, v, g2, 2, lr)
    m2 = 0.9 * 0.1 * g1 + 0.1 * g2
    v2 = 0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2
    exp = p1 - lr * (m2 / (1 - b1 ** 2)) / (np.sqrt(v2 / (1 - b2 ** 2)) + eps)
    np.testing.assert_allclose(p2, exp, rtol=0, atol=1e-15)


def test_bf16_round_to_nearest_even_bit_patterns():
    u = np.array([0x3F800000,   # 1.0 exact
                  0x3F808000,   # 1 + 2^-8: tie, even mantissa stays
                  0x3F818000,   # tie, odd mantissa rounds up
                  0x3F808001,   # just above the tie rounds up
                  0x3F807FFF,   # just below the tie rounds down
                  0x3FFFFFFF,   # carries into the exponent: 2.0
                  0xBF818000,   # sign kept
                  0x00000000], np.uint32)
    out = M.bf16(u.view(np.float32)).view(np.uint32)
    exp = np.array([0x3F800000, 0x3F800000, 0x3F820000, 0x3F810000, 0x3F800000, 0x40000000, 0xBF820000, 0],
                   np.uint32)
    assert np.array_equal(out, exp)


def test_bf16_emulation_differs_from_exact_at_bf16_scale():
    p, X, t = _problem(5, 64)
    ye, yb = M.forward(p, X, True), M.forward(p, X, False)
    d = np.abs(ye - yb).max() / np.abs(ye).max()
    assert 1e-5 < d < 5e-2


def test_glorot_init_from_pcg32_stream():
    p = M.glorot_init(0xA22, oracle.pcg32)
    W1, b1, W2, b2, w3, b3 = M.unpack(p)
    assert (b1 == 0).all() and (b2 == 0).all() and b3 == 0
    a1 = np.sqrt(6 / 73)
    assert np.abs(W1).max() <= a1 and np.abs(W1).max() > 0.9 * a1
    # first entry from the first stream value by hand
    x0 = int(oracle.pcg32(0xA22, 0, 1)[0])
    u0 = np.float32(((x0 >> 8) + 1) / 16777216.0)
    assert p[0] == (np.float32(2) * u0 - np.float32(1)) * np.float32(np.sqrt(np.float32(6) / np.float32(73)))
    assert abs(W1.mean()) < 0.05 and abs(W2.mean()) < 0.02


def test_feature_target_and_signal_formulas():
    r = np.array([[1.0, 1.0, 1.0], [0.0, 0.0, 0.0]])
    f = np.array([[np.e - 1, np.e - 1, np.e - 1], [0.0, 0.0, 0.0]])
    e = np.array([1.0, 0.0])
    a = np.array([[0.5, 0.5, 0.5], [1.0, 0.0, 0.0]])
    n = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, -1.0]])
    pos = np.array([[3.0, 4.0, 0.0], [0.0, 0.0, 0.0]])
    N = np.array([np.e ** 5, 0.0])
    sq = np.array([[1.0 + np.e ** 10, 1.0 + np.e ** 10, 1.0 + np.e ** 10], [0.0, 0.0, 0.0]])
    sqf = np.array([[(np.e - 1) ** 2 + np.e ** 20] * 3, [0.0, 0.0, 0.0]])
    x = M.features(r, f, e, sq, sqf, pos, n, a, N, (0.0, 0.0, 0.0))
    np.testing.assert_allclose(x[0], [np.log(2), 1.0, 0.0, 1.0, 0.5, 2.0, np.log(6) * 0.2, 1.0, 1.0], atol=1e-12)
    np.testing.assert_allclose(x[1, :2], [0, 0])
    assert abs(x[1, 2] - np.log(1e-12) * 0.1) < 1e-12 and abs(x[1, 3] - np.log(1e-12) * 0.1) < 1e-12
    assert x[1, 4] == 0.2126 and x[1, 7] == 0.0 and x[1, 8] == 0.0   # N = 0 counts as 1
    t = M.target(np.array([[1.0, 2.0, 3.0]]), np.array([[0.0, 2.0, 0.0]]), np.array([3]))
    assert abs(t[0] - np.log(10.0 + 1e-12)) < 1e-12
    t = M.target(np.array([[1.0, 2.0, 3.0]]), np.array([[0.0, 2.0, 0.0]]), np.array([3]), clamp=1.5)
    assert abs(t[0] - np.log(1.0 + 1.5 ** 2 + 1e-12)) < 1e-12   # r -> (1, 1.5, 1.5), ref -> (0, 1.5, 0)
    s = M.learned_error(np.array([0.0, np.log(8.0)]), np.array([5, 2]), 5)
    np.testing.assert_allclose(s, [0.5 / 5, 5 / 7 * 8 / 2])
    np.testing.assert_allclose(M.learned_error(np.array([0.0, np.log(8.0)]), np.array([5, 2]), 5, signal=1),
                               [1 / 5, 8 / 2])
    np.testing.assert_allclose(M.learned_error(np.array([0.0, np.log(8.0)]), np.array([5, 2]), 5, signal=2),
                               [1 / 25, 8 / 4])


def test_learned_error_geometric_mean_signal():
    """Signal 3 is the geometric mean of exp(y) / N and the mode's own error map:
    equal inputs give that value back, and it scales as the square root of each."""
    rng = np.random.default_rng(3)
    y = rng.normal(-3.0, 1.0, 64)
    N = rng.integers(1, 50, 64)
    v = M.learned_error(y, N, 5, signal=1)
    assert np.allclose(M.learned_error(y, N, 5, signal=3, plain=v), v, rtol=1e-12)
    plain = rng.uniform(0.0, 0.1, 64)
    assert np.allclose(M.learned_error(y, N, 5, signal=3, plain=4 * plain),
                       2 * M.learned_error(y, N, 5, signal=3, plain=plain), rtol=1e-12)
    assert (M.learned_error(y, N, 5, signal=3, plain=-plain) == 0).all()

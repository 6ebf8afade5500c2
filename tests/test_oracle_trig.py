"""Pins of the oracle's FP32 trigonometry, the operation sequence the device
library shares (devmath.h): sin / cos of the sampling angles 2 pi u by Cephes'
(S. L. Moshier) octant reduction and polynomials, atan2 by Cephes' atanf core
and its octants, acos(y) = atan2(sqrt((1 - y)(1 + y)), y).  Checked against
float64 libm: a swapped quadrant, a missing pi/4 shift or a wrong polynomial
coefficient lands far outside these bounds."""
import numpy as np

import oracle


def _ulps(v, t):
    u = np.spacing(np.abs(t).astype(np.float32)).astype(np.float64)
    return np.abs(v.astype(np.float64) - t) / u


def test_sincos_on_the_sampling_range():
    rng = np.random.default_rng(11)
    x = rng.uniform(0.0, 2.0 * np.pi, 200000).astype(np.float32)
    s, c = oracle.sincos_def(x)
    xd = x.astype(np.float64)
    assert _ulps(s, np.sin(xd)).max() <= 2.0
    assert _ulps(c, np.cos(xd)).max() <= 2.0
    np.testing.assert_allclose(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2, 1.0, atol=4e-7)


def test_sincos_octant_points():
    # the octant boundaries and zeros: absolute error below 1e-7, signs right
    k = np.arange(0, 9)
    x = (k * np.pi / 4).astype(np.float32)
    s, c = oracle.sincos_def(x)
    xd = x.astype(np.float64)
    np.testing.assert_allclose(s, np.sin(xd), atol=1e-7)
    np.testing.assert_allclose(c, np.cos(xd), atol=1e-7)
    assert s[0] == 0.0 and c[0] == 1.0
    assert s[2] > 0.99 and s[6] < -0.99 and c[4] < -0.99


def test_atan2_all_quadrants():
    rng = np.random.default_rng(12)
    y = rng.normal(size=200000).astype(np.float32)
    x = rng.normal(size=200000).astype(np.float32)
    a = oracle.atan2_def(y, x)
    t = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.abs(a - t).max() <= 4e-7
    axes = oracle.atan2_def(np.float32([0, 1, 0, -1, 1, -1]), np.float32([1, 0, -1, 0, 1, -1]))
    np.testing.assert_allclose(axes, [0, np.pi / 2, np.pi, -np.pi / 2, np.pi / 4, -3 * np.pi / 4], atol=3e-7)


def test_acos_range():
    rng = np.random.default_rng(13)
    y = np.concatenate([rng.uniform(-1, 1, 200000), [-1.0, 1.0, 0.0, 0.5, -0.5]]).astype(np.float32)
    a = oracle.acos_def(y)
    assert np.abs(a - np.arccos(y.astype(np.float64))).max() <= 4e-7
    assert oracle.acos_def(np.float32([1.0]))[0] == 0.0

"""FP64 restatement of the per-pixel error MLP (test infrastructure only).

Imported by tests/ and the cpu_baseline leg of bench.py, never by the product.
The model has no counterpart in the reference (CudaPT links FANN but calls no
fann_* function), so parity is pinned by mathematics, not by reference output:

* ``loss_grad(exact=True)``: analytic FP64 backward, pinned against central
  finite differences of ``loss`` over every parameter (tests/test_oracle_mlp.py);
* ``adam``: one Adam step (Kingma & Ba 2015, Algorithm 1), pinned against the
  hand-expanded first and second steps;
* ``bf16``: round-to-nearest-even to bfloat16, pinned against hand-worked
  bit patterns (ties, carries into the exponent, signed values);
* ``forward/loss_grad(exact=False)``: the same network with operands rounded to
  bfloat16 where the device kernel rounds them (X, W1, W2, H1, dZ2, dZ1) and all
  sums in FP64; the device result (FP32 accumulation, the GELU through the
  hardware exponential and reciprocal as x sigmoid(2u)) is compared
  with it under stated tolerances.

Network: 9 -> 64 -> 64 -> 1, GELU (tanh form, Hendrycks & Gimpel 2016) after
both hidden layers, loss = mean (y - t)^2.  Flat parameter layout, row-major
with the input index first: W1[9][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 9, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID
N_PARAMS = OFF_B3 + 1
assert N_PARAMS == 4865

_C = 0.7978845608028654   # sqrt(2 / pi)
_A = 0.044715


def bf16(a) -> np.ndarray:
    """Rounds FP32 values to bfloat16 (nearest, ties to even), returned as FP32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(N_IN, N_HID)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(N_HID, N_HID)
    b2 = p[OFF_B2:OFF_W3]
    w3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, b1, W2, b2, w3, b3


def gelu(z):
    th = np.tanh(_C * (z + _A * z ** 3))
    g = 0.5 * z * (1.0 + th)
    dg = 0.5 * (1.0 + th) + 0.5 * z * (1.0 - th * th) * _C * (1.0 + 3.0 * _A * z * z)
    return g, dg


def _r(a, exact):
    """Operand rounding of the device kernel (identity in exact mode)."""
    if exact:
        return np.asarray(a, np.float64)
    return bf16(np.asarray(a, np.float32)).astype(np.float64)


def _forward(p, X, exact):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    Xr = _r(X, exact)
    z1 = Xr @ _r(W1, exact) + b1
    h1, dg1 = gelu(z1)
    h1r = _r(h1, exact)
    z2 = h1r @ _r(W2, exact) + b2
    h2, dg2 = gelu(z2)
    y = h2 @ w3 + b3
    return y, (Xr, dg1, h1r, dg2, h2)


def forward(p, X, exact: bool = True) -> np.ndarray:
    return _forward(p, X, exact)[0]


def loss(p, X, t, exact: bool = True) -> float:
    y = forward(p, X, exact)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def loss_grad(p, X, t, exact: bool = True):
    """(loss, flat gradient) of mean (y - t)^2."""
    y, (Xr, dg1, h1r, dg2, h2) = _forward(p, X, exact)
    W1, b1, W2, b2, w3, b3 = unpack(p)
    n = len(y)
    d = y - np.asarray(t, np.float64)
    dy = 2.0 * d / n
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    dz2 = dy[:, None] * w3[None, :] * dg2
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz2r = _r(dz2, exact)
    g[OFF_W2:OFF_B2] = (h1r.T @ dz2r).ravel()
    dh1 = dz2r @ _r(W2, exact).T
    dz1 = dh1 * dg1
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W1:OFF_B1] = (Xr.T @ _r(dz1, exact)).ravel()
    return float(np.mean(d * d)), g


def adam(p, m, v, g, step: int, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
    """Adam step `step` (1-based): returns (p, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def glorot_init(seed: int, pcg32) -> np.ndarray:
    """Initial parameters of cpt_mlp_create(seed): Glorot-uniform weights from
    the PCG32 stream (initstate = seed, initseq = 0), W1 then W2 then w3, each
    entry (2u - 1) * sqrt(6 / (fan_in + fan_out)) in FP32, u = ((x >> 8) + 1) / 2^24;
    zero biases.  `pcg32(initstate, initseq, n)` supplies the raw stream."""
    n = N_IN * N_HID + N_HID * N_HID + N_HID
    x = pcg32(seed, 0, n)
    u = ((x >> 8) + 1).astype(np.float32) * np.float32(1.0 / 16777216.0)
    p = np.zeros(N_PARAMS, np.float32)
    k = 0
    for off, fin, fout in ((OFF_W1, N_IN, N_HID), (OFF_W2, N_HID, N_HID), (OFF_W3, N_HID, 1)):
        a = np.sqrt(np.float32(6.0) / np.float32(fin + fout), dtype=np.float32)
        cnt = fin * fout
        p[off:off + cnt] = (np.float32(2.0) * u[k:k + cnt] - np.float32(1.0)) * a
        k += cnt
    return p


def luminance(c):
    c = np.asarray(c, np.float32)
    return (np.float32(0.2126) * c[..., 0] + np.float32(0.7152) * c[..., 1]) + np.float32(0.0722) * c[..., 2]


def features(result, filtered, err, sq, sq_f, gpos, gnrm, gdif, sample_n, cam_pos) -> np.ndarray:
    """The 9 per-pixel inputs of the model (FP64 evaluation of the defined formulas)."""
    r = np.asarray(result, np.float64).reshape(-1, 3)
    f = np.asarray(filtered, np.float64).reshape(-1, 3)
    e = np.asarray(err, np.float64).ravel()
    s = np.asarray(sq, np.float64).reshape(-1, 3)
    a = np.asarray(gdif, np.float64).reshape(-1, 3)
    sf = np.asarray(sq_f, np.float64).reshape(-1, 3)
    nr = np.asarray(gnrm, np.float64).reshape(-1, 3)
    ps = np.asarray(gpos, np.float64).reshape(-1, 3)
    N = np.asarray(sample_n, np.float64).ravel()
    w = np.array([0.2126, 0.7152, 0.0722])
    v = np.maximum((s - r * r).mean(1), 0.0)
    x = np.empty((len(e), N_IN))
    x[:, 0] = np.log1p(r @ w)
    x[:, 1] = np.log1p(f @ w)
    x[:, 2] = np.log(e + 1e-12) * 0.1
    x[:, 3] = np.log(v + 1e-12) * 0.1
    x[:, 4] = a @ w
    x[:, 5] = np.log(np.maximum((sf - f * f).mean(1), 0.0) + 1e-12) * 0.1
    x[:, 6] = np.log1p(np.linalg.norm(ps - np.asarray(cam_pos, np.float64), axis=1)) * 0.2
    x[:, 7] = np.log(np.maximum(N, 1.0)) * 0.2
    x[:, 8] = nr[:, 1]
    return x


def target(result, ref, sample_n, clamp: float = 0.0) -> np.ndarray:
    """log(N * |r - ref|^2 / 3 + 1e-12): log per-sample variance estimate
    (r and ref clamped to `clamp` per channel when clamp > 0)."""
    c = clamp if clamp > 0 else np.inf
    d = (np.minimum(np.asarray(result, np.float64).reshape(-1, 3), c) -
         np.minimum(np.asarray(ref, np.float64).reshape(-1, 3), c))
    return np.log(np.asarray(sample_n, np.float64).ravel() * (d * d).sum(1) / 3.0 + 1e-12)


def learned_error(y, sample_n, nd: int, signal: int = 0, plain=None) -> np.ndarray:
    """Allocation signal nd / (nd + N) * exp(y) / N (signal 0), exp(y) / N (1),
    exp(y) / N^2 (2) or sqrt(exp(y) / N * plain) (3, plain = the mode's own
    error map)."""
    N = np.maximum(np.asarray(sample_n, np.float64).ravel(), 1.0)
    v = np.exp(np.clip(np.asarray(y, np.float64), -60, 60)) / N
    if signal == 3:
        return np.sqrt(v * np.maximum(np.asarray(plain, np.float64).ravel(), 0.0))
    return v / N if signal == 2 else v if signal else nd / (nd + N) * v

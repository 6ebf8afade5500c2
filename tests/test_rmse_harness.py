"""RMSE harness (cudapt_amd/rmse.py): the reference's single-precision
sequential RMSE, the capture modes and the "sec, frames, rmse" file."""
import math

import numpy as np
import pytest

from cudapt_amd import rmse as R


def test_rmse_closed_forms():
    a = np.zeros((2, 3, 3), np.float32)
    b = np.full((2, 3, 3), 2.0, np.float32)
    assert R.rmse(a, b) == 2.0
    b[0, 0, 0] = 0.0
    assert abs(R.rmse(a, b) - math.sqrt(4 * 17 / 18)) < 1e-6
    assert R.rmse(a, a) == 0.0
    with pytest.raises(ValueError):
        R.rmse(a, np.zeros(5, np.float32))


def test_rmse_sums_in_float_in_order():
    """1e8 + many ones: a float accumulator in index order loses the ones
    (a pairwise or double sum would not)."""
    d = np.ones(1 << 16, np.float32)
    d[0] = 1e4                                       # (1e4)^2 = 1e8 first, then 65535 ones
    got = R.rmse(d, np.zeros_like(d))
    seq = np.float32(1e8)
    for _ in range(3):                                # each +1 is below half an ulp of 1e8 (4.0)
        seq = np.float32(seq + np.float32(1.0))
    assert seq == np.float32(1e8)
    assert got == float(np.sqrt(np.float32(1e8) / np.float32(d.size), dtype=np.float32))
    assert got != float(np.sqrt((1e8 + 65535) / d.size))


class _FakeRenderer:
    """Image converging as 1/frames towards 1: RMSE = 1/f after f frames."""
    def __init__(self, shape, dt=0.01):
        self.f, self.shape, self.dt = 0, shape, dt

    def render(self, scene, camera, frames=1):
        import time
        time.sleep(self.dt * frames)
        self.f += frames
        return np.full(self.shape, 1.0 + 1.0 / self.f, np.float32)


def test_capture_modes():
    conv = np.ones((4, 4, 3), np.float32)
    h = R.RmseHarness(_FakeRenderer(conv.shape), None, {}, conv)
    s = h.capture(frames=4)
    assert s.frames == 4 and abs(s.rmse - 0.25) < 1e-6
    h = R.RmseHarness(_FakeRenderer(conv.shape), None, {}, conv)
    s = h.capture(limit=0.1)
    assert s.frames == 10 and s.rmse <= 0.1
    h = R.RmseHarness(_FakeRenderer(conv.shape), None, {}, conv)
    s = h.capture(seconds=0.05)
    assert s.sec >= 0.05 and s.frames >= 5
    with pytest.raises(ValueError):
        h.capture()


def test_collect_and_csv_round_trip(tmp_path):
    conv = np.ones((4, 4, 3), np.float32)
    h = R.RmseHarness(_FakeRenderer(conv.shape, dt=0.005), None, {}, conv, frames_per_step=2)
    rows = h.collect(interval=0.02, count=3)
    assert len(rows) == 3
    assert all(b.sec > a.sec and b.frames > a.frames and b.rmse < a.rmse for a, b in zip(rows, rows[1:]))
    assert all(r.frames % 2 == 0 for r in rows)
    p = tmp_path / "rmse.txt"
    h.write_csv(str(p))
    lines = p.read_text().splitlines()
    assert len(lines) == 3 and all(len(l.split(", ")) == 3 for l in lines)
    back = R.read_csv(str(p))
    assert [(r.frames, round(r.rmse, 6)) for r in back] == [(r.frames, round(r.rmse, 6)) for r in rows]

"""Host-side pieces of the learned-vs-plain adaptive protocol (cudapt_amd/learned.py)."""
import math

import numpy as np

from cudapt_amd import learned


def test_time_to_reach_interpolates_log_mse():
    ts = np.array([1.0, 2.0, 3.0])
    ms = np.array([1.0, 0.1, 0.01])
    assert learned.time_to_reach(ts, ms, 1.0) == 1.0
    assert abs(learned.time_to_reach(ts, ms, 0.1) - 2.0) < 1e-12
    assert abs(learned.time_to_reach(ts, ms, math.sqrt(0.1)) - 1.5) < 1e-12
    assert learned.time_to_reach(ts, ms, 1e-3) is None


def test_mse_clamps_both_images():
    a = np.array([[[5.0, 0.5, 0.0]]])
    b = np.array([[[1.0, 0.0, 0.0]]])
    assert learned.mse(a, b, clamp=1.0) == 0.25 / 3
    assert learned.mse(a, b, clamp=0) == (16 + 0.25) / 3


def test_poses_are_seeded_and_aim_at_the_ground():
    a = learned.sphere_grid_poses(4)
    b = learned.sphere_grid_poses(4)
    assert a == b
    for c in a:
        d = np.asarray(c["dir"])
        assert abs(np.linalg.norm(d) - 1) < 1e-12 and d[1] < 0 and c["pos"][1] > 0

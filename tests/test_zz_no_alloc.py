"""cpt_filter_cross_bilateral_device_noalloc against cpt_filter_cross_bilateral_device.

The no-allocation entry point runs the same filter kernel with the spatial
table staged in the library's static device scratch instead of a per-call
allocation, so both must return the same bits.  Two calls in a row, the second
on a smaller image with a smaller radius and another colour sigma: the second
call must not see anything of the first call's table.  A radius beyond the
scratch table's 63 taps is refused.
"""
import numpy as np
import pytest

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


CASES = [  # (seed, W, H, radius, sigmas, albedo term)
    (11, 96, 72, 9, (0.7, 0.4, 0.25, 0.1), True),
    (12, 41, 23, 3, (0.3, 1.0, 0.5, 0.2), False),
]


def test_noalloc_equals_allocating_entry_point(cpt):
    print("no-alloc filter: two calls, second smaller", flush=True)
    for seed, W, H, radius, sig, use_diff in CASES:
        col, pos, nrm, dif, N = _images(seed, H, W)
        d = dif if use_diff else None
        f_na, e_na = cpt.filter_cross_bilateral(col, pos, nrm, d, N, *sig, radius, nd=5, noalloc=True)
        f_ref, e_ref = cpt.filter_cross_bilateral(col, pos, nrm, d, N, *sig, radius, nd=5)
        assert np.array_equal(f_na.view(np.uint32), f_ref.view(np.uint32)), (seed, W, H, radius)
        assert np.array_equal(e_na.view(np.uint32), e_ref.view(np.uint32)), (seed, W, H, radius)


def test_noalloc_refuses_radius_beyond_scratch(cpt):
    col, pos, nrm, dif, N = _images(13, 8, 8)
    with pytest.raises(cpt.CptError, match="scratch table"):
        cpt.filter_cross_bilateral(col, pos, nrm, None, N, 0.5, 0.4, 0.25, 0.1, 32, noalloc=True)

"""The oracle's restatement of the BVH debug renderer's depth queries
(TraceDepth / TraceDepthParent, src/cudaRTCommon.cu:129-231), pinned by cases
worked by hand on a two-leaf tree: root 0 over [0, 11] x [0, 1] x {0}, left
leaf 1 over [0, 1] x [0, 1] (triangle 0), right leaf 2 over [10, 11] x [0, 1]
(triangle 1).  The walk pops the right child first and reports, for
TraceDepth, stack pointer + 2 at a leaf box entered and at least stack
pointer + 1 at each box missed."""
import numpy as np
import pytest

import oracle
from cudapt_amd import scenes

POS = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [10, 0, 0], [11, 0, 0], [10, 1, 0]])
IDX = np.uint32([[0, 1, 2], [3, 4, 5]])


@pytest.fixture(scope="module")
def two_leaves():
    O = oracle.Scene(scenes.mesh_scene(POS, IDX))
    bounds, off, order = O.bvh()
    assert off.tolist() == [[2, 0], [0, 1], [1, 1]]       # the tree the cases assume
    return O


def _down(x, y):
    return np.float32([[x, y, 5.0, 1e20, 0.0, 0.0, -1.0, 1e-4]])


# (x, y): TraceDepth {hit, depth, leaf}
DEPTH_CASES = {
    (0.5, 0.5): (1, 1, 1),     # right leaf missed at sp 0 (depth 1), left leaf entered at sp -1 (-1 + 2)
    (10.5, 0.5): (1, 2, 1),    # right leaf entered at sp 0 (0 + 2), left missed (max(2, 0))
    (5.0, 0.5): (1, 1, 0),     # root entered, both leaves missed: max(1, 0)
    (5.0, 5.0): (0, 0, 0),     # root missed at sp -1: depth 0, no hit
}


@pytest.mark.parametrize("xy", list(DEPTH_CASES))
def test_trace_depth_hand_cases(two_leaves, xy):
    assert tuple(two_leaves.trace_depth(_down(*xy))[0]) == DEPTH_CASES[xy]


# (x, y, spec): TraceDepthParent {hit, depth, parent}
NONE = 0xFFFFFFFF
PARENT_CASES = {
    (10.5, 0.5, 0): (1, 0, 0),      # the root itself, parent slot 0
    (10.5, 0.5, 1): (1, 1, 0),      # right leaf first, parent the root
    (0.5, 0.5, 1): (1, 1, 0),       # right missed, left entered
    (5.0, 0.5, 1): (1, NONE, 0),    # root entered, no box at depth 1
    (10.5, 0.5, 2): (1, NONE, 0),   # leaves do not push: depth 2 never exists
    (5.0, 5.0, 0): (0, NONE, 0),    # nothing entered
}


@pytest.mark.parametrize("case", list(PARENT_CASES))
def test_trace_depth_parent_hand_cases(two_leaves, case):
    x, y, spec = case
    assert tuple(two_leaves.trace_depth(_down(x, y), spec)[0]) == PARENT_CASES[case]


def test_max_distance_bounds_the_walk(two_leaves):
    r = _down(10.5, 0.5)
    r[0, 3] = 4.0                   # boxes are entered at t = 5
    # the root counts as missed at stack pointer -1: depth max(0, 0), no hit
    assert tuple(two_leaves.trace_depth(r)[0]) == (0, 0, 0)
    assert tuple(two_leaves.trace_depth(r, 0)[0]) == (0, NONE, 0)

"""Traversal-cost heat maps (the BVH debug renderer's TraceCost mode) over the
device traversal: per-ray counts agree with the aggregate counters, rays that
miss the scene's box visit only the root, and the device tree's 4-wide nodes
take fewer visits than the reference's binary tree.  Depth maps (its "Normal"
and "BVHDepth" modes) equal the oracle's restatement for every ray."""
import numpy as np
import pytest

import oracle
from cudapt_amd import heatmap, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grid(cpt):
    sc = scenes.sphere_grid(grid=16)
    return sc, cpt.Scene(sc)


@pytest.mark.parametrize("layout", [0, 1, 2])
def test_per_ray_costs_sum_to_the_counters(grid, layout):
    sc, S = grid
    S.counters(reset=True)
    cost, hit = heatmap.cost_map(S, sc["camera"], 96, 64, layout)
    S.counters(reset=True)
    c2 = S.trace_cost(heatmap.primary_rays(sc["camera"], 96, 64), layout=layout)
    c = S.counters(reset=True)
    assert c["nodes"] == int(c2[:, 0].sum()) and c["tris"] == int(c2[:, 1].sum())
    assert np.array_equal(c2.reshape(64, 96, 2), cost)
    assert (cost[..., 0] >= 1).all()
    assert (cost[hit][:, 1] >= 1).all()           # a hit needs a triangle test


def test_rays_missing_the_scene_visit_the_root_only(grid):
    sc, S = grid
    up = dict(sc["camera"], pos=(0.0, 500.0, 0.0), dir=(0.0, 1.0, 0.0), up=(0.0, 0.0, 1.0))
    for layout in (0, 2):
        cost, hit = heatmap.cost_map(S, up, 16, 16, layout)
        assert not hit.any() and (cost[..., 0] == 1).all() and (cost[..., 1] == 0).all()


def test_device_tree_is_cheaper_than_reference_tree(grid):
    sc, S = grid
    ref, _ = heatmap.cost_map(S, sc["camera"], 96, 64, 1)
    dev, _ = heatmap.cost_map(S, sc["camera"], 96, 64, 2)
    # 4-wide nodes of the device tree against binary nodes of the reference tree
    assert dev[..., 0].sum() < ref[..., 0].sum()
    rgb = heatmap.to_rgb(dev)
    assert rgb.shape == (64, 96, 3) and np.isfinite(rgb).all() and rgb.max() > 0


@pytest.mark.parametrize("name", ["cornell", "grid"])
@pytest.mark.parametrize("spec", [None, 0, 1, 3, 7, 12, 40])
def test_depth_maps_equal_the_oracle(grid, name, spec):
    sc, S = grid if name == "grid" else (scenes.cornell(), None)
    if S is None:
        import cudapt_amd
        S = cudapt_amd.Scene(sc)
    O = oracle.Scene(sc)
    rays = heatmap.primary_rays(sc["camera"], 96, 64)
    g = S.trace_depth(rays, spec)
    o = O.trace_depth(rays, spec)
    assert np.array_equal(g, o)
    assert g[:, 0].any()
    if spec is None:
        assert (g[:, 1] > 0).any() and g[:, 2].any()
    elif spec <= 7:
        assert (g[:, 1] == spec).any()


def test_depth_map_images(grid):
    sc, S = grid
    depth, leaf, hit = heatmap.depth_map(S, sc["camera"], 48, 32)
    img = heatmap.depth_rgb(depth, leaf)
    assert img.shape == (32, 48, 3) and np.isfinite(img).all() and img[..., 1].max() > 0
    d, parent, hit = heatmap.depth_map(S, sc["camera"], 48, 32, spec_depth=4)
    img2 = heatmap.parent_rgb(d, parent, hit, 4)
    assert img2[..., 1].max() > 0 and len(np.unique(parent[d == 4])) > 1

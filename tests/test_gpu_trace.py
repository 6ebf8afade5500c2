"""First-hit parity on the GPU: the HIP traversal against the oracle.

The product promises the tie-defined closest hit (minimum of t, then original
triangle id, over every triangle at 0 <= t < maxDist, each tested with
IntersectTri's arithmetic, src/cudaRTCommon.cu:42-64).  Records (triangle, t,
u, v) must match the oracle bit for bit.
"""
import numpy as np
import pytest

import oracle
from cudapt_amd import scenes
from tests.helpers import camera_rays, hits_equal, random_rays, vertex_rays

pytestmark = pytest.mark.gpu


def _gpu_hits(cpt, S, rays, layout=0):
    dr = cpt.DeviceBuffer.from_array(np.ascontiguousarray(rays, np.float32))
    dh = cpt.DeviceBuffer(len(rays) * 16)
    S.trace_device(dr.ptr, len(rays), dh.ptr, layout=layout)
    return dh.download(oracle.HIT_DTYPE, len(rays))


@pytest.fixture(scope="module")
def cornell(cpt):
    sc = scenes.cornell()
    return sc, cpt.Scene(sc), oracle.Scene(sc)


def test_cornell_primary_rays_bit_exact(cornell):
    sc, S, O = cornell
    rays = camera_rays(sc["camera"], 256, 256)
    g = S.trace(rays)
    ref, _ = O.trace(rays, "brute")
    eq = hits_equal(g, ref)
    assert eq.all(), f"{(~eq).sum()} of {len(eq)} primary hits differ"
    assert (g["tri"] != 0xFFFFFFFF).mean() > 0.95


@pytest.mark.parametrize("seed", [1, 2])
def test_cornell_random_rays_bit_exact(cornell, seed):
    sc, S, O = cornell
    rays = random_rays(40000, [-0.99, 0.01, -0.99], [0.99, 1.99, 0.99], seed=seed)
    g = S.trace(rays)
    ref, _ = O.trace(rays, "brute")
    eq = hits_equal(g, ref)
    assert eq.all(), f"{(~eq).sum()} hits differ"


def test_cornell_edge_and_vertex_rays(cornell):
    sc, S, O = cornell
    rays = vertex_rays(sc, 20000, seed=3)
    g = S.trace(rays)
    ref, _ = O.trace(rays, "brute")
    assert hits_equal(g, ref).all()


def test_backface_cull_and_max_dist(cornell):
    sc, S, O = cornell
    rays = random_rays(20000, [-0.9, 0.1, -0.9], [0.9, 1.9, 0.9], seed=4)
    rays[:, 3] = np.random.default_rng(5).uniform(0.05, 2.0, len(rays)).astype(np.float32)
    for cull in (False, True):
        g = S.trace(rays, cull=cull)
        ref, _ = O.trace(rays, "brute", cull=cull)
        assert hits_equal(g, ref).all()


def test_layouts_agree(cpt, cornell):
    sc, S, O = cornell
    rays = random_rays(30000, [-0.9, 0.1, -0.9], [0.9, 1.9, 0.9], seed=6)
    a = _gpu_hits(cpt, S, rays, layout=0)
    b = _gpu_hits(cpt, S, rays, layout=1)
    assert hits_equal(a, b).all()


@pytest.fixture(scope="module")
def sphere_grid(cpt):
    sc = scenes.sphere_grid()
    return sc, cpt.Scene(sc), oracle.Scene(sc)


def test_sphere_grid_bit_exact_brute(sphere_grid):
    sc, S, O = sphere_grid
    cam = sc["camera"]
    rays = camera_rays(cam, 64, 36)
    rays = np.concatenate([rays, vertex_rays(sc, 1500, seed=7)])
    g = S.trace(rays)
    ref, _ = O.trace(rays, "brute")
    assert hits_equal(g, ref).all()


def test_sphere_grid_bit_exact_bvh(sphere_grid):
    sc, S, O = sphere_grid
    rays = np.concatenate([camera_rays(sc["camera"], 320, 180, jitter=True, seed=8),
                           random_rays(60000, [-64, 0.0, -64], [64, 6.0, 64], seed=9)])
    g = S.trace(rays)
    ref, _ = O.trace(rays, "defined")
    eq = hits_equal(g, ref)
    assert eq.all(), f"{(~eq).sum()} hits differ"


def test_occlusion_matches_closest(cpt, sphere_grid):
    sc, S, O = sphere_grid
    rays = random_rays(50000, [-64, 0.0, -64], [64, 6.0, 64], seed=10)
    rays[:, 3] = np.random.default_rng(11).uniform(0.1, 20.0, len(rays)).astype(np.float32)
    ref, _ = O.trace(rays, "defined")
    dr = cpt.DeviceBuffer.from_array(rays)
    occ = cpt.DeviceBuffer(len(rays) * 4)
    S.occluded_device(dr.ptr, len(rays), occ.ptr)
    assert np.array_equal(occ.download(np.uint32, len(rays)) != 0, ref["tri"] != 0xFFFFFFFF)


def test_empty_batch(cornell):
    sc, S, O = cornell
    out = S.trace(np.zeros((0, 8), np.float32))
    assert len(out) == 0


def test_quantized_nodes_cornell_bit_exact(cpt, cornell):
    """64-byte quantized nodes: conservative boxes, the same records."""
    sc, S, O = cornell
    rays = np.concatenate([camera_rays(sc["camera"], 128, 128), vertex_rays(sc, 20000, seed=12),
                           random_rays(30000, [-0.99, 0.01, -0.99], [0.99, 1.99, 0.99], seed=13)])
    ref, _ = O.trace(rays, "brute")
    q = _gpu_hits(cpt, S, rays, layout=2)
    assert hits_equal(q, ref).all()


def test_quantized_nodes_sphere_grid_bit_exact(cpt, sphere_grid):
    sc, S, O = sphere_grid
    rays = np.concatenate([camera_rays(sc["camera"], 320, 180, jitter=True, seed=14),
                           random_rays(60000, [-64, 0.0, -64], [64, 6.0, 64], seed=15),
                           vertex_rays(sc, 3000, seed=16)])
    ref, _ = O.trace(rays, "defined")
    q = _gpu_hits(cpt, S, rays, layout=2)
    eq = hits_equal(q, ref)
    assert eq.all(), f"{(~eq).sum()} hits differ"


@pytest.mark.parametrize("leaf", [1, 4, 16])
def test_leaf_sizes_bit_exact(cpt, sphere_grid, leaf):
    """Leaves of any size hold the same triangles: identical hits in both node layouts."""
    sc, _, O = sphere_grid
    S = cpt.Scene(sc, upload=False)
    S.set_leaf_size(leaf)
    S.upload()
    rays = np.concatenate([camera_rays(sc["camera"], 160, 90, jitter=True, seed=17),
                           random_rays(30000, [-64, 0.0, -64], [64, 6.0, 64], seed=18)])
    ref, _ = O.trace(rays, "defined")
    for layout in (0, 2):
        eq = hits_equal(_gpu_hits(cpt, S, rays, layout=layout), ref)
        assert eq.all(), f"leaf {leaf} layout {layout}: {(~eq).sum()} hits differ"

"""CPU checks of the product's host side: the fast SAH build equals the
oracle's literal bvhhelper restatement, the 4-wide re-layout covers every
triangle, and the C ABI library exports every entry point of include/cpt.h."""
import ctypes
import os
import re

import numpy as np
import pytest

import cudapt_amd as ca
import oracle
from cudapt_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "cpt.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(cpt_\w+)\(", hdr, re.M)))
    assert len(names) >= 30
    L = ctypes.CDLL(ca.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def _same(sc, **kw):
    S = ca.Scene(sc, upload=False, **kw)
    O = oracle.Scene(sc, **kw)
    for a, b in zip(S.bvh(), O.bvh()):
        np.testing.assert_array_equal(a, b)
    return S


@pytest.mark.parametrize("name", ["test_cube", "test_sphere", "cornell_box_modded", "test_prism",
                                  "test_oriented_cubes", "cornell_box_with_sphere", "test_spheres"])
def test_sah_build_equals_literal_restatement_meshes(name):
    m = np.load(os.path.join(GOLDEN, "meshes.npz"))
    _same(scenes.mesh_scene(m[name + "_pos"], m[name + "_idx"]))


@pytest.mark.parametrize("seed", range(12))
def test_sah_build_equals_literal_restatement_random(seed):
    rng = np.random.default_rng(seed)
    nt = int(rng.integers(1, 2500))
    pos = rng.normal(size=(nt * 3, 3)).astype(np.float32) * np.float32(rng.uniform(0.01, 100))
    if seed % 3 == 0:
        pos = np.round(pos, 1)          # ties and duplicate centroids
    if seed % 4 == 0:
        pos[:, 1] = 0                   # flat
    sc = scenes.mesh_scene(pos, np.arange(nt * 3, dtype=np.uint32).reshape(-1, 3))
    _same(sc, max_depth=int(rng.integers(4, 33)))


def test_sah_build_equals_literal_restatement_cornell():
    _same(scenes.cornell())


def test_bvh4_layout_summary():
    S = ca.Scene(scenes.cornell(), upload=False)
    info = S.bvh4_info()
    assert info["nodes"] >= 1 and 1 <= info["depth"] <= 31 and info["leaves"] >= 1


def test_errors_are_reported_not_fatal():
    S = ca.Scene()
    with pytest.raises(ca.CptError, match="empty mesh"):
        S.build_bvh()
    with pytest.raises(ca.CptError, match="out of range"):
        S.set_mesh(np.zeros((3, 3), np.float32), None, np.array([[0, 1, 5]], np.uint32), np.zeros(1, np.uint32))
    with pytest.raises(ca.CptError):
        ca.Renderer(0, 10)


def test_leaf_size_changes_the_4wide_tree_only():
    """A smaller leaf bound: no fewer 4-wide nodes (the collapse may choose leaves
    below a larger bound), one triangle per leaf at bound 1, and the binary SAH
    tree unchanged."""
    S = ca.Scene(scenes.sphere_grid(grid=8), upload=False)
    b0 = S.bvh()
    info = {}
    for L in (1, 2, 4, 8, 64):
        S.set_leaf_size(L)
        info[L] = S.bvh4_info()
        assert all(np.array_equal(x, y) for x, y in zip(S.bvh(), b0))
    assert info[1]["nodes"] > info[2]["nodes"] >= info[4]["nodes"] >= info[8]["nodes"] >= info[64]["nodes"]
    assert info[2]["nodes"] > info[64]["nodes"]
    assert info[1]["leaves"] >= S.nt                       # one triangle per leaf
    S.set_leaf_size(2)
    assert S.bvh4_info() == info[2]
    for bad in (0, 65):
        with pytest.raises(ca.CptError, match="not in 1..64"):
            S.set_leaf_size(bad)


def test_collapse_opens_small_subtrees_when_cheaper():
    """SAH-optimal collapse: four unit triangles 10 apart along x.  The binary
    tree pairs them; a pair's box is about 11 long, so a 2-triangle leaf costs
    about 2 x 11 against 2 x 1 for two 1-triangle leaves, and the four
    triangles hang as four leaves of the root.  (Taking every subtree within
    the leaf bound as one leaf, as a greedy collapse does, gives two leaves of
    two triangles.)  With the bound at 1 the tree is the same."""
    pos = []
    for k in range(4):
        x = 10.0 * k
        pos += [(x, 0.0, 0.0), (x + 1.0, 0.0, 0.0), (x, 1.0, 0.0)]
    S = ca.Scene()
    S.set_mesh(np.array(pos, np.float32), None, np.arange(12, dtype=np.uint32).reshape(4, 3),
               np.zeros(4, np.uint32))
    S.build_bvh()
    for L in (2, 1):
        S.set_leaf_size(L)
        assert S.bvh4_info() == {"nodes": 1, "depth": 1, "leaves": 4}


def test_collapse_keeps_a_leaf_of_coincident_triangles():
    """Two triangles with one bounding box cost the same as one 2-triangle leaf
    or two 1-triangle leaves; the collapse keeps the single leaf (ties go to
    the entry that is already one), so four such pairs far apart give four
    leaves under the root."""
    pos = []
    for k in range(4):
        x = 10.0 * k
        pos += [(x, 0.0, 0.0), (x + 1.0, 0.0, 0.0), (x, 1.0, 0.0),
                (x + 1.0, 1.0, 0.0), (x, 1.0, 0.0), (x + 1.0, 0.0, 0.0)]
    S = ca.Scene()
    S.set_mesh(np.array(pos, np.float32), None, np.arange(24, dtype=np.uint32).reshape(8, 3),
               np.zeros(8, np.uint32))
    S.build_bvh()
    S.set_leaf_size(2)
    assert S.bvh4_info() == {"nodes": 1, "depth": 1, "leaves": 4}

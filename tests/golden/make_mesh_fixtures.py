"""Writes tests/golden/meshes.npz from the reference's small OBJ assets.

Reads /root/reference/asset/model/*.obj (data files, not source) with the
package's own OBJ reader and stores positions and triangle indices, so the
fixtures travel without the reference checkout.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from cudapt_amd.scenes import read_obj  # noqa: E402

SRC = "/root/reference/asset/model"
NAMES = ["test_cube", "test_sphere", "cornell_box_modded", "test_prism", "test_prismC", "test_oriented_cubes",
         "cornell_box_with_sphere", "test_spheres"]

if __name__ == "__main__":
    out = {}
    for n in NAMES:
        with open(os.path.join(SRC, n + ".obj")) as f:
            pos, idx = read_obj(f.read())
        out[n + "_pos"] = pos
        out[n + "_idx"] = idx
        print(n, pos.shape, idx.shape)
    np.savez_compressed(os.path.join(HERE, "meshes.npz"), **out)

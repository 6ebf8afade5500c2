"""Copies the reference's preset and model data files used by the scene-ingest
tests into tests/golden/presets/, so that they travel without the reference
checkout.

Data only: two .NPRTENV presets (src/cudapt.cpp:623-737 reads them) and the
OBJ / MTL models they name, from /root/reference/asset/.
"""
import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/asset"
FILES = ["preset/cornellwtranball.NPRTENV", "preset/prismA.NPRTENV", "model/cornell_box_with_sphere.obj",
         "model/cornell_box_with_sphere.mtl", "model/test_prism.obj", "model/test_prism.mtl"]

if __name__ == "__main__":
    out = os.path.join(HERE, "presets")
    os.makedirs(out, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(SRC, f), os.path.join(out, os.path.basename(f)))
        print(f)

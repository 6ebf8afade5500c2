"""The C++ facade include/cpt_rt.hpp compiles against cpt.h, links against the
built library, and drives the host-side scene path (mesh + SAH BVH export)
without a GPU; Init reports failure cleanly when no device is present."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import cudapt_amd
from cudapt_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include <cstdio>
#include <vector>
#include "cpt_rt.hpp"
int main() {
  // two triangles side by side
  float pos[] = {0,0,0, 1,0,0, 0,1,0,  2,0,0, 3,0,0, 2,1,0};
  uint32_t idx[] = {0,1,2, 3,4,5};
  uint32_t mat[] = {0, 0};
  cpt::RTScene scene;
  if (!scene.SetMesh(pos, nullptr, 6, idx, mat, 2)) { std::printf("mesh %s\n", cpt::LastError().c_str()); return 1; }
  std::vector<float> b; std::vector<uint32_t> off, reorder;
  if (!scene.GetCompactBVH(b, off, reorder, 2)) return 2;
  std::printf("nodes %zu first %u %u\n", b.size() / 6, reorder[0], reorder[1]);
  int n = 0;
  cpt_device_count(&n);
  cpt::RTRenderer r;
  r.SetRendererMode(cpt::RTRenderer::BDPT_STREAM_ADAP_BILAT_ATOMIC);
  bool ok = r.Init(64, 48);
  std::printf("devices %d init %d attribs %u\n", n, ok ? 1 : 0, r.GetAttribsN());
  return (n == 0 && ok) ? 3 : 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_facade_compiles_links_and_builds_a_bvh(tmp_path):
    lib = cudapt_amd.LIB_PATH
    if not os.path.exists(lib):
        pytest.skip("library not built")
    src = tmp_path / "f.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "f"
    cmd = ["g++", "-std=c++17", "-O1", str(src), "-I", os.path.join(ROOT, "include"), lib,
           f"-Wl,-rpath,{os.path.dirname(lib)}", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].startswith("nodes 3 first")      # root + two single-triangle leaves
    n_dev = int(lines[1].split()[1])
    if n_dev == 0:
        assert "init 0" in lines[1]
    else:
        assert "init 1" in lines[1] and int(lines[1].split()[-1]) > 15   # the adaptive file's 15 + common ones

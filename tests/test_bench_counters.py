"""bench.py's HBM-traffic arithmetic on a synthetic rocprofv3 counter CSV (no GPU).

The extension is averaged over every instance whose demangled name starts with
the bench's prefix (the first bounce's eye-ray instance and the later bounces'
one), the shadow instance of the same kernel is left out, and the scale comes
from the calibration copy alone."""
import csv
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    argv = sys.argv
    try:
        sys.argv = ["bench.py"]
        spec.loader.exec_module(m)
    finally:
        sys.argv = argv
    return m


def _csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for i, (name, kib) in enumerate(rows):
            w.writerow({"Dispatch_Id": i + 1, "Kernel_Name": name, "Counter_Name": "FETCH_SIZE", "Counter_Value": kib})


class _Args:
    lds_stack, node_format, refill = 806, 1, 64


def test_extension_prefix_covers_both_instances(bench):
    name = bench.extend_kernel_name(_Args, demangled=True)
    assert "void cpt::k_trace_dyn<1, 8, 6, true, true>(cpt::DevScene, ...)".find(name) > 0
    assert "void cpt::k_trace_dyn<1, 8, 6, true, false>(cpt::DevScene, ...)".find(name) > 0
    assert "void cpt::k_trace_dyn<3, 8, 6, true, false>(cpt::DevScene, ...)".find(name) < 0


def test_counter_per_launch_scales_by_the_calibration_copy(bench, tmp_path):
    name = bench.extend_kernel_name(_Args, demangled=True)
    p = tmp_path / "run_counter_collection.csv"
    calib = 256 << 20
    _csv(p, [("__amd_rocclr_copyBuffer", 3.0),                        # a small upload
             ("__amd_rocclr_copyBuffer", 131072.0),                   # 256 MiB copy read as 128 MiB
             ("void cpt::k_trace_dyn<1, 8, 6, true, true>(cpt::DevScene)", 300.0),
             ("void cpt::k_trace_dyn<1, 8, 6, true, false>(cpt::DevScene)", 150.0),
             ("void cpt::k_trace_dyn<1, 8, 6, true, false>(cpt::DevScene)", 150.0),
             ("void cpt::k_trace_dyn<3, 8, 6, true, false>(cpt::DevScene)", 9999.0),
             ("cpt::k_shade(cpt::DevScene)", 7777.0)])
    mean, scale = bench.counter_per_launch(str(p), name, calib)
    assert mean == pytest.approx(200.0 * 1024)
    assert scale == pytest.approx(2.0)


def test_counter_per_launch_missing_rows(bench, tmp_path):
    name = bench.extend_kernel_name(_Args, demangled=True)
    p = tmp_path / "c.csv"
    _csv(p, [("__amd_rocclr_copyBuffer", 131072.0)])
    assert bench.counter_per_launch(str(p), name, 256 << 20) is None
    _csv(p, [("void cpt::k_trace_dyn<1, 8, 6, true, false>(x)", 1.0)])
    assert bench.counter_per_launch(str(p), name, 256 << 20) is None

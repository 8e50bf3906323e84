"""Host planning of the histogram kernels and of the batched matrix-core pair
sums, and the FP64 kernel's fixed-point floor (no GPU).  Histograms: the
work items keep every item below 2^32 pairs, the uint32 LDS slots' limit,
and visit every tile pair exactly once.  The checks are C++ (tests/cpp/)
compiled with g++ against the library's own headers."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "softx-d-16-00083_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_item_tables_bound_and_coverage(tmp_path):
    exe = tmp_path / "hist_plan_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", CSRC,
                    os.path.join(HERE, "cpp", "hist_plan_check.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")
    # the longest paired runs are capped at 2047 tiles: 2 * 1024 * 2047 * 1024 pairs
    assert "max_pairs=4292870144" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_mfma_batch_plan_tables(tmp_path):
    """The batched matrix-core pair sums' column groups (csrc/xs_mfma_plan.h):
    every pair type is one column in exactly one group of one slice count
    that covers its pair count, every group's k range covers its columns'
    bins, the count slices do not overlap, and bounds of 2^35 are refused."""
    exe = tmp_path / "mfma_plan_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", CSRC,
                    os.path.join(HERE, "cpp", "mfma_plan_check.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_fixed_point_floor_worst_rsq(tmp_path):
    """The FP64 kernel's fixed-point floor (csrc/xs_fixbin.h) with reciprocal
    square roots at v_rsq_f64's documented worst accuracy (2^-23): wherever
    the low word passes, the high word gives (unsigned)(sqrt(s)/bin) in IEEE
    double, on uniform and bin-edge distances for grids of 64 to 2^19 bins;
    two unsound bands (no margin, half the test threshold) are caught; zero,
    negative and non-finite squared distances go to the exact evaluation."""
    exe = tmp_path / "fixbin_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", CSRC,
                    os.path.join(HERE, "cpp", "fixbin_check.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")

"""bench.py's roofline bookkeeping, on the CPU: the kernels it names for each stage of the
headline workload exist in the built library, and the committed rocprof summary holds their
HBM-traffic rows, so the bench line's `traffic` is filled in whichever stage dominates."""
import os
import shutil
import subprocess

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "rainbow_amd", "librbw.so")


@pytest.fixture(scope="module")
def kernel_symbols():
    if shutil.which("nm") is None:
        pytest.skip("nm is not installed")
    return subprocess.run(["nm", "-C", LIB], capture_output=True, text=True, check=True).stdout


@pytest.mark.parametrize("stage", ["extend", "shade", "connect"])
def test_headline_stage_kernels_exist_and_are_profiled(stage, kernel_symbols):
    names = bench.profiled_kernels(stage, 4)
    for name in names:
        assert f"::{name}(" in kernel_symbols, name
    traffic = bench.profiled_traffic(names)
    assert traffic is not None and traffic > 0


def test_traffic_is_none_without_a_profile_row():
    assert bench.profiled_traffic(("k_no_such_kernel",)) is None

"""Host-side BVH builders (rainbow_amd/csrc/bvh.cpp): the eight-wide tree collapsed
from the binary SAH tree keeps every primitive exactly once, its quantised child
boxes contain their subtrees in the traversal's float arithmetic, and a closest-hit
walk over it finds the brute-force nearest triangle (tests/cpp/bvh8_check.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bvh8_check_exe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bvh8") / "bvh8_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-o", exe, os.path.join(ROOT, "tests", "cpp", "bvh8_check.cpp"),
                    os.path.join(ROOT, "rainbow_amd", "csrc", "bvh.cpp")], check=True)
    return exe


@pytest.mark.parametrize("ntris", [1, 2, 5, 9, 300, 20000])
def test_bvh8_structure_and_closest_hits(bvh8_check_exe, ntris):
    r = subprocess.run([bvh8_check_exe, str(ntris)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")

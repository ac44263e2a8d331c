"""Host-side BVH builders (rainbow_amd/csrc/bvh.cpp): the four- and eight-wide trees
collapsed from the binary SAH tree keep every primitive exactly once, their child
boxes contain their subtrees (the eight-wide tree's quantised ones in the traversal's
float arithmetic), the four-wide tree's recorded stack bound holds, and a closest-hit
walk over each finds the brute-force nearest triangle (tests/cpp/bvh4_check.cpp,
tests/cpp/bvh8_check.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build_check(tmp_path_factory, name):
    exe = str(tmp_path_factory.mktemp(name) / name)
    subprocess.run(["g++", "-std=c++17", "-O2", "-o", exe, os.path.join(ROOT, "tests", "cpp", name + ".cpp"),
                    os.path.join(ROOT, "rainbow_amd", "csrc", "bvh.cpp")], check=True)
    return exe


@pytest.fixture(scope="module")
def bvh8_check_exe(tmp_path_factory):
    return _build_check(tmp_path_factory, "bvh8_check")


@pytest.fixture(scope="module")
def bvh4_check_exe(tmp_path_factory):
    return _build_check(tmp_path_factory, "bvh4_check")


@pytest.mark.parametrize("ntris", [1, 2, 5, 9, 300, 20000])
def test_bvh4_structure_and_closest_hits(bvh4_check_exe, ntris):
    r = subprocess.run([bvh4_check_exe, str(ntris)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")


@pytest.mark.parametrize("ntris", [1, 2, 5, 9, 300, 20000])
def test_bvh8_structure_and_closest_hits(bvh8_check_exe, ntris):
    r = subprocess.run([bvh8_check_exe, str(ntris)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")

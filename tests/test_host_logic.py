"""CPU test of the serving loop's host bookkeeping (range pool, request retirement when writing
out a request's results throws), compiled with g++ from tests/cpp/pool_test.cpp."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_range_pool_and_retirement(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "pool_test")
    src = os.path.join(REPO, "tests", "cpp", "pool_test.cpp")
    inc = os.path.join(REPO, "cpm-ibc_amd", "csrc")
    # host code only, so the sanitizers apply to all of it
    build = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Werror", "-fsanitize=address,undefined",
                            "-I", inc, src, "-o", exe], capture_output=True, text=True, timeout=180)
    assert build.returncode == 0, build.stderr
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "pool ok" in out.stdout


def test_refinement_step_rules(tmp_path):
    """mx_update: a refinement step that does not reduce the true residual sends the lane to the
    FP64 re-solve whatever asked for the step; delta and the failure count judge delta-test steps."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "refine_test")
    src = os.path.join(REPO, "tests", "cpp", "refine_test.cpp")
    inc = os.path.join(REPO, "cpm-ibc_amd", "csrc")
    build = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", inc, src, "-o", exe],
                           capture_output=True, text=True, timeout=180)
    assert build.returncode == 0, build.stderr
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "refine ok" in out.stdout


def test_fp64_routing_rule(tmp_path):
    """The FP64 routing decision (cpm-ibc_amd/csrc/route.h): an operator whose FP32+refine systems never
    fell back keeps FP32; one whose expected FP32+refine cost (FP32 iterations at kFp32IterCost plus the
    re-solves of the fraction that falls back) exceeds the FP64 iterations of its re-solves goes to FP64;
    the boundary case either side."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "route_test")
    src = os.path.join(REPO, "tests", "cpp", "route_test.cpp")
    inc = os.path.join(REPO, "cpm-ibc_amd", "csrc")
    build = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", inc, src, "-o", exe],
                           capture_output=True, text=True, timeout=180)
    assert build.returncode == 0, build.stderr
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "route ok" in out.stdout

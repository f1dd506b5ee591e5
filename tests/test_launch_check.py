"""The host-side check that runs before every stage-kernel launch
(hifiles-solver_amd/csrc/fr_launch_check.hpp), compiled with g++ on the CPU.

The case it exists for: a learned-closure context (LES model 3) whose network
weight pointers were null -- a kernel evaluating the network then reads the
weights through null pointers.  The check must refuse that launch, and every
launch whose arrays are all set must pass.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "hifiles-solver_amd", "csrc")
MAIN = os.path.join(ROOT, "tests", "native", "launch_check_main.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_launch_check(tmp_path):
    exe = str(tmp_path / "launch_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", CSRC, MAIN, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "learned LES, no weights, gradient traces" in r.stdout

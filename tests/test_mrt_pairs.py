"""The FP32-compute MRT transform over opposite-direction pairs
(hippolbm_amd/csrc/mrt_pairs.hpp) against the moment matrix and its exact
inverse that the FP64 kernels use (those are pinned bit-exact to the oracle):
compiled for the host with g++ and run in double precision on random vectors."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pair_transforms_equal_the_basis_tables(tmp_path):
    exe = tmp_path / "mrt_pairs_check"
    src = os.path.join(HERE, "csrc", "mrt_pairs_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-o", str(exe), src], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    worst = [float(v) for v in r.stdout.split()[1::2]]
    assert len(worst) == 3 and max(worst) < 1e-13

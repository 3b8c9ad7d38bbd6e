"""The batch's first-fit allocator (hippolbm_amd/csrc/free_list.hpp), which
places solves in the population and flag arenas and results in the read-back
pool: checked against a unit-by-unit model on random allocation and release
sequences.  Compiled for the host with g++."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_first_fit_matches_the_unit_model(tmp_path):
    exe = tmp_path / "free_list_check"
    src = os.path.join(HERE, "csrc", "free_list_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-o", str(exe), src], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("failures 0 ")

"""Host-side constants of the FP32-compute forms (hippolbm_amd/csrc/fp32_constants.hpp):
the BGK constants keep mass and momentum exactly in exact arithmetic over the
practical tau range, and the MRT rates default to hippoLBM's (mrt.hpp:43-47).
Compiled for the host with g++."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_bgk_constants_conserve_and_mrt_rates_default(tmp_path):
    exe = tmp_path / "fp32_constants_check"
    src = os.path.join(HERE, "csrc", "fp32_constants_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-o", str(exe), src], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("failures 0 ")

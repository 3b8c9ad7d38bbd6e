"""The C++ operator layer end to end, as the reference's regression tests run:
`hlbm_run deck.msp` executes initialize / compute_loop / checker, the
check_macro_quantities pin of every regression deck passes (exit status 0,
"Test Pass."), and the final simulation_state equals the oracle's."""
import glob
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as orc
from hippolbm_amd import deck as dk

from conftest import DECKS

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "hippolbm_amd", "hlbm_run")
DECK_FILES = sorted(glob.glob(os.path.join(DECKS, "regression_*.msp")))


def run(args, cwd):
    # output operators write under the working directory, as the reference does
    print("hlbm_run", *args, flush=True)
    sys.stdout.flush()
    return subprocess.run([RUN] + args, capture_output=True, text=True, timeout=120, cwd=cwd)


def final_state(out):
    line = [l for l in out.splitlines() if l.startswith("final step")][-1].split()
    return int(line[2]), float(line[4]), float(line[6]), float(line[8])


@pytest.mark.parametrize("path", DECK_FILES, ids=[os.path.basename(p) for p in DECK_FILES])
def test_regression_deck_passes_its_checker(path, tmp_path):
    r = run([path, "--quiet"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Test Pass." in r.stdout
    # the first-timestep and final paraview dumps of main-config.msp
    pv = tmp_path / "hippoLBMOutputDir" / "ParaviewOutput"
    d0 = dk.load_msp(path)
    assert (pv / "hippoLBM_0000000000" / "000000.vtr").exists()
    assert (pv / ("hippoLBM_%010d.pvtr" % (d0.end_iteration + 1))).exists()
    step, s, vmin, vmax = final_state(r.stdout)
    d = dk.load_msp(path)
    ref = orc.OracleRun(d)
    ref.step(d.end_iteration)
    o = ref.final_state()
    assert step == d.end_iteration
    assert abs(s - o[0]) <= 1e-12 * abs(o[0])
    assert abs(vmin - o[1]) <= 1e-12 * max(abs(o[1]), 1e-300) + 1e-300
    assert abs(vmax - o[2]) <= 1e-12 * abs(o[2])


def test_log_table_and_aa_pattern(tmp_path):
    path = os.path.join(DECKS, "regression_lbm_cavite_wall.msp")
    r = run([path, "--pattern", "aa", "--steps", "7", "--no-output"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert "== Grid size: (LX: 31, LY: 31, LZ 31)" in lines  # print_domain first
    lines = lines[[l.split()[:1] for l in lines].index(["Step"]):]
    assert lines[0].split() == ["Step", "Time", "Mesh", "Size", "Sum(density)", "min(||V||)", "max(||V||)", "MLUPS"]
    steps = [int(l.split()[0]) for l in lines[1:] if l.strip() and l.split()[0].isdigit()]
    assert steps == list(range(0, 8))
    assert "Test Pass." in r.stdout


def test_failing_checker_exits_with_failure(tmp_path):
    txt = open(os.path.join(DECKS, "regression_lbm_cavite.msp")).read().replace("density: 29791", "density: 29000")
    deck = tmp_path / "bad_density.msp"
    deck.write_text(txt)
    r = run([str(deck), "--quiet"], tmp_path)
    assert r.returncode != 0
    assert "Regression test checker [macro quantities]" in r.stdout and "Sum of densities" in r.stdout


def test_quadric_deck_runs_through_the_operator_layer(tmp_path):
    from test_gpu_parity import QUADRIC_DECK
    deck = tmp_path / "quadrics.msp"
    deck.write_text(QUADRIC_DECK)
    r = run([str(deck), "--quiet", "--no-output"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    step, s, vmin, vmax = final_state(r.stdout)
    d = dk.loads_msp(QUADRIC_DECK, "quadrics")
    ref = orc.OracleRun(d)
    ref.step(d.end_iteration)
    o = ref.final_state()
    assert step == d.end_iteration
    assert abs(s - o[0]) <= 1e-12 * abs(o[0])
    assert vmin == o[1] and vmax == o[2]


def test_start_iteration_deck_runs_through_the_operator_layer(tmp_path):
    from test_gpu_parity import START_DECK
    deck = tmp_path / "start.msp"
    deck.write_text(START_DECK)
    r = run([str(deck), "--quiet", "--no-output", "--pattern", "aa"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    step, s, vmin, vmax = final_state(r.stdout)
    d = dk.loads_msp(START_DECK, "start")
    ref = orc.OracleRun(d)
    ref.step(d.end_iteration)
    o = ref.final_state()
    assert step == d.end_iteration
    assert abs(s - o[0]) <= 1e-12 * abs(o[0])
    assert vmin == o[1] and vmax == o[2]


def test_double_couette_deck_runs_through_the_operator_layer(tmp_path):
    from test_gpu_parity import DOUBLE_COUETTE_DECK
    deck = tmp_path / "dc.msp"
    deck.write_text(DOUBLE_COUETTE_DECK)
    r = run([str(deck), "--no-output"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Prepro double couette starting ... dim Z" in r.stdout
    r = run([str(deck), "--quiet", "--no-output"], tmp_path)
    step, s, vmin, vmax = final_state(r.stdout)
    d = dk.loads_msp(DOUBLE_COUETTE_DECK, "dc")
    ref = orc.OracleRun(d)
    ref.step(d.end_iteration)
    o = ref.final_state()
    assert abs(s - o[0]) <= 1e-12 * abs(o[0])
    assert vmin == o[1] and vmax == o[2]

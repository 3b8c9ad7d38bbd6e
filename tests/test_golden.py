"""The oracle against the committed golden fixtures (tests/golden/): the
populations and flags of six small cases (every boundary operator, MRT with
the reference's and custom rates, obstacles of the three kinds,
start_iteration) and the SHA-256 of the 64^3 moving-lid Couette populations
after 1, 10 and 100 steps.  The fixtures were written by
tests/golden/make_golden.py; a change of the oracle's arithmetic shows here."""
import json
import os
import sys

import numpy as np
import pytest

import oracle as orc
from hippolbm_amd import deck as dk
from hippolbm_amd import workloads

from conftest import GOLDEN

sys.path.insert(0, GOLDEN)
import cases  # noqa: E402

FIX = np.load(os.path.join(GOLDEN, "populations.npz"))
META = json.load(open(os.path.join(GOLDEN, "golden.json")))


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_oracle_reproduces_the_small_fixtures(name):
    txt, nodes, seed, steps = cases.CASES[name]
    d = dk.loads_msp(txt, name)
    ref = orc.OracleRun(d, init_f=cases.seeded_equilibrium(nodes, seed))
    assert np.array_equal(ref.o.real(ref.o.obst()), FIX[f"flags_{name}"])
    ref.step(steps)
    f = ref.o.node_f()
    assert np.array_equal(f, FIX[f"f_{name}"])
    assert cases.f_hash(f) == META["cases"][name]["sha256"]


def test_fixtures_are_not_trivial():
    # every case moved away from its initial state; obstacle cases have walls
    for name, (txt, nodes, seed, steps) in cases.CASES.items():
        f0 = cases.seeded_equilibrium(nodes, seed)
        assert np.abs(FIX[f"f_{name}"] - f0).max() > 1e-5, name
    assert META["cases"]["ball_wall_bounce_back"]["walls"] > 0
    assert META["cases"]["quadric_start_iteration"]["walls"] > 0


def test_oracle_reproduces_the_couette64_hashes():
    d = workloads.couette(cases.HASH_NODES)
    ref = orc.OracleRun(d, init_f=cases.seeded_equilibrium(cases.HASH_NODES, cases.HASH_SEED))
    done = 0
    for s in cases.HASH_STEPS:
        ref.step(s - done)
        done = s
        assert cases.f_hash(ref.o.node_f()) == META["couette64"][str(s)]["sha256"], s

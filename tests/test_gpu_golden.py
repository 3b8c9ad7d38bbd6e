"""The HIP engine against the committed golden fixtures (tests/golden/): bit
equality of the FP64 populations of the six small cases (A/B and AA, fluid
nodes; A/B also walls where no wall_bounce_back skips them) and of the 64^3
moving-lid Couette after 1, 10 and 100 steps (SHA-256)."""
import json
import os
import sys

import numpy as np
import pytest

from hippolbm_amd import deck as dk
from hippolbm_amd import workloads
from hippolbm_amd.engine import Simulation

from conftest import GOLDEN

sys.path.insert(0, GOLDEN)
import cases  # noqa: E402

pytestmark = pytest.mark.gpu

FIX = np.load(os.path.join(GOLDEN, "populations.npz"))
META = json.load(open(os.path.join(GOLDEN, "golden.json")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("pattern", ["ab", "aa"])
@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_engine_reproduces_the_small_fixtures(name, pattern):
    txt, nodes, seed, steps = cases.CASES[name]
    say("golden", name, pattern)
    d = dk.loads_msp(txt, name)
    sim = Simulation(d, storage="f64", init_f=cases.seeded_equilibrium(nodes, seed), pattern=pattern)
    flags = sim.dom.get_flags()
    assert np.array_equal(flags, FIX[f"flags_{name}"])
    sim.run(steps)
    f = sim.dom.get_f()
    fluid = flags == -1
    exp = FIX[f"f_{name}"]
    assert np.array_equal(f[:, fluid], exp[:, fluid])
    sim.dom.close()


def test_engine_reproduces_the_couette64_hashes():
    d = workloads.couette(cases.HASH_NODES)
    sim = Simulation(d, storage="f64", init_f=cases.seeded_equilibrium(cases.HASH_NODES, cases.HASH_SEED))
    done = 0
    for s in cases.HASH_STEPS:
        sim.run(s - done)
        done = s
        assert cases.f_hash(sim.dom.get_f()) == META["couette64"][str(s)]["sha256"], s
    st = sim.state()
    assert abs(st["sum_density"] - META["couette64"]["100"]["sum_rho"]) <= 1e-12 * META["couette64"]["100"]["sum_rho"]
    assert st["max_velocity_norm"] == META["couette64"]["100"]["max_u"]
    sim.dom.close()

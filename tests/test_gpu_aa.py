"""AA-pattern (single lattice, in place) parity: the same FP64 populations on
fluid nodes as the oracle and as the A/B kernel, for odd and even step counts,
across read-backs, and for every boundary operator combination."""
import glob
import os

import numpy as np
import pytest

import oracle as orc
from hippolbm_amd import deck as dk
from hippolbm_amd.engine import Simulation

from conftest import DECKS
from test_gpu_parity import _couette_deck, _porous_deck, seeded_equilibrium

pytestmark = pytest.mark.gpu

DECK_FILES = sorted(glob.glob(os.path.join(DECKS, "regression_*.msp")))


def fluid_equal(sim, ref):
    fo = ref.o.node_f()
    fg = sim.dom.get_f()
    fluid = ref.o.real(ref.o.obst()) == -1
    bad = int((fg[:, fluid] != fo[:, fluid]).sum())
    return bad, float(np.abs(fg[:, fluid] - fo[:, fluid]).max())


@pytest.mark.parametrize("path", DECK_FILES, ids=[os.path.basename(p) for p in DECK_FILES])
def test_aa_regression_decks_bitwise(path):
    d = dk.load_msp(path)
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64", pattern="aa")
    for n in (1, 6, 31):  # odd and even counts, read-backs in between
        ref.step(n)
        sim.run(n)
        bad, mx = fluid_equal(sim, ref)
        assert bad == 0, (n, bad, mx)
    s = sim.state()
    assert abs(s["sum_density"] - ref.final_state()[0]) <= 1e-12 * abs(s["sum_density"])
    sim.dom.close()


@pytest.mark.parametrize("collision", ["bgk", "mrt"])
def test_aa_porous_bitwise(collision):
    d = _porous_deck(24, 20, 18, 8, 3, collision)
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64", pattern="aa")
    ref.step(41)
    sim.run(41)
    bad, mx = fluid_equal(sim, ref)
    assert bad == 0, (bad, mx)
    sim.dom.close()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_aa_equals_ab_low_precision(storage):
    d = _couette_deck(24)
    f0 = seeded_equilibrium(d.nodes, 3)
    a = Simulation(d, storage=storage, init_f=f0, pattern="ab")
    b = Simulation(d, storage=storage, init_f=f0, pattern="aa")
    a.run(57)
    b.run(57)
    fa, fb = a.dom.get_f(), b.dom.get_f()
    # same arithmetic, possibly contracted differently in the two kernels
    assert np.abs(fa - fb).max() < (1e-6 if storage == "f32" else 5e-5)
    # one population lattice instead of two
    assert b.dom.bytes() < 0.7 * a.dom.bytes()
    a.dom.close()
    b.dom.close()


REPEATED_BB = """
do_domain:
  - domain:
     bounds: [[0,0,0],[10,9,12]]
     cell_dims: [10, 9, 12]
     periodic: [true, false, false]
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.8
     nuth: 0.1
pre_stream_bcs:
  - pre_bounce_back
  - cavity_z_l:
     U: [0.04, 0.01, 0]
  - cavity_z_0:
     U: [0.0, -0.02, 0]
post_stream_bcs:
  - post_bounce_back
  - lid_driven_cavity:
     U: [0.02, 0.0, 0]
     regions: [plan_xz_l]
  - post_bounce_back
"""


@pytest.mark.parametrize("pattern", ["ab", "aa"])
def test_repeated_post_bounce_back_with_moving_walls(pattern):
    # post_bounce_back first (the gather adds the moving-wall term) and again after
    # the lid: the re-applied bounce-back must add the moving-wall term once
    d = dk.loads_msp(REPEATED_BB, "repeated_bb")
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64", pattern=pattern)
    for n in (1, 4, 7, 12):  # odd and even totals, read-backs in between
        ref.step(n)
        sim.run(n)
        bad, mx = fluid_equal(sim, ref)
        assert bad == 0, (pattern, n, bad, mx)
    sim.dom.close()

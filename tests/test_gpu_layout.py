"""Node layouts in HBM (hlbm_set_layout): every axis order gives the same
results, bit for bit, as the oracle.

The layout only moves nodes in memory; the per-node arithmetic is the same, so
FP64 populations on every fluid node must equal the oracle's for all six axis
orders, for a layout set before stepping and for one changed between steps
(A/B and AA), with the boundary operators on x-normal planes that the layout is
meant for (density inlet, outflow), wall bounce-back on balls, walls across a
periodic face and a moving lid.
"""
import itertools
import sys

import numpy as np
import pytest

import oracle as orc
from hippolbm_amd import deck as dk
from hippolbm_amd import workloads
from hippolbm_amd.engine import Simulation

pytestmark = pytest.mark.gpu

ORDERS = list(itertools.permutations((0, 1, 2)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CHANNEL = workloads.porous_channel(nx=40, ny=18, nz=22, n_spheres=6, r_range=(2.0, 3.5), seed=5)

PERIODIC_WALLS = """
do_domain:
  - domain:
     bounds: [[0,0,0],[0.4,0.12,0.2]]
     cell_dims: [40, 12, 20]
     periodic: [true, true, false]
set_lbm_parameters:
  - lbm_parameters:
     Fext: [1.3, 0, 0]
     nuth: 1e-3
     tau: 0.65
set_obstacles:
  - register_quadrics:
     id: 0
     quadrics: cyly
     transform:
        - scale: [ 0.04, 1, 0.04 ]
        - translate: [ 0.1, 0.06, 0.1 ]
boundary_conditions:
  - neumann:
     U: [0,0,0]
     regions: [plan_xy_l]
  - neumann:
     U: [0,0,0]
     regions: [plan_xy_0]
pre_stream_bcs:
  - wall_bounce_back
"""

LID = """
do_domain:
  - domain:
     bounds: [[0,0,0],[13,9,11]]
     cell_dims: [13, 9, 11]
     periodic: [false, true, false]
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.75
     nuth: 0.0833333333333333333
collision: mrt
pre_stream_bcs:
  - pre_bounce_back
  - cavity_z_l:
     U: [0.04, 0.01, 0]
post_stream_bcs:
  - post_bounce_back
  - lid_driven_cavity:
     U: [0.02, 0, 0]
     regions: [plan_yz_0]
  - post_bounce_back
"""


def _deck(name):
    if name == "channel":
        return CHANNEL
    return dk.loads_msp(PERIODIC_WALLS if name == "periodic_walls" else LID, name)


def _fluid_equal(sim, ref):
    o = ref.o
    fluid = o.real(o.obst()) == -1
    fo = o.node_f()
    fg = sim.dom.get_f()
    bad = int((fg[:, fluid] != fo[:, fluid]).sum())
    return bad, float(np.abs(fg[:, fluid] - fo[:, fluid]).max())


@pytest.mark.parametrize("order", ORDERS, ids=["".join("xyz"[a] for a in o) for o in ORDERS])
@pytest.mark.parametrize("name", ["channel", "periodic_walls", "lid"])
@pytest.mark.parametrize("pattern", ["ab", "aa"])
def test_every_layout_is_bit_equal_to_the_oracle(name, order, pattern):
    d = _deck(name)
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64", pattern=pattern)
    sim.dom.set_layout(order)
    assert sim.dom.get_layout() == order
    assert np.array_equal(sim.dom.get_flags(), ref.o.real(ref.o.obst()))
    ref.step(13)
    sim.run(13)
    bad, mx = _fluid_equal(sim, ref)
    assert bad == 0, (name, order, pattern, bad, mx)
    # change the layout between steps (the pending streaming is applied first)
    other = ORDERS[(ORDERS.index(order) + 3) % 6]
    sim.dom.set_layout(other)
    ref.step(8)
    sim.run(8)
    bad, mx = _fluid_equal(sim, ref)
    assert bad == 0, (name, order, other, pattern, bad, mx)
    # rho/u of the last step, in node layout whatever the device layout
    rho, u = sim.dom.get_macro()
    o = ref.o
    fluid = o.real(o.obst()) == -1
    assert np.array_equal(rho[fluid], o.real(o.m0())[fluid])
    sim.dom.close()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_reduced_precision_is_layout_independent(storage):
    runs = []
    for order in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        sim = Simulation(CHANNEL, storage=storage)
        sim.dom.set_layout(order)
        sim.run(17)
        runs.append(sim.dom.get_f())
        s, c, lo, hi = sim.dom.plane_profile(0)
        runs.append(np.concatenate([s, c, lo, hi]))
        sim.dom.close()
    assert np.array_equal(runs[0], runs[2]) and np.array_equal(runs[0], runs[4])
    assert np.array_equal(runs[1], runs[3]) and np.array_equal(runs[1], runs[5])

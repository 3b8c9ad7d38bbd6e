"""wall_bounce_back across periodic faces, FP64 against the oracle.

hippoLBM bounces a wall node on the face of a periodic axis off its ghost
neighbour, which still holds the previous step's populations and whose flag is
rasterized at the ghost's own position
(src/bcs/include/hippoLBM/bcs/bounce_back.hpp:163-185, src/collision/streaming.cu:107-111,
src/obstacle/update_obstacles.cu:58-67).  The oracle restates that literally
(ghost layers, swap streaming); the engine keeps the previous-step values of
the affected nodes.  Cases:

* the shape of example/lbm_quadrics.msp: a y-cylinder across the periodic y
  faces, a sphere and a cone, Zou/He walls on z, a body force (ghost flags equal
  the periodic images); with a wider cone, nodes next to walls across the y
  faces also lie on the Zou/He planes (as in the example itself);
* a wall box from y = -0.5 to y = 2.5 with dx = 1: the strict AABB test leaves
  the ghost at y = -1 fluid while its image y = ny - 1 is fluid too, and the
  Real walls y = 0..2 border the periodic face;
* a ball across a periodic face, whose ghosts are walls over fluid images: the
  engine refuses it (hippoLBM then streams the wall's own populations).

A/B and AA, odd and even step counts, every fluid population bit-equal.
"""
import sys

import numpy as np
import pytest

import oracle as orc
from hippolbm_amd import _lib
from hippolbm_amd import deck as dk
from hippolbm_amd.engine import Simulation

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


QUADRICS_ACROSS = """
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
  - register_quadrics:
     id: 1
     quadrics: sphere
     transform:
        - scale: [ 0.03, 0.04, 0.03 ]
        - translate: [ 0.22, 0.06, 0.1 ]
  - register_quadrics:
     id: 2
     quadrics: conez
     transform:
        - scale: [ 0.01, 0.01, 0.05 ]
        - translate: [ 0.32, 0.06, 0.1 ]
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

# the y-cylinder moved down to cut the z = 0 face: nodes next to it across the
# periodic y faces lie on the neumann plane (as the cone does in
# example/lbm_quadrics.msp)
QUADRICS_ON_PLANES = QUADRICS_ACROSS.replace("""        - scale: [ 0.04, 1, 0.04 ]
        - translate: [ 0.1, 0.06, 0.1 ]""", """        - scale: [ 0.04, 1, 0.04 ]
        - translate: [ 0.1, 0.06, 0.02 ]""")

WALL_AT_FACE = """
do_domain:
  - domain:
     bounds: [[0,0,0],[17,13,11]]
     cell_dims: [17, 13, 11]
     periodic: [true, true, false]
set_lbm_parameters:
  - lbm_parameters:
     Fext: [0, 2e-5, 0]
     tau: 0.7
     nuth: 0.0666666666666666667
set_obstacles:
  - register_solid_wall:
     id: 0
     bounds: [[4.5, -0.5, 2.5], [9.5, 2.5, 7.5]]
collision: mrt
pre_stream_bcs:
  - pre_bounce_back
  - wall_bounce_back
  - cavity_z_l:
     U: [0.03, 0.01, 0]
post_stream_bcs:
  - post_bounce_back
"""

BALL_ACROSS = """
do_domain:
  - domain:
     bounds: [[0,0,0],[15,15,11]]
     cell_dims: [15, 15, 11]
     periodic: [true, true, false]
set_lbm_parameters:
  - lbm_parameters:
     Fext: [1e-5, 0, 0]
     tau: 0.8
     nuth: 0.1
set_obstacles:
  - register_solid_ball:
     id: 0
     center: [0.0, 7.0, 5.0]
     radius: 3.2
pre_stream_bcs:
  - wall_bounce_back
"""


def _run(text, name, pattern, steps, storage="f64"):
    d = dk.loads_msp(text, name)
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage=storage, pattern=pattern)
    flags = sim.dom.get_flags()
    assert np.array_equal(flags, ref.o.real(ref.o.obst()))
    ref.step(steps)
    sim.run(steps)
    return sim, ref, flags


def _fluid_bitwise(sim, ref):
    o = ref.o
    fluid = o.real(o.obst()) == -1
    fo = o.node_f()
    fg = sim.dom.get_f()
    bad = fg[:, fluid] != fo[:, fluid]
    return int(bad.sum()), float(np.abs(fg[:, fluid] - fo[:, fluid]).max())


@pytest.mark.parametrize("pattern", ["ab", "aa"])
@pytest.mark.parametrize("steps", [30, 31])
def test_quadrics_across_periodic_faces_bitwise(pattern, steps):
    say("quadrics across periodic y faces", pattern, steps)
    sim, ref, flags = _run(QUADRICS_ACROSS, "quadrics_across", pattern, steps)
    # the cylinder crosses both y faces: walls on y = 0 and y = ny - 1
    assert (flags[:, 0, :] == -2).sum() > 10 and (flags[:, -1, :] == -2).sum() > 10
    nbad, maxdiff = _fluid_bitwise(sim, ref)
    assert nbad == 0, (nbad, maxdiff)
    rho, u = sim.dom.macro()
    o = ref.o
    o.macro_variables()
    fluid = o.real(o.obst()) == -1
    assert np.array_equal(rho[fluid], o.real(o.m0())[fluid])
    sim.dom.close()


@pytest.mark.parametrize("pattern", ["ab", "aa"])
@pytest.mark.parametrize("steps", [20, 21])
def test_walls_across_periodic_faces_on_neumann_planes_bitwise(pattern, steps):
    say("quadrics across periodic y faces, on the neumann planes", pattern, steps)
    assert QUADRICS_ON_PLANES != QUADRICS_ACROSS
    sim, ref, flags = _run(QUADRICS_ON_PLANES, "quadrics_on_planes", pattern, steps)
    assert (flags[0, 0, :] == -2).sum() > 0  # walls on the y face at z = 0
    nbad, maxdiff = _fluid_bitwise(sim, ref)
    assert nbad == 0, (nbad, maxdiff)
    ref.step(3)
    sim.run(3)
    nbad, maxdiff = _fluid_bitwise(sim, ref)
    assert nbad == 0, (nbad, maxdiff)
    sim.dom.close()


@pytest.mark.parametrize("pattern", ["ab", "aa"])
def test_wall_box_at_periodic_face_bitwise(pattern):
    say("wall box at the periodic y face", pattern)
    sim, ref, flags = _run(WALL_AT_FACE, "wall_at_face", pattern, 41)
    assert (flags[:, 0, :] == -2).sum() > 0 and (flags[:, -1, :] == -2).sum() == 0
    nbad, maxdiff = _fluid_bitwise(sim, ref)
    assert nbad == 0, (nbad, maxdiff)
    # a read in the middle (stream-only / AA read-out kernels) and more steps
    ref.step(6)
    sim.run(6)
    nbad, maxdiff = _fluid_bitwise(sim, ref)
    assert nbad == 0, (nbad, maxdiff)
    sim.dom.close()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_quadrics_across_periodic_faces_reduced_precision(storage):
    say("quadrics across periodic y faces", storage)
    sim64, _, _ = _run(QUADRICS_ACROSS, "quadrics_across", "ab", 30)
    d = dk.loads_msp(QUADRICS_ACROSS, "quadrics_across")
    sim = Simulation(d, storage=storage)
    sim.run(30)
    f64 = sim64.dom.get_f()
    f = sim.dom.get_f()
    fluid = sim64.dom.get_flags() == -1
    err = float(np.abs(f[:, fluid] - f64[:, fluid]).max())
    say(f"  max |f - f64| {err:.3e}")
    assert err < {"f32": 1e-6, "f16": 1e-3}[storage]
    sim.dom.close()
    sim64.dom.close()


def test_ball_across_periodic_face_refused():
    d = dk.loads_msp(BALL_ACROSS, "ball_across")
    sim = Simulation(d, storage="f64")
    with pytest.raises(_lib.HLBMError) as e:
        sim.run(2)
    assert "periodic face" in str(e.value)
    sim.dom.close()


def test_batch_refuses_obstacles_at_periodic_faces():
    from hippolbm_amd import batch as bt
    b = bt.Batch(storage="f64", budget_bytes=64 << 20, chunk_steps=4)
    ball = {"kind": "ball", "center": (0.5, 6.0, 6.0), "radius": 2.0}
    with pytest.raises(_lib.HLBMError) as e:
        b.submit(bt.Solve(nodes=(12, 12, 12), steps=4, tau=0.8, periodic=(True, False, False), obstacles=[ball]))
    assert "periodic face" in str(e.value)
    # the same ball in a closed box, or away from the periodic face, is accepted
    b.submit(bt.Solve(nodes=(12, 12, 12), steps=4, tau=0.8, obstacles=[ball]))
    far = {"kind": "ball", "center": (6.0, 6.0, 6.0), "radius": 2.0}
    b.submit(bt.Solve(nodes=(12, 12, 12), steps=4, tau=0.8, periodic=(True, False, False), obstacles=[far]))
    b.close()

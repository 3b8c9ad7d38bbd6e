"""FP64 parity of the HIP engine (through the C ABI) with the CPU oracle.

The FP64 kernel keeps the reference's evaluation order without contraction, so
the bar is bit equality of every population on every fluid node, of rho/u and
of the obstacle flags.  Cases: the seven regression decks, a seeded Couette,
a porous MRT channel with rasterized balls and wall bounce-back, and
operator-order variants.
"""
import glob
import os
import sys

import numpy as np
import pytest

import oracle as orc
from hippolbm_amd import _lib
from hippolbm_amd import deck as dk
from hippolbm_amd.engine import Domain, Simulation

from conftest import DECKS, GOLDEN

sys.path.insert(0, GOLDEN)
from cases import seeded_equilibrium  # noqa: E402,F401  (also used by the other GPU tests)

pytestmark = pytest.mark.gpu

DECK_FILES = sorted(glob.glob(os.path.join(DECKS, "regression_*.msp")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def compare(sim, ref, fluid_only=False):
    """Populations f^N and rho/u of the GPU simulation vs the oracle run."""
    o = ref.o
    fo = o.node_f()
    fg = sim.dom.get_f()
    obst = o.real(o.obst())
    fluid = obst == -1
    mask = fluid if fluid_only else np.ones_like(fluid)
    diff = fg[:, mask] != fo[:, mask]
    return int(diff.sum()), float(np.abs(fg[:, mask] - fo[:, mask]).max()), fluid


@pytest.mark.parametrize("path", DECK_FILES, ids=[os.path.basename(p) for p in DECK_FILES])
def test_regression_deck_bitwise(path):
    d = dk.load_msp(path)
    say("deck", os.path.basename(path))
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64")
    # flags are bit-exact
    assert np.array_equal(sim.dom.get_flags(), ref.o.real(ref.o.obst()))
    steps = d.end_iteration
    ref.step(steps)
    sim.run(steps)
    nbad, maxdiff, fluid = compare(sim, ref, fluid_only=True)
    assert nbad == 0, f"{nbad} fluid populations differ, max |diff| {maxdiff}"
    # walls too (no obstacle touches a periodic face in these decks)
    nbad_all, maxdiff_all, _ = compare(sim, ref)
    assert nbad_all == 0, f"{nbad_all} populations differ (walls), max {maxdiff_all}"
    # macro fields written by the last step (macro_variables of f^{N-1}) and the checker
    rho, u = sim.dom.get_macro()
    assert np.array_equal(rho[fluid], ref.o.real(ref.o.m0())[fluid])
    assert np.array_equal(u[:, fluid], ref.o.real(ref.o.m1())[:, fluid])
    s = sim.state()
    so = ref.final_state()
    # sum order differs (sequential on the CPU, tree on the GPU): relative 1e-12
    assert abs(s["sum_density"] - so[0]) <= 1e-12 * abs(so[0])
    assert s["min_velocity_norm"] == so[1] and s["max_velocity_norm"] == so[2]
    assert sim.check() == []
    sim.dom.close()


def _couette_deck(n, tau=0.8, U=0.05):
    txt = f"""
do_domain:
  - domain:
     bounds: [[0,0,0],[{n - 1},{n - 1},{n - 1}]]
     cell_dims: [{n - 1},{n - 1},{n - 1}]
     periodic: [true, true, false]
set_lbm_parameters:
  - lbm_parameters:
     tau: {tau}
     nuth: {(tau - 0.5) / 3}
pre_stream_bcs:
  - pre_bounce_back
  - cavity_z_l:
     U: [{U}, 0, 0]
post_stream_bcs:
  - post_bounce_back
"""
    return dk.loads_msp(txt, "couette")


def test_seeded_couette_bitwise():
    d = _couette_deck(20)
    f0 = seeded_equilibrium(d.nodes, 12345)
    say("couette 20^3 seeded")
    ref = orc.OracleRun(d, init_f=f0)
    sim = Simulation(d, storage="f64", init_f=f0)
    ref.step(150)
    sim.run(150)
    nbad, maxdiff, _ = compare(sim, ref)
    assert nbad == 0, (nbad, maxdiff)
    sim.dom.close()


def test_couette_64_cubed_2000_steps():
    # the 64^3 Couette configuration: rho = 1 + U(-1e-3, 1e-3) seeded, u = 0,
    # f = f_eq, 2000 steps; populations bit-identical to the oracle and the
    # plane-averaged profile on the startup series (walls half a spacing
    # outside the end nodes, moving at U (1 + 0.5/(L-1)))
    from test_oracle import couette_startup
    d = _couette_deck(64)
    f0 = seeded_equilibrium(d.nodes, 20240601)
    say("couette 64^3 seeded, 2000 steps")
    ref = orc.OracleRun(d, init_f=f0)
    sim = Simulation(d, storage="f64", init_f=f0)
    ref.step(2000)
    sim.run(2000)
    nbad, maxdiff, _ = compare(sim, ref)
    assert nbad == 0, (nbad, maxdiff)
    rho, u = sim.dom.macro()
    ux = u[0].mean(axis=(1, 2))
    ana = couette_startup(np.arange(64) + 0.5, 2000, 64, 0.05 * (1 + 0.5 / 63), 0.1)
    say("max |ux - startup series| / U", float(np.abs(ux - ana).max() / 0.05))
    assert np.abs(ux - ana).max() < 1e-3 * 0.05
    sim.dom.close()


def _porous_deck(nx, ny, nz, nballs, seed, collision="mrt"):
    rng = np.random.default_rng(seed)
    lines = [f"""
do_domain:
  - domain:
     bounds: [[0,0,0],[{nx - 1},{ny - 1},{nz - 1}]]
     cell_dims: [{nx - 1},{ny - 1},{nz - 1}]
     periodic: [false, true, false]
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.7
     nuth: {0.2 / 3}
     Fext: [2.0e-5, 0, 0]
collision: {collision}
pre_stream_bcs:
  - pre_bounce_back
  - wall_bounce_back
post_stream_bcs:
  - post_bounce_back
set_obstacles:
"""]
    for k in range(nballs):
        # balls stay off the periodic y faces: there the reference's wall_bounce_back
        # reads ghost nodes that still hold the previous step's values
        r = rng.uniform(2.0, 3.5)
        c = (rng.uniform(0, nx - 1), rng.uniform(r + 2, ny - 3 - r), rng.uniform(0, nz - 1))
        lines.append(f"  - register_solid_ball:\n     id: {k}\n     center: [{c[0]},{c[1]},{c[2]}]\n     radius: {r}\n")
    return dk.loads_msp("".join(lines), "porous")


@pytest.mark.parametrize("collision", ["mrt", "bgk"])
def test_porous_channel_bitwise(collision):
    d = _porous_deck(24, 20, 18, 8, 3, collision)
    say("porous", collision)
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64")
    assert np.array_equal(sim.dom.get_flags(), ref.o.real(ref.o.obst()))
    assert (sim.dom.get_flags() == -2).sum() > 100
    ref.step(60)
    sim.run(60)
    nbad, maxdiff, _ = compare(sim, ref, fluid_only=True)
    assert nbad == 0, (nbad, maxdiff)
    nbad, maxdiff, _ = compare(sim, ref)
    assert nbad == 0, (nbad, maxdiff)
    sim.dom.close()


def _channel_deck(nx, ny, nz, nballs, seed, collision="mrt", outlet="outflow"):
    """Porous channel: density inlet at x=0, outflow (or a density outlet) at
    x=L, bounce-back y/z walls, seeded non-overlapping balls, MRT with
    s_nu = 1/0.7 and the other rates 1.2 (reference weps/wepsj/wxx)."""
    rng = np.random.default_rng(seed)
    balls = []
    while len(balls) < nballs:
        r = rng.uniform(1.5, 3.0)
        c = np.array([rng.uniform(r + 3, nx - 4 - r), rng.uniform(0, ny - 1), rng.uniform(0, nz - 1)])
        if all(np.linalg.norm(c - b[0]) > r + b[1] for b in balls):
            balls.append((c, r))
    out = ("  - outflow:\n     regions: [plan_yz_l]\n" if outlet == "outflow" else
           "  - pressure_inlet:\n     rho: 1.0\n     regions: [plan_yz_l]\n")
    lines = [f"""
do_domain:
  - domain:
     bounds: [[0,0,0],[{nx - 1},{ny - 1},{nz - 1}]]
     cell_dims: [{nx - 1},{ny - 1},{nz - 1}]
     periodic: [false, false, false]
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.7
     nuth: {0.2 / 3}
collision: {collision}
mrt_rates: [1.2, 1.2, 1.2, {1 / 0.7}, 1.2, {1 / 0.7}, 1.2, 0.0, {-475.0 / 63.0}, 0.0]
pre_stream_bcs:
  - pre_bounce_back
  - wall_bounce_back
post_stream_bcs:
  - post_bounce_back
boundary_conditions:
  - pressure_inlet:
     rho: 1.001
     regions: [plan_yz_0]
{out}set_obstacles:
"""]
    for k, (c, r) in enumerate(balls):
        lines.append(f"  - register_solid_ball:\n     id: {k}\n     center: [{c[0]},{c[1]},{c[2]}]\n     radius: {r}\n")
    return dk.loads_msp("".join(lines), "channel")


@pytest.mark.parametrize("collision,outlet,pattern", [("mrt", "outflow", "ab"), ("bgk", "outflow", "ab"),
                                                      ("mrt", "pressure", "ab"), ("mrt", "outflow", "aa")])
def test_porous_channel_open_boundaries_bitwise(collision, outlet, pattern):
    d = _channel_deck(40, 18, 16, 6, 21, collision, outlet)
    say("channel", collision, outlet, pattern)
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64", pattern=pattern)
    for n in (1, 12, 47):
        ref.step(n)
        sim.run(n)
        nbad, maxdiff, _ = compare(sim, ref, fluid_only=True)
        assert nbad == 0, (n, nbad, maxdiff)
    if pattern == "ab":
        nbad, maxdiff, _ = compare(sim, ref)
        assert nbad == 0, (nbad, maxdiff)
    rho, u = sim.dom.macro()
    assert np.abs(rho[1:-1, 1:-1, 0][sim.dom.get_flags()[1:-1, 1:-1, 0] == -1] - 1.001).max() < 1e-12
    assert u[0].max() > 0  # the inlet drives a flow
    sim.dom.close()


@pytest.mark.parametrize("pattern", ["ab", "aa"])
def test_wide_address_variant_bitwise(pattern, monkeypatch):
    # the 64-bit-address bulk variant (domains above 2^32 bytes per population
    # plane) forced onto a small porous channel: the same populations
    monkeypatch.setenv("HLBM_FORCE_WIDE", "1")
    d = _porous_deck(30, 20, 18, 6, 4, "bgk")
    say("wide", pattern)
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64", pattern=pattern)
    monkeypatch.delenv("HLBM_FORCE_WIDE")
    ref.step(23)
    sim.run(23)
    nbad, maxdiff, _ = compare(sim, ref, fluid_only=True)
    assert nbad == 0, (nbad, maxdiff)
    sim.dom.close()


# quadric obstacles (register_quadrics): a cylinder across the closed z axis,
# an ellipsoid and a narrow double cone, all clear of the periodic x/y faces
QUADRIC_DECK = """
do_domain:
  - domain:
     bounds: [[0,0,0],[0.5,0.2,0.3]]
     cell_dims: [50, 20, 30]
     periodic: [true, true, false]
set_lbm_parameters:
  - lbm_parameters:
     Fext: [1.3, 0, 0]
     nuth: 1e-3
     tau: 0.65
set_obstacles:
  - register_quadrics:
     id: 0
     quadrics: cylz
     transform:
        - scale: [ 0.03, 0.03, 1 ]
        - translate: [ 0.12, 0.1, 0 ]
  - register_quadrics:
     id: 1
     quadrics: sphere
     transform:
        - scale: [ 0.04, 0.05, 0.04 ]
        - translate: [ 0.27, 0.1, 0.15 ]
  - register_quadrics:
     id: 2
     quadrics: conez
     transform:
        - scale: [ 0.01, 0.01, 0.05 ]
        - zrot: 0.3
        - translate: [ 0.4, 0.1, 0.2 ]
boundary_conditions:
  - neumann:
     U: [0,0,0]
     regions: [plan_xy_l]
  - neumann:
     U: [0,0,0]
     regions: [plan_xy_0]
pre_stream_bcs:
  - wall_bounce_back
global:
   simulation_end_iteration: 30
   simulation_print_log_freq: -1
"""


@pytest.mark.parametrize("pattern", ["ab", "aa"])
def test_quadric_obstacles_bitwise(pattern):
    d = dk.loads_msp(QUADRIC_DECK, "quadrics")
    say("quadrics", pattern)
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64", pattern=pattern)
    flags = sim.dom.get_flags()
    assert np.array_equal(flags, ref.o.real(ref.o.obst()))
    assert (flags == -2).sum() > 1000
    ref.step(30)
    sim.run(30)
    nbad, maxdiff, _ = compare(sim, ref, fluid_only=True)
    assert nbad == 0, (nbad, maxdiff)
    sim.dom.close()


# start_iteration: set_distribution on a slab at the start of every step
# (example/ball.msp:23-26), an MRT channel with a ball off the periodic faces
START_DECK = """
do_domain:
  - domain:
     bounds: [[0,0,0],[0.1,0.03,0.04]]
     cell_dims: [50, 15, 20]
     periodic: [false, true, true]
set_lbm_parameters:
  - lbm_parameters:
     nuth: 1e-4
set_obstacles:
  - register_solid_ball:
     id: 0
     center: [0.03, 0.015, 0.02]
     radius: 0.005
set_distributions:
  - set_distribution:
     value: 1.0
     bounds: [[0,0,0],[0.1,0.03,0.04]]
start_iteration:
  - set_distribution:
     value: 1.2
     bounds: [[0,0,0.0], [0.004,0.03,0.04]]
collision: mrt
pre_stream_bcs:
  - wall_bounce_back
global:
   simulation_end_iteration: 40
   simulation_print_log_freq: -1
"""


@pytest.mark.parametrize("pattern", ["ab", "aa"])
def test_start_iteration_set_distribution_bitwise(pattern):
    d = dk.loads_msp(START_DECK, "start")
    assert len(d.start_iteration) == 1
    say("start_iteration", pattern)
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64", pattern=pattern)
    for n in (1, 39):  # one step, then a multi-step call
        ref.step(n)
        sim.run(n)
        nbad, maxdiff, _ = compare(sim, ref, fluid_only=True)
        assert nbad == 0, (n, nbad, maxdiff)
    # the slab drives a density wave down the channel
    rho, _ = sim.dom.macro()
    assert rho[:, :, 1].mean() > 1.05 and abs(rho[:, :, 45].mean() - 1.0) < 0.05
    sim.dom.close()


def test_start_iteration_reduced_precision_tracks_fp64():
    d = dk.loads_msp(START_DECK, "start")
    ref = Simulation(d, storage="f64")
    lo = Simulation(d, storage="f32")
    ref.run(40)
    lo.run(40)
    r64, u64 = ref.dom.macro()
    r32, u32 = lo.dom.macro()
    fluid = ref.dom.get_flags() == -1
    assert np.abs(r32 - r64)[fluid].max() < 1e-5
    assert np.linalg.norm((u32 - u64)[:, fluid]) <= 1e-4 * np.linalg.norm(u64[:, fluid])
    ref.dom.close()
    lo.dom.close()


# init_double_couette (example/lbm_prepro_double_couette.msp at 31^3)
DOUBLE_COUETTE_DECK = """
do_domain:
  - domain:
     bounds:    [ [0,0,0],[0.03,0.03,0.03] ]
     cell_dims: [ 30 , 30 , 30 ]
     periodic:  [ true, true, false ]
set_distributions:
  - init_double_couette:
     U: [0.001,0,0]
     dimension: Z
set_lbm_parameters:
  - lbm_parameters:
     Fext: [0,0,0]
     nuth: 1e-2
     celerity: 1.0
boundary_conditions:
  - neumann:
     U: [-0.001,0,0]
     regions: [plan_xy_l]
  - neumann:
     U: [0.001,0,0]
     regions: [plan_xy_0]
global:
   simulation_end_iteration: 20
   simulation_print_log_freq: -1
"""


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_init_double_couette(storage):
    d = dk.loads_msp(DOUBLE_COUETTE_DECK, "double_couette")
    say("double couette", storage)
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage=storage)
    ref.o.macro_variables()
    rho, u = sim.dom.macro()
    if storage == "f64":
        assert np.array_equal(sim.dom.get_f(), ref.o.node_f())
        ref.step(20)
        sim.run(20)
        nbad, maxdiff, _ = compare(sim, ref)
        assert nbad == 0, (nbad, maxdiff)
    else:
        uo = ref.o.real(ref.o.m1())
        assert np.abs(u - uo).max() < 1e-9 and np.abs(rho - 1.0).max() < 1e-6
    sim.dom.close()


def test_outflow_after_an_operator_on_its_interior_layer_is_rejected():
    d = _channel_deck(20, 10, 10, 0, 1, "bgk")
    d.boundary_conditions.insert(0, ("lid_driven_cavity", {"U": (0.01, 0.0, 0.0), "regions": ["plan_xy_l"]}))
    with pytest.raises(_lib.HLBMError):
        Simulation(d, storage="f64")


def test_operator_order_lid_before_post_bounce_back():
    # lid_driven_cavity listed before post_bounce_back: post_bounce_back then
    # overwrites the lid's incoming populations on the closed faces
    txt = """
do_domain:
  - domain:
     bounds: [[0,0,0],[11,9,13]]
     cell_dims: [11, 9, 13]
     periodic: [false, false, false]
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.9
     nuth: 0.13333333333333333
pre_stream_bcs:
  - pre_bounce_back
  - cavity_z_0:
     U: [0.0, 0.02, 0]
post_stream_bcs:
  - lid_driven_cavity:
     U: [0.03, 0.0, 0]
     regions: [plan_xy_l, plan_yz_0]
  - post_bounce_back
boundary_conditions:
  - neumann:
     U: [0.0, 0.01, 0]
     regions: [plan_xz_0]
"""
    d = dk.loads_msp(txt, "order")
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64")
    ref.step(40)
    sim.run(40)
    nbad, maxdiff, _ = compare(sim, ref)
    assert nbad == 0, (nbad, maxdiff)
    sim.dom.close()


def test_reference_layout_ghost_fill_matches_update_ghost():
    d = _couette_deck(9)
    f0 = seeded_equilibrium(d.nodes, 99)
    ref = orc.OracleRun(d, init_f=f0)
    sim = Simulation(d, storage="f64", init_f=f0)
    ref.step(7)
    sim.run(7)
    ref.o.update_ghost()
    fr = ref.o.f()
    g = np.zeros_like(fr)
    sim.dom.get_f(layout=_lib.LAYOUT_REF, out=g)
    # Real nodes and the periodic (x, y) ghost layers of the Real z range
    sl = (slice(None), slice(2, -2), slice(None), slice(None))
    assert np.array_equal(g[sl], fr[sl])
    sim.dom.close()


def test_macro_variables_on_current_state():
    d = _couette_deck(12)
    f0 = seeded_equilibrium(d.nodes, 5)
    ref = orc.OracleRun(d, init_f=f0)
    sim = Simulation(d, storage="f64", init_f=f0)
    ref.step(11)
    sim.run(11)
    ref.o.macro_variables()
    rho, u = sim.dom.macro()
    assert np.array_equal(rho, ref.o.real(ref.o.m0()))
    assert np.array_equal(u, ref.o.real(ref.o.m1()))
    sim.dom.close()


def test_step_chunking_is_invariant():
    # step(a) + step(b) == step(a + b), also across population read-backs
    d = _couette_deck(10)
    f0 = seeded_equilibrium(d.nodes, 77)
    a = Simulation(d, storage="f64", init_f=f0)
    b = Simulation(d, storage="f64", init_f=f0)
    a.run(30)
    for k in (1, 4, 10, 15):
        b.run(k)
        b.dom.get_f()
    assert np.array_equal(a.dom.get_f(), b.dom.get_f())
    assert a.dom.timestep() == b.dom.timestep() == 30
    a.dom.close()
    b.dom.close()


def test_unsupported_operator_placement_is_rejected():
    dom = Domain((8, 8, 8), (True, True, False), 0.8)
    with pytest.raises(_lib.HLBMError):
        dom.set_bcs([("pre_stream_bcs", "neumann", {"U": (0, 0, 0), "regions": ["plan_xy_0"]})])
    with pytest.raises(_lib.HLBMError):
        dom.set_flags(np.full(8 * 8 * 8, 3, dtype=np.int32))
    dom.close()

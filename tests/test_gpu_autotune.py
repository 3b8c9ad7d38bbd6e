"""Launch-shape autotuner: it leaves the domain's state alone, every shape it
can pick gives the same populations bit for bit, and a second domain of the
same shape takes the cached choice."""
import sys

import numpy as np
import pytest

from hippolbm_amd.engine import Simulation

from test_gpu_parity import _channel_deck, _couette_deck, _porous_deck, seeded_equilibrium

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("storage,pattern", [("f64", "ab"), ("f16", "ab"), ("f32", "aa")])
def test_autotune_keeps_state_and_results(storage, pattern):
    d = _couette_deck(40)
    f0 = seeded_equilibrium(d.nodes, 5)
    ref = Simulation(d, storage=storage, init_f=f0, pattern=pattern)
    ref.run(9)
    sim = Simulation(d, storage=storage, init_f=f0, pattern=pattern)
    sim.run(4)
    info = sim.dom.autotune(reps=2)
    say(storage, pattern, info)
    assert info["block"][0] * info["block"][1] % 64 == 0 and info["npt"] in (1, 2) and info["variant"] in (0, 1)
    assert info["candidates"] == 36
    sim.run(5)
    assert np.array_equal(sim.dom.get_f(), ref.dom.get_f())
    # the same shape again: taken from the cache
    other = Simulation(d, storage=storage, init_f=f0, pattern=pattern)
    again = other.dom.autotune(reps=2)
    assert again["candidates"] == 0 and again["block"] == info["block"] and again["npt"] == info["npt"]
    assert again["variant"] == info["variant"]
    for s in (ref, sim, other):
        s.dom.close()


@pytest.mark.parametrize("shape", [((64, 1), 1, 0), ((64, 4), 2, 1), ((256, 1), 2, 0), ((128, 1), 1, 1),
                                   ((128, 2), 1, 1, 1), ((64, 4), 1, 0, 1)])
def test_every_launch_shape_gives_the_same_populations(shape):
    d = _porous_deck(37, 21, 19, 5, 7, "bgk")
    a = Simulation(d, storage="f64")
    b = Simulation(d, storage="f64")
    b.dom.set_launch(*shape)
    a.run(13)
    b.run(13)
    assert np.array_equal(a.dom.get_f(), b.dom.get_f())
    a.dom.close()
    b.dom.close()


@pytest.mark.parametrize("pattern", ["ab", "aa"])
@pytest.mark.parametrize("shape", [((64, 1), 1, 0), ((64, 4), 2, 1), ((256, 1), 2, 0), ((128, 2), 1, 1, 1)])
def test_launch_shapes_with_boundary_program_planes(shape, pattern):
    # density inlet + outflow planes (the boundary kernel) beside bulk tiles of
    # one or two rows per thread
    d = _channel_deck(40, 18, 14, 4, 3, "mrt")
    a = Simulation(d, storage="f64", pattern=pattern)
    b = Simulation(d, storage="f64", pattern=pattern)
    b.dom.set_launch(*shape)
    a.run(11)
    b.run(11)
    assert np.array_equal(a.dom.get_f(), b.dom.get_f())
    a.dom.close()
    b.dom.close()


@pytest.mark.parametrize("pattern,steps_before", [("ab", 4), ("aa", 4), ("aa", 3)])
def test_placement_relocation_keeps_the_state(monkeypatch, pattern, steps_before):
    # 128^3 FP64: 319 MB of populations per buffer, so the autotuner checks the
    # placement; forced, it moves the state to three fresh buffers in turn (AA:
    # after an even and an odd step count, the two layouts of the lattice)
    d = _couette_deck(128)
    f0 = seeded_equilibrium(d.nodes, 9)
    ref = Simulation(d, storage="f64", init_f=f0, pattern=pattern)
    ref.run(9)
    sim = Simulation(d, storage="f64", init_f=f0, pattern=pattern)
    sim.run(steps_before)
    monkeypatch.setenv("HLBM_PLACEMENT", "force")
    info = sim.dom.autotune(reps=1)
    monkeypatch.delenv("HLBM_PLACEMENT")
    say("placement", info)
    assert info["relocated"] == 3
    sim.run(9 - steps_before)
    assert np.array_equal(sim.dom.get_f(), ref.dom.get_f())
    rho, u = sim.dom.get_macro()
    rho_r, u_r = ref.dom.get_macro()
    assert np.array_equal(rho, rho_r) and np.array_equal(u, u_r)
    ref.dom.close()
    sim.dom.close()

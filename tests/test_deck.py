"""Deck front end (no GPU): the open-boundary operators and the MRT-rate
extension parse and validate, and the seeded workload generators are
deterministic and well formed."""
import numpy as np
import pytest

from hippolbm_amd import deck as dk
from hippolbm_amd import workloads
from hippolbm_amd.engine import bc_entries

BASE = """
do_domain:
  - domain:
     bounds: [[0,0,0],[9,7,5]]
     cell_dims: [9,7,5]
     periodic: [false, true, false]
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.7
     nuth: 0.0666666666666667
collision: mrt
"""


def test_open_boundary_operators_parse():
    d = dk.loads_msp(BASE + """
mrt_rates: [1.2, 1.2, 1.2, 1.4285714285714286, 1.2, 1.4285714285714286, 1.2, 0, -7.5396825396825395, 0]
boundary_conditions:
  - pressure_inlet:
     rho: 1.001
     regions: [plan_yz_0]
  - outflow:
     regions: [plan_yz_l]
""")
    assert d.boundary_conditions == [("pressure_inlet", {"rho": 1.001, "regions": ["plan_yz_0"]}),
                                     ("outflow", {"regions": ["plan_yz_l"]})]
    assert d.mrt_rates[3] == pytest.approx(1 / 0.7)
    e = bc_entries(d)
    assert e[0] == ("boundary_conditions", "pressure_inlet",
                    {"rho": 1.001, "regions": ["plan_yz_0"], "U": (1.001, 0.0, 0.0)})


@pytest.mark.parametrize("extra,msg", [
    ("boundary_conditions:\n  - pressure_inlet:\n     regions: [plan_yz_0]\n", "rho"),
    ("boundary_conditions:\n  - outflow:\n     regions: [plan_xx_0]\n", "Traversal"),
    ("mrt_rates: [1, 2, 3]\n", "mrt_rates"),
])
def test_bad_open_boundary_decks_are_rejected(extra, msg):
    with pytest.raises(dk.DeckError, match=msg):
        dk.loads_msp(BASE + extra)


def test_sphere_packing_is_seeded_and_non_overlapping():
    a = workloads.pack_spheres((64, 32, 32), 20, (2.0, 5.0), seed=3, margin=(2.0, 0.0, 0.0))
    b = workloads.pack_spheres((64, 32, 32), 20, (2.0, 5.0), seed=3, margin=(2.0, 0.0, 0.0))
    assert all(np.array_equal(p[0], q[0]) and p[1] == q[1] for p, q in zip(a, b))
    for i, (c, r) in enumerate(a):
        assert 2.0 + r <= c[0] <= 63 - 2.0 - r and r <= c[1] <= 31 - r and r <= c[2] <= 31 - r
        for c2, r2 in a[i + 1:]:
            assert np.linalg.norm(c - c2) >= r + r2


def test_porous_channel_deck_matches_the_configuration():
    d = workloads.porous_channel()
    assert d.nodes == (256, 128, 128) and len(d.obstacles) == 150
    radii = [o["radius"] for o in d.obstacles.values()]
    assert 5.0 <= min(radii) and max(radii) <= 12.0
    assert d.collision == "mrt" and d.mrt_rates[3] == pytest.approx(1 / 0.7) and d.mrt_rates[0] == 1.2
    assert d.pre_stream_bcs == [("pre_bounce_back", {}), ("wall_bounce_back", {})]
    assert [n for n, _ in d.boundary_conditions] == ["pressure_inlet", "outflow"]
    p = d.params
    assert p["tau"] == pytest.approx(0.7) and p["celerity"] == pytest.approx(1.0)


def test_poiseuille_deck_force_in_lattice_units():
    d = workloads.poiseuille((32, 8, 16))
    assert d.params["Fext"][0] == pytest.approx(1e-6)
    assert d.periodic == (True, True, False)


def test_python_reader_refuses_dangling_includes_and_graph_blocks(tmp_path):
    from conftest import REGRESSION_DECKS
    base = open(REGRESSION_DECKS[0]).read()
    p = tmp_path / "d.msp"
    p.write_text("includes: [nope.msp]\n" + base)
    with pytest.raises(dk.DeckError, match="not found"):
        dk.load_msp(str(p))
    # an included block is used when the deck has none of that name; the deck's own block replaces it otherwise
    import yaml
    doc = yaml.safe_load(base)
    doc.pop("global")
    (tmp_path / "inc.msp").write_text("global:\n   simulation_end_iteration: 7\n")
    p.write_text("includes: [inc.msp]\n" + yaml.safe_dump(doc, default_flow_style=None))
    assert dk.load_msp(str(p)).end_iteration == 7
    p.write_text("includes: [inc.msp]\n" + base)
    assert dk.load_msp(str(p)).end_iteration == 100
    for block in ("compute_loop:\n  loop: true\n  body:\n    - macro_variables\n",
                  "my_batch:\n  condition: x\n  body:\n    - log\n", "mystery_block: 3\n"):
        with pytest.raises(dk.DeckError):
            dk.loads_msp(base + "\n" + block)


def test_engine_parameters_equal_the_oracle_restatement(decks_dir):
    """Deck.params (the engine's lbm_parameters) against the oracle's own C
    restatement of lbm_parameters.cpp:73-99, bit for bit, for the regression
    decks and for each of the three definitions (dt, tau, celerity)."""
    import glob
    import os

    import oracle as orc
    from hippolbm_amd import deck as dk
    cases = [dk.load_msp(p) for p in sorted(glob.glob(os.path.join(decks_dir, "regression_*.msp")))]
    assert len(cases) == 7
    for lbm in ({"dt": 2.5e-4, "nuth": 1e-3, "Fext": [0.3, 0, -1e-2]},
                {"tau": 0.65, "nuth": 1e-3, "Fext": [1.3, 0, 0]},
                {"celerity": 10.0, "nuth": 1e-3, "Fext": [0, 2.0, 0]},
                {"nuth": 1e-4}):
        d = dk.loads_msp("do_domain:\n  - domain:\n     bounds: [[0,0,0],[0.3,0.2,0.1]]\n"
                         "     cell_dims: [30, 20, 10]\n     periodic: [true, false, false]\n", "params")
        d.lbm = dict(lbm)
        cases.append(d)
    for d in cases:
        lb = d.lbm
        want = orc.lbm_parameters(d.dx, lb.get("nuth", 1e-4), lb.get("celerity", 1.0), lb.get("tau"),
                                  lb.get("dt", 0.0), lb.get("Fext", (0.0, 0.0, 0.0)))
        got = d.params
        for k in ("tau", "nu", "dtLB", "celerity"):
            assert got[k] == want[k], (d.name, k, got[k], want[k])
        assert tuple(got["Fext"]) == tuple(want["Fext"]), d.name

"""Cases of the golden fixtures (tests/golden/populations.npz, golden.json).

Each small case is a deck on a grid of a few hundred nodes, started from a
seeded equilibrium and run for a few steps; the fixture holds the oracle's
FP64 populations (node layout, (19, nz, ny, nx)) and obstacle flags after
those steps.  The hash cases run the 64^3 moving-lid Couette (the reference's
CPU-runnable configuration) and keep a SHA-256 of the populations after 1, 10
and 100 steps.  make_golden.py writes the fixtures from the oracle; the tests
check the oracle (CPU) and the HIP engine (GPU) against them.
"""
import hashlib

import numpy as np


def seeded_equilibrium(nodes, seed, amp=1e-3):
    """rho = 1 + U(-amp, amp) (splitmix64 -> 53-bit uniform), u = 0, f = f_eq."""
    n = nodes[0] * nodes[1] * nodes[2]
    with np.errstate(over="ignore"):
        # state_k = seed + (k+1)*gamma (mod 2^64), then the splitmix64 finalizer
        z = np.uint64(seed) + (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    out = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    rho = 1.0 + (2.0 * out - 1.0) * amp
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    return (w[:, None] * rho[None, :]).reshape(19, nodes[2], nodes[1], nodes[0])


def f_hash(f):
    """SHA-256 of FP64 populations, (19, nz, ny, nx) C order, little endian."""
    a = np.ascontiguousarray(f, dtype="<f8")
    return hashlib.sha256(a.tobytes()).hexdigest()


def _domain(n, periodic, bounds_max=None):
    hi = bounds_max or [v - 1 for v in n]
    per = ", ".join("true" if p else "false" for p in periodic)
    return f"""
do_domain:
  - domain:
     bounds: [[0,0,0],[{hi[0]},{hi[1]},{hi[2]}]]
     cell_dims: [{n[0] - 1},{n[1] - 1},{n[2] - 1}]
     periodic: [{per}]
"""


# name -> (deck text, nodes, seed, steps)
CASES = {
    "couette_moving_lid": (_domain((6, 6, 8), (True, True, False)) + """
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.8
     nuth: 0.1
pre_stream_bcs:
  - pre_bounce_back
  - cavity_z_l:
     U: [0.05, 0.01, 0]
post_stream_bcs:
  - post_bounce_back
""", (6, 6, 8), 11, 12),
    "neumann_shear": (_domain((6, 5, 7), (True, True, False)) + """
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.7
     nuth: 0.2
     Fext: [1.0e-5, 0, 0]
boundary_conditions:
  - neumann:
     U: [0.02, 0.005, 0]
     regions: [plan_xy_l]
  - neumann:
     U: [0, 0, 0]
     regions: [plan_xy_0]
""", (6, 5, 7), 12, 12),
    "lid_cavity_mrt": (_domain((7, 6, 6), (False, False, False)) + """
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.9
     nuth: 0.1
collision: mrt
boundary_conditions:
  - lid_driven_cavity:
     U: [0.03, 0, 0]
     regions: [plan_xy_l]
""", (7, 6, 6), 13, 12),
    "ball_wall_bounce_back": (_domain((9, 7, 7), (False, False, False)) + """
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.75
     nuth: 0.1
     Fext: [2.0e-5, 0, 0]
collision: mrt
pre_stream_bcs:
  - pre_bounce_back
  - wall_bounce_back
post_stream_bcs:
  - post_bounce_back
set_obstacles:
  - register_solid_ball:
     id: 0
     center: [4.2, 3.1, 2.9]
     radius: 1.7
  - register_solid_wall:
     id: 1
     bounds: [[6.5, -1, -1], [7.5, 2.5, 8]]
""", (9, 7, 7), 14, 12),
    "pressure_outflow_channel": (_domain((10, 5, 5), (False, False, False)) + """
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.7
     nuth: 0.2
collision: mrt
mrt_rates: [1.2, 1.2, 1.2, 1.4285714285714286, 1.2, 1.4285714285714286, 1.2, 0.0, -7.5396825396825395, 0.0]
pre_stream_bcs:
  - pre_bounce_back
post_stream_bcs:
  - post_bounce_back
boundary_conditions:
  - pressure_inlet:
     rho: 1.01
     regions: [plan_yz_0]
  - outflow:
     regions: [plan_yz_l]
""", (10, 5, 5), 15, 12),
    "quadric_start_iteration": (_domain((9, 6, 6), (True, False, False)) + """
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.8
     nuth: 0.1
set_obstacles:
  - register_quadrics:
     id: 0
     quadrics: cylz
     transform:
        - scale: [1.6, 1.4, 1]
        - translate: [4.3, 2.6, 0]
start_iteration:
  - set_distribution:
     value: 1.05
     bounds: [[0, 0, 0], [1.5, 5, 5]]
pre_stream_bcs:
  - wall_bounce_back
""", (9, 6, 6), 16, 12),
}

# the 64^3 moving-lid Couette, seeded with splitmix64 seed 1: hashes after these steps
HASH_NODES = (64, 64, 64)
HASH_SEED = 1
HASH_STEPS = (1, 10, 100)

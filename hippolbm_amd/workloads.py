"""Seeded synthetic workloads: the porous-channel deck (a density inlet, an
outflow, bounce-back side walls and packed spheres), the plane Couette deck
and the plane Poiseuille deck, as hippoLBM operator lists (``Deck`` objects).

Sphere placement follows the reference's ``register_solid_ball`` semantics
(``src/obstacle/register_obstacles.cpp``, rasterized by ``update_obstacles.cu:32-72``):
a node is solid when ``|p - c|^2 <= r^2`` with ``p = node * dx``.
"""
from __future__ import annotations

import numpy as np

from . import deck as dk

# MRT rates [s1, s2, s4, s9, s10, s13, s16, weps, wepsj, wxx]: the viscous rates
# s9 = s13 = 1/tau_nu, every other relaxation rate `other`; weps, wepsj, wxx keep
# the reference's values (mrt.hpp:44-47)
def mrt_rates(tau_nu=0.7, other=1.2):
    s_nu = 1.0 / tau_nu
    return [other, other, other, s_nu, other, s_nu, other, 0.0, -475.0 / 63.0, 0.0]


def pack_spheres(dims, n, r_range, seed, margin=(0.0, 0.0, 0.0), max_tries=200000):
    """n non-overlapping spheres, radii uniform in r_range, centres uniform in the
    box [margin + r, dims - 1 - margin - r] per axis; deterministic for a seed.
    Returns [(centre (3,), radius)]."""
    rng = np.random.default_rng(seed)
    out = []
    tries = 0
    while len(out) < n:
        tries += 1
        if tries > max_tries:
            raise ValueError(f"could not place {n} spheres in {dims} (placed {len(out)})")
        r = float(rng.uniform(*r_range))
        lo = np.array(margin) + r
        hi = np.array(dims, dtype=float) - 1 - np.array(margin) - r
        if np.any(hi < lo):
            continue
        c = lo + rng.uniform(size=3) * (hi - lo)
        if all(np.linalg.norm(c - c2) >= r + r2 for c2, r2 in out):
            out.append((c, r))
    return out


def _ball_lines(spheres):
    return "".join(f"  - register_solid_ball:\n     id: {k}\n     center: [{float(c[0])!r},{float(c[1])!r},"
                   f"{float(c[2])!r}]\n     radius: {float(r)!r}\n" for k, (c, r) in enumerate(spheres))


def porous_channel(nx=256, ny=128, nz=128, n_spheres=150, r_range=(5.0, 12.0), seed=2, collision="mrt",
                   rho_in=1.001, tau=0.7, outlet="outflow"):
    """Porous channel: density inlet rho_in on plan_yz_0, zero-gradient outflow
    (or a density-1 outlet) on plan_yz_l, bounce-back y/z walls
    (pre/post_bounce_back), wall_bounce_back on the packed spheres, MRT with
    s_nu = 1/tau and the other rates 1.2."""
    spheres = pack_spheres((nx, ny, nz), n_spheres, r_range, seed, margin=(2.0, 0.0, 0.0))
    out = ("  - outflow:\n     regions: [plan_yz_l]\n" if outlet == "outflow" else
           "  - pressure_inlet:\n     rho: 1.0\n     regions: [plan_yz_l]\n")
    rates = ", ".join(repr(v) for v in mrt_rates(tau))
    txt = f"""
do_domain:
  - domain:
     bounds: [[0,0,0],[{nx - 1},{ny - 1},{nz - 1}]]
     cell_dims: [{nx - 1},{ny - 1},{nz - 1}]
     periodic: [false, false, false]
set_lbm_parameters:
  - lbm_parameters:
     tau: {tau!r}
     nuth: {(tau - 0.5) / 3!r}
collision: {collision}
mrt_rates: [{rates}]
pre_stream_bcs:
  - pre_bounce_back
  - wall_bounce_back
post_stream_bcs:
  - post_bounce_back
boundary_conditions:
  - pressure_inlet:
     rho: {rho_in!r}
     regions: [plan_yz_0]
{out}set_obstacles:
{_ball_lines(spheres)}"""
    return dk.loads_msp(txt, f"porous_channel_{nx}x{ny}x{nz}")


def couette(n=(64, 64, 64), U=0.05, tau=0.8):
    """Plane Couette: periodic x/y, on-node bounce-back at z=0, moving wall
    u_x = U at z = nz-1 (pre_bounce_back + cavity_z_l + post_bounce_back)."""
    nx, ny, nz = n
    txt = f"""
do_domain:
  - domain:
     bounds: [[0,0,0],[{nx - 1},{ny - 1},{nz - 1}]]
     cell_dims: [{nx - 1},{ny - 1},{nz - 1}]
     periodic: [true, true, false]
set_lbm_parameters:
  - lbm_parameters:
     tau: {tau!r}
     nuth: {(tau - 0.5) / 3!r}
collision: bgk
pre_stream_bcs:
  - pre_bounce_back
  - cavity_z_l:
     U: [{U!r}, 0, 0]
post_stream_bcs:
  - post_bounce_back
"""
    return dk.loads_msp(txt, f"couette_{nx}x{ny}x{nz}")


def poiseuille(n=(512, 256, 256), F=1e-6, tau=0.6):
    """Plane Poiseuille: periodic x/y, bounce-back plates at z=0 and z=nz-1,
    body force F along x (lattice units: celerity 1)."""
    nx, ny, nz = n
    txt = f"""
do_domain:
  - domain:
     bounds: [[0,0,0],[{nx - 1},{ny - 1},{nz - 1}]]
     cell_dims: [{nx - 1},{ny - 1},{nz - 1}]
     periodic: [true, true, false]
set_lbm_parameters:
  - lbm_parameters:
     tau: {tau!r}
     nuth: {(tau - 0.5) / 3!r}
     Fext: [{F!r}, 0, 0]
collision: bgk
pre_stream_bcs:
  - pre_bounce_back
post_stream_bcs:
  - post_bounce_back
"""
    return dk.loads_msp(txt, f"poiseuille_{nx}x{ny}x{nz}")

"""TEST INFRASTRUCTURE ONLY -- ctypes binding and deck driver for the CPU FP64 oracle.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module; the product package (``hippolbm_amd``)
never does.

The driver restates hippoLBM's default operator graph:

* initialize: ``data/config/config_initialize.msp:14-32``
  (global, domain, lbm_parameters, build_traversal, define_grid_3dq19,
  init_obst, set_distribution{do_update: false}, set_distributions,
  init_obstacles + set_obstacles, update_obstacles, macro_variables,
  simulation_state, update_ghost + macro_variables (first-timestep dump,
  ``data/config/config_io.msp:20-27``), log);
* one time step: ``data/config/config_compute_loop.msp:23-38``
  (macro_variables, collision, pre_stream_bcs, streaming, post_stream_bcs,
  boundary_conditions);
* the end: ``data/config/main-config.msp:29-32`` + ``config_verification.msp:6-8``
  (simulation_state, check_macro_quantities).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhlbm_oracle.so")
_lib = None

# LBMGridRegion plane order (src/grid/include/hippoLBM/grid/grid_region.hpp:75-77)
PLANES = {"plan_xy_0": 0, "plan_xy_l": 1, "plan_xz_0": 2, "plan_xz_l": 3, "plan_yz_0": 4, "plan_yz_l": 5}

D = ctypes.c_double
I = ctypes.c_int
P = ctypes.c_void_p
DP = ctypes.POINTER(ctypes.c_double)
IP = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_create": (P, [I, I, I, I, I, I, D]),
            "orc_destroy": (None, [P]),
            "orc_dims": (None, [P, IP]),
            "orc_f": (DP, [P]),
            "orc_m0": (DP, [P]),
            "orc_m1": (DP, [P]),
            "orc_obst": (IP, [P]),
            "orc_level": (IP, [P]),
            "orc_threads": (I, []),
            "orc_set_threads": (None, [I]),
            "orc_lbm_parameters": (None, [D, D, D, I, D, D, DP, DP]),
            "orc_init_obst": (None, [P]),
            "orc_set_distribution": (None, [P, D, I, DP, DP, I]),
            "orc_set_node_populations": (None, [P, DP]),
            "orc_get_node_populations": (None, [P, DP]),
            "orc_update_obstacle": (None, [P, I, DP, DP, D]),
            "orc_update_quadric": (None, [P, DP]),
            "orc_init_double_couette": (None, [P, I, DP]),
            "orc_macro_variables": (None, [P]),
            "orc_bgk": (None, [P, D, D, D, D]),
            "orc_mrt": (None, [P, D, D, D, D, DP]),
            "orc_update_ghost": (None, [P]),
            "orc_streaming": (None, [P]),
            "orc_pre_bounce_back": (None, [P]),
            "orc_post_bounce_back": (None, [P]),
            "orc_cavity": (None, [P, I, D, D, D]),
            "orc_cavity_coeff": (None, [I, D, D, D, D, DP]),
            "orc_wall_bounce_back": (None, [P]),
            "orc_lid_driven_cavity": (None, [P, I, D, D, D]),
            "orc_neumann": (None, [P, I, D, D, D]),
            "orc_pressure_inlet": (None, [P, I, D]),
            "orc_outflow": (None, [P, I]),
            "orc_simulation_state": (None, [P, DP]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(DP)


def lbm_parameters(dx, nuth=1e-4, celerity=1.0, tau=None, dt=0.0, Fext=(0.0, 0.0, 0.0)):
    """src/grid/lbm_parameters.cpp:73-99 -> dict(tau, nu, dtLB, celerity, Fext)."""
    out = np.zeros(7)
    F = np.asarray(Fext, dtype=np.float64)
    lib().orc_lbm_parameters(dx, nuth, celerity, 0 if tau is None else 1, 0.0 if tau is None else tau,
                             dt, _dp(F), _dp(out))
    return dict(tau=out[0], nu=out[1], dtLB=out[2], celerity=out[3], Fext=tuple(out[4:7]))


class Oracle:
    """One single-rank hippoLBM subdomain: SoA f over the All box with 2 ghost layers."""

    def __init__(self, nodes, periodic, dx):
        self.L = lib()
        nx, ny, nz = (int(v) for v in nodes)
        self.h = self.L.orc_create(nx, ny, nz, int(bool(periodic[0])), int(bool(periodic[1])),
                                   int(bool(periodic[2])), float(dx))
        dims = (ctypes.c_int * 6)()
        self.L.orc_dims(self.h, dims)
        self.Lb = tuple(dims[0:3])
        self.n = tuple(dims[3:6])
        self.N = self.Lb[0] * self.Lb[1] * self.Lb[2]
        self.periodic = tuple(bool(p) for p in periodic)
        self.dx = dx

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.orc_destroy(h)
            self.h = None

    # raw views over the reference-layout arrays (z, y, x order, ghosts included)
    def f(self):
        a = np.ctypeslib.as_array(self.L.orc_f(self.h), shape=(19 * self.N,))
        return a.reshape(19, self.Lb[2], self.Lb[1], self.Lb[0])

    def m0(self):
        return np.ctypeslib.as_array(self.L.orc_m0(self.h), shape=(self.N,)).reshape(self.Lb[::-1])

    def m1(self):
        return np.ctypeslib.as_array(self.L.orc_m1(self.h), shape=(3 * self.N,)).reshape(3, *self.Lb[::-1])

    def obst(self):
        return np.ctypeslib.as_array(self.L.orc_obst(self.h), shape=(self.N,)).reshape(self.Lb[::-1])

    def real(self, a):
        """Slice the Real nodes out of a (..., Lz, Ly, Lx) array."""
        return a[..., 2:-2, 2:-2, 2:-2]

    def node_f(self):
        return np.ascontiguousarray(self.real(self.f()))

    def set_node_f(self, fnode):
        fnode = np.ascontiguousarray(fnode, dtype=np.float64)
        self.L.orc_set_node_populations(self.h, _dp(fnode))

    # operators -------------------------------------------------------------
    def init_obst(self):
        self.L.orc_init_obst(self.h)

    def set_distribution(self, value=1.0, bounds=None, do_update=False):
        if bounds is None:
            z = np.zeros(3)
            self.L.orc_set_distribution(self.h, value, 0, _dp(z), _dp(z), int(do_update))
        else:
            lo = np.asarray(bounds[0], dtype=np.float64)
            hi = np.asarray(bounds[1], dtype=np.float64)
            self.L.orc_set_distribution(self.h, value, 1, _dp(lo), _dp(hi), int(do_update))

    def update_obstacle(self, kind, p0, p1=(0.0, 0.0, 0.0), r=0.0):
        a = np.asarray(p0, dtype=np.float64)
        b = np.asarray(p1, dtype=np.float64)
        self.L.orc_update_obstacle(self.h, {"ball": 0, "wall": 1}[kind], _dp(a), _dp(b), float(r))

    def init_double_couette(self, dim, U):
        u = np.ascontiguousarray(U, dtype=np.float64)
        self.L.orc_init_double_couette(self.h, int(dim), _dp(u))

    def update_quadric(self, q):
        a = np.ascontiguousarray(q, dtype=np.float64).reshape(16)
        self.L.orc_update_quadric(self.h, _dp(a))

    def macro_variables(self):
        self.L.orc_macro_variables(self.h)

    def bgk(self, tau, F=(0.0, 0.0, 0.0)):
        self.L.orc_bgk(self.h, tau, F[0], F[1], F[2])

    def mrt(self, tau, F=(0.0, 0.0, 0.0), rates=None):
        if rates is None:
            self.L.orc_mrt(self.h, tau, F[0], F[1], F[2], None)
        else:
            r = np.asarray(rates, dtype=np.float64)
            self.L.orc_mrt(self.h, tau, F[0], F[1], F[2], _dp(r))

    def update_ghost(self):
        self.L.orc_update_ghost(self.h)

    def streaming(self):
        self.L.orc_streaming(self.h)

    def pre_bounce_back(self):
        self.L.orc_pre_bounce_back(self.h)

    def post_bounce_back(self):
        self.L.orc_post_bounce_back(self.h)

    def cavity(self, side, U):
        self.L.orc_cavity(self.h, side, U[0], U[1], U[2])

    def wall_bounce_back(self):
        self.L.orc_wall_bounce_back(self.h)

    def lid_driven_cavity(self, plane, U):
        self.L.orc_lid_driven_cavity(self.h, PLANES[plane], U[0], U[1], U[2])

    def neumann(self, plane, U):
        self.L.orc_neumann(self.h, PLANES[plane], U[0], U[1], U[2])

    def pressure_inlet(self, plane, rho):
        self.L.orc_pressure_inlet(self.h, PLANES[plane], float(rho))

    def outflow(self, plane):
        self.L.orc_outflow(self.h, PLANES[plane])

    def simulation_state(self):
        out = np.zeros(3)
        self.L.orc_simulation_state(self.h, _dp(out))
        return out


def cavity_coeff(side, L, U):
    out = np.zeros(5)
    lib().orc_cavity_coeff(side, float(L), U[0], U[1], U[2], _dp(out))
    return out


class OracleRun:
    """Runs a parsed deck (hippolbm_amd.deck.Deck) through the oracle, operator by operator."""

    def __init__(self, deck, init_f=None, threads=None):
        if threads:
            lib().orc_set_threads(int(threads))
        self.deck = deck
        # lbm_parameters from the deck's raw slot values through the oracle's own
        # restatement (slot defaults: lbm_parameters.cpp:48-56), not the engine's
        lb = deck.lbm
        self.p = lbm_parameters(deck.dx, lb.get("nuth", 1e-4), lb.get("celerity", 1.0), lb.get("tau"),
                                lb.get("dt", 0.0), lb.get("Fext", (0.0, 0.0, 0.0)))
        self.o = Oracle(deck.nodes, deck.periodic, deck.dx)
        o = self.o
        o.init_obst()
        o.set_distribution(1.0, None, do_update=False)
        for sd in deck.set_distributions:
            if sd.get("kind") == "double_couette":
                o.init_double_couette(sd["dim"], [u / self.p["celerity"] for u in sd["U"]])
            else:
                o.set_distribution(sd.get("value", 1.0), sd.get("bounds"), bool(sd.get("do_update", False)))
        for ob in deck.obstacles_in_id_order():
            if ob["kind"] == "ball":
                o.update_obstacle("ball", ob["center"], r=ob["radius"])
            elif ob["kind"] == "quadric":
                o.update_quadric(ob["q"])
            else:
                o.update_obstacle("wall", ob["bounds"][0], ob["bounds"][1])
        if init_f is not None:
            o.set_node_f(init_f)
        o.macro_variables()
        o.update_ghost()
        o.macro_variables()
        self.steps_done = 0

    def step(self, n=1):
        o, d, p = self.o, self.deck, self.p
        for _ in range(n):
            for sd in d.start_iteration:  # config_compute_loop.msp:27
                o.set_distribution(sd.get("value", 1.0), sd.get("bounds"), False)
            o.macro_variables()
            if d.collision == "mrt":
                o.mrt(p["tau"], p["Fext"], d.mrt_rates)
            else:
                o.bgk(p["tau"], p["Fext"])
            for name, args in d.pre_stream_bcs:
                self._apply(name, args)
            o.streaming()
            for name, args in d.post_stream_bcs:
                self._apply(name, args)
            for name, args in d.boundary_conditions:
                self._apply(name, args)
            self.steps_done += 1

    def _apply(self, name, args):
        o, p = self.o, self.p
        c = p["celerity"]
        if name == "pre_bounce_back":
            o.pre_bounce_back()
        elif name == "post_bounce_back":
            o.post_bounce_back()
        elif name == "wall_bounce_back":
            o.wall_bounce_back()
        elif name in ("cavity_z_0", "cavity_z_l"):
            o.cavity(0 if name.endswith("_0") else 1, args["U"])
        elif name == "lid_driven_cavity":
            U = [u / c for u in args["U"]]
            for r in args["regions"]:
                o.lid_driven_cavity(r, U)
        elif name == "neumann":
            U = [u / c for u in args["U"]]
            for r in args["regions"]:
                o.neumann(r, U)
        elif name == "pressure_inlet":
            for r in args["regions"]:
                o.pressure_inlet(r, args["rho"])
        elif name == "outflow":
            for r in args["regions"]:
                o.outflow(r)
        elif name in ("nop", "update_ghost"):
            if name == "update_ghost":
                o.update_ghost()
        else:
            raise ValueError(f"oracle: unsupported operator {name!r}")

    def final_state(self):
        """simulation_state after the loop: [sum rho, min |u|, max |u|]."""
        return self.o.simulation_state()

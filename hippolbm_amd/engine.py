"""Python host API over the C ABI: a device-resident D3Q19 domain and a runner
that drives it through hippoLBM's default operator graph.

``Domain`` is a thin owner of an ``hlbm_domain*``; every compute call goes to
``libhlbm.so`` (HIP kernels on gfx950).  ``Simulation`` replays the reference's
``initialize`` / ``compute_loop`` / ``checker`` sequence
(``data/config/config_initialize.msp:14-32``,
``data/config/config_compute_loop.msp:23-38``, ``data/config/main-config.msp:29-32``)
for a parsed deck.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import BC, DomainDesc, Obstacle, check, lib

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Domain:
    """One single-GPU subdomain (replaces domain + define_grid_3dq19 + fields)."""

    def __init__(self, nodes, periodic, tau, fext=(0.0, 0.0, 0.0), storage="f64", collision="bgk", dx=1.0,
                 mrt_rates=None, pattern="ab", rounding="stochastic"):
        d = DomainDesc()
        if rounding not in ("stochastic", "nearest"):
            raise ValueError("rounding is 'stochastic' or 'nearest'")
        d.round_nearest = 1 if rounding == "nearest" else 0
        for k in range(3):
            d.nodes[k] = int(nodes[k])
            d.periodic[k] = 1 if periodic[k] else 0
            d.fext[k] = float(fext[k])
        d.storage = _lib.STORAGE[storage]
        d.pattern = {"ab": 0, "aa": 1}[pattern]
        d.collision = _lib.COLLISION[collision]
        d.tau = float(tau)
        d.dx = float(dx)
        if mrt_rates is not None:
            d.mrt_custom = 1
            for k in range(10):
                d.mrt_rates[k] = float(mrt_rates[k])
        self.nodes = tuple(int(n) for n in nodes)
        self.N = self.nodes[0] * self.nodes[1] * self.nodes[2]
        self.storage = storage
        self.periodic = tuple(bool(p) for p in periodic)
        h = ctypes.c_void_p()
        check("hlbm_domain_create", lib().hlbm_domain_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().hlbm_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- plan ---------------------------------------------------------------
    def set_bcs(self, entries):
        """entries: [(stage, operator_name, {U, regions})] in list order; U in LBM units
        for lid_driven_cavity / neumann, raw for cavity_z_*."""
        arr = []
        for stage, name, args in entries:
            if name in ("nop", "update_ghost"):
                continue
            if name == "set_distribution":  # start_iteration: U[0] = value, plane = bounded
                b = BC()
                b.kind = _lib.BC_KIND[name]
                b.stage = _lib.STAGE[stage]
                b.U[0] = float(args.get("value", 1.0))
                if args.get("bounds") is not None:
                    b.plane = 1
                    for k in range(3):
                        b.bounds[k] = float(args["bounds"][0][k])
                        b.bounds[3 + k] = float(args["bounds"][1][k])
                arr.append(b)
                continue
            regions = args.get("regions") or [None]
            for r in regions:
                b = BC()
                b.kind = _lib.BC_KIND[name]
                b.stage = _lib.STAGE[stage]
                b.plane = _lib.PLANE[r] if r is not None else 0
                U = args.get("U", (0.0, 0.0, 0.0))
                for k in range(3):
                    b.U[k] = float(U[k])
                arr.append(b)
        buf = (BC * max(1, len(arr)))(*arr)
        check("hlbm_set_bcs", lib().hlbm_set_bcs(self.h, buf, len(arr)))

    def set_collision(self, collision, mrt_rates=None):
        """The collision slot: "bgk" or "mrt" (mrt_rates None = hippoLBM's rates)."""
        r = None if mrt_rates is None else np.ascontiguousarray(mrt_rates, dtype=np.float64)
        check("hlbm_set_collision", lib().hlbm_set_collision(self.h, _lib.COLLISION[collision],
                                                             _dp(r) if r is not None else None))

    def set_layout(self, order):
        """Node layout in HBM: the axes (0 x, 1 y, 2 z) from fastest to slowest."""
        o = np.ascontiguousarray(order, dtype=np.int32)
        check("hlbm_set_layout", lib().hlbm_set_layout(self.h, o.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))

    def get_layout(self):
        o = np.zeros(3, dtype=np.int32)
        check("hlbm_get_layout", lib().hlbm_get_layout(self.h, o.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return tuple(int(v) for v in o)

    def set_params(self, tau, fext):
        F = np.ascontiguousarray(fext, dtype=np.float64)
        check("hlbm_set_params", lib().hlbm_set_params(self.h, float(tau), _dp(F)))

    # ---- state --------------------------------------------------------------
    def init_obst(self):
        check("hlbm_init_obst", lib().hlbm_init_obst(self.h))

    def add_obstacles(self, obstacles):
        arr = [_lib.make_obstacle(ob) for ob in obstacles]
        if not arr:
            return
        buf = (Obstacle * len(arr))(*arr)
        check("hlbm_add_obstacles", lib().hlbm_add_obstacles(self.h, buf, len(arr)))

    def set_flags(self, obst, layout=_lib.LAYOUT_NODES):
        a = np.ascontiguousarray(obst, dtype=np.int32)
        check("hlbm_set_flags", lib().hlbm_set_flags(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                     layout))

    def get_flags(self, layout=_lib.LAYOUT_NODES):
        n = self.N if layout == _lib.LAYOUT_NODES else self._nall()
        a = np.empty(n, dtype=np.int32)
        check("hlbm_get_flags", lib().hlbm_get_flags(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                     layout))
        return a.reshape(self._shape(layout))

    def set_distribution(self, value=1.0, bounds=None):
        if bounds is None:
            check("hlbm_set_distribution", lib().hlbm_set_distribution(self.h, float(value), None))
        else:
            b = np.array(list(bounds[0]) + list(bounds[1]), dtype=np.float64)
            check("hlbm_set_distribution", lib().hlbm_set_distribution(self.h, float(value), _dp(b)))

    def init_double_couette(self, dim, U):
        """init_double_couette with U in LBM units (divided by celerity)."""
        u = np.ascontiguousarray(U, dtype=np.float64)
        check("hlbm_init_double_couette", lib().hlbm_init_double_couette(self.h, int(dim), _dp(u)))

    def _nall(self):
        return (self.nodes[0] + 4) * (self.nodes[1] + 4) * (self.nodes[2] + 4)

    def _shape(self, layout):
        n = self.nodes if layout == _lib.LAYOUT_NODES else tuple(v + 4 for v in self.nodes)
        return (n[2], n[1], n[0])

    def set_f(self, f, layout=_lib.LAYOUT_NODES):
        a = np.ascontiguousarray(f, dtype=np.float64)
        check("hlbm_set_populations_f64", lib().hlbm_set_populations_f64(self.h, _dp(a), layout))

    def get_f(self, layout=_lib.LAYOUT_NODES, out=None):
        shape = (19,) + self._shape(layout)
        a = out if out is not None else np.zeros(shape, dtype=np.float64)
        check("hlbm_get_populations_f64", lib().hlbm_get_populations_f64(self.h, _dp(a), layout))
        return a

    # ---- stepping -----------------------------------------------------------
    def step(self, n=1, write_macro=False, stream=None):
        flags = _lib.STEP_WRITE_MACRO if write_macro else 0
        s = ctypes.c_void_p(stream) if stream else None
        check("hlbm_step", lib().hlbm_step(self.h, int(n), flags, s))

    def synchronize(self):
        check("hlbm_synchronize", lib().hlbm_synchronize(self.h))

    def timestep(self):
        t = ctypes.c_int64(0)
        check("hlbm_timestep", lib().hlbm_timestep(self.h, ctypes.byref(t)))
        return t.value

    def autotune(self, reps=5, stream=None):
        """Pick the fastest node layout and launch shape for this domain (cached
        per shape); returns the choice and the per-step times."""
        t = _lib.TuneInfo()
        s = ctypes.c_void_p(stream) if stream else None
        check("hlbm_autotune", lib().hlbm_autotune(self.h, int(reps), s, ctypes.byref(t)))
        order = (t.order // 100, t.order // 10 % 10, t.order % 10)
        return {"block": (t.block_x, t.block_y), "npt": t.npt, "xcd_remap": t.xcd_remap, "variant": t.variant,
                "order": "".join("xyz"[a] for a in order), "candidates": t.candidates, "relocated": t.relocated,
                "ms_best": t.ms_best, "ms_default": t.ms_default}

    def set_launch(self, block=(128, 2), npt=1, xcd_remap=1, variant=0):
        """Launch shape by hand; variant 0 = scalar plane bases + 32-bit offsets,
        1 = 64-bit per-lane addresses."""
        check("hlbm_set_launch", lib().hlbm_set_launch(self.h, int(block[0]), int(block[1]), int(npt),
                                                       int(xcd_remap), int(variant)))

    def bytes(self):
        b = ctypes.c_uint64(0)
        check("hlbm_domain_bytes", lib().hlbm_domain_bytes(self.h, ctypes.byref(b)))
        return b.value

    # ---- macro / diagnostics -------------------------------------------------
    def macro(self, layout=_lib.LAYOUT_NODES):
        """macro_variables on the current state; returns (rho, u)."""
        rho = np.zeros(self._shape(layout))
        u = np.zeros((3,) + self._shape(layout))
        check("hlbm_macro", lib().hlbm_macro(self.h, _dp(rho), _dp(u), layout))
        return rho, u

    def get_macro(self, layout=_lib.LAYOUT_NODES):
        rho = np.zeros(self._shape(layout))
        u = np.zeros((3,) + self._shape(layout))
        check("hlbm_get_macro", lib().hlbm_get_macro(self.h, _dp(rho), _dp(u), layout))
        return rho, u

    def plane_profile(self, dim, scale=1.0):
        """plane_velocity_profile statistics of |scale*u| per plane perpendicular to
        axis `dim` (0, 1, 2) over the fluid nodes of the held u field:
        (sum, count, min, max) arrays."""
        n = self.nodes[dim]
        s, lo, hi = np.zeros(n), np.zeros(n), np.zeros(n)
        c = np.zeros(n, dtype=np.int32)
        check("hlbm_plane_profile", lib().hlbm_plane_profile(self.h, int(dim), float(scale), _dp(s),
                                                             c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                             _dp(lo), _dp(hi)))
        return s, c, lo, hi

    def line_velocity(self, start, axis, n, scale=1.0):
        """plot_line_velocity samples: scale*u at start + k*e_axis, shape (n, 3)."""
        st = np.asarray(start, dtype=np.int32)
        u = np.zeros((int(n), 3))
        check("hlbm_line_velocity", lib().hlbm_line_velocity(self.h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                             int(axis), int(n), float(scale), _dp(u)))
        return u

    def stats(self):
        """simulation_state over the rho/u fields: sum rho, min/max |u|, momentum, fluid count."""
        out = np.zeros(7)
        check("hlbm_reduce_stats", lib().hlbm_reduce_stats(self.h, _dp(out)))
        return {"sum_density": out[0], "min_velocity_norm": out[1], "max_velocity_norm": out[2],
                "momentum": tuple(out[3:6]), "fluid_nodes": int(out[6])}


def bc_entries(deck):
    """The three operator lists of a deck as (stage, name, args) with LBM-unit velocities
    (lid_driven_cavity.cu:75 and neuman.cu:74 divide by celerity; cavity.cu:70 does not)."""
    c = deck.params["celerity"]
    out = [("start_iteration", "set_distribution", dict(sd)) for sd in deck.start_iteration]
    for stage in ("pre_stream_bcs", "post_stream_bcs", "boundary_conditions"):
        for name, args in getattr(deck, stage):
            a = dict(args)
            if name in ("lid_driven_cavity", "neumann"):
                a["U"] = tuple(u / c for u in args["U"])
            if name == "pressure_inlet":  # the density travels in U[0]
                a["U"] = (float(args["rho"]), 0.0, 0.0)
            out.append((stage, name, a))
    return out


class Simulation:
    """A deck run on the GPU engine with the reference's operator order."""

    def __init__(self, deck, storage="f64", init_f=None, pattern="ab", rounding="stochastic"):
        p = deck.params
        self.deck = deck
        self.params = p
        self.dom = Domain(deck.nodes, deck.periodic, p["tau"], p["Fext"], storage=storage,
                          collision=deck.collision, dx=deck.dx, mrt_rates=deck.mrt_rates, pattern=pattern,
                          rounding=rounding)
        self.dom.set_bcs(bc_entries(deck))
        # initialize (config_initialize.msp:21-27)
        self.dom.init_obst()
        self.dom.set_distribution(1.0)
        for sd in deck.set_distributions:
            if sd.get("kind") == "double_couette":
                self.dom.init_double_couette(sd["dim"], [u / p["celerity"] for u in sd["U"]])
            else:
                self.dom.set_distribution(sd.get("value", 1.0), sd.get("bounds"))
        self.dom.add_obstacles(deck.obstacles_in_id_order())
        if init_f is not None:
            self.dom.set_f(init_f)
        self.dom.macro(layout=_lib.LAYOUT_NODES)

    def run(self, n):
        """n iterations of compute_loop; rho/u left as macro_variables wrote them last."""
        self.dom.step(n, write_macro=True)

    def state(self):
        return self.dom.stats()

    def check(self):
        """check_macro_quantities (src/check/check_marco_quantities.cpp:71-114)."""
        c = self.deck.checker or {}
        s = self.state()
        errs = []
        if s["min_velocity_norm"] < c.get("vmin", 0.0) - 1e-15:
            errs.append(f"min vel {s['min_velocity_norm']} < {c.get('vmin', 0.0)}")
        if "vmax" in c and s["max_velocity_norm"] > c["vmax"] + 1e-15:
            errs.append(f"max vel {s['max_velocity_norm']} > {c['vmax']}")
        if "density" in c:
            if "tol" not in c:
                raise ValueError("Please add a tolerance if you use the density INPUT slot.")
            if abs(s["sum_density"] - c["density"]) > c["tol"]:
                errs.append(f"sum density {s['sum_density']} != {c['density']} +/- {c['tol']}")
        return errs

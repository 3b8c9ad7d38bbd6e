"""Deck: the subset of hippoLBM's ``.msp`` operator graph that configures the hot path.

A deck is what one ``.msp`` file (plus hippoLBM's defaults, ``data/config/*.msp``)
says about the D3Q19 step: the domain, the LBM parameters, the collision model,
the three boundary-condition lists and the initial state.  Blocks are read by the
names the reference's default graph uses:

* ``do_domain: [domain: {cell_dims, bounds, periodic}]`` -- ``src/grid/domain_lbm.cpp:44-49``;
* ``set_lbm_parameters: [lbm_parameters: {Fext, nuth, celerity, avg_rho, tau, dt}]``
  -- ``src/grid/lbm_parameters.cpp:46-56``;
* ``collision: bgk|mrt`` -- ``data/config/config_compute_loop.msp:21``;
* ``pre_stream_bcs``, ``post_stream_bcs``, ``boundary_conditions`` -- operator
  lists run in order around ``streaming`` (``config_compute_loop.msp:28-36``);
* ``set_distributions: [set_distribution: {value, bounds, do_update}]`` and
  ``set_obstacles: [register_solid_ball|register_solid_wall: ...]``
  (``data/config/config_initialize.msp:3-26``);
* ``global: {simulation_end_iteration, ...}`` and ``checker_op:
  [check_macro_quantities: {...}]``.

Name resolution follows what the reference's own regression pins imply: a user
block that reuses an operator's name (``set_distribution:`` at top level,
``regression-test/regression_lbm_diff_pressure.msp:24-27``) does not replace the
operator, so it has no effect on the run (the pin sum(rho) = 29124 holds only
with f_i = w_i everywhere).

Physical quantities may carry a unit suffix (``0.01 m``); the suffix is
dropped, as every value in the reference's decks is already SI.
"""
from __future__ import annotations

import dataclasses
import os
import re
from typing import Any

import yaml

from . import quadrics

_NUM_RE = re.compile(r"^\s*([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)\s*[A-Za-z./^0-9-]*\s*$")

HOT_OPERATORS = (
    "bgk", "mrt", "macro_variables", "streaming", "update_ghost", "pre_bounce_back", "post_bounce_back",
    "wall_bounce_back", "cavity_z_0", "cavity_z_l", "lid_driven_cavity", "neumann",
)
BC_OPERATORS = ("pre_bounce_back", "post_bounce_back", "wall_bounce_back", "cavity_z_0", "cavity_z_l",
                "lid_driven_cavity", "neumann", "update_ghost", "nop",
                # open boundaries with no counterpart in the reference (porous channel inlet / outlet)
                "pressure_inlet", "outflow")
PLANE_NAMES = ("plan_xy_0", "plan_xy_l", "plan_xz_0", "plan_xz_l", "plan_yz_0", "plan_yz_l")


class DeckError(ValueError):
    """Malformed or unsupported deck (the reference exits with EXIT_FAILURE or throws)."""


# top-level blocks the deck model reads; a block named after the set_distribution
# operator is read and has no effect (see the module docstring), `configuration`
# holds Onika's logging / profiling switches
KNOWN_BLOCKS = ("do_domain", "domain", "set_lbm_parameters", "collision", "pre_stream_bcs", "post_stream_bcs",
                "boundary_conditions", "set_distributions", "set_obstacles", "start_iteration", "checker_op",
                "checker", "analysis", "global", "mrt_rates", "includes", "set_distribution", "configuration")
# batches of the default graph (data/config/*.msp): the run follows their fixed
# sequence, so a deck that redefines one is refused
GRAPH_BLOCKS = ("compute_loop", "compute_loop_stop", "initialize", "simulation", "end_iteration", "setup_obstacles",
                "grid_flavor", "default_grid", "trigger_dump_paraview", "trigger_print_log", "trigger_analysis",
                "dump_paraview_if_triggered", "print_log_if_triggered", "print_analysis_if_triggered",
                "write_paraview_first_timestep", "streaming")


def check_blocks(doc: dict):
    for k, v in doc.items():
        if k in GRAPH_BLOCKS:
            raise DeckError(f"the deck redefines {k!r} of the default graph (data/config/*.msp); "
                            "redefinitions are not executed")
        if isinstance(v, dict):
            for c in ("rebind", "condition", "loop", "body"):
                if c in v:
                    raise DeckError(f"block {k!r} uses Onika graph control ({c!r}), which is not executed")
        if k not in KNOWN_BLOCKS:
            raise DeckError(f"unsupported top-level block {k!r}")


def _num(v):
    if isinstance(v, bool):
        return v
    if isinstance(v, (int, float)):
        return float(v)
    if isinstance(v, str):
        m = _NUM_RE.match(v)
        if m:
            return float(m.group(1))
    raise DeckError(f"not a number: {v!r}")


def _vec(v, n=3):
    if not isinstance(v, (list, tuple)) or len(v) != n:
        raise DeckError(f"expected a list of {n} numbers, got {v!r}")
    return tuple(_num(x) for x in v)


def _aabb(v):
    if isinstance(v, dict) and "bmin" in v:
        return (_vec(v["bmin"]), _vec(v["bmax"]))
    if not isinstance(v, (list, tuple)) or len(v) != 2:
        raise DeckError(f"expected [[xmin,ymin,zmin],[xmax,ymax,zmax]], got {v!r}")
    return (_vec(v[0]), _vec(v[1]))


def _items(block):
    """Normalise an operator list into [(name, args)]."""
    if block is None:
        return []
    if isinstance(block, str):
        return [] if block == "nop" else [(block, {})]
    if isinstance(block, dict):
        block = [block]
    out = []
    for it in block:
        if isinstance(it, str):
            if it != "nop":
                out.append((it, {}))
        elif isinstance(it, dict):
            for k, v in it.items():
                out.append((k, v if isinstance(v, dict) else {}))
        else:
            raise DeckError(f"bad operator entry {it!r}")
    return out


@dataclasses.dataclass
class Deck:
    cell_dims: tuple
    bounds: tuple
    periodic: tuple
    lbm: dict
    collision: str = "bgk"
    pre_stream_bcs: list = dataclasses.field(default_factory=list)
    post_stream_bcs: list = dataclasses.field(default_factory=list)
    boundary_conditions: list = dataclasses.field(default_factory=list)
    set_distributions: list = dataclasses.field(default_factory=list)
    start_iteration: list = dataclasses.field(default_factory=list)  # set_distribution entries
    obstacles: dict = dataclasses.field(default_factory=dict)
    end_iteration: int = 100
    checker: dict | None = None
    mrt_rates: Any = None
    name: str = ""

    # --- derived quantities (src/grid/domain_lbm.cpp:66-106) -----------------
    @property
    def nodes(self):
        return tuple(int(c) + 1 for c in self.cell_dims)

    @property
    def dx(self):
        lo, hi = self.bounds
        return (hi[0] - lo[0]) / float(self.nodes[0] - 1)

    def validate(self):
        lo, hi = self.bounds
        n = self.nodes
        res = [(hi[d] - lo[d]) / float(n[d] - 1) for d in range(3)]
        if res[0] != res[1] or res[0] != res[2]:
            raise DeckError(f"[Error, domain], Dx is not the same for all dimension: {res}")
        for d in range(3):
            if lo[d] + (n[d] - 1) * res[0] != hi[d]:
                raise DeckError("[Error, domain], The resolution slot and bounds slot mismatch.")
        if self.collision not in ("bgk", "mrt"):
            raise DeckError(f"unsupported collision {self.collision!r}")
        for lst in (self.pre_stream_bcs, self.post_stream_bcs, self.boundary_conditions):
            for name, args in lst:
                if name not in BC_OPERATORS:
                    raise DeckError(f"unsupported boundary operator {name!r}")
                if name in ("lid_driven_cavity", "neumann"):
                    if "U" not in args or "regions" not in args:
                        raise DeckError(f"{name}: U and regions are required")
                    for r in args["regions"]:
                        if r not in PLANE_NAMES:
                            raise DeckError(f"Traversal name not allowed: {r}")
                if name in ("pressure_inlet", "outflow"):
                    if "regions" not in args or (name == "pressure_inlet" and "rho" not in args):
                        raise DeckError(f"{name}: regions{' and rho' if name == 'pressure_inlet' else ''} required")
                    for r in args["regions"]:
                        if r not in PLANE_NAMES:
                            raise DeckError(f"Traversal name not allowed: {r}")
                if name in ("cavity_z_0", "cavity_z_l") and "U" not in args:
                    raise DeckError(f"{name}: U is required")
        ids = sorted(self.obstacles)
        if ids and ids != list(range(len(ids))):
            raise DeckError("Internal error: unsupported obstacle type encountered (obstacle ids must be dense)")
        return self

    @property
    def params(self):
        """lbm_parameters (src/grid/lbm_parameters.cpp:73-99)."""
        dx = self.dx
        nuth = self.lbm.get("nuth", 1e-4)
        c = self.lbm.get("celerity", 1.0)
        tau = self.lbm.get("tau")
        dt = self.lbm.get("dt", 0.0)
        F = self.lbm.get("Fext", (0.0, 0.0, 0.0))
        if dt > 0.0:
            dtLB = dt
            tau_ = 3 * nuth * dtLB / (dx * dx) + 0.5
            nu = (tau_ - 0.5) / 3.0
            c_ = dx / dtLB
        elif tau is not None:
            tau_ = tau
            dtLB = dx * dx * (tau_ - 0.5) / (3.0 * nuth)
            nu = nuth * dtLB / (dx * dx)
            c_ = dx / dtLB
        else:
            c_ = c
            dtLB = dx / c_
            nu = nuth * dtLB / (dx * dx)
            tau_ = 3. * nu + 0.5
        dxc = c_ * dtLB
        Fl = tuple(F[d] * dtLB * dtLB / dxc for d in range(3))
        return dict(tau=tau_, nu=nu, dtLB=dtLB, celerity=c_, Fext=Fl, nuth=nuth)

    def obstacles_in_id_order(self):
        return [self.obstacles[i] for i in sorted(self.obstacles)]

    def __post_init__(self):
        self.cell_dims = tuple(int(c) for c in self.cell_dims)


def from_dict(doc: dict, name: str = "") -> Deck:
    """Build a Deck from a parsed .msp mapping."""
    check_blocks(doc)
    dom = None
    for k, v in _items(doc.get("do_domain")):
        if k == "domain":
            dom = v
    if dom is None and isinstance(doc.get("domain"), dict):
        dom = doc["domain"]
    if dom is None:
        raise DeckError("no domain operator in do_domain")
    if "cell_dims" not in dom:
        # the domain operator requires cell_dims (src/grid/domain_lbm.cpp:47-48)
        raise DeckError("domain: slot cell_dims is REQUIRED")
    cell = tuple(int(_num(c)) for c in dom["cell_dims"])
    bounds = _aabb(dom["bounds"])
    per = tuple(bool(p) for p in dom["periodic"])
    if len(per) != 3:
        raise DeckError("periodic must have 3 entries")
    lbm = {"Fext": (0.0, 0.0, 0.0), "nuth": 1e-4, "celerity": 1.0, "avg_rho": 1000.0, "dt": 0.0}
    for k, v in _items(doc.get("set_lbm_parameters")):
        if k == "lbm_parameters":
            for key, val in v.items():
                if key == "Fext":
                    lbm["Fext"] = _vec(val)
                elif key in ("nuth", "celerity", "avg_rho", "tau", "dt"):
                    lbm[key] = _num(val)
    coll = doc.get("collision", "bgk")
    if isinstance(coll, list):
        coll = _items(coll)[0][0]
    d = Deck(cell_dims=cell, bounds=bounds, periodic=per, lbm=lbm, collision=str(coll), name=name)
    # extension (no reference slot: mrt.hpp hard-codes its rates): a top-level
    # `mrt_rates: [s1, s2, s4, s9, s10, s13, s16, weps, wepsj, wxx]`
    if "mrt_rates" in doc:
        rates = [_num(v) for v in doc["mrt_rates"]]
        if len(rates) != 10:
            raise DeckError("mrt_rates needs 10 values: s1 s2 s4 s9 s10 s13 s16 weps wepsj wxx")
        d.mrt_rates = rates
    for blk, attr in (("pre_stream_bcs", "pre_stream_bcs"), ("post_stream_bcs", "post_stream_bcs"),
                      ("boundary_conditions", "boundary_conditions")):
        lst = []
        for k, v in _items(doc.get(blk)):
            args = dict(v)
            if "U" in args:
                args["U"] = _vec(args["U"])
            if "regions" in args:
                args["regions"] = list(args["regions"])
            if "rho" in args:
                args["rho"] = _num(args["rho"])
            lst.append((k, args))
        setattr(d, attr, lst)
    for k, v in _items(doc.get("set_distributions")):
        if k == "set_distribution":
            sd = {"value": _num(v.get("value", 1.0)), "do_update": bool(v.get("do_update", False))}
            if "bounds" in v:
                sd["bounds"] = _aabb(v["bounds"])
            d.set_distributions.append(sd)
        elif k == "init_double_couette":
            dim = str(v.get("dimension", ""))
            if dim not in ("X", "Y", "Z"):
                raise DeckError('[init_double_couette] Please, select a valid dimension "X", "Y", or "Z".')
            if "U" not in v:
                raise DeckError("init_double_couette: the U slot is required")
            d.set_distributions.append({"kind": "double_couette", "dim": "XYZ".index(dim), "U": _vec(v["U"])})
        else:
            raise DeckError(f"unsupported set_distributions operator {k!r}")
    # start_iteration (config_compute_loop.msp:27): operators at the start of every step
    for k, v in _items(doc.get("start_iteration")):
        if k == "set_distribution":
            sd = {"value": _num((v or {}).get("value", 1.0))}
            if v and "bounds" in v:
                sd["bounds"] = _aabb(v["bounds"])
            d.start_iteration.append(sd)
        elif k != "nop":
            raise DeckError(f"unsupported start_iteration operator {k!r}")
    for k, v in _items(doc.get("set_obstacles")):
        if k == "register_solid_ball":
            d.obstacles[int(v["id"])] = {"kind": "ball", "center": _vec(v["center"]), "radius": _num(v["radius"])}
        elif k == "register_solid_wall":
            d.obstacles[int(v["id"])] = {"kind": "wall", "bounds": _aabb(v["bounds"])}
        elif k == "register_quadrics":
            if "quadrics" not in v or "id" not in v:
                raise DeckError("register_quadrics: the id and quadrics slots are required")
            try:
                q = quadrics.from_slots(v["quadrics"], v.get("transform"))
            except ValueError as e:
                raise DeckError(str(e)) from None
            d.obstacles[int(v["id"])] = {"kind": "quadric", "q": q}
        elif k != "init_obstacles":
            raise DeckError(f"unsupported obstacle operator {k!r}")
    g = doc.get("global") or {}
    d.end_iteration = int(g.get("simulation_end_iteration", 100))
    for k, v in _items(doc.get("checker_op")):
        if k == "check_macro_quantities":
            d.checker = {key: _num(val) for key, val in v.items()}
    return d.validate()


def _load_with_includes(path: str, depth: int = 0) -> dict:
    """A deck file with its ``includes`` (data/config/main-config.msp:11): the listed
    files, relative to the including file, first and in order; a later file's
    top-level block replaces an earlier one's; a missing include is an error."""
    if depth > 16:
        raise DeckError(f"{path}: includes nested deeper than 16 files")
    with open(path) as fh:
        doc = yaml.safe_load(fh) or {}
    if not isinstance(doc, dict):
        raise DeckError(f"{path}: a deck is a map of blocks")
    incs = doc.get("includes") or []
    if isinstance(incs, str):
        incs = [incs]
    merged: dict = {}
    for inc in incs:
        ip = inc if os.path.isabs(inc) else os.path.join(os.path.dirname(path), inc)
        if not os.path.exists(ip):
            raise DeckError(f"{path}: included file {inc!r} not found (looked for {ip})")
        merged.update(_load_with_includes(ip, depth + 1))
    merged.update(doc)
    merged.pop("includes", None)
    return merged


def load_msp(path: str) -> Deck:
    """Parse one .msp file (with its ``includes``) into a Deck."""
    return from_dict(_load_with_includes(path), name=os.path.basename(path))


def loads_msp(text: str, name: str = "") -> Deck:
    return from_dict(yaml.safe_load(text) or {}, name=name)

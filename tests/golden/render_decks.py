"""Render the regression cases of regression_cases.json into .msp deck text.

Each case is data (domain, LBM parameters, collision, operator lists,
obstacles, initial distribution, checker pins, globals); this module writes the
deck that configures it, block by block, in the layout hippoLBM's default graph
reads (``data/config/config_initialize.msp``, ``config_compute_loop.msp``).
The decks are generated at test time (``write_decks``), so no deck text is
stored in the repository.

``initial_distribution`` names both the top-level block and the operator inside
it, because one case writes its distribution under the operator's own name
(block ``set_distribution``, operator ``set_distributions``), which leaves the
populations at rest in the reference (its density pin 29124 shows it).
"""
from __future__ import annotations

import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
CASES_FILE = os.path.join(HERE, "regression_cases.json")


def cases() -> dict:
    with open(CASES_FILE) as fh:
        return json.load(fh)["cases"]


def _num(v, unit=None) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        s = str(v)
    else:
        s = repr(float(v))
        mant, _, exp = s.partition("e")
        if exp and "." not in mant:      # YAML 1.1 reads a float only with a '.'
            s = f"{mant}.0e{exp}"
    return f"{s} {unit}" if unit else s


def _flow(v, unit=None) -> str:
    if isinstance(v, (list, tuple)):
        return "[" + ", ".join(_flow(x, unit) for x in v) + "]"
    if isinstance(v, str):
        return v
    return _num(v, unit)


def _op_list(lines, block, items, unit=None):
    lines.append(f"{block}:")
    for name, args in items:
        if not args:
            lines.append(f"  - {name}")
            continue
        lines.append(f"  - {name}:")
        for k, v in args.items():
            lines.append(f"      {k}: {_flow(v, unit if k in ('bounds', 'center', 'radius') else None)}")
    lines.append("")


def render(name: str, case: dict) -> str:
    unit = case.get("length_unit")
    out = [f"# regression case {name}, rendered from tests/golden/regression_cases.json", ""]
    dom = case["domain"]
    out += ["do_domain:", "  - domain:",
            f"      bounds: {_flow(dom['bounds'], unit)}",
            f"      cell_dims: {_flow(dom['cell_dims'])}",
            f"      periodic: {_flow(dom['periodic'])}", ""]
    out += ["set_lbm_parameters:", "  - lbm_parameters:"]
    out += [f"      {k}: {_flow(v)}" for k, v in case["lbm_parameters"].items()]
    out.append("")
    for block in ("pre_stream_bcs", "post_stream_bcs", "boundary_conditions"):
        if block in case:
            _op_list(out, block, case[block])
    if "obstacles" in case:
        _op_list(out, "set_obstacles", case["obstacles"], unit)
    if "initial_distribution" in case:
        init = dict(case["initial_distribution"])
        blk, op = init.pop("block"), init.pop("operator")
        _op_list(out, blk, [(op, init)], unit)
    if "collision" in case:
        out += [f"collision: {case['collision']}", ""]
    _op_list(out, "checker_op", [("check_macro_quantities", case["checker"])])
    out.append("global:")
    out += [f"  {k}: {_flow(v)}" for k, v in case["global"].items()]
    return "\n".join(out) + "\n"


def write_decks(dirpath: str) -> list:
    """Write every case as <name>.msp under dirpath; returns the sorted paths."""
    os.makedirs(dirpath, exist_ok=True)
    paths = []
    for name, case in cases().items():
        p = os.path.join(dirpath, name + ".msp")
        with open(p, "w") as fh:
            fh.write(render(name, case))
        paths.append(p)
    return sorted(paths)

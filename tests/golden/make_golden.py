"""Writes the golden fixtures from the oracle: populations.npz (small cases)
and golden.json (64^3 Couette hashes, and per-case sums).  Run from the
repository root after `make -C oracle`:  python3 tests/golden/make_golden.py"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle"), HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import oracle as orc  # noqa: E402
from hippolbm_amd import deck as dk  # noqa: E402
from hippolbm_amd import workloads  # noqa: E402

import cases  # noqa: E402


def main():
    arrays, meta = {}, {"cases": {}, "couette64": {}}
    for name, (txt, nodes, seed, steps) in cases.CASES.items():
        d = dk.loads_msp(txt, name)
        assert d.nodes == nodes, (name, d.nodes)
        ref = orc.OracleRun(d, init_f=cases.seeded_equilibrium(nodes, seed))
        ref.step(steps)
        f = ref.o.node_f()
        flags = ref.o.real(ref.o.obst()).astype(np.int32)
        arrays[f"f_{name}"] = f
        arrays[f"flags_{name}"] = flags
        meta["cases"][name] = {"nodes": list(nodes), "seed": seed, "steps": steps, "sha256": cases.f_hash(f),
                               "walls": int((flags == -2).sum()), "sum_f": float(f.sum())}
        print(name, meta["cases"][name], flush=True)
    d = workloads.couette(cases.HASH_NODES)
    ref = orc.OracleRun(d, init_f=cases.seeded_equilibrium(cases.HASH_NODES, cases.HASH_SEED))
    done = 0
    for s in cases.HASH_STEPS:
        ref.step(s - done)
        done = s
        st = ref.final_state()
        meta["couette64"][str(s)] = {"sha256": cases.f_hash(ref.o.node_f()), "sum_rho": float(st[0]),
                                     "max_u": float(st[2])}
        print("couette64", s, meta["couette64"][str(s)], flush=True)
    np.savez_compressed(os.path.join(HERE, "populations.npz"), **arrays)
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(meta, fh, indent=1, sort_keys=True)
        fh.write("\n")


if __name__ == "__main__":
    main()

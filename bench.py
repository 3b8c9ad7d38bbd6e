#!/usr/bin/env python3
"""Benchmark: D3Q19 Couette flow on one MI355X, MLUPS and % of the HBM roofline
per storage type.

Workload (one "step" = one lattice Boltzmann time step of the whole domain):
plane Couette 512^3 nodes, BGK, periodic x/y, bounce-back walls in z with a
moving lid (pre_bounce_back + cavity_z_l U=(0.05,0,0) + post_bounce_back),
tau = 0.8, synthetic initial state f_i = w_i (rho = 1, u = 0).  `value` is
the FP64 drop-in's MLUPS (nodes x steps / s / 1e6, the definition of
hippoLBM's log operator); the FP32/FP16/BF16 shifted-storage forms are
reported beside it under "per_dtype".

Timing: HIP events on the stream the step kernels run on, W untimed warm-up
steps, then exactly K timed steps, device synchronised before and after.
Roofline: algorithmic bytes per node update = 2*19*sizeof(storage) + 1
(flag byte) against MI355X's 8 TB/s HBM peak.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES = {"f64": 8, "f32": 4, "f16": 2, "bf16": 2}


def couette_domain(n, storage, collision="bgk", pattern="ab"):
    from hippolbm_amd.engine import Domain
    dom = Domain((n, n, n), (True, True, False), 0.8, (0.0, 0.0, 0.0), storage=storage, collision=collision,
                 dx=1.0, pattern=pattern)
    dom.set_bcs([("pre_stream_bcs", "pre_bounce_back", {}), ("pre_stream_bcs", "cavity_z_l", {"U": (0.05, 0, 0)}),
                 ("post_stream_bcs", "post_bounce_back", {})])
    dom.init_obst()
    dom.set_distribution(1.0)
    return dom


def time_steps(dom, steps, warmup):
    from hippolbm_amd import hip
    s = hip.Stream()
    e0, e1 = hip.Event(), hip.Event()
    dom.step(warmup, stream=s.handle)
    hip.device_synchronize()
    t0 = time.perf_counter()
    e0.record(s)
    dom.step(steps, stream=s.handle)
    e1.record(s)
    e1.synchronize()
    hip.device_synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1)
    for o in (e0, e1, s):
        o.close()
    return ms, wall


def measure(n, storage, steps, warmup, collision="bgk", tune=True, pattern="ab"):
    dom = couette_domain(n, storage, collision, pattern)
    tuned = dom.autotune(reps=3) if tune else None
    ms, wall = time_steps(dom, steps, warmup)
    # sanity: the run stayed finite and mass-conserving
    dom.macro()
    st = dom.stats()
    dom.close()
    nodes = n ** 3
    ms_step = ms / steps
    mlups = nodes * steps / (ms * 1e-3) / 1e6
    bpn = 2 * 19 * BYTES[storage] + 1
    gbs = nodes * bpn / (ms_step * 1e-3) / 1e9
    return {"mlups": mlups, "ms_per_step": ms_step, "wall_s": wall, "bytes_per_node": bpn, "achieved_gbs": gbs,
            "frac": gbs / HBM_PEAK_GBS, "sum_density": st["sum_density"], "max_u": st["max_velocity_norm"],
            "launch": tuned}


def porous_channel(steps=100, warmup=10, storage="f64", tune=True):
    """The porous channel 256x128x128 (MRT, s_nu = 1/0.7, other rates 1.2;
    density 1.001 inlet, zero-gradient outflow, bounce-back y/z walls, 150 seeded
    non-overlapping spheres r ~ U[5,12]) on the A/B kernel."""
    from hippolbm_amd import workloads
    from hippolbm_amd.engine import Simulation
    d = workloads.porous_channel()
    sim = Simulation(d, storage=storage)
    tuned = sim.dom.autotune(reps=3) if tune else None
    ms, _ = time_steps(sim.dom, steps, warmup)
    flags = sim.dom.get_flags()
    nodes = sim.dom.N
    sim.dom.close()
    ms_step = ms / steps
    bpn = 2 * 19 * BYTES[storage] + 1
    gbs = nodes * bpn / (ms_step * 1e-3) / 1e9
    return {"workload": "porous_channel_256x128x128_mrt", "storage": storage, "nodes": nodes,
            "solid_nodes": int((flags == -2).sum()), "mlups": nodes * steps / (ms * 1e-3) / 1e6,
            "ms_per_step": ms_step, "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS, "launch": tuned}


def batched_cavities(n_dom=2048, chunk=200, seed=4, storage="f16", use_graph=True):
    """2048 independent lid-driven cavities, edges ~U{16..96} per axis, lid speed
    ~U[0.01,0.1], tau ~U[0.52,1.0], BGK, one launch per step (one workgroup per
    1024-node tile), chunks of `chunk` steps replayed from a captured HIP graph
    (use_graph=False: the same launches issued one by one on the stream)."""
    import numpy as np
    from hippolbm_amd.batch import Batch, Solve
    rng = np.random.default_rng(seed)
    solves = []
    for _ in range(n_dom):
        nodes = tuple(int(v) for v in rng.integers(16, 97, size=3))
        solves.append(Solve(nodes=nodes, steps=1 + 2 * chunk, tau=float(rng.uniform(0.52, 1.0)),
                            cavity_zl=(float(rng.uniform(0.01, 0.1)), 0.0, 0.0)))
    nodes_total = sum(s.node_count() for s in solves)
    b = Batch(storage=storage, budget_bytes=nodes_total * (38 * BYTES[storage] + 1) + (2 << 30),
              chunk_steps=chunk, use_graph=use_graph)
    for s in solves:
        b.submit(s)
    b.run()  # admission + first chunk (warm-up)
    i0 = b.info()
    b.run()  # timed chunk (+ retirement)
    i1 = b.info()
    b.close()
    ms = i1["gpu_ms"] - i0["gpu_ms"]
    mlups = nodes_total * chunk / (ms * 1e-3) / 1e6
    bpn = 2 * 19 * BYTES[storage] + 1
    return {"domains": n_dom, "nodes": nodes_total, "chunk_steps": chunk, "storage": storage, "mlups": mlups,
            "ms_per_step": ms / chunk, "frac": mlups * 1e6 * bpn / 1e9 / HBM_PEAK_GBS,
            "arenas_relocated": i1["relocated"]}


# arrival rates of the served legs: 8 Hz loads the FP32 path to about 85 % of its
# capacity (mean request 3.4e9 node updates at ~32 GLUPS), 16 Hz saturates both
# storages, 0 = every request at t = 0
SERVED_RATES = (8.0, 16.0, 0.0)
SERVED_N = 200   # requests per rate point: p99 is the 198th of 200 ordered samples, not the maximum


def served_sweep(rates=SERVED_RATES, n=SERVED_N, seed=5, storages=("f32", "f16"), chunk=200, budget_gb=96,
                 use_graph=True):
    """served_stream at several arrival rates (0 = every request at t = 0), per storage."""
    out = {}
    for st in storages:
        out[st] = []
        for r in rates:
            out[st].append(served_stream(n, r, seed, st, chunk, budget_gb, use_graph))
            x = out[st][-1]
            print(f"# served {st} @ {r} Hz: {x['n']} requests, p50 {x['p50_s']:.2f} s p99 {x['p99_s']:.2f} s, "
                  f"{x['mlups']:.0f} MLUPS ({100 * x['frac']:.1f}%), GPU busy {100 * x['gpu_busy']:.0f}%, "
                  f"{len(x['failed'])} failed", file=sys.stderr, flush=True)
    return out


def served_stream(n=SERVED_N, rate_hz=8.0, seed=5, storage="f16", chunk=200, budget_gb=96, use_graph=True):
    """Poisson arrivals of cavity / channel / porous solves (sizes log-uniform
    32^3..256^3, 500..5000 steps, BGK or MRT), HBM-budget admission, continuous
    batching; time-to-result percentiles and aggregate MLUPS."""
    from hippolbm_amd.batch import Batch, poisson_requests, serve
    reqs = poisson_requests(n, rate_hz, seed)
    b = Batch(storage=storage, budget_bytes=budget_gb << 30, chunk_steps=chunk, use_graph=use_graph)
    r = serve(b, reqs)
    r["arenas_relocated"] = b.info()["relocated"]
    b.close()
    bpn = 2 * 19 * BYTES[storage] + 1
    r.pop("results")
    r.update({"rate_hz": rate_hz, "storage": storage, "chunk_steps": chunk,
              "frac": r["mlups"] * 1e6 * bpn / 1e9 / HBM_PEAK_GBS})
    return r


def precision_sweep(steps=10000, which=("couette", "poiseuille")):
    """FP64 vs FP32 compute on FP32/FP16/BF16 shifted storage after `steps`
    steps: rho and u errors (relative L2 and max) per dtype, for Couette 512^3
    and Poiseuille 512x256x256 (F = 1e-6, tau = 0.6)."""
    import numpy as np
    from hippolbm_amd import workloads
    from hippolbm_amd.engine import Simulation
    out = {}
    decks = {"couette": lambda: workloads.couette((512, 512, 512)),
             "poiseuille": lambda: workloads.poiseuille((512, 256, 256))}
    for name in which:
        d = decks[name]()
        ref = None
        res = {}
        for dt in ("f64", "f32", "f16", "bf16"):
            t0 = time.perf_counter()
            sim = Simulation(d, storage=dt)
            sim.run(steps)
            rho, u = sim.dom.macro()
            sim.dom.close()
            if ref is None:
                ref = (rho, u)
                res[dt] = {"u_max": float(np.sqrt((u * u).sum(0)).max()), "rho_mean": float(rho.mean())}
            else:
                du = u - ref[1]
                dr = rho - ref[0]
                res[dt] = {"u_rel_l2": float(np.sqrt((du * du).sum() / (ref[1] * ref[1]).sum())),
                           "u_max_abs": float(np.abs(du).max()),
                           "rho_rel_l2": float(np.sqrt((dr * dr).sum() / (ref[0] * ref[0]).sum())),
                           "rho_max_abs": float(np.abs(dr).max())}
            print(f"# precision {name} {dt}: {res[dt]} ({time.perf_counter() - t0:.0f} s)", file=sys.stderr,
                  flush=True)
            del rho, u
        out[name] = {"nodes": list(d.nodes), "steps": steps, "errors_vs_f64": res}
    return out


def output_operators(n=512, reps=5):
    """The output path: plane_velocity_profile per axis and plot_line_velocity on the
    Couette n^3 FP64 state (device reductions, timed per call including the copy
    back), and write_paraview through hlbm_run on a 128^3 deck (host text)."""
    import subprocess
    import tempfile
    dom = couette_domain(n, "f64")
    dom.step(20, write_macro=True)
    dom.synchronize()
    N = n ** 3
    out = {"nodes": [n, n, n], "bytes_per_node": 3 * 8 + 1}
    for dim, name in enumerate("xyz"):
        dom.plane_profile(dim)  # warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            dom.plane_profile(dim, 1.0)
        ms = 1e3 * (time.perf_counter() - t0) / reps
        out[f"plane_profile_{name}_ms"] = ms
        out[f"plane_profile_{name}_gbs"] = 25.0 * N / (ms * 1e-3) / 1e9
    t0 = time.perf_counter()
    for _ in range(reps):
        dom.line_velocity((n // 2, n // 2, 0), 2, n)
    out["line_velocity_ms"] = 1e3 * (time.perf_counter() - t0) / reps
    dom.close()
    m = 128
    deck = f"""
do_domain:
  - domain:
     bounds: [[0,0,0],[{m - 1},{m - 1},{m - 1}]]
     cell_dims: [{m - 1},{m - 1},{m - 1}]
     periodic: [true, true, false]
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.8
     nuth: 0.1
pre_stream_bcs:
  - pre_bounce_back
  - cavity_z_l:
     U: [0.05, 0, 0]
post_stream_bcs:
  - post_bounce_back
global:
   simulation_end_iteration: 10
   simulation_print_log_freq: -1
"""
    run = os.path.join(ROOT, "hippolbm_amd", "hlbm_run")
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "io.msp")
        with open(path, "w") as fh:
            fh.write(deck)
        walls = []
        for extra in (["--no-output"], []):
            t0 = time.perf_counter()
            subprocess.run([run, path, "--quiet"] + extra, check=True, capture_output=True, cwd=tmp)
            walls.append(time.perf_counter() - t0)
        pv = os.path.join(tmp, "hippoLBMOutputDir", "ParaviewOutput")
        size = sum(os.path.getsize(os.path.join(dp, f)) for dp, _, fs in os.walk(pv) for f in fs)
    dumps = 2  # first time step + the final write
    out["paraview"] = {"nodes": [m, m, m], "dumps": dumps, "bytes_per_dump": size / dumps,
                       "s_per_dump": (walls[1] - walls[0]) / dumps,
                       "mb_per_s": size / 1e6 / max(walls[1] - walls[0], 1e-9)}
    return out


def copy_bandwidth():
    import ctypes
    from hippolbm_amd import _lib
    ms = ctypes.c_double()
    nbytes = 4 << 30
    _lib.check("hlbm_copy_bench", _lib.lib().hlbm_copy_bench(nbytes, 20, None, ctypes.byref(ms)))
    return 2 * nbytes / (ms.value * 1e-3) / 1e9


def _oracle_couette(orc, n):
    o = orc.Oracle((n, n, n), (True, True, False), 1.0)
    o.init_obst()
    o.set_distribution(1.0)
    o.macro_variables()

    def step():
        o.macro_variables()
        o.bgk(0.8)
        o.pre_bounce_back()
        o.cavity(1, (0.05, 0.0, 0.0))
        o.streaming()
        o.post_bounce_back()
    return step


def _time_sample(step, nodes, budget_s, max_steps, min_steps=2):
    step()  # first touch of the lattice outside the timed region
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps and (steps < min_steps or time.perf_counter() - t0 < budget_s):
        step()
        steps += 1
    dt = time.perf_counter() - t0
    return nodes * steps / dt / 1e6, steps, dt


def cpu_baseline(n=512, budget_s=10.0):
    """The FP64 oracle (a CPU restatement of the reference step chain, OpenMP
    static, the reference's SoA layout with two ghost layers) timed on this host:
    the headline figure on a bounded sample of the same Couette n^3 workload, on
    all host cores; beside it the 64^3 Couette and the porous MRT channel
    configurations on all cores, and the 64^3 Couette on one core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from hippolbm_amd import workloads
    cores = orc.lib().orc_threads()
    mlups, steps, dt = _time_sample(_oracle_couette(orc, n), n ** 3, budget_s, 400)
    legs = {}
    m, k, t = _time_sample(_oracle_couette(orc, 64), 64 ** 3, 3.0, 2000)
    legs["couette_64^3_bgk"] = {"mlups": m, "steps": k, "s": t, "cores": cores}
    run = orc.OracleRun(workloads.porous_channel())
    m, k, t = _time_sample(lambda: run.step(1), 256 * 128 * 128, 5.0, 5000)
    legs["porous_channel_256x128x128_mrt"] = {"mlups": m, "steps": k, "s": t, "cores": cores}
    del run
    orc.lib().orc_set_threads(1)
    try:
        m, k, t = _time_sample(_oracle_couette(orc, 64), 64 ** 3, 3.0, 2000)
    finally:
        orc.lib().orc_set_threads(cores)
    legs["couette_64^3_bgk_1core"] = {"mlups": m, "steps": k, "s": t, "cores": 1}
    return {"value": mlups, "unit": "MLUPS", "cores": cores, "kind": "port",
            "sample": f"oracle FP64 Couette {n}^3 BGK (the same workload), {steps} steps ({dt:.1f} s), "
                      "OpenMP static",
            "legs": legs}


def read_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    vals = {}
    for line in open(p):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--dtypes", default="f64,f32,f16,bf16")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only", default=None, help="run a single storage type (profiling)")
    ap.add_argument("--pattern", choices=("ab", "aa"), default="ab",
                    help="with --only: A/B buffers or the in-place AA pattern")
    ap.add_argument("--copy", action="store_true", help="run only the streaming-copy reference kernel (profiling)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-cavity and served-stream legs")
    ap.add_argument("--batch-only", nargs="?", const="both", choices=("both", "batched", "served"), default=None,
                    help="run only the batched and/or served legs")
    ap.add_argument("--batch-launch", choices=("graph", "stream"), default="graph",
                    help="batched and served legs: replay each chunk from a HIP graph, or launch its steps "
                         "one by one on the stream (per-dispatch profiling)")
    ap.add_argument("--batch-domains", type=int, default=2048,
                    help="batched leg: the first N cavities of the seeded mix (counter passes use a few hundred)")
    ap.add_argument("--batch-chunk", type=int, default=200, help="batched leg: steps per chunk")
    ap.add_argument("--precision", type=int, default=0, metavar="STEPS",
                    help="run only the precision sweep (Couette 512^3, Poiseuille 512x256x256) for STEPS steps")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape (no autotuning)")
    ap.add_argument("--io", action="store_true", help="run only the output-operator measurements")
    ap.add_argument("--porous", action="store_true", help="run only the porous-channel leg (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the domain-decomposed multi-GPU path is a design note, not code)", file=sys.stderr)
        return 2
    if args.copy:
        print(json.dumps({"copy_kernel_gbs": copy_bandwidth(), "bytes_read_per_launch": 4 << 30,
                          "bytes_written_per_launch": 4 << 30}), flush=True)
        return 0
    if args.io:
        print(json.dumps({"output_operators": output_operators(args.size)}), flush=True)
        return 0
    if args.porous:
        print(json.dumps({"porous_channel": porous_channel(args.steps, args.warmup, tune=not args.no_tune)}), flush=True)
        return 0
    if args.precision:
        print(json.dumps({"precision_sweep": precision_sweep(args.precision)}), flush=True)
        return 0
    if args.only:
        r = measure(args.size, args.only, args.steps, args.warmup, tune=not args.no_tune, pattern=args.pattern)
        print(json.dumps({"storage": args.only, "pattern": args.pattern, **r}), flush=True)
        return 0
    if args.batch_only:
        g = args.batch_launch == "graph"
        out = {"batch_launch": args.batch_launch}
        if args.batch_only in ("both", "batched"):
            out["batched"] = batched_cavities(n_dom=args.batch_domains, chunk=args.batch_chunk, use_graph=g)
        if args.batch_only in ("both", "served"):
            out["served"] = served_sweep(use_graph=g)
        print(json.dumps(out), flush=True)
        return 0
    n = args.size
    per = {}
    for dt in [d for d in args.dtypes.split(",") if d]:
        per[dt] = measure(n, dt, args.steps, args.warmup, tune=not args.no_tune)
        print(f"# {dt}: {per[dt]['mlups']:.0f} MLUPS, {per[dt]['achieved_gbs']:.0f} GB/s "
              f"({100 * per[dt]['frac']:.1f}% of 8 TB/s)", file=sys.stderr, flush=True)
    main_r = per.get("f64") or next(iter(per.values()))
    # the AA pattern (one lattice, in place): same bytes per update, half the footprint
    aa = {}
    for dt in ("f64", "f16"):
        r = measure(n, dt, args.steps, args.warmup, tune=not args.no_tune, pattern="aa")
        aa[dt] = {"mlups": r["mlups"], "frac": r["frac"], "ms_per_step": r["ms_per_step"], "launch": r["launch"]}
        print(f"# AA {dt}: {r['mlups']:.0f} MLUPS ({100 * r['frac']:.1f}%)", file=sys.stderr, flush=True)
    copy_gbs = copy_bandwidth()
    porous = porous_channel()
    print(f"# porous channel MRT f64: {porous['mlups']:.0f} MLUPS ({100 * porous['frac']:.1f}%)", file=sys.stderr,
          flush=True)
    batched = served = None
    if not args.no_batch:
        batched = batched_cavities()
        print(f"# batched cavities: {batched['mlups']:.0f} MLUPS ({100 * batched['frac']:.1f}%)", file=sys.stderr,
              flush=True)
        served = served_sweep()
    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": "MLUPS (D3Q19 BGK FP64 Couette 512^3, one MI355X) + % HBM roofline per storage dtype",
        "value": main_r["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (f_i = w_i, rho = 1)",
        "config": {"workload": f"couette_{n}^3_bgk_f64", "nodes": [n, n, n], "periodic": [True, True, False],
                   "tau": 0.8, "bcs": "pre_bounce_back + cavity_z_l(U=0.05) + post_bounce_back",
                   "pattern": "AB pull, fused collide+stream"},
        "roofline": {"bound": "hbm", "achieved": main_r["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": main_r["frac"], "traffic": read_traffic(f"traffic_couette{n}_f64.txt"),
                     "bytes_per_node": main_r["bytes_per_node"], "copy_kernel_gbs": copy_gbs,
                     "launch": main_r["launch"]},
        "per_dtype": {k: {"mlups": v["mlups"], "frac": v["frac"], "achieved_gbs": v["achieved_gbs"],
                          "bytes_per_node": v["bytes_per_node"], "ms_per_step": v["ms_per_step"]}
                      for k, v in per.items()},
        "latency": ({"p50_s": served["f32"][0]["p50_s"], "p99_s": served["f32"][0]["p99_s"],
                     "requests": served["f32"][0]["n"], "failed": len(served["f32"][0]["failed"]),
                     "aggregate_mlups": served["f32"][0]["mlups"], "frac": served["f32"][0]["frac"],
                     "workload": f"served stream, FP32 storage, Poisson arrivals at {served['f32'][0]['rate_hz']} Hz"}
                    if served else None),
        "aa_pattern": aa,
        "porous_channel": porous,
        "batched": batched,
        "served": served,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

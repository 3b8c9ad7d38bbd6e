"""Batched domains and the solve server (``hlbm_batch_*`` in include/hlbm.h).

``Batch`` owns an HBM budget in which many independent solves of uneven sizes
run together: one kernel launch per collision model and time step over all
running solves (one workgroup per 1024-node tile), chunks of ``chunk_steps``
replayed from a captured HIP graph, solves admitted as memory frees up and
retired (rho/u copied out) as they finish.
``serve`` drives a Poisson stream of requests through it and reports
time-to-result percentiles beside the aggregate MLUPS.  A solve whose final
state is not finite ends FAILED (its own status; the other solves are not
affected).
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import _lib
from ._lib import check, lib

P = ctypes.c_void_p


class BatchDesc(ctypes.Structure):
    _fields_ = [("storage", ctypes.c_int32), ("chunk_steps", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("round_nearest", ctypes.c_int32), ("budget_bytes", ctypes.c_uint64)]


class SolveDesc(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.c_int32 * 3), ("periodic", ctypes.c_int32 * 3), ("collision", ctypes.c_int32),
        ("steps", ctypes.c_int32), ("wall_bb", ctypes.c_int32), ("n_obstacles", ctypes.c_int32),
        ("cavity_on", ctypes.c_int32 * 2), ("keep_populations", ctypes.c_int32), ("mrt_custom", ctypes.c_int32),
        ("tau", ctypes.c_double), ("fext", ctypes.c_double * 3), ("dx", ctypes.c_double),
        ("init_rho", ctypes.c_double), ("cavity_u", (ctypes.c_double * 3) * 2), ("mrt_rates", ctypes.c_double * 10),
        ("obstacles", ctypes.POINTER(_lib.Obstacle)), ("init_f", ctypes.POINTER(ctypes.c_double)),
    ]


# enum hlbm_solve_state
QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3


class BatchInfo(ctypes.Structure):
    _fields_ = [("queued", ctypes.c_int64), ("running", ctypes.c_int64), ("done", ctypes.c_int64),
                ("chunks", ctypes.c_int64), ("bytes_in_use", ctypes.c_uint64), ("budget_bytes", ctypes.c_uint64),
                ("node_updates", ctypes.c_double), ("gpu_ms", ctypes.c_double),
                ("peak_bytes_in_use", ctypes.c_uint64), ("relocated", ctypes.c_int64)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = lib()
    I64P = ctypes.POINTER(ctypes.c_int64)
    DP = ctypes.POINTER(ctypes.c_double)
    sig = {
        "hlbm_batch_create": ([ctypes.POINTER(BatchDesc), ctypes.POINTER(P)], ctypes.c_int),
        "hlbm_batch_destroy": ([P], ctypes.c_int),
        "hlbm_batch_submit": ([P, ctypes.POINTER(SolveDesc), I64P], ctypes.c_int),
        "hlbm_batch_run": ([P, ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
        "hlbm_batch_status": ([P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32), I64P], ctypes.c_int),
        "hlbm_batch_times": ([P, ctypes.c_int64, DP, DP, DP], ctypes.c_int),
        "hlbm_batch_result": ([P, ctypes.c_int64, DP, DP, DP], ctypes.c_int),
        "hlbm_batch_stats": ([P, ctypes.c_int64, DP], ctypes.c_int),
        "hlbm_batch_info_get": ([P, ctypes.POINTER(BatchInfo)], ctypes.c_int),
        "hlbm_clock": ([], ctypes.c_double),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _bound = True


@dataclasses.dataclass
class Solve:
    """One request: geometry, parameters, boundary conditions, step count."""
    nodes: tuple
    steps: int
    tau: float
    periodic: tuple = (False, False, False)
    collision: str = "bgk"
    fext: tuple = (0.0, 0.0, 0.0)
    cavity_z0: tuple | None = None      # moving wall velocity on the low z face (cavity_z_0)
    cavity_zl: tuple | None = None      # moving wall velocity on the high z face (cavity_z_l)
    obstacles: list = dataclasses.field(default_factory=list)   # dicts as in deck.obstacles
    wall_bb: bool = True
    dx: float = 1.0
    init_rho: float = 1.0
    keep_populations: bool = False
    init_f: object = None   # optional FP64 initial populations, shape (19, nz, ny, nx)

    def node_count(self):
        return int(self.nodes[0]) * int(self.nodes[1]) * int(self.nodes[2])

    def to_desc(self):
        d = SolveDesc()
        for k in range(3):
            d.nodes[k] = int(self.nodes[k])
            d.periodic[k] = 1 if self.periodic[k] else 0
            d.fext[k] = float(self.fext[k])
        d.collision = _lib.COLLISION[self.collision]
        d.steps = int(self.steps)
        d.tau = float(self.tau)
        d.dx = float(self.dx)
        d.init_rho = float(self.init_rho)
        d.keep_populations = 1 if self.keep_populations else 0
        for side, u in ((0, self.cavity_z0), (1, self.cavity_zl)):
            if u is not None:
                d.cavity_on[side] = 1
                for k in range(3):
                    d.cavity_u[side][k] = float(u[k])
        obs = [_lib.make_obstacle(ob) for ob in self.obstacles]
        keep = None
        if obs:
            keep = (_lib.Obstacle * len(obs))(*obs)
            d.obstacles = ctypes.cast(keep, ctypes.POINTER(_lib.Obstacle))
            d.n_obstacles = len(obs)
        d.wall_bb = 1 if (self.wall_bb and obs) else 0
        f0 = None
        if self.init_f is not None:
            f0 = np.ascontiguousarray(self.init_f, dtype=np.float64).reshape(-1)
            if f0.size != 19 * self.node_count():
                raise ValueError("init_f must hold 19 populations per node")
            d.init_f = f0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return d, (keep, f0)

    def run_single(self, storage="f64"):
        """The same solve on the single-domain path (reference operator order)."""
        from .engine import Domain
        dom = Domain(self.nodes, self.periodic, self.tau, self.fext, storage=storage, collision=self.collision,
                     dx=self.dx)
        bcs = [("pre_stream_bcs", "pre_bounce_back", {})]
        if self.cavity_z0 is not None:
            bcs.append(("pre_stream_bcs", "cavity_z_0", {"U": self.cavity_z0}))
        if self.cavity_zl is not None:
            bcs.append(("pre_stream_bcs", "cavity_z_l", {"U": self.cavity_zl}))
        if self.wall_bb and self.obstacles:
            bcs.append(("pre_stream_bcs", "wall_bounce_back", {}))
        bcs.append(("post_stream_bcs", "post_bounce_back", {}))
        dom.set_bcs(bcs)
        dom.init_obst()
        dom.set_distribution(self.init_rho)
        dom.add_obstacles(self.obstacles)
        if self.init_f is not None:
            dom.set_f(np.asarray(self.init_f, dtype=np.float64).reshape((19,) + tuple(self.nodes[::-1])))
        dom.step(self.steps)
        rho, u = dom.macro()
        f = dom.get_f() if self.keep_populations else None
        dom.close()
        return rho, u, f


class Batch:
    def __init__(self, storage="f16", budget_bytes=8 << 30, chunk_steps=200, use_graph=True, rounding="stochastic"):
        _bind()
        d = BatchDesc()
        d.storage = _lib.STORAGE[storage]
        d.round_nearest = 1 if rounding == "nearest" else 0
        d.chunk_steps = int(chunk_steps)
        d.use_graph = 1 if use_graph else 0
        d.budget_bytes = int(budget_bytes)
        h = P()
        check("hlbm_batch_create", lib().hlbm_batch_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.storage = storage
        self.specs = {}

    def close(self):
        if getattr(self, "h", None):
            lib().hlbm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, solve: Solve):
        d, keep = solve.to_desc()
        sid = ctypes.c_int64()
        check("hlbm_batch_submit", lib().hlbm_batch_submit(self.h, ctypes.byref(d), ctypes.byref(sid)))
        del keep  # copied by the library
        self.specs[sid.value] = solve
        return sid.value

    def run(self):
        n = ctypes.c_int32()
        check("hlbm_batch_run", lib().hlbm_batch_run(self.h, ctypes.byref(n)))
        return n.value

    def state(self, sid):
        st, done = ctypes.c_int32(), ctypes.c_int64()
        check("hlbm_batch_status", lib().hlbm_batch_status(self.h, sid, ctypes.byref(st), ctypes.byref(done)))
        return st.value, done.value

    def stats(self, sid):
        """simulation_state of a finished solve: sum rho, min |u|, max |u|, sum rho u (3), fluid nodes."""
        out = np.zeros(7)
        check("hlbm_batch_stats", lib().hlbm_batch_stats(self.h, sid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def times(self, sid):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check("hlbm_batch_times", lib().hlbm_batch_times(self.h, sid, ctypes.byref(a), ctypes.byref(b),
                                                          ctypes.byref(c)))
        return a.value, b.value, c.value

    def result(self, sid):
        s = self.specs.pop(sid)
        n = s.node_count()
        rho = np.zeros(n)
        u = np.zeros(3 * n)
        f = np.zeros(19 * n) if s.keep_populations else None
        fp = f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if f is not None else None
        check("hlbm_batch_result", lib().hlbm_batch_result(
            self.h, sid, rho.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), fp))
        shp = (s.nodes[2], s.nodes[1], s.nodes[0])
        return rho.reshape(shp), u.reshape((3,) + shp), (f.reshape((19,) + shp) if f is not None else None)

    def release(self, sid):
        """Forget a finished solve without copying its fields."""
        self.specs.pop(sid)
        check("hlbm_batch_result", lib().hlbm_batch_result(self.h, sid, None, None, None))

    def info(self):
        i = BatchInfo()
        check("hlbm_batch_info_get", lib().hlbm_batch_info_get(self.h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in BatchInfo._fields_}


def clock():
    _bind()
    return lib().hlbm_clock()


def poisson_requests(n, rate_hz, seed, size_range=(32, 256), steps_range=(500, 5000), storage_kinds=None):
    """A seeded request stream: arrival offsets (s) and solves.

    Kinds cycle through a lid-driven cavity (closed box, moving lid), a periodic
    channel driven by a body force, and a porous box (seeded non-overlapping-ish
    spheres, wall bounce-back, force); edge lengths are log-uniform in
    size_range per axis, step counts uniform in steps_range, BGK or MRT.
    """
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += rng.exponential(1.0 / rate_hz) if rate_hz > 0 else 0.0
        edges = np.exp(rng.uniform(np.log(size_range[0]), np.log(size_range[1]), 3)).astype(int)
        nodes = tuple(int(max(e, 8)) for e in edges)
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        coll = "bgk" if rng.uniform() < 0.5 else "mrt"
        kind = k % 3
        if kind == 0:
            s = Solve(nodes=nodes, steps=steps, tau=float(rng.uniform(0.52, 1.0)), collision=coll,
                      cavity_zl=(float(rng.uniform(0.01, 0.1)), 0.0, 0.0))
        elif kind == 1:
            s = Solve(nodes=nodes, steps=steps, tau=float(rng.uniform(0.55, 1.0)), periodic=(True, True, False),
                      fext=(1e-6, 0.0, 0.0), collision=coll)
        else:
            m = min(nodes)
            balls = []
            for _ in range(8):
                r = float(rng.uniform(0.05, 0.12) * m)
                balls.append({"kind": "ball", "center": tuple(float(rng.uniform(r, v - 1 - r)) for v in nodes),
                              "radius": r})
            s = Solve(nodes=nodes, steps=steps, tau=0.7, fext=(2e-6, 0.0, 0.0), obstacles=balls, collision=coll)
        out.append((t, s))
    return out


def serve(batch, requests, keep_results=False, poll_sleep=0.0005):
    """Continuous batching: requests join at chunk boundaries as they arrive and as
    memory frees; returns per-request time-to-result (s) and aggregate figures."""
    import time
    t0 = clock()
    gpu0 = batch.info()["gpu_ms"] * 1e-3
    pending = list(requests)
    inflight = {}
    lat = []
    results = {}
    failed = []
    node_updates = 0.0
    while pending or inflight:
        now = clock() - t0
        while pending and pending[0][0] <= now:
            arr, spec = pending.pop(0)
            sid = batch.submit(spec)
            inflight[sid] = (arr, spec)
        if not inflight:
            time.sleep(min(poll_sleep, max(0.0, pending[0][0] - now)))
            continue
        batch.run()
        for sid in list(inflight):
            st, _ = batch.state(sid)
            if st in (DONE, FAILED):
                arr, spec = inflight.pop(sid)
                _, _, t_done = batch.times(sid)
                lat.append(t_done - t0 - arr)
                node_updates += spec.node_count() * spec.steps
                if st == FAILED:
                    failed.append(sid)
                r = batch.result(sid)
                if keep_results:
                    results[sid] = (spec, r, st)
    makespan = clock() - t0
    lat = np.array(lat)
    gpu_s = batch.info()["gpu_ms"] * 1e-3 - gpu0
    return {"n": len(lat), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "makespan_s": makespan, "node_updates": node_updates,
            "mlups": node_updates / makespan / 1e6, "gpu_busy": gpu_s / makespan, "failed": failed,
            "results": results}

"""Batched domains / solve server: each solve's result is the same alone (the
single-domain path) and in any mix, with any chunk length, with or without the
captured graph, and under a memory budget that forces queued admission."""
import sys

import numpy as np
import pytest

from hippolbm_amd.batch import Batch, Solve

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mix(seed, n=6, keep=True):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nodes = tuple(int(v) for v in rng.integers(6, 23, size=3))
        kind = k % 3
        if kind == 0:  # closed cavity with a moving lid
            out.append(Solve(nodes=nodes, steps=int(rng.integers(5, 40)), tau=float(rng.uniform(0.52, 1.0)),
                             cavity_zl=(float(rng.uniform(0.01, 0.1)), 0.0, 0.0), keep_populations=keep))
        elif kind == 1:  # periodic channel driven by a body force, closed in z
            out.append(Solve(nodes=nodes, steps=int(rng.integers(5, 40)), tau=float(rng.uniform(0.6, 1.0)),
                             periodic=(True, True, False), fext=(1e-5, 0.0, 0.0), collision="mrt",
                             keep_populations=keep))
        else:  # porous box: balls away from the faces, wall bounce-back
            balls = [{"kind": "ball", "center": tuple(float(rng.uniform(2, m - 3)) for m in nodes),
                      "radius": float(rng.uniform(1.0, 2.0))} for _ in range(2)]
            out.append(Solve(nodes=nodes, steps=int(rng.integers(5, 40)), tau=0.7, fext=(0.0, 2e-5, 0.0),
                             obstacles=balls, cavity_z0=(0.0, 0.02, 0.0), keep_populations=keep))
    return out


def run_all(b, ids, max_chunks=10000):
    done = {}
    chunks = 0
    while len(done) < len(ids) and chunks < max_chunks:
        b.run()
        chunks += 1
        for sid in ids:
            if sid not in done and b.state(sid)[0] == 2:  # DONE
                done[sid] = b.result(sid)
    assert len(done) == len(ids), "batch did not finish"
    return done


@pytest.mark.parametrize("storage", ["f64", "f16"])
def test_batch_equals_single_domain(storage):
    solves = mix(11)
    b = Batch(storage=storage, budget_bytes=512 << 20, chunk_steps=8, use_graph=True)
    ids = [b.submit(s) for s in solves]
    res = run_all(b, ids)
    for sid, s in zip(ids, solves):
        rho, u, f = res[sid]
        r1, u1, f1 = s.run_single(storage)
        say(storage, s.nodes, s.steps, "max|df| vs single-domain path", float(np.abs(f - f1).max()))
        if storage == "f64":  # both paths are built without contraction: bit-identical
            assert np.array_equal(f, f1)
            assert np.array_equal(rho, r1)
            assert np.array_equal(u, u1)
        else:  # FP32 arithmetic may contract differently in the two kernels, which can move a
            # stored value by one FP16 step (1.5e-5 at |g| ~ 0.03)
            assert np.abs(f - f1).max() < 5e-5
    b.close()


@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "bf16"])
def test_alone_equals_in_the_mix(storage):
    solves = mix(13)
    b = Batch(storage=storage, budget_bytes=512 << 20, chunk_steps=8)
    ids = [b.submit(s) for s in solves]
    res = run_all(b, ids)
    b.close()
    for k in (0, 2, 4):
        a = Batch(storage=storage, budget_bytes=512 << 20, chunk_steps=8)
        sid = a.submit(solves[k])
        alone = run_all(a, [sid])[sid]
        a.close()
        assert np.array_equal(alone[2], res[ids[k]][2])
        assert np.array_equal(alone[1], res[ids[k]][1])


@pytest.mark.parametrize("storage", ["f64", "f16"])
def test_arena_placement_move_does_not_change_results(storage, monkeypatch):
    """HLBM_PLACEMENT=force makes hlbm_batch_create take every fresh arena pair it
    probes; the solves then run in the moved arenas with the same results."""
    solves = mix(14)
    outs = []
    for mode in ("0", "force"):
        monkeypatch.setenv("HLBM_PLACEMENT", mode)
        b = Batch(storage=storage, budget_bytes=512 << 20, chunk_steps=8)
        moved = b.info()["relocated"]
        say(storage, "HLBM_PLACEMENT", mode, "arena pairs moved", moved)
        assert moved == (0 if mode == "0" else 3)
        ids = [b.submit(s) for s in solves]
        res = run_all(b, ids)
        outs.append([res[i] for i in ids])
        b.close()
    for a, c in zip(*outs):
        for x, y in zip(a, c):
            assert np.array_equal(x, y)


def test_chunk_length_and_graph_do_not_change_results():
    solves = mix(12)
    outs = []
    for chunk, graph in ((8, True), (5, False), (64, True)):
        b = Batch(storage="f32", budget_bytes=512 << 20, chunk_steps=chunk, use_graph=graph)
        ids = [b.submit(s) for s in solves]
        res = run_all(b, ids)
        outs.append([res[i][2] for i in ids])
        b.close()
    for other in outs[1:]:
        for a, c in zip(outs[0], other):
            assert np.array_equal(a, c)


def test_admission_under_a_small_budget():
    # room for about two solves at a time: the rest wait in the queue
    solves = [Solve(nodes=(20, 20, 20), steps=17 + 3 * k, tau=0.8, cavity_zl=(0.05, 0, 0), keep_populations=True)
              for k in range(6)]
    per_solve = 20 ** 3 * (38 * 2 + 1)
    b = Batch(storage="f16", budget_bytes=int(2.5 * per_solve), chunk_steps=10)
    ids = [b.submit(s) for s in solves]
    b.run()
    info = b.info()
    assert info["running"] <= 2 and info["queued"] >= 4
    res = run_all(b, ids)
    for sid, s in zip(ids, solves):
        _, _, f1 = s.run_single("f16")
        assert np.abs(res[sid][2] - f1).max() < 5e-5
    b.close()


def test_oversized_solve_is_rejected():
    b = Batch(storage="f16", budget_bytes=1 << 20, chunk_steps=4)
    from hippolbm_amd._lib import HLBMError
    with pytest.raises(HLBMError):
        b.submit(Solve(nodes=(64, 64, 64), steps=1, tau=0.8))
    b.close()


# ---------------------------------------------------------------------------
# batch path against the FP64 oracle
# ---------------------------------------------------------------------------
# tau values for which nuth = (tau - 1/2)/3 and 3*nuth are exact, so that the
# deck's unit conversion (lbm_parameters.cpp:73-99) gives dtLB = 1 and the
# oracle runs on exactly the solve's force and relaxation time
EXACT_TAU = (0.546875, 0.59375, 0.6875, 0.875, 1.25)


def solve_deck(s):
    """The hippoLBM deck that runs solve `s`: lattice units (dx = 1), every
    closed face with pre/post_bounce_back, cavity_z_* for moving walls and
    wall_bounce_back for obstacles, in the order the batch applies them."""
    from hippolbm_amd import deck as dk
    n = s.nodes
    pre = ["  - pre_bounce_back"]
    for name, u in (("cavity_z_0", s.cavity_z0), ("cavity_z_l", s.cavity_zl)):
        if u is not None:
            pre.append(f"  - {name}:\n     U: [{u[0]!r}, {u[1]!r}, {u[2]!r}]")
    if s.obstacles and s.wall_bb:
        pre.append("  - wall_bounce_back")
    obs = []
    for k, ob in enumerate(s.obstacles):
        c = ob["center"]
        obs.append(f"  - register_solid_ball:\n     id: {k}\n     center: [{c[0]!r}, {c[1]!r}, {c[2]!r}]\n"
                   f"     radius: {ob['radius']!r}")
    per = ", ".join("true" if p else "false" for p in s.periodic)
    txt = f"""
do_domain:
  - domain:
     bounds: [[0,0,0],[{n[0] - 1},{n[1] - 1},{n[2] - 1}]]
     cell_dims: [{n[0] - 1},{n[1] - 1},{n[2] - 1}]
     periodic: [{per}]
set_lbm_parameters:
  - lbm_parameters:
     tau: {s.tau!r}
     nuth: {(s.tau - 0.5) / 3!r}
     Fext: [{s.fext[0]!r}, {s.fext[1]!r}, {s.fext[2]!r}]
collision: {s.collision}
pre_stream_bcs:
{chr(10).join(pre)}
post_stream_bcs:
  - post_bounce_back
"""
    if obs:
        txt += "set_obstacles:\n" + "\n".join(obs) + "\n"
    d = dk.loads_msp(txt, "solve")
    p = d.params
    assert p["tau"] == s.tau and p["Fext"] == tuple(float(v) for v in s.fext), "inexact unit conversion"
    return d


def oracle_mix(seed, n=8):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nodes = tuple(int(v) for v in rng.integers(6, 19, size=3))
        tau = float(EXACT_TAU[int(rng.integers(len(EXACT_TAU)))])
        steps = int(rng.integers(5, 30))
        kind = k % 4
        if kind == 0:  # regression_lbm_cavite's operator list: pre_bounce_back + cavity_z_l + post_bounce_back
            out.append(Solve(nodes=nodes, steps=steps, tau=tau, cavity_zl=(float(rng.uniform(0.01, 0.1)), 0.0, 0.0),
                             keep_populations=True))
        elif kind == 1:  # periodic channel, body force, MRT
            out.append(Solve(nodes=nodes, steps=steps, tau=tau, periodic=(True, True, False),
                             fext=(2.0 ** -17, 0.0, 0.0), collision="mrt", keep_populations=True))
        elif kind == 2:  # porous box: balls off the faces, wall bounce-back, both moving walls
            balls = [{"kind": "ball", "center": tuple(float(rng.uniform(2.5, m - 3.5)) for m in nodes),
                      "radius": float(rng.uniform(1.0, 2.0))} for _ in range(2)]
            out.append(Solve(nodes=nodes, steps=steps, tau=tau, fext=(0.0, 2.0 ** -16, 0.0), obstacles=balls,
                             cavity_z0=(0.0, 0.02, 0.0), cavity_zl=(0.03, 0.0, 0.0), keep_populations=True))
        else:  # x-periodic duct with MRT and a ball
            balls = [{"kind": "ball", "center": tuple(float(rng.uniform(2.5, m - 3.5)) for m in nodes),
                      "radius": 1.5}]
            out.append(Solve(nodes=nodes, steps=steps, tau=tau, periodic=(True, False, False),
                             fext=(2.0 ** -16, 0.0, 0.0), obstacles=balls, collision="mrt", keep_populations=True))
    return out


def test_batch_fp64_matches_the_oracle():
    import oracle as orc
    solves = oracle_mix(21)
    b = Batch(storage="f64", budget_bytes=256 << 20, chunk_steps=7, use_graph=True)
    ids = [b.submit(s) for s in solves]
    res = run_all(b, ids)
    b.close()
    for sid, s in zip(ids, solves):
        ref = orc.OracleRun(solve_deck(s))
        ref.step(s.steps)
        fo = ref.o.node_f()
        fluid = ref.o.real(ref.o.obst()) == -1
        rho, u, f = res[sid]
        nbad = int((f[:, fluid] != fo[:, fluid]).sum())
        say(s.nodes, s.steps, s.collision, "populations differing from the oracle:", nbad)
        assert nbad == 0
        # rho of the final post-stream state against the oracle's populations (sum in order)
        rho_o = fo[0].copy()
        for i in range(1, 19):
            rho_o = rho_o + fo[i]
        assert np.array_equal(rho[fluid], rho_o[fluid])


@pytest.mark.parametrize("storage", ["f64", "f16"])
def test_non_finite_solve_fails_alone(storage):
    from hippolbm_amd import batch as bt
    good = mix(17, n=3, keep=True)
    # a solve whose initial populations (built on the host) hold one NaN
    nodes = (10, 9, 8)
    f0 = np.empty((19,) + nodes[::-1])
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    f0[:] = w[:, None, None, None]
    f0[5, 4, 4, 4] = np.nan
    bad = Solve(nodes=nodes, steps=12, tau=0.8, cavity_zl=(0.05, 0, 0), init_f=f0)
    b = Batch(storage=storage, budget_bytes=256 << 20, chunk_steps=5)
    ids = [b.submit(s) for s in good[:2]] + [b.submit(bad)] + [b.submit(good[2])]
    done = {}
    while len(done) < len(ids):
        b.run()
        for sid in ids:
            if sid not in done:
                st = b.state(sid)[0]
                if st in (bt.DONE, bt.FAILED):
                    done[sid] = (st, b.stats(sid), b.result(sid))
    b.close()
    st, stats, _ = done[ids[2]]
    say(storage, "bad solve state", st, "stats", stats)
    assert st == bt.FAILED and not np.isfinite(stats[0])
    for k, sid in enumerate(ids[:2] + ids[3:]):
        st, stats, (rho, u, f) = done[sid]
        assert st == bt.DONE and np.all(np.isfinite(stats))
        a = Batch(storage=storage, budget_bytes=256 << 20, chunk_steps=5)
        s = (good[:2] + good[2:])[k]
        aid = a.submit(s)
        alone = run_all(a, [aid])[aid]
        a.close()
        assert np.array_equal(alone[2], f) and np.array_equal(alone[0], rho)


def test_initial_populations_are_used():
    # a seeded non-equilibrium start: the batch and the single-domain path agree bit for bit
    rng = np.random.default_rng(3)
    nodes = (9, 8, 7)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    f0 = w[:, None, None, None] * (1.0 + 1e-3 * rng.uniform(-1, 1, (19,) + nodes[::-1]))
    s = Solve(nodes=nodes, steps=9, tau=0.7, cavity_zl=(0.04, 0, 0), init_f=f0, keep_populations=True)
    b = Batch(storage="f64", budget_bytes=64 << 20, chunk_steps=4)
    sid = b.submit(s)
    rho, u, f = run_all(b, [sid])[sid]
    b.close()
    r1, u1, f1 = s.run_single("f64")
    assert np.array_equal(f, f1) and np.array_equal(rho, r1)

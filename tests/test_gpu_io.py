"""Output operators against the oracle.

* plane_velocity_profile statistics (hlbm_plane_profile) and plot_line_velocity
  samples (hlbm_line_velocity) computed on the device from the engine's u
  field, vs the same quantities taken from the oracle's m1/obst after the same
  steps: counts, minima, maxima and samples bit-equal, sums to 1e-12 relative
  (device tree order vs the reference's atomics);
* the files `hlbm_run` writes for a deck with an analysis block and a
  paraview frequency: names and directory layout of write_paraview (.pvtr and
  .vtr), plane_velocity_profile and plot_line_velocity, and their text, built
  here from the oracle's state at the same time steps
  (src/io/include/hippoLBM/io/write_paraview.hpp:139-327,
  src/check/plane_velocity_profile.cu:228-250, src/check/plot_over_line.cu:186-210).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as orc
from hippolbm_amd import deck as dk
from hippolbm_amd.engine import Simulation

from test_gpu_parity import _porous_deck

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "hippolbm_amd", "hlbm_run")
DBL_MAX = np.finfo(np.float64).max


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_profile(u, fluid, dim, scale):
    """Per-plane sum, count, min, max of |scale*u| over fluid nodes; u (3, nz, ny, nx)."""
    ax = {0: 2, 1: 1, 2: 0}[dim]  # array axis of the node coordinate
    a, b, c = scale * u[0], scale * u[1], scale * u[2]
    v = np.sqrt(a * a + b * b + c * c)
    n = u.shape[1 + ax]
    out = np.zeros(n), np.zeros(n, dtype=np.int32), np.full(n, DBL_MAX), np.full(n, -DBL_MAX)
    for p in range(n):
        sl = [slice(None)] * 3
        sl[ax] = p
        vp, fp = v[tuple(sl)], fluid[tuple(sl)]
        sel = vp[fp]
        out[1][p] = sel.size
        if sel.size:
            out[0][p] = np.sum(sel)
            out[2][p] = sel.min()
            out[3][p] = sel.max()
    return out


@pytest.mark.parametrize("dim", [0, 1, 2])
def test_plane_profile_matches_oracle(dim):
    d = _porous_deck(30, 20, 18, 6, 4, "bgk")
    say("plane profile dim", dim)
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64")
    ref.step(17)
    sim.run(17)
    # held u (macro of f^16) first, then the current state's macro
    for recompute in (False, True):
        if recompute:
            sim.dom.macro()
            ref.o.macro_variables()
        o = ref.o
        fluid = o.real(o.obst()) == -1
        scale = 0.37
        s, c, lo, hi = sim.dom.plane_profile(dim, scale)
        es, ec, elo, ehi = _oracle_profile(o.real(o.m1()), fluid, dim, scale)
        assert np.array_equal(c, ec)
        assert np.array_equal(lo, elo) and np.array_equal(hi, ehi)
        assert np.allclose(s, es, rtol=1e-12, atol=0.0)
    sim.dom.close()


def test_plane_profile_all_wall_plane_reports_reset_values():
    # a wall slab filling the plane x = 3: count 0, min DBL_MAX, max -DBL_MAX
    txt = """
do_domain:
  - domain:
     bounds: [[0,0,0],[11,9,7]]
     cell_dims: [11,9,7]
     periodic: [true, true, true]
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.8
     nuth: 0.1
     Fext: [1.0e-5, 0, 0]
set_obstacles:
  - register_solid_wall:
     id: 0
     bounds: [[2.5,-1,-1],[3.5,10,8]]
"""
    d = dk.loads_msp(txt, "slab")
    sim = Simulation(d, storage="f64")
    sim.run(3)
    s, c, lo, hi = sim.dom.plane_profile(0, 1.0)
    assert c[3] == 0 and lo[3] == DBL_MAX and hi[3] == -DBL_MAX and s[3] == 0.0
    assert all(c[x] == 10 * 8 for x in range(12) if x != 3)
    sim.dom.close()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_line_velocity_matches_oracle(axis):
    d = _porous_deck(30, 20, 18, 6, 4, "mrt")
    say("line axis", axis)
    ref = orc.OracleRun(d)
    sim = Simulation(d, storage="f64")
    ref.step(11)
    sim.run(11)
    n = d.nodes
    start = [n[0] // 3, n[1] // 2, n[2] // 4]
    start[axis] = -2  # the line starts and ends outside the domain
    count = n[axis] + 4
    scale = 2.5
    got = sim.dom.line_velocity(start, axis, count, scale)
    u = ref.o.real(ref.o.m1())
    exp = np.zeros((count, 3))
    for k in range(count):
        p = list(start)
        p[axis] += k
        if all(0 <= p[i] < n[i] for i in range(3)):
            exp[k] = scale * u[:, p[2], p[1], p[0]]
    assert np.array_equal(got, exp)
    sim.dom.close()


# ---- hlbm_run output files ----------------------------------------------------

IO_DECK = """
do_domain:
  - domain:
     bounds: [[0,0,0],[0.019,0.019,0.019]]
     cell_dims: [19,19,19]
     periodic: [true, false, false]
set_lbm_parameters:
  - lbm_parameters:
     nuth: 1e-3
     tau: 0.7
     avg_rho: 998.0
boundary_conditions:
  - neumann:
     U: [0.02,0,0]
     regions: [plan_xy_l]
  - neumann:
     U: [0,0,0]
     regions: [plan_xy_0]
set_obstacles:
  - register_solid_ball:
     id: 0
     center: [0.009, 0.009, 0.008]
     radius: 0.004
analysis:
  - plane_velocity_profile:
     dimension: Z
  - plot_line_velocity:
     line: [[0.01,0.01,0],[0.01,0.01,0.019]]
  - plot_line_velocity:
     line_lbm: [[-2,5,12],[21,5,12]]
     dump_file: "xline_%05d.csv"
global:
   simulation_paraview_freq: 4
   simulation_end_iteration: 8
   simulation_analysis_freq: 4
   simulation_print_log_freq: 4
   output_directory: "IoTestDir"
"""


def _g(v):
    return "%g" % v


def _vtr_expected(d, o, ts):
    """Text of write_paraview's .vtr for the oracle's Real-node state."""
    p = d.params
    n = d.nodes
    L = [v + 4 for v in n]
    dx = d.dx
    pc = 1. / 3. * p["celerity"] * p["celerity"] * d.lbm["avg_rho"]
    uc = dx / p["dtLB"]
    m0, m1, obst = o.real(o.m0()), o.real(o.m1()), o.real(o.obst())
    # untouched ghosts on the closed axes: f = w_i, rho summed in direction order
    rho_g = 0.0
    for wd in (3, 18, 18, 18, 18, 18, 18, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36):
        rho_g += 1.0 / wd
    P, U, OB = [], [], []
    for Z in range(L[2]):
        for Y in range(L[1]):
            for X in range(L[0]):
                r = []
                for dd, c in enumerate((X - 2, Y - 2, Z - 2)):
                    if not 0 <= c < n[dd]:
                        if not d.periodic[dd]:
                            r = None
                            break
                        c %= n[dd]
                    r.append(c)
                if r is None:
                    P.append(_g(pc * (rho_g - 1)))
                    U.append("0 0 0")
                    OB.append("-1")
                    continue
                x, y, z = r
                if obst[z, y, x] == -1:
                    P.append(_g(pc * (m0[z, y, x] - 1)))
                    U.append(" ".join(_g(float(np.float32(uc * m1[c, z, y, x]))) for c in range(3)))
                    OB.append("-1")
                else:
                    P.append("0")
                    U.append("0 0 0")
                    OB.append("-2")
    t = ['<VTKFile type="RectilinearGrid">\n',
         f' <RectilinearGrid WholeExtent=" 0 {n[0] - 1} 0 {n[1] - 1} 0 {n[2] - 1}">\n',
         '      <Piece Extent="' + "".join(f"-2 {n[k] + 1} " for k in range(3)) + '">\n',
         "      <Coordinates>\n"]
    for k, name in enumerate("XYZ"):
        t.append(f'          <DataArray type="Float32" Name="{name}" format="ascii">\n')
        t.append("".join(_g(x * dx) + " " for x in range(-2, n[k] + 2)) + "\n")
        t.append("          </DataArray>\n")
    t.append("      </Coordinates>\n      <PointData>\n")
    t.append('          <DataArray type="Float32" Name="P" format="ascii">\n')
    t.append("".join(v + " " for v in P) + "\n          </DataArray>\n")
    t.append('          <DataArray type="Float32" Name="U" format="ascii" NumberOfComponents="3">\n')
    t.append("".join(v + " " for v in U) + "\n          </DataArray>\n")
    t.append('          <DataArray type="Float32" Name="OBST" format="ascii">\n')
    t.append("".join(v + " " for v in OB) + "\n          </DataArray>\n")
    t.append("      </PointData>\n      </Piece>\n </RectilinearGrid>\n</VTKFile>\n")
    return "".join(t)


def _pvtr_expected(n, ts):
    name = "hippoLBM_%010d" % ts
    return (' <VTKFile type="PRectilinearGrid"> \n'
            f'   <PRectilinearGrid WholeExtent="0 {n[0] - 1} 0 {n[1] - 1} 0 {n[2] - 1}"\n'
            '                     GhostLevel="#">\n'
            f'     <Piece Extent=" -2 {n[0] + 1} -2 {n[1] + 1} -2 {n[2] + 1}" Source="{name}/000000.vtr"/>\n'
            "    <PCoordinates>\n"
            '      <PDataArray type="Float32" Name="X"/>\n'
            '      <PDataArray type="Float32" Name="Y"/>\n'
            '      <PDataArray type="Float32" Name="Z"/>\n'
            "    </PCoordinates>\n"
            '     <PPointData Scalars="P OBST"  Vectors="U" >\n'
            '       <PDataArray Name="P" type="Float32" NumberOfComponents="1"/>\n'
            '       <PDataArray Name="OBST" type="Float32" NumberOfComponents="1"/>\n'
            '       <PDataArray Name="U" type="Float32" NumberOfComponents="3"/>\n'
            "     </PPointData> \n"
            "   </PRectilinearGrid>\n"
            " </VTKFile>\n")


def _line_expected(d, o, start, axis, count):
    dx = d.dx
    uc = dx / d.params["dtLB"]
    n = d.nodes
    m1 = o.real(o.m1())
    rows = ["%16s%16s%16s%16s%16s%16s\n" % ("rx", "ry", "rz", "ux", "uy", "uz")]
    for k in range(count):
        p = list(start)
        p[axis] += k
        u = [0.0, 0.0, 0.0]
        if all(0 <= p[i] < n[i] for i in range(3)):
            u = [uc * m1[c, p[2], p[1], p[0]] for c in range(3)]
        rows.append(("%16.8g" * 6 + "\n") % (float(p[0]) * dx, float(p[1]) * dx, float(p[2]) * dx, *u))
    return "".join(rows)


def _check_profile(path, d, o):
    dx = d.dx
    scale = dx / d.params["dtLB"]
    fluid = o.real(o.obst()) == -1
    s, c, lo, hi = _oracle_profile(o.real(o.m1()), fluid, 2, scale)
    lines = open(path).read().splitlines()
    assert lines[0] == "%16s%16s%16s%16s" % ("position", "avg", "min", "max")
    assert len(lines) == 1 + d.nodes[2]
    for i, line in enumerate(lines[1:]):
        cols = [line[16 * k:16 * (k + 1)] for k in range(4)]
        assert cols[0] == "%16.8g" % (0.0 + i * dx)
        assert cols[2] == "%16.8g" % lo[i] and cols[3] == "%16.8g" % hi[i]
        mean = s[i] / c[i] if c[i] else 0.0
        assert abs(float(cols[1]) - mean) <= 1e-7 * abs(mean)


def test_hlbm_run_writes_the_reference_output_files(tmp_path):
    deck_path = tmp_path / "io.msp"
    deck_path.write_text(IO_DECK)
    say("hlbm_run io deck")
    r = subprocess.run([RUN, str(deck_path)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    base = tmp_path / "IoTestDir"
    assert "writing paraview file: IoTestDir/ParaviewOutput/hippoLBM_0000000000" in out
    assert "writing profile file: IoTestDir/Profile/profile_0000000004.csv" in out

    d = dk.loads_msp(IO_DECK, "io")
    n = d.nodes
    ref = orc.OracleRun(d)
    g_line = [int(np.floor((0.01 - 0.0) / d.dx)), int(np.floor(0.01 / d.dx)), int(np.floor(0.0 / d.dx))]
    g_end = int(np.floor(0.019 / d.dx))

    def check_step(ts, o):
        pv = base / "ParaviewOutput"
        assert (pv / ("hippoLBM_%010d.pvtr" % ts)).read_text() == _pvtr_expected(n, ts)
        vtr = (pv / ("hippoLBM_%010d" % ts) / "000000.vtr").read_text()
        assert vtr == _vtr_expected(d, o, ts), f"vtr at step {ts} differs"

    def check_analysis(ts, o):
        prof = base / "Profile"
        _check_profile(prof / ("profile_%010d.csv" % ts), d, o)
        assert (prof / ("line_%010d.csv" % ts)).read_text() == _line_expected(d, o, g_line, 2, g_end - g_line[2] + 1)
        assert (prof / ("xline_%05d.csv" % ts)).read_text() == _line_expected(d, o, [-2, 5, 12], 0, 24)

    # step 0: the initial state (write_paraview_first_timestep, then analysis)
    check_step(0, ref.o)
    check_analysis(0, ref.o)
    for t in range(1, 9):
        ref.step(1)
        if t % 4 == 0:  # dump_paraview_if_triggered recomputes macro_variables first
            ref.o.macro_variables()
            check_step(t, ref.o)
            check_analysis(t, ref.o)
    # the write_paraview after compute_loop, at timestep end + 1, of the fields as held
    check_step(9, ref.o)
    names = sorted(os.listdir(base / "ParaviewOutput"))
    assert names == sorted([f"hippoLBM_{t:010d}{e}" for t in (0, 4, 8, 9) for e in ("", ".pvtr")])


def test_no_output_flag_writes_nothing(tmp_path):
    deck_path = tmp_path / "io.msp"
    deck_path.write_text(IO_DECK)
    r = subprocess.run([RUN, str(deck_path), "--no-output", "--quiet"], capture_output=True, text=True, timeout=120,
                       cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert not (tmp_path / "IoTestDir").exists()

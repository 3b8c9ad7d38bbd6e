"""The C ABI library: it loads, exports every entry point include/hlbm.h
declares, and its struct layouts match the header (checked with a C compiler).
No compute call is made here (no GPU in the CPU suite)."""
import ctypes
import os
import re
import subprocess

import pytest

from hippolbm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hlbm.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hlbm_[a-z0-9_]+)\s*\(", txt)))


def test_library_loads():
    L = _lib.lib()
    assert L.hlbm_version().decode().startswith("hippolbm")


def test_every_declared_symbol_is_exported():
    declared = header_functions()
    assert declared == sorted(_lib.SYMBOLS)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (hlbm_[a-z0-9_]+)", out))
    missing = [s for s in declared if s not in exported]
    assert not missing, missing


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "hlbm.h"\n'
        "int main(void){\n"
        'printf("%zu %zu %zu %zu %zu\\n", sizeof(hlbm_domain_desc), offsetof(hlbm_domain_desc, tau),'
        " offsetof(hlbm_domain_desc, dx), offsetof(hlbm_domain_desc, mrt_rates), offsetof(hlbm_domain_desc, fext));\n"
        'printf("%zu %zu %zu\\n", sizeof(hlbm_bc), offsetof(hlbm_bc, U), offsetof(hlbm_bc, bounds));\n'
        'printf("%zu %zu %zu %zu\\n", sizeof(hlbm_obstacle), offsetof(hlbm_obstacle, p1), offsetof(hlbm_obstacle, radius),'
        ' offsetof(hlbm_obstacle, q));\n'
        'printf("%zu %zu %zu %zu\\n", sizeof(hlbm_solve_desc), offsetof(hlbm_solve_desc, tau),'
        ' offsetof(hlbm_solve_desc, cavity_u), offsetof(hlbm_solve_desc, obstacles));\n'
        'printf("%zu\\n", offsetof(hlbm_solve_desc, init_f));\n'
        'printf("%zu %zu %zu\\n", sizeof(hlbm_batch_desc), sizeof(hlbm_batch_info), offsetof(hlbm_batch_info, gpu_ms));\n'
        'printf("%zu %zu\\n", sizeof(hlbm_tune_info), offsetof(hlbm_tune_info, ms_default));\n'
        "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    d = [int(v) for v in lines[0].split()]
    D = _lib.DomainDesc
    assert d == [ctypes.sizeof(D), D.tau.offset, D.dx.offset, D.mrt_rates.offset, D.fext.offset]
    b = [int(v) for v in lines[1].split()]
    assert b == [ctypes.sizeof(_lib.BC), _lib.BC.U.offset, _lib.BC.bounds.offset]
    o = [int(v) for v in lines[2].split()]
    assert o == [ctypes.sizeof(_lib.Obstacle), _lib.Obstacle.p1.offset, _lib.Obstacle.radius.offset,
                 _lib.Obstacle.q.offset]
    from hippolbm_amd import batch as bt
    sd = [int(v) for v in lines[3].split()]
    S = bt.SolveDesc
    assert sd == [ctypes.sizeof(S), S.tau.offset, S.cavity_u.offset, S.obstacles.offset]
    assert int(lines[4]) == S.init_f.offset
    bd = [int(v) for v in lines[5].split()]
    assert bd == [ctypes.sizeof(bt.BatchDesc), ctypes.sizeof(bt.BatchInfo), bt.BatchInfo.gpu_ms.offset]
    tu = [int(v) for v in lines[6].split()]
    assert tu == [ctypes.sizeof(_lib.TuneInfo), _lib.TuneInfo.ms_default.offset]


def test_null_arguments_are_rejected_without_a_device():
    L = _lib.lib()
    assert L.hlbm_domain_create(None, None) == -1
    assert L.hlbm_device_count(None) == -1
    assert b"NULL" in L.hlbm_last_error()
    assert L.hlbm_domain_destroy(None) == 0


def test_tau_at_or_below_one_half_is_rejected_without_a_device():
    # nu = (tau - 1/2)/3 <= 0: refused before any device call
    L = _lib.lib()
    d = _lib.DomainDesc()
    for k in range(3):
        d.nodes[k] = 8
    d.tau = 0.5
    h = ctypes.c_void_p()
    assert L.hlbm_domain_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"tau must be > 0.5" in L.hlbm_last_error()


def test_enums_match_header():
    txt = open(HEADER).read()
    for name, val in (("HLBM_F64", 0), ("HLBM_F32", 1), ("HLBM_F16", 2), ("HLBM_BF16", 3)):
        assert re.search(rf"{name} = {val}\b", txt)
    for name, val in _lib.BC_KIND.items():
        assert re.search(rf"= {val},?\s*/\* \"{name}\"", txt), name
    for name, val in _lib.PLANE.items():
        assert re.search(rf"HLBM_{name.upper()} = {val}\b", txt), name
    for name, val in (("START", -1), ("PRE_STREAM", 0), ("POST_STREAM", 1), ("BOUNDARY", 2)):
        assert re.search(rf"HLBM_STAGE_{name} = {val}\b", txt), name
    assert sorted(_lib.STAGE.values()) == [-1, 0, 1, 2]

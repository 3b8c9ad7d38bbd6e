"""hlbm_plane_profile_noalloc: the allocation-free plane_velocity_profile.

Checked against hlbm_plane_profile (same kernels; the bar is bit equality of
every sum, count, min and max) on two calls in a row, the second on a smaller
domain, plus the refusal of a domain with more columns than the library's
static partials hold.  The test allocates the device output itself through the
HIP runtime; the entry point allocates nothing.
"""
import ctypes
import sys

import numpy as np
import pytest

from hippolbm_amd import _lib
from hippolbm_amd import deck as dk
from hippolbm_amd.engine import Simulation
from hippolbm_amd.hip import Stream, hip

pytestmark = pytest.mark.gpu

REC = np.dtype([("sum", "<f8"), ("vmin", "<f8"), ("vmax", "<f8"), ("count", "<i4"), ("pad", "<i4")])


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DeviceArray:
    """Device memory owned by the test (hipMalloc / hipMemcpy through ctypes)."""

    def __init__(self, nbytes):
        h = hip()
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMalloc.restype = ctypes.c_int
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipFree.restype = ctypes.c_int
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemcpy.restype = ctypes.c_int
        self.p = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def fill(self, byte):
        host = np.full(self.nbytes, byte, dtype=np.uint8)
        assert hip().hipMemcpy(self.p, host.ctypes.data, self.nbytes, 1) == 0  # host to device

    def to_host(self, dtype, n):
        out = np.zeros(n, dtype=dtype)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = ctypes.c_void_p()


def _sim(nodes, seed):
    """A moving-lid box with a wall slab across y = 3, 4 (planes without fluid give
    the reset records)."""
    nx, ny, nz = nodes
    text = f"""
do_domain:
  - domain:
     bounds: [[0,0,0],[{nx - 1},{ny - 1},{nz - 1}]]
     cell_dims: [{nx - 1}, {ny - 1}, {nz - 1}]
     periodic: [false, false, false]
set_lbm_parameters:
  - lbm_parameters:
     tau: 0.7
     nuth: 0.0666666666666666667
     Fext: [{2e-5 * (1 + seed)}, 0, 0]
set_obstacles:
  - register_solid_wall:
     id: 0
     bounds: [[-1, 2.5, -1], [{nx}, 4.5, {nz}]]
pre_stream_bcs:
  - pre_bounce_back
  - wall_bounce_back
  - cavity_z_l:
     U: [0.04, 0.01, 0]
post_stream_bcs:
  - post_bounce_back
"""
    sim = Simulation(dk.loads_msp(text, f"profile_{seed}"), storage="f64")
    sim.run(17 + seed)
    sim.dom.macro()
    return sim


def _check(sim, stream, scale):
    for dim in range(3):
        n = sim.dom.nodes[dim]
        s, c, lo, hi = sim.dom.plane_profile(dim, scale)  # the allocating entry point
        out = DeviceArray(n * REC.itemsize)
        out.fill(0xA5)
        _lib.check("hlbm_plane_profile_noalloc",
                   _lib.lib().hlbm_plane_profile_noalloc(sim.dom.h, dim, scale, out.p, stream.handle))
        stream.synchronize()
        r = out.to_host(REC, n)
        out.free()
        assert np.array_equal(r["sum"], s) and np.array_equal(r["count"], c), dim
        assert np.array_equal(r["vmin"], lo) and np.array_equal(r["vmax"], hi), dim
        assert (c > 0).any() and ((c == 0).sum() == 2) == (dim == 1), dim  # the slab's y planes hold no fluid


def test_noalloc_profile_equals_the_allocating_entry_point():
    stream = Stream()
    big = _sim((40, 24, 32), 0)
    say("noalloc profile, 40x24x32")
    _check(big, stream, 1.0)
    small = _sim((20, 12, 16), 1)   # second call, smaller input: stale partials must not leak in
    say("noalloc profile, 20x12x16")
    _check(small, stream, 2.5)
    big.dom.close()
    small.dom.close()
    stream.close()


def test_noalloc_profile_refuses_more_columns_than_its_scratch():
    from hippolbm_amd.engine import Domain
    dom = Domain((1024, 2, 520), (False, False, False), 0.8)   # x-planes: 1024 * 520 > 2^19 columns
    out = DeviceArray(1024 * REC.itemsize)
    rc = _lib.lib().hlbm_plane_profile_noalloc(dom.h, 0, 1.0, out.p, None)
    assert rc == -1  # HLBM_ERR_ARG
    assert b"HLBM_PROFILE_SCRATCH_COLUMNS" in _lib.lib().hlbm_last_error()
    # y-planes of the same domain fit (1024 * 2 columns)
    out2 = DeviceArray(2 * REC.itemsize)
    assert _lib.lib().hlbm_plane_profile_noalloc(dom.h, 1, 1.0, out2.p, None) == 0
    out.free()
    out2.free()
    dom.close()

"""ctypes binding of the C ABI in ``include/hlbm.h`` (``hippolbm_amd/libhlbm.so``).

The library is the product: there is no Python or CPU fallback for any compute
entry point.  Loading fails loudly when ``libhlbm.so`` is missing; build it
with ``__graft_entry__.build()`` (or ``make -C hippolbm_amd/csrc``).
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhlbm.so")

HLBM_OK = 0
ERRORS = {-1: "HLBM_ERR_ARG", -2: "HLBM_ERR_HIP", -3: "HLBM_ERR_UNSUPPORTED", -4: "HLBM_ERR_NOMEM",
          -5: "HLBM_ERR_NODEVICE"}

STORAGE = {"f64": 0, "f32": 1, "f16": 2, "bf16": 3}
STORAGE_BYTES = {"f64": 8, "f32": 4, "f16": 2, "bf16": 2}
COLLISION = {"bgk": 0, "mrt": 1}
LAYOUT_NODES, LAYOUT_REF = 0, 1
BC_KIND = {"pre_bounce_back": 1, "post_bounce_back": 2, "wall_bounce_back": 3, "cavity_z_0": 4, "cavity_z_l": 5,
           "lid_driven_cavity": 6, "neumann": 7, "pressure_inlet": 8, "outflow": 9, "set_distribution": 10}
STAGE = {"start_iteration": -1, "pre_stream_bcs": 0, "post_stream_bcs": 1, "boundary_conditions": 2}
PLANE = {"plan_xy_0": 0, "plan_xy_l": 1, "plan_xz_0": 2, "plan_xz_l": 3, "plan_yz_0": 4, "plan_yz_l": 5}
STEP_WRITE_MACRO = 1

# every entry point include/hlbm.h declares
SYMBOLS = (
    "hlbm_last_error", "hlbm_version", "hlbm_device_count", "hlbm_domain_create", "hlbm_domain_destroy",
    "hlbm_domain_bytes", "hlbm_set_bcs", "hlbm_set_params", "hlbm_init_obst", "hlbm_add_obstacles",
    "hlbm_set_flags", "hlbm_get_flags", "hlbm_set_distribution", "hlbm_set_populations_f64",
    "hlbm_get_populations_f64", "hlbm_step", "hlbm_timestep", "hlbm_synchronize", "hlbm_macro", "hlbm_get_macro",
    "hlbm_reduce_stats", "hlbm_ghost_update_ref", "hlbm_copy_bench", "hlbm_autotune", "hlbm_set_launch",
    "hlbm_set_collision", "hlbm_set_layout", "hlbm_get_layout", "hlbm_plane_profile", "hlbm_plane_profile_noalloc",
    "hlbm_line_velocity", "hlbm_init_double_couette",
    "hlbm_batch_create", "hlbm_batch_destroy", "hlbm_batch_submit", "hlbm_batch_run", "hlbm_batch_status",
    "hlbm_batch_times", "hlbm_batch_result", "hlbm_batch_stats", "hlbm_batch_info_get", "hlbm_clock",
)


class DomainDesc(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.c_int32 * 3),
        ("periodic", ctypes.c_int32 * 3),
        ("storage", ctypes.c_int32),
        ("pattern", ctypes.c_int32),
        ("collision", ctypes.c_int32),
        ("mrt_custom", ctypes.c_int32),
        ("tau", ctypes.c_double),
        ("fext", ctypes.c_double * 3),
        ("dx", ctypes.c_double),
        ("mrt_rates", ctypes.c_double * 10),
        ("round_nearest", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class BC(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("stage", ctypes.c_int32),
        ("plane", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("U", ctypes.c_double * 3),
        ("bounds", ctypes.c_double * 6),
    ]


class Obstacle(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("p0", ctypes.c_double * 3),
        ("p1", ctypes.c_double * 3),
        ("radius", ctypes.c_double),
        ("q", ctypes.c_double * 16),
    ]


def make_obstacle(ob):
    """hlbm_obstacle from a deck obstacle dict (kind ball / wall / quadric)."""
    o = Obstacle()
    if ob["kind"] == "ball":
        o.kind = 0
        for k in range(3):
            o.p0[k] = float(ob["center"][k])
        o.radius = float(ob["radius"])
    elif ob["kind"] == "wall":
        o.kind = 1
        for k in range(3):
            o.p0[k] = float(ob["bounds"][0][k])
            o.p1[k] = float(ob["bounds"][1][k])
    elif ob["kind"] == "quadric":
        o.kind = 2
        for k in range(16):
            o.q[k] = float(ob["q"][k])
    else:
        raise ValueError(f"unknown obstacle kind {ob['kind']!r}")
    return o


class TuneInfo(ctypes.Structure):
    _fields_ = [
        ("block_x", ctypes.c_int32),
        ("block_y", ctypes.c_int32),
        ("npt", ctypes.c_int32),
        ("xcd_remap", ctypes.c_int32),
        ("candidates", ctypes.c_int32),
        ("variant", ctypes.c_int32),
        ("relocated", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("ms_best", ctypes.c_double),
        ("ms_default", ctypes.c_double),
    ]


class HLBMError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with {ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libhlbm.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(or `make -C hippolbm_amd/csrc`) -- there is no fallback")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    DP = ctypes.POINTER(ctypes.c_double)
    I32P = ctypes.POINTER(ctypes.c_int32)
    sig = {
        "hlbm_last_error": (ctypes.c_char_p, []),
        "hlbm_version": (ctypes.c_char_p, []),
        "hlbm_device_count": (ctypes.c_int, [I32P]),
        "hlbm_domain_create": (ctypes.c_int, [ctypes.POINTER(DomainDesc), ctypes.POINTER(P)]),
        "hlbm_domain_destroy": (ctypes.c_int, [P]),
        "hlbm_domain_bytes": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_uint64)]),
        "hlbm_set_bcs": (ctypes.c_int, [P, ctypes.POINTER(BC), ctypes.c_int32]),
        "hlbm_set_params": (ctypes.c_int, [P, ctypes.c_double, DP]),
        "hlbm_init_obst": (ctypes.c_int, [P]),
        "hlbm_add_obstacles": (ctypes.c_int, [P, ctypes.POINTER(Obstacle), ctypes.c_int32]),
        "hlbm_set_flags": (ctypes.c_int, [P, I32P, ctypes.c_int32]),
        "hlbm_get_flags": (ctypes.c_int, [P, I32P, ctypes.c_int32]),
        "hlbm_set_distribution": (ctypes.c_int, [P, ctypes.c_double, DP]),
        "hlbm_set_populations_f64": (ctypes.c_int, [P, DP, ctypes.c_int32]),
        "hlbm_get_populations_f64": (ctypes.c_int, [P, DP, ctypes.c_int32]),
        "hlbm_step": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_uint32, P]),
        "hlbm_timestep": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int64)]),
        "hlbm_synchronize": (ctypes.c_int, [P]),
        "hlbm_macro": (ctypes.c_int, [P, DP, DP, ctypes.c_int32]),
        "hlbm_get_macro": (ctypes.c_int, [P, DP, DP, ctypes.c_int32]),
        "hlbm_reduce_stats": (ctypes.c_int, [P, DP]),
        "hlbm_ghost_update_ref": (ctypes.c_int, [P, DP]),
        "hlbm_copy_bench": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int32, P, DP]),
        "hlbm_autotune": (ctypes.c_int, [P, ctypes.c_int32, P, ctypes.POINTER(TuneInfo)]),
        "hlbm_set_launch": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_int32]),
        "hlbm_set_collision": (ctypes.c_int, [P, ctypes.c_int32, DP]),
        "hlbm_set_layout": (ctypes.c_int, [P, I32P]),
        "hlbm_get_layout": (ctypes.c_int, [P, I32P]),
        "hlbm_plane_profile": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_double, DP, I32P, DP, DP]),
        "hlbm_plane_profile_noalloc": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_double, P, P]),
        "hlbm_line_velocity": (ctypes.c_int, [P, I32P, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, DP]),
        "hlbm_init_double_couette": (ctypes.c_int, [P, ctypes.c_int32, DP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(fn_name, rc):
    if rc != HLBM_OK:
        msg = lib().hlbm_last_error()
        raise HLBMError(fn_name, rc, msg.decode() if msg else "")
    return rc


def device_count():
    n = ctypes.c_int32(0)
    check("hlbm_device_count", lib().hlbm_device_count(ctypes.byref(n)))
    return n.value

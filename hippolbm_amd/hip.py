"""Minimal ctypes access to the HIP runtime that libhlbm.so links against:
streams and events for timing.  Loaded after libhlbm.so so both share one
runtime instance."""
from __future__ import annotations

import ctypes

from . import _lib

_hip = None


def hip():
    global _hip
    if _hip is None:
        _lib.lib()  # make sure libhlbm.so (and its libamdhip64) is loaded first
        h = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        P = ctypes.c_void_p
        h.hipStreamCreate.argtypes = [ctypes.POINTER(P)]
        h.hipStreamDestroy.argtypes = [P]
        h.hipStreamSynchronize.argtypes = [P]
        h.hipEventCreate.argtypes = [ctypes.POINTER(P)]
        h.hipEventDestroy.argtypes = [P]
        h.hipEventRecord.argtypes = [P, P]
        h.hipEventSynchronize.argtypes = [P]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), P, P]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        for fn in ("hipStreamCreate", "hipStreamDestroy", "hipStreamSynchronize", "hipEventCreate", "hipEventDestroy",
                   "hipEventRecord", "hipEventSynchronize", "hipEventElapsedTime", "hipDeviceSynchronize"):
            getattr(h, fn).restype = ctypes.c_int
        _hip = h
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


class Stream:
    def __init__(self):
        self.s = ctypes.c_void_p()
        _ok(hip().hipStreamCreate(ctypes.byref(self.s)), "hipStreamCreate")

    @property
    def handle(self):
        return self.s.value

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.s), "hipStreamSynchronize")

    def close(self):
        if self.s:
            hip().hipStreamDestroy(self.s)
            self.s = ctypes.c_void_p()


class Event:
    def __init__(self):
        self.e = ctypes.c_void_p()
        _ok(hip().hipEventCreate(ctypes.byref(self.e)), "hipEventCreate")

    def record(self, stream=None):
        _ok(hip().hipEventRecord(self.e, stream.s if stream is not None else None), "hipEventRecord")

    def synchronize(self):
        _ok(hip().hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _ok(hip().hipEventElapsedTime(ctypes.byref(ms), self.e, end.e), "hipEventElapsedTime")
        return ms.value

    def close(self):
        if self.e:
            hip().hipEventDestroy(self.e)
            self.e = ctypes.c_void_p()


def device_synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")

"""ctypes binding of libpamo_hip.so (include/pamo_hip.h).

The library is the only implementation of the hot path: if it is missing or
fails to load, every entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpamo_hip.so")

c_int, c_float, c_double, c_ll, c_vp = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_longlong, ctypes.c_void_p

PAMO_DBG_EDGES = 1
PAMO_DBG_EDGE_COST = 2
PAMO_DBG_RECORDS = 3
PAMO_DBG_RING_OFF = 4
PAMO_DBG_RING = 5
PAMO_DBG_QUADRICS = 6
PAMO_DBG_UNDO_ITERS = 7
PAMO_DBG_N_REGULAR = 8

# (name, restype, argtypes) for every symbol declared in include/pamo_hip.h
SIGNATURES = [
    ("pamo_ctx_create", c_int, [ctypes.POINTER(c_vp), c_int]),
    ("pamo_ctx_destroy", c_int, [c_vp]),
    ("pamo_last_error", ctypes.c_char_p, []),
    ("pamo_round", c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_int, c_float, c_float, c_int, c_int,
                           c_vp, c_vp, c_vp, c_vp, c_vp, c_int, ctypes.POINTER(c_int)]),
    ("pamo_centroid", c_int, [c_vp, c_vp, c_int, c_vp, c_int, ctypes.POINTER(c_float)]),
    ("pamo_round_debug", c_int, [c_vp, c_int, c_vp, c_ll, ctypes.POINTER(c_ll)]),
    ("pamo_simplify", c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_double, c_int, c_float, c_ll, c_ll, c_vp, c_vp,
                              ctypes.POINTER(c_int), ctypes.POINTER(c_int), c_vp, c_int, ctypes.POINTER(c_int)]),
]

_lib = None


class PamoError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PamoError(f"{LIB_PATH} not built: run `python -m pamo_amd.build` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().pamo_last_error()
        raise PamoError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


class Context:
    """Owns one pamo_ctx (persistent device state + one HIP stream)."""

    def __init__(self, device: int = 0):
        p = c_vp()
        check(lib().pamo_ctx_create(ctypes.byref(p), int(device)), "pamo_ctx_create")
        self.ptr = p
        self.device = device

    def close(self):
        if getattr(self, "ptr", None):
            lib().pamo_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

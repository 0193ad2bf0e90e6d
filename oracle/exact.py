"""ctypes wrapper of oracle/exact_check.cpp: exact orientation predicates and
an exhaustive exact self-intersection check of triangle meshes.

Test infrastructure: imported only by tests/ to check outputs.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import oracle as _oracle

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libexact_check.so")
_lib = None
c_int, c_ll, c_vp = ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            _oracle.build(force=True)
        L = ctypes.CDLL(_LIB_PATH)
        L.ex_orient3d.argtypes = [c_vp] * 4
        L.ex_orient2d.argtypes = [c_vp] * 3
        L.ex_faces_intersect.argtypes = [c_vp] * 4
        L.ex_self_intersections.restype = c_ll
        L.ex_self_intersections.argtypes = [c_vp, c_int, c_vp, c_int, c_vp, c_ll, ctypes.POINTER(c_ll),
                                            ctypes.POINTER(c_ll)]
        _lib = L
    return _lib


def _d(a, n):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    assert a.size == n
    return a


def orient3d(a, b, c, d) -> int:
    a, b, c, d = (_d(x, 3) for x in (a, b, c, d))
    return int(lib().ex_orient3d(a.ctypes.data, b.ctypes.data, c.ctypes.data, d.ctypes.data))


def orient2d(a, b, c) -> int:
    a, b, c = (_d(x, 2) for x in (a, b, c))
    return int(lib().ex_orient2d(a.ctypes.data, b.ctypes.data, c.ctypes.data))


def faces_intersect(A, ia, B, ib) -> bool:
    """Faces with corners A, B (3x3) and vertex ids ia, ib intersect beyond
    what their shared vertices allow."""
    A, B = _d(A, 9), _d(B, 9)
    ia = np.ascontiguousarray(np.asarray(ia, np.int32))
    ib = np.ascontiguousarray(np.asarray(ib, np.int32))
    return bool(lib().ex_faces_intersect(A.ctypes.data, ia.ctypes.data, B.ctypes.data, ib.ctypes.data))


def self_intersections(V, F, cap: int = 1 << 16):
    """Exact check of a mesh: returns (intersecting pairs (k,2) with a < b,
    number of box-overlap candidates, number of degenerate faces)."""
    V = np.ascontiguousarray(np.asarray(V, np.float32).reshape(-1, 3))
    F = np.ascontiguousarray(np.asarray(F, np.int32).reshape(-1, 3))
    out = np.empty((max(cap, 1), 2), np.int32)
    nc, nd = c_ll(0), c_ll(0)
    n = lib().ex_self_intersections(V.ctypes.data, len(V), F.ctypes.data, len(F), out.ctypes.data, cap,
                                    ctypes.byref(nc), ctypes.byref(nd))
    return out[: min(n, cap)].copy(), int(nc.value), int(nd.value)

"""ctypes wrapper of oracle/point_mesh.cpp: brute-force point-to-mesh squared
distances and the Hausdorff distance with libigl 2.5.1's semantics
(simp_cuda/safe_project/src/pamo_safe_project/metrics.py:41-49).

Test infrastructure: imported only by tests/ to check the GPU metrics.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import oracle as _oracle

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libpoint_mesh.so")
_lib = None
c_int, c_ll, c_vp, c_double = ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_double


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            _oracle.build(force=True)
        L = ctypes.CDLL(_LIB_PATH)
        for name in ("pm_point_tri_sqdist", "pm_point_tri_sqdist_alt"):
            getattr(L, name).restype = c_double
            getattr(L, name).argtypes = [c_vp] * 4
        L.pm_point_mesh_sqdist.restype = None
        L.pm_point_mesh_sqdist.argtypes = [c_vp, c_ll, c_vp, c_vp, c_int, c_vp, c_vp]
        L.pm_hausdorff.restype = c_double
        L.pm_hausdorff.argtypes = [c_vp, c_int, c_vp, c_int, c_vp, c_int, c_vp, c_int]
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, 3))


def _i(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1, 3))


def point_tri_sqdist(p, a, b, c, alt: bool = False) -> float:
    """Squared distance of p to triangle abc (Ericson; alt: the plane /
    edge construction)."""
    p, a, b, c = (_d(x) for x in (p, a, b, c))
    fn = lib().pm_point_tri_sqdist_alt if alt else lib().pm_point_tri_sqdist
    return float(fn(p.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data))


def point_mesh_sqdist(P, V, F):
    """(sqrD, I): nearest face of each point by brute force (first face in
    order at the minimum)."""
    P, V, F = _d(P), _d(V), _i(F)
    sq = np.empty(len(P), np.float64)
    idx = np.empty(len(P), np.int32)
    lib().pm_point_mesh_sqdist(P.ctypes.data, len(P), V.ctypes.data, F.ctypes.data, len(F), sq.ctypes.data,
                               idx.ctypes.data)
    return sq, idx


def hausdorff(VA, FA, VB, FB) -> float:
    VA, FA, VB, FB = _d(VA), _i(FA), _d(VB), _i(FB)
    return float(lib().pm_hausdorff(VA.ctypes.data, len(VA), FA.ctypes.data, len(FA), VB.ctypes.data, len(VB),
                                     FB.ctypes.data, len(FB)))

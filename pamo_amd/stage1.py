"""Stage-1 remeshing, PaMO.remesh (simp_cuda/pamo/__init__.py:60-120), on the
device through pamo_s1_* of libpamo_hip.so.

The input triangles are normalised into the unit cube as preprocess_mesh
does, a narrow-band distance field on a resolution^3 grid is signed by a
flood fill of the outside from the grid border, and the level 0.9 /
resolution is extracted by marching cubes with one consistent rule for
ambiguous faces (each inside corner cut off on its own), which keeps the
surface closed, manifold and oriented.  The vertices come back in the centred frame of
PaMO.run (input minus the numpy float32 face-centroid mean), the frame in
which stage 2 runs.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _hip


def resolution_for(target_faces: int) -> int:
    """pamo/__init__.py:136-141: 256, 128 for targets of at most 1000 faces,
    64 for at most 50."""
    R = 256
    if target_faces <= 1000:
        R = 128
    if target_faces <= 50:
        R = 64
    return R


class Remesher:
    """One stage-1 context (device buffers kept between calls)."""

    def __init__(self, device: int | None = None):
        if device is None:
            device = torch.cuda.current_device()
        self.device = device
        p = ctypes.c_void_p()
        _hip.check(_hip.lib().pamo_s1_create(ctypes.byref(p), int(device)), "pamo_s1_create")
        self.ptr = p

    def __del__(self):
        try:
            if self.ptr:
                _hip.lib().pamo_s1_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass

    def remesh(self, points: torch.Tensor, faces: torch.Tensor, mean, resolution: int):
        """Device float32 (n,3) points, int32 (m,3) faces with checked
        indices, and the centring offset `mean` (3 floats); returns the
        remeshed (vertices, faces) as device tensors, vertices centred."""
        if points.dtype != torch.float32 or faces.dtype != torch.int32:
            raise TypeError("remesh: points must be float32 and faces int32")
        points = points.contiguous()
        faces = faces.contiguous()
        m = (ctypes.c_float * 3)(*[float(x) for x in np.asarray(mean, np.float32)])
        nv, nf = ctypes.c_int(0), ctypes.c_int(0)
        torch.cuda.current_stream(points.device).synchronize()
        _hip.check(_hip.lib().pamo_s1_remesh(self.ptr, points.data_ptr(), int(points.shape[0]), faces.data_ptr(),
                                             int(faces.shape[0]), m, int(resolution), ctypes.byref(nv),
                                             ctypes.byref(nf)), "pamo_s1_remesh")
        v = torch.empty((nv.value, 3), dtype=torch.float32, device=points.device)
        f = torch.empty((nf.value, 3), dtype=torch.int32, device=points.device)
        if nv.value and nf.value:
            _hip.check(_hip.lib().pamo_s1_get(self.ptr, v.data_ptr(), f.data_ptr()), "pamo_s1_get")
        return v, f

    # -- test access (no counterpart in the reference's API)
    def debug(self, what: int, dtype, cap: int) -> np.ndarray:
        out = np.zeros(cap, dtype=dtype)
        n = ctypes.c_longlong(0)
        _hip.check(_hip.lib().pamo_s1_debug(self.ptr, int(what), out.ctypes.data, int(cap), ctypes.byref(n)),
                   "pamo_s1_debug")
        return out[: n.value]

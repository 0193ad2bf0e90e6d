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

PAMO_OPT_GUARD_SMALL_STACK = 1
PAMO_OPT_EXACT_GUARD = 2
PAMO_OPT_WALK_LIMIT = 3
PAMO_OPT_PLACEMENT = 4
PAMO_OPT_TREE_KEEP = 5
PAMO_OPT_GUARD_SPLIT_FACES = 6
PAMO_OPT_GUARD_SPLIT_LOG = 7
PAMO_OPT_GUARD_WAVES = 8
PAMO_OPT_EDGE_LANES = 9
PAMO_OPT_RING_LANES = 10
PAMO_DBG_EDGES = 1
PAMO_DBG_EDGE_COST = 2
PAMO_DBG_RECORDS = 3
PAMO_DBG_RING_OFF = 4
PAMO_DBG_RING = 5
PAMO_DBG_QUADRICS = 6
PAMO_DBG_UNDO_ITERS = 7
PAMO_DBG_N_REGULAR = 8
PAMO_DBG_BVH_N = 9
PAMO_DBG_BVH_CODES = 10
PAMO_DBG_BVH_FACES = 11
PAMO_DBG_BVH_LEFT = 12
PAMO_DBG_BVH_RIGHT = 13
PAMO_DBG_BVH_BOXES = 14

PAMO_S3_E_M2GT = 1
PAMO_S3_E_ELASTIC = 2
PAMO_S3_E_HINGE = 4
PAMO_S3_E_COLLISION = 8
PAMO_S3_E_DENSE = 16
PAMO_S3_E_DEFAULT = 15

PAMO_S3_DBG_SET_Q = 1
PAMO_S3_DBG_TARGET = 2
PAMO_S3_DBG_CONTACTS = 3
PAMO_S3_DBG_ENERGY = 4
PAMO_S3_DBG_DIFF = 5
PAMO_S3_DBG_HVP = 6
PAMO_S3_DBG_SOLVE = 7
PAMO_S3_DBG_CCD = 8
PAMO_S3_DBG_HINGE = 9
PAMO_S3_DBG_RESOLVE = 10
PAMO_S3_OPT_CG_GRAPH = 1
PAMO_S1_DBG_FRAME = 1
PAMO_S1_DBG_CASES = 2
PAMO_S1_DBG_FIELD = 3
PAMO_S1_DBG_OUTSIDE = 4
PAMO_S1_DBG_NORMALIZED = 5
PAMO_S1_DBG_VERTS = 6


class S3Params(ctypes.Structure):
    """pamo_s3_params (include/pamo_hip.h)."""
    _fields_ = [
        ("energy_mask", c_int), ("n_newton_iters", c_int), ("n_cg_iters", c_int), ("n_line_search_iters", c_int),
        ("max_blocks", c_int), ("ccd_max_iters", c_int), ("d_hat", c_double), ("coll_stiffness", c_double),
        ("contact_detection_radius", c_double), ("ee_classify_thres", c_double), ("ccd_slackness", c_double),
        ("ccd_thickness", c_double), ("elas_young_modulus", c_double), ("elas_poisson_ratio", c_double),
        ("hinge_stiffness", c_double), ("mesh2gt_dist_stiffness", c_double), ("min_angle_thres", c_double),
    ]


# (name, restype, argtypes) for every symbol declared in include/pamo_hip.h
SIGNATURES = [
    ("pamo_ctx_create", c_int, [ctypes.POINTER(c_vp), c_int]),
    ("pamo_ctx_destroy", c_int, [c_vp]),
    ("pamo_ctx_reset", c_int, [c_vp]),
    ("pamo_ctx_set_stream", c_int, [c_vp, c_vp, c_int]),
    ("pamo_ctx_set_option", c_int, [c_vp, c_int, c_ll]),
    ("pamo_last_error", ctypes.c_char_p, []),
    ("pamo_round", c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_int, c_float, c_float, c_int, c_int,
                           c_vp, c_vp, c_vp, c_vp, c_vp, c_int, ctypes.POINTER(c_int)]),
    ("pamo_centroid", c_int, [c_vp, c_vp, c_int, c_vp, c_int, ctypes.POINTER(c_float)]),
    ("pamo_s1_create", c_int, [ctypes.POINTER(c_vp), c_int]),
    ("pamo_s1_destroy", c_int, [c_vp]),
    ("pamo_s1_remesh", c_int, [c_vp, c_vp, c_int, c_vp, c_int, ctypes.POINTER(c_float), c_int, ctypes.POINTER(c_int),
                               ctypes.POINTER(c_int)]),
    ("pamo_s1_get", c_int, [c_vp, c_vp, c_vp]),
    ("pamo_s1_debug", c_int, [c_vp, c_int, c_vp, c_ll, ctypes.POINTER(c_ll)]),
    ("pamo_round_debug", c_int, [c_vp, c_int, c_vp, c_ll, ctypes.POINTER(c_ll)]),
    ("pamo_simplify", c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_double, c_int, c_float, c_ll, c_ll, c_vp, c_vp,
                              ctypes.POINTER(c_int), ctypes.POINTER(c_int), c_vp, c_int, ctypes.POINTER(c_int)]),
    ("pamo_prof_enable", c_int, [c_vp, ctypes.c_ulonglong]),
    ("pamo_prof_reset", c_int, [c_vp]),
    ("pamo_prof_count", c_int, []),
    ("pamo_prof_get", c_int, [c_vp, c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(c_ll),
                              ctypes.POINTER(c_double), ctypes.POINTER(c_double)]),
    ("pamo_s3_create", c_int, [ctypes.POINTER(c_vp), c_int, ctypes.POINTER(S3Params)]),
    ("pamo_s3_destroy", c_int, [c_vp]),
    ("pamo_s3_set_target", c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_double, ctypes.POINTER(c_double)]),
    ("pamo_s3_set_mesh", c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_double, ctypes.POINTER(c_double), c_vp, c_int,
                                 c_vp, c_int, c_vp, c_int]),
    ("pamo_s3_step", c_int, [c_vp, c_int]),
    ("pamo_s3_step_eval", c_int, [c_vp, c_vp, c_vp]),
    ("pamo_s3_get_vertices", c_int, [c_vp, c_vp, c_double, ctypes.POINTER(c_double)]),
    ("pamo_s3_get_normalized", c_int, [c_vp, c_vp]),
    ("pamo_s3_set_dense", c_int, [c_vp, c_vp, c_vp, c_double, c_int]),
    ("pamo_s3_debug", c_int, [c_vp, c_int, c_vp, c_ll, c_vp, c_ll, ctypes.POINTER(c_ll)]),
    ("pamo_s3_stats", c_int, [c_vp, c_vp]),
    ("pamo_s3_active_contacts", c_int, [c_vp, ctypes.POINTER(c_ll)]),
    ("pamo_s3_set_option", c_int, [c_vp, c_int, c_ll]),
    ("pamo_s3_prepare_host", c_int, [c_vp, c_int, c_vp, c_int, c_double, ctypes.POINTER(c_double), c_vp,
                                     ctypes.POINTER(c_int), c_vp, ctypes.POINTER(c_int), c_vp, ctypes.POINTER(c_double)]),
    ("pamo_s3_normalization", c_int, [c_vp, c_vp, c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_double)]),
    ("pamo_s3_set_mesh_device", c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_double, ctypes.POINTER(c_double)]),
    ("pamo_s3_merge_vertices", c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_vp, ctypes.POINTER(c_int)]),
    ("pamo_batch_run", c_int, [c_int, c_int, ctypes.POINTER(S3Params), c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                               c_int, c_float, c_ll, c_ll, c_int, c_vp, c_vp, c_vp, c_vp,
                               ctypes.POINTER(c_double), c_vp, c_vp]),
    ("pamo_serve_create", c_int, [ctypes.POINTER(c_vp), c_int, c_int, c_ll, ctypes.POINTER(S3Params), c_int]),
    ("pamo_serve_destroy", c_int, [c_vp]),
    ("pamo_serve_submit", c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_double, c_int, c_int, c_float, c_ll, c_ll,
                                  ctypes.POINTER(c_ll)]),
    ("pamo_serve_poll", c_int, [c_vp, c_ll, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                                ctypes.POINTER(c_int), ctypes.POINTER(c_double), ctypes.POINTER(c_double)]),
    ("pamo_serve_fetch", c_int, [c_vp, c_ll, c_vp, c_vp]),
    ("pamo_blender_simplify", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_float]),
    ("pamo_pasp_round", c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_float, c_float, c_int, c_vp, c_vp, c_vp, c_vp]),
    ("pamo_pasp_run", c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_float, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("pamo_metrics_create", c_int, [ctypes.POINTER(c_vp), c_int]),
    ("pamo_metrics_destroy", c_int, [c_vp]),
    ("pamo_point_mesh_sqdist", c_int, [c_vp, c_vp, c_ll, c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_vp]),
    ("pamo_hausdorff", c_int, [c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_int, c_vp, c_int,
                               ctypes.POINTER(ctypes.c_double)]),
    ("pamo_point_mesh_sqdist_noalloc", c_int, [c_vp, c_ll, c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_vp]),
]

# pamo_batch_run's per-mesh completion callback: (mesh, status, user)
BATCH_DONE_FN = ctypes.CFUNCTYPE(None, c_int, c_int, c_vp)
# pamo_s3_step_eval's per-Newton-iteration callback: (newton_iter, user)
S3_NEWTON_FN = ctypes.CFUNCTYPE(None, c_int, c_vp)

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


class Profiler:
    """HIP-event timing of selected kernels inside the library."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def enable(self, on: bool = True, kernels=None):
        if not on:
            mask = 0
        elif kernels is None:
            mask = (1 << lib().pamo_prof_count()) - 1
        else:
            names = self.names()
            mask = 0
            for k in kernels:
                mask |= 1 << names.index(k)
        check(lib().pamo_prof_enable(self.ctx.ptr, mask), "pamo_prof_enable")

    def reset(self):
        check(lib().pamo_prof_reset(self.ctx.ptr), "pamo_prof_reset")

    def names(self):
        out = []
        for k in range(lib().pamo_prof_count()):
            nm = ctypes.c_char_p()
            check(lib().pamo_prof_get(self.ctx.ptr, k, ctypes.byref(nm), None, None, None), "pamo_prof_get")
            out.append(nm.value.decode())
        return out

    def stats(self):
        res = []
        for k in range(lib().pamo_prof_count()):
            nm = ctypes.c_char_p()
            n = c_ll()
            ms = c_double()
            by = c_double()
            check(lib().pamo_prof_get(self.ctx.ptr, k, ctypes.byref(nm), ctypes.byref(n), ctypes.byref(ms),
                                      ctypes.byref(by)), "pamo_prof_get")
            if n.value:
                res.append({"name": nm.value.decode(), "launches": n.value, "total_ms": round(ms.value, 4),
                            "bytes": by.value})
        return res

"""ctypes wrapper of the CPU oracle (oracle/pamo_oracle.cpp).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker of the HIP path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libpamo_oracle.so")
_lib = None

c_int, c_float, c_double, c_ll, c_vp = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_longlong, ctypes.c_void_p


def build(force: bool = False):
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_ctx_create.restype = c_vp
        L.orc_ctx_destroy.argtypes = [c_vp]
        L.orc_num_threads.restype = c_int
        L.orc_shuffle_perm.argtypes = [c_ll, c_vp]
        L.orc_minstd.restype = ctypes.c_uint
        L.orc_minstd.argtypes = [c_ll]
        L.orc_morton3.restype = ctypes.c_uint
        L.orc_morton3.argtypes = [c_float, c_float, c_float]
        L.orc_feistel.restype = ctypes.c_ulonglong
        L.orc_feistel.argtypes = [c_ll, ctypes.c_ulonglong]
        L.orc_forward.restype = c_int
        L.orc_forward.argtypes = [c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_int, c_float, c_float, c_int, c_int, c_int,
                                  c_vp, c_vp, c_vp, c_vp, c_vp, c_int]
        L.orc_n_edges.restype = c_int
        L.orc_n_edges.argtypes = [c_vp]
        L.orc_get_edges.argtypes = [c_vp, c_vp]
        L.orc_get_edge_cost.argtypes = [c_vp, c_vp]
        L.orc_get_quadrics.argtypes = [c_vp, c_vp]
        L.orc_get_ring.restype = c_int
        L.orc_get_ring.argtypes = [c_vp, c_vp, c_vp]
        L.orc_get_collapsed.restype = c_int
        L.orc_get_collapsed.argtypes = [c_vp, c_vp, c_vp]
        L.orc_get_undo_iters.restype = c_int
        L.orc_get_undo_iters.argtypes = [c_vp, c_vp]
        L.orc_guard_pairs.restype = c_ll
        L.orc_guard_pairs.argtypes = [c_vp]
        L.orc_lbvh.restype = c_int
        L.orc_lbvh.argtypes = [c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_vp]
        L.orc_karras.restype = None
        L.orc_karras.argtypes = [c_vp, c_vp, c_int, c_vp, c_vp]
        L.orc_candidates.restype = c_ll
        L.orc_candidates.argtypes = [c_vp, c_int, c_vp, c_int, c_vp, c_ll]
        L.orc_ref_guard.restype = c_int
        L.orc_ref_guard.argtypes = [c_vp, c_vp, c_vp, c_vp, c_float]
        L.orc_pasp_forward.restype = None
        L.orc_pasp_forward.argtypes = [c_vp, c_vp, c_int, c_vp, c_int, c_float, c_float, c_int, c_vp, c_vp, c_vp,
                                       c_vp]
        L.orc_pasp_run.restype = c_int
        L.orc_pasp_run.argtypes = [c_vp, c_vp, c_int, c_vp, c_int, c_float, c_ll, c_vp, c_vp, c_vp, c_vp]
        L.orc_set_placement.restype = None
        L.orc_set_placement.argtypes = [c_vp, c_int]
        L.orc_place_optimal.restype = None
        L.orc_place_optimal.argtypes = [c_vp, c_vp, c_vp, c_vp]
        L.orc_run.restype = c_int
        L.orc_run.argtypes = [c_vp, c_vp, c_int, c_vp, c_int, c_double, c_int, c_float, c_ll, c_ll, c_int, c_vp, c_vp,
                              c_vp, c_vp, c_vp, c_int]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def shuffle_perm(m: int) -> np.ndarray:
    out = np.empty(m, dtype=np.int64)
    lib().orc_shuffle_perm(m, out.ctypes.data)
    return out


def minstd(n: int) -> int:
    return int(lib().orc_minstd(n))


def morton3(x: float, y: float, z: float) -> int:
    return int(lib().orc_morton3(x, y, z))


def feistel(m: int, idx: int) -> int:
    return int(lib().orc_feistel(m, idx))


class Round:
    """Stateful mirror of CUDSP_Free (one per simplification run)."""

    def __init__(self):
        self.ctx = lib().orc_ctx_create()

    def __del__(self):
        if getattr(self, "ctx", None):
            lib().orc_ctx_destroy(self.ctx)
            self.ctx = None

    def set_placement(self, mode: int):
        """0: midpoint collapses (the reference); 1: optimal placement."""
        lib().orc_set_placement(self.ctx, int(mode))

    def forward(self, points, triangles, verts_undo, n_verts_undo, scale, threshold, is_stuck, init, pred_mode=0):
        P = _f32(points).reshape(-1, 3)
        T = _i32(triangles).reshape(-1, 3)
        U = _i32(verts_undo).reshape(-1)
        nP, nT = len(P), len(T)
        op = np.empty((nP, 3), np.float32)
        ot = np.empty((nT, 3), np.int32)
        oo = np.empty((nP, 1), np.int32)
        om = np.empty((nP, 1), np.int32)
        cap = 8 * (nP + 1) + len(U)
        ou = np.empty(cap, np.int32)
        n = lib().orc_forward(self.ctx, P.ctypes.data, nP, T.ctypes.data, nT, U.ctypes.data, int(n_verts_undo),
                              float(scale), float(threshold), int(bool(is_stuck)), int(bool(init)), int(pred_mode),
                              op.ctypes.data, ot.ctypes.data, oo.ctypes.data, om.ctypes.data, ou.ctypes.data, cap)
        if n < 0:
            raise RuntimeError("undo list overflow")
        return op, ot, oo, om, ou[:n].copy()

    # debug views of the last round
    def edges(self):
        n = lib().orc_n_edges(self.ctx)
        out = np.empty((n, 2), np.int32)
        lib().orc_get_edges(self.ctx, out.ctypes.data)
        return out

    def edge_cost(self):
        n = lib().orc_n_edges(self.ctx)
        out = np.empty(n, np.uint32)
        lib().orc_get_edge_cost(self.ctx, out.ctypes.data)
        return out

    def quadrics(self, n_pts):
        out = np.empty((n_pts, 10), np.float32)
        lib().orc_get_quadrics(self.ctx, out.ctypes.data)
        return out

    def ring(self, n_pts):
        off = np.empty(n_pts + 1, np.int32)
        n = lib().orc_get_ring(self.ctx, off.ctypes.data, None)
        ring = np.empty(n, np.int32)
        lib().orc_get_ring(self.ctx, None, ring.ctypes.data)
        return off, ring

    def collapsed(self):
        n = lib().orc_get_collapsed(self.ctx, None, None)
        out = np.empty(n, np.int32)
        nreg = c_int()
        lib().orc_get_collapsed(self.ctx, out.ctypes.data, ctypes.byref(nreg))
        return out, nreg.value

    def undo_iters(self):
        n = lib().orc_get_undo_iters(self.ctx, None)
        out = np.empty(max(n, 1), np.int32)
        lib().orc_get_undo_iters(self.ctx, out.ctypes.data)
        return out[:n]


def lbvh(points, triangles):
    P = _f32(points).reshape(-1, 3)
    T = _i32(triangles).reshape(-1, 3)
    n = len(T)
    codes = np.empty(n, np.uint32)
    faces = np.empty(n, np.int32)
    left = np.empty(max(n, 1), np.int32)
    right = np.empty(max(n, 1), np.int32)
    m = lib().orc_lbvh(P.ctypes.data, len(P), T.ctypes.data, n, codes.ctypes.data, faces.ctypes.data,
                       left.ctypes.data, right.ctypes.data)
    return codes[:m], faces[:m], left[:max(m - 1, 0)], right[:max(m - 1, 0)]


def karras(codes, faces):
    """Karras children (left, right) of leaves with sorted keys (code << 32 | face)."""
    codes = np.ascontiguousarray(codes, np.uint32)
    faces = np.ascontiguousarray(faces, np.int32)
    n = len(codes)
    left = np.empty(max(n - 1, 1), np.int32)
    right = np.empty(max(n - 1, 1), np.int32)
    lib().orc_karras(codes.ctypes.data, faces.ctypes.data, n, left.ctypes.data, right.ctypes.data)
    return left[:max(n - 1, 0)], right[:max(n - 1, 0)]


def candidates(points, triangles, cap=None):
    P = _f32(points).reshape(-1, 3)
    T = _i32(triangles).reshape(-1, 3)
    if cap is None:
        cap = 64 * len(T) + 16
    out = np.empty((cap, 2), np.int32)
    n = lib().orc_candidates(P.ctypes.data, len(P), T.ctypes.data, len(T), out.ctypes.data, cap)
    if n > cap:
        return candidates(points, triangles, cap=int(n))
    return out[:n].copy()


def ref_guard(cur9, ci, qry9, qi, epsilon=1e-3) -> bool:
    a = _f32(cur9).reshape(-1)
    b = _f32(qry9).reshape(-1)
    ia = _i32(ci).reshape(-1)
    ib = _i32(qi).reshape(-1)
    return bool(lib().orc_ref_guard(a.ctypes.data, ia.ctypes.data, b.ctypes.data, ib.ctypes.data, float(epsilon)))


def place_optimal(M10, p0, p1):
    """The oracle's optimal collapse position for M = Q_u + Q_v (10 floats:
    00 01 02 03 11 12 13 22 23 33) and the edge's end points."""
    M = np.ascontiguousarray(M10, np.float32).reshape(10)
    a = _f32(p0).reshape(3)
    b = _f32(p1).reshape(3)
    out = np.empty(3, np.float32)
    lib().orc_place_optimal(M.ctypes.data, a.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out


def run(verts, faces, ratio, tolerance=4, threshold=1e-3, iters=1000000, min_verts=10000000000, pred_mode=0,
        ctx: Round | None = None, placement: int = 0):
    """The PaMO.run stage-2 loop on already-centred vertices."""
    V = _f32(verts).reshape(-1, 3)
    F = _i32(faces).reshape(-1, 3)
    own = ctx is None
    r = Round() if own else ctx
    if own:
        r.set_placement(placement)
    ov = np.empty_like(V)
    of = np.empty_like(F)
    nv = c_int()
    nf = c_int()
    hist = np.zeros(100000, np.int32)
    rounds = lib().orc_run(r.ctx, V.ctypes.data, len(V), F.ctypes.data, len(F), float(ratio), int(tolerance),
                           float(threshold), int(iters), int(min_verts), int(pred_mode), ov.ctypes.data,
                           ctypes.byref(nv), of.ctypes.data, ctypes.byref(nf), hist.ctypes.data, len(hist))
    return ov[:nv.value].copy(), of[:nf.value].copy(), hist[:min(rounds, len(hist))].copy()


def pasp_forward(verts, faces, scale, threshold, init, ctx: Round | None = None):
    """One legacy round (cusimp::CUDSP.forward); returns (points, faces, occ, map)."""
    V = _f32(verts).reshape(-1, 3)
    F = _i32(faces).reshape(-1, 3)
    r = Round() if ctx is None else ctx
    op = np.empty_like(V)
    of = np.empty_like(F)
    oo = np.empty(len(V), np.int32)
    om = np.empty(len(V), np.int32)
    lib().orc_pasp_forward(r.ctx, V.ctypes.data, len(V), F.ctypes.data, len(F), float(scale), float(threshold),
                           int(bool(init)), op.ctypes.data, of.ctypes.data, oo.ctypes.data, om.ctypes.data)
    return op, of, oo, om


def pasp_run(verts, faces, threshold=0.001, iters=1000, ctx: Round | None = None):
    """PaSP.run on the oracle; returns (vertices, faces, rounds)."""
    V = _f32(verts).reshape(-1, 3)
    F = _i32(faces).reshape(-1, 3)
    r = Round() if ctx is None else ctx
    ov = np.empty_like(V)
    of = np.empty_like(F)
    nv = c_int()
    nf = c_int()
    rounds = lib().orc_pasp_run(r.ctx, V.ctypes.data, len(V), F.ctypes.data, len(F), float(threshold), int(iters),
                                ov.ctypes.data, ctypes.byref(nv), of.ctypes.data, ctypes.byref(nf))
    return ov[:nv.value].copy(), of[:nf.value].copy(), rounds

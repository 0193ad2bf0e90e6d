"""Drop-in host API: ``PaMO`` and the per-round operator ``CUDSP_Free``.

Mirrors simp_cuda/pamo/__init__.py:18-205 (class PaMO) and
simp_cuda/src/pybind.cpp:16-103 (class CUDSP_Free) of the reference, on top
of libpamo_hip.so.  Tensors stay on the GPU and are handed to the library by
pointer.  A call runs on the caller's current stream when that is a created
stream, so it is ordered after the caller's queued work without a host wait;
from the null stream it runs on the context's own stream after that stream
has drained.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _hip


def _check_input(t: torch.Tensor, name: str):
    # pybind.cpp:8-14 (AT_ASSERTM -> RuntimeError)
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _check_mesh_types(points: torch.Tensor, triangles: torch.Tensor):
    """The dtype and shape checks of the per-round operator (pybind.cpp:54-63):
    float32 (V, 3) points and int32 (F, 3) triangles, else RuntimeError.
    The dtypes are checked first, so that a float64 or int64 array is named
    as such wherever it lives."""
    if not isinstance(points, torch.Tensor) or points.dtype != torch.float32:
        raise RuntimeError("points type must match the pamo class")
    if not isinstance(triangles, torch.Tensor) or triangles.dtype != torch.int32:
        raise RuntimeError("triangles type must match the pamo class")
    _check_input(points, "points")
    _check_input(triangles, "triangles")
    if points.dim() != 2 or points.shape[1] != 3:
        raise RuntimeError("points must have shape (n, 3)")
    if triangles.dim() != 2 or triangles.shape[1] != 3:
        raise RuntimeError("triangles must have shape (n, 3)")


def _check_faces(triangles: torch.Tensor, n_pts: int, name: str = "triangles"):
    """Every face row is removed (all -1) or indexes the points; the kernels
    read the points of every other row."""
    t = triangles.reshape(-1, 3)
    ok = (t == -1).all(1) | ((t >= 0) & (t < n_pts)).all(1)
    if not bool(ok.all()):
        raise RuntimeError(f"{name}: face indices outside [0, {n_pts}) (rows must be all -1 or valid)")


def _order_after_caller(ctx: _hip.Context, device) -> None:
    """Orders the context's next call after the work the caller has queued:
    on the caller's current stream when it is a created stream (no host
    wait), else on the context's own stream once the caller's null stream
    has drained."""
    s = torch.cuda.current_stream(device)
    if s.cuda_stream:
        _hip.check(_hip.lib().pamo_ctx_set_stream(ctx.ptr, ctypes.c_void_p(s.cuda_stream), 0), "pamo_ctx_set_stream")
    else:
        _hip.check(_hip.lib().pamo_ctx_set_stream(ctx.ptr, None, 1), "pamo_ctx_set_stream")
        s.synchronize()


class CUDSP_Free:
    """One collapse round per ``forward`` call (pybind.cpp:52-103).

    State (persistent buffers, the stale-buffer semantics of the reference)
    lives in the device context and carries over between calls, exactly as
    one ``_C.CUDSP_Free()`` object does.
    """

    def __init__(self, device: int | None = None):
        if device is None:
            device = torch.cuda.current_device()
        self.device = device
        self.ctx = _hip.Context(device)

    def forward(self, points, triangles, verts_undo, n_verts_undo, scale, threshold, is_stuck, init):
        _check_mesh_types(points, triangles)
        _check_input(verts_undo, "verts_undo")
        if verts_undo.dtype != torch.int32:
            raise RuntimeError("verts_undo must be int32")
        n_pts, n_tris = int(points.shape[0]), int(triangles.shape[0])
        n_undo = int(n_verts_undo)
        _check_faces(triangles, n_pts)
        if n_undo < 0 or n_undo > verts_undo.numel():
            raise RuntimeError(f"n_verts_undo={n_undo} outside [0, {verts_undo.numel()}]")
        # the entries are vertex ids of the previous call's numbering (the
        # reference maps them through that call's stale map); the library
        # rejects ids outside the context's points
        dev = points.device
        out_pts = torch.empty((n_pts, 3), dtype=torch.float32, device=dev)
        out_tris = torch.empty((n_tris, 3), dtype=torch.int32, device=dev)
        out_occ = torch.empty((n_pts, 1), dtype=torch.int32, device=dev)
        out_map = torch.empty((n_pts, 1), dtype=torch.int32, device=dev)
        cap = 4 * (n_pts + 1) + n_undo + 16
        out_undo = torch.empty(cap, dtype=torch.int32, device=dev)
        n_out = ctypes.c_int(0)
        _order_after_caller(self.ctx, dev)
        _hip.check(_hip.lib().pamo_round(
            self.ctx.ptr, points.data_ptr(), n_pts, triangles.data_ptr(), n_tris,
            verts_undo.data_ptr() if n_undo > 0 else None, n_undo, float(scale), float(threshold),
            int(bool(is_stuck)), int(bool(init)), out_pts.data_ptr(), out_tris.data_ptr(), out_occ.data_ptr(),
            out_map.data_ptr(), out_undo.data_ptr(), cap, ctypes.byref(n_out)), "pamo_round")
        return out_pts, out_tris, out_occ, out_map, out_undo[: n_out.value].clone()

    __call__ = forward

    def debug(self, what: int, dtype=torch.int32, width: int = 1, cap: int | None = None):
        """Intermediate arrays of the last round (parity tests)."""
        lib = _hip.lib()
        n = ctypes.c_longlong(0)
        if what in (_hip.PAMO_DBG_UNDO_ITERS, _hip.PAMO_DBG_N_REGULAR, _hip.PAMO_DBG_BVH_N):
            _hip.check(lib.pamo_round_debug(self.ctx.ptr, what, None, 0, ctypes.byref(n)), "pamo_round_debug")
            return n.value
        if cap is None:
            cap = 1 << 26
        buf = torch.empty(cap, dtype=dtype, device=torch.device("cuda", self.device))
        _hip.check(lib.pamo_round_debug(self.ctx.ptr, what, buf.data_ptr(), cap, ctypes.byref(n)), "pamo_round_debug")
        return buf[: n.value * (2 if what == _hip.PAMO_DBG_EDGES else 1)].clone()


def face_centroid_mean(ctx: _hip.Context, points: torch.Tensor, triangles: torch.Tensor) -> np.ndarray:
    """tris_mean of PaMO.preprocess_mesh (pamo/__init__.py:60-72): numpy's
    points[triangles].mean(axis=1).mean(axis=0) in float32, bit for bit."""
    _check_mesh_types(points, triangles)
    _check_faces(triangles, int(points.shape[0]))
    out = (ctypes.c_float * 3)()
    _order_after_caller(ctx, points.device)
    _hip.check(_hip.lib().pamo_centroid(ctx.ptr, points.data_ptr(), int(points.shape[0]), triangles.data_ptr(),
                                        int(triangles.shape[0]), out), "pamo_centroid")
    return np.array([out[0], out[1], out[2]], dtype=np.float32)


def simplify_device(ctx: _hip.Context, verts: torch.Tensor, faces: torch.Tensor, ratio: float, tolerance: int = 4,
                    threshold: float = 1e-3, iters: int = 1000000, min_verts: int = 10000000000):
    """The stage-2 round loop of PaMO.run on device; verts must be centred.
    Returns (vertices, faces, face count after each round)."""
    _check_mesh_types(verts, faces)
    nv, nf = int(verts.shape[0]), int(faces.shape[0])
    _check_faces(faces, nv)
    out_v = torch.empty_like(verts)
    out_f = torch.empty_like(faces)
    onv, onf, rounds = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    # one entry per round: at most `iters` rounds, and a round that changes
    # the face count removes at least one face
    hist = np.zeros(max(1, min(int(iters), nf + int(tolerance) * (nf + 1) + 1, 1 << 26)), dtype=np.int32)
    _order_after_caller(ctx, verts.device)
    _hip.check(_hip.lib().pamo_simplify(
        ctx.ptr, verts.data_ptr(), nv, faces.data_ptr(), nf, float(ratio), int(tolerance), float(threshold),
        int(iters), int(min_verts), out_v.data_ptr(), out_f.data_ptr(), ctypes.byref(onv), ctypes.byref(onf),
        hist.ctypes.data, len(hist), ctypes.byref(rounds)), "pamo_simplify")
    r = rounds.value
    return out_v[: onv.value], out_f[: onf.value], hist[: min(r, len(hist))].copy()


class _Bounds:
    def __init__(self, v):
        v = np.asarray(v)
        self.bounds = np.stack([v.min(0), v.max(0)])


class PaMO(torch.nn.Module):
    """``PaMO(input_mesh, use_stage1=True, use_stage3=True)`` (pamo/__init__.py:18-53).

    ``input_mesh`` is a trimesh-like object (``vertices``, ``faces`` and,
    optionally, ``bounding_box.bounds``) or a ``(vertices, faces)`` pair.
    Stage 1 (``use_stage1=True``, the reference's default) remeshes the input
    as the iso-surface of an offset distance field (``pamo_amd.stage1``) at
    the reference's resolution rule, and stage 2 then simplifies the
    remeshed surface towards the input's target face count.
    ``exact_guard=True`` (not in the reference's signature) replaces the
    reference's guard predicate by the exact intersection test, so that no
    collapse that makes two faces intersect survives; the default keeps the
    reference's predicate and results.  ``placement="optimal"`` (not in the
    reference, which always collapses an edge to its midpoint) moves a
    collapsed edge to the minimiser of its summed quadric, solved in FP64,
    with the midpoint kept where that system is singular; the default
    "midpoint" keeps the reference's results.  ``tuning`` takes launch-shape
    settings found by ``pamo_amd.autotune.tune_stage2`` (they change timings,
    never results).
    """

    def __init__(self, input_mesh, use_stage1: bool = True, use_stage3: bool = True, exact_guard: bool = False,
                 placement: str = "midpoint", tuning: dict | None = None):
        super().__init__()
        self.use_stage1 = use_stage1
        self.use_stage3 = use_stage3
        if isinstance(input_mesh, (tuple, list)):
            V, F = input_mesh
        else:
            V, F = input_mesh.vertices, input_mesh.faces
        self.gt_vertices = np.asarray(V.cpu() if isinstance(V, torch.Tensor) else V, dtype=np.float64).copy()
        self.gt_faces = np.asarray(F.cpu() if isinstance(F, torch.Tensor) else F, dtype=np.int64).copy()
        bb = getattr(input_mesh, "bounding_box", None)
        self.bbox = bb.bounds if bb is not None else _Bounds(self.gt_vertices).bounds
        self.ctx = _hip.Context(torch.cuda.current_device())
        if exact_guard:
            _hip.check(_hip.lib().pamo_ctx_set_option(self.ctx.ptr, _hip.PAMO_OPT_EXACT_GUARD, 1), "set_option")
        if placement not in ("midpoint", "optimal"):
            raise ValueError(f"placement must be 'midpoint' or 'optimal', not {placement!r}")
        self.placement = placement
        if placement == "optimal":
            _hip.check(_hip.lib().pamo_ctx_set_option(self.ctx.ptr, _hip.PAMO_OPT_PLACEMENT, 1), "set_option")
        if tuning:
            # launch-shape settings from pamo_amd.autotune (results unchanged)
            from . import autotune
            autotune.apply(self.ctx, tuning)
        if use_stage3:
            from . import stage3
            self.config = stage3.Stage3Config()
            self.system = stage3.Stage3System(self.config)
            # the input surface, normalised from its float64 vertices as
            # process() does (processing.py:15-20,57), stays resident on the
            # device for stage 3
            self.s3_target = stage3.normalized_target(self.gt_vertices)
            self.gt_v_dev = torch.as_tensor(self.gt_vertices, dtype=torch.float32).cuda()
            self.gt_f_dev = torch.as_tensor(self.gt_faces, dtype=torch.int32).cuda()
        if use_stage1:
            # the remeshing context (pamo/__init__.py:45-50: mesh2vol and
            # vol2mesh set up once per PaMO object)
            from .stage1 import Remesher
            self.remesher = Remesher(torch.cuda.current_device())
        self.R = 256
        self.band = 3 / self.R
        self.margin = self.band * 2 + 1
        self.target_faces = None
        self.timings = {}

    def run(self, points, triangles, ratio, tolerance=4, threshold=1e-3, iter=1000000, min_verts=10000000000):
        """pamo/__init__.py:122-205; returns numpy (vertices, faces)."""
        _check_mesh_types(points, triangles)
        self.target_faces = max(int(ratio * len(triangles)), min_verts)
        self.timings = {}
        t0 = time.perf_counter()
        tris_mean = face_centroid_mean(self.ctx, points, triangles)
        if self.use_stage1:
            # pamo/__init__.py:130-143: remesh at a resolution chosen by the
            # target, then stage 2 on the remeshed (centred) mesh towards
            # the target of the input face count
            from .stage1 import resolution_for
            self.R = resolution_for(self.target_faces)
            self.band = 3 / self.R
            self.margin = self.band * 2 + 1
            v1, f1 = self.remesher.remesh(points, triangles, tris_mean, self.R)
            torch.cuda.synchronize()
            self.timings["stage1_s"] = time.perf_counter() - t0
            self.stage1_out = (v1, f1)
            v2, f2, hist = simplify_device(self.ctx, v1, f1, 0.0, tolerance, threshold, iter, self.target_faces)
        else:
            verts = points - torch.from_numpy(tris_mean).to(points.device)
            v2, f2, hist = simplify_device(self.ctx, verts.contiguous(), triangles, ratio, tolerance, threshold, iter,
                                           min_verts)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        self.rounds = len(hist)
        self.face_hist = hist
        # the stage-2 result (device, centred) and the centring offset
        self.stage2_out = (v2, f2)
        self.tris_mean = tris_mean
        self.timings["stage2_s"] = t1 - t0 - self.timings.get("stage1_s", 0.0)
        if self.use_stage3:
            # pamo/__init__.py:191-203: the stage-2 result (+ centring offset,
            # a float32 add) goes through trimesh.Trimesh(process=True)'s
            # vertex merge and is projected back onto the input surface
            from . import stage3
            v3 = v2 + torch.from_numpy(tris_mean).to(v2.device)
            v3, f3 = self.system.merge_vertices(v3, f2)
            self.stage3_in = (v3, f3)
            verts_np, faces_np = stage3.process(self.gt_v_dev, self.gt_f_dev, v3, f3, 5, system=self.system,
                                                config=self.config, target=self.s3_target)
            self.timings["stage3_s"] = time.perf_counter() - t1
            return verts_np, faces_np.astype(np.int64)
        verts_np = v2.cpu().numpy() + tris_mean
        faces_np = f2.cpu().numpy()
        return verts_np, faces_np


class CUDSP:
    """The legacy per-round operator behind ``PaSP`` (cusimp::CUDSP,
    pybind.cpp:117-183): ``forward(points, triangles, scale, threshold,
    init)`` returns (points, triangles with removed rows -1, occupancy (n,1),
    map (n,1)).  One object keeps its buffers across calls like the
    reference's."""

    def __init__(self, device: int | None = None):
        if device is None:
            device = torch.cuda.current_device()
        self.device = device
        self.ctx = _hip.Context(device)

    def forward(self, points, triangles, scale, threshold, init):
        _check_input(points, "points")
        _check_input(triangles, "triangles")
        if points.dtype != torch.float32:
            raise RuntimeError("points type must match the sp class")
        if triangles.dtype != torch.int32:
            raise RuntimeError("triangles type must match the sp class")
        n_pts, n_tris = int(points.shape[0]), int(triangles.shape[0])
        _check_faces(triangles, n_pts)
        dev = points.device
        out_pts = torch.empty((n_pts, 3), dtype=torch.float32, device=dev)
        out_tris = torch.empty((n_tris, 3), dtype=torch.int32, device=dev)
        out_occ = torch.empty((n_pts, 1), dtype=torch.int32, device=dev)
        out_map = torch.empty((n_pts, 1), dtype=torch.int32, device=dev)
        _order_after_caller(self.ctx, dev)
        _hip.check(_hip.lib().pamo_pasp_round(
            self.ctx.ptr, points.data_ptr(), n_pts, triangles.data_ptr(), n_tris, float(scale), float(threshold),
            int(bool(init)), out_pts.data_ptr(), out_tris.data_ptr(), out_occ.data_ptr(), out_map.data_ptr()),
            "pamo_pasp_round")
        return out_pts, out_tris, out_occ, out_map

    __call__ = forward


class PaSP(torch.nn.Module):
    """The legacy simplifier (pamo/__init__.py:208-253): rounds of midpoint
    collapses of edges whose legacy cost is at most ``threshold``, chosen as
    independent sets, until a round removes no face; no intersection guard.
    ``run`` keeps the loop on the device (pamo_pasp_run) and returns CUDA
    tensors (vertices, faces)."""

    def __init__(self, device: int | None = None):
        super().__init__()
        if device is None:
            device = torch.cuda.current_device()
        self.ctx = _hip.Context(device)
        self.rounds = 0

    def run(self, points, triangles, threshold=0.001, iter=1000):
        _check_input(points, "points")
        _check_input(triangles, "triangles")
        if points.dtype != torch.float32 or triangles.dtype != torch.int32:
            raise RuntimeError("points must be float32 and triangles int32")
        nv, nf = int(points.shape[0]), int(triangles.shape[0])
        _check_faces(triangles, nv)
        out_v = torch.empty_like(points)
        out_f = torch.empty_like(triangles)
        onv, onf, rounds = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        _order_after_caller(self.ctx, points.device)
        _hip.check(_hip.lib().pamo_pasp_run(
            self.ctx.ptr, points.data_ptr(), nv, triangles.data_ptr(), nf, float(threshold), int(iter),
            out_v.data_ptr(), out_f.data_ptr(), ctypes.byref(onv), ctypes.byref(onf), ctypes.byref(rounds)),
            "pamo_pasp_run")
        self.rounds = rounds.value
        if self.rounds < int(iter):
            print("Converged at iteration {}".format(self.rounds - 1))
        return out_v[: onv.value], out_f[: onf.value]

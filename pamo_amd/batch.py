"""Batched and served forms of PaMO.run on one GPU (libpamo_hip.so,
pamo_batch_run / pamo_serve_*).

Each mesh is one whole ``PaMO(mesh, use_stage1=False, use_stage3=...).run``
(simp_cuda/pamo/__init__.py:122-205) and gives the same result as running it
alone; meshes run concurrently on a pool of slots with their own HIP streams.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _hip
from .stage3 import Stage3Config


def _s3_params(use_stage3, config: Stage3Config | None):
    if not any(use_stage3):
        return None
    return (config or Stage3Config()).to_params()


def pack_csr(meshes):
    """[(V (n,3) f32, F (m,3) i32)] -> device (verts, faces), host offsets."""
    vo = np.zeros(len(meshes) + 1, np.int64)
    fo = np.zeros(len(meshes) + 1, np.int64)
    for i, (V, F) in enumerate(meshes):
        vo[i + 1] = vo[i] + len(V)
        fo[i + 1] = fo[i] + len(F)
    V = torch.empty((int(vo[-1]), 3), dtype=torch.float32)
    F = torch.empty((int(fo[-1]), 3), dtype=torch.int32)
    for i, (v, f) in enumerate(meshes):
        V[vo[i]:vo[i + 1]] = torch.as_tensor(np.asarray(v, np.float32))
        F[fo[i]:fo[i + 1]] = torch.as_tensor(np.asarray(f, np.int32))
    return V.cuda(), F.cuda(), vo, fo


def simplify_batch(meshes, ratio, use_stage3=False, n_slots: int = 4, tolerance: int = 4, threshold: float = 1e-3,
                   iters: int = 100000, min_verts: int = 0, config: Stage3Config | None = None, packed=None,
                   on_done=None, exact_guard: bool = False):
    """Runs every mesh through the PaMO pipeline; returns ([(V, F)], seconds).

    ``ratio`` and ``use_stage3`` are scalars or per-mesh sequences.  Outputs
    are numpy float32 vertices and int32 faces.  ``packed`` may hold the
    result of ``pack_csr(meshes)`` (device inputs already resident).
    ``on_done(mesh, status)``, if given, is called from a worker thread as each
    mesh finishes.  ``exact_guard`` selects the exact guard predicate for
    every mesh, as ``PaMO(..., exact_guard=True)`` does."""
    n = len(meshes) if meshes is not None else len(packed[2]) - 1
    ratio = np.broadcast_to(np.asarray(ratio, np.float64), (n,)).copy()
    s3 = np.broadcast_to(np.asarray(use_stage3, np.int32), (n,)).copy()
    Vd, Fd, vo, fo = packed if packed is not None else pack_csr(meshes)
    ov = torch.empty_like(Vd)
    of = torch.empty_like(Fd)
    onv = np.zeros(n, np.int32)
    onf = np.zeros(n, np.int32)
    secs = ctypes.c_double(0)
    params = _s3_params(s3, config)
    cb = _hip.BATCH_DONE_FN(lambda mesh, status, user: on_done(mesh, status)) if on_done else None
    torch.cuda.current_stream().synchronize()
    _hip.check(_hip.lib().pamo_batch_run(
        torch.cuda.current_device(), int(n_slots), ctypes.byref(params) if params is not None else None, n,
        Vd.data_ptr(), Fd.data_ptr(), vo.ctypes.data, fo.ctypes.data, ratio.ctypes.data, s3.ctypes.data,
        int(tolerance), float(threshold), int(iters), int(min_verts), int(bool(exact_guard)), ov.data_ptr(),
        of.data_ptr(),
        onv.ctypes.data, onf.ctypes.data, ctypes.byref(secs), ctypes.cast(cb, ctypes.c_void_p) if cb else None, None),
        "pamo_batch_run")
    ovh, ofh = ov.cpu().numpy(), of.cpu().numpy()
    out = [(ovh[vo[i]:vo[i] + onv[i]].copy(), ofh[fo[i]:fo[i] + onf[i]].copy()) for i in range(n)]
    return out, secs.value


class Server:
    """Request loop (pamo_serve_*): jobs are admitted in arrival order by
    free slot and estimated HBM footprint and run concurrently."""

    def __init__(self, n_slots: int = 4, hbm_budget_bytes: int = 0, stage3: bool = True,
                 config: Stage3Config | None = None, exact_guard: bool = False):
        self.ptr = ctypes.c_void_p()
        self._params = (config or Stage3Config()).to_params() if stage3 else None
        _hip.check(_hip.lib().pamo_serve_create(
            ctypes.byref(self.ptr), torch.cuda.current_device(), int(n_slots), int(hbm_budget_bytes),
            ctypes.byref(self._params) if self._params is not None else None, int(bool(exact_guard))),
            "pamo_serve_create")
        self._keep = {}

    def close(self):
        if self.ptr.value:
            _hip.lib().pamo_serve_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, V, F, ratio, use_stage3=False, tolerance=4, threshold=1e-3, iters=100000, min_verts=0):
        v = V if isinstance(V, torch.Tensor) else torch.as_tensor(np.asarray(V, np.float32))
        f = F if isinstance(F, torch.Tensor) else torch.as_tensor(np.asarray(F, np.int32))
        v = v.to(torch.float32).cuda().contiguous()
        f = f.to(torch.int32).cuda().contiguous()
        jid = ctypes.c_longlong(0)
        torch.cuda.current_stream().synchronize()
        _hip.check(_hip.lib().pamo_serve_submit(
            self.ptr, v.data_ptr(), int(v.shape[0]), f.data_ptr(), int(f.shape[0]), float(ratio), int(bool(use_stage3)),
            int(tolerance), float(threshold), int(iters), int(min_verts), ctypes.byref(jid)), "pamo_serve_submit")
        self._keep[jid.value] = (v, f)  # inputs stay alive until the job is fetched
        return jid.value

    def poll(self, job_id, wait=False):
        st, nv, nf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lat, qd = ctypes.c_double(), ctypes.c_double()
        _hip.check(_hip.lib().pamo_serve_poll(self.ptr, int(job_id), int(bool(wait)), ctypes.byref(st),
                                              ctypes.byref(nv), ctypes.byref(nf), ctypes.byref(lat),
                                              ctypes.byref(qd)), "pamo_serve_poll")
        return {"state": ("queued", "running", "done", "failed")[st.value], "n_verts": nv.value, "n_faces": nf.value,
                "latency_s": lat.value, "queue_s": qd.value}

    def result(self, job_id):
        """Waits for the job; returns (V, F, info) and releases it."""
        info = self.poll(job_id, wait=True)
        v = torch.empty((max(info["n_verts"], 1), 3), dtype=torch.float32, device="cuda")
        f = torch.empty((max(info["n_faces"], 1), 3), dtype=torch.int32, device="cuda")
        _hip.check(_hip.lib().pamo_serve_fetch(self.ptr, int(job_id), v.data_ptr(), f.data_ptr()), "pamo_serve_fetch")
        self._keep.pop(job_id, None)
        return v[: info["n_verts"]].cpu().numpy(), f[: info["n_faces"]].cpu().numpy(), info

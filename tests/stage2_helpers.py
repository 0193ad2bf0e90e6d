"""Shared drivers for stage-2 parity tests."""
from __future__ import annotations

import numpy as np


def centred(V, F):
    mean = V[F].mean(axis=1).mean(axis=0)
    return (V - mean).astype(np.float32)


def scale_of(V):
    ext = V.max(0) - V.min(0)
    s = ext[0]
    if ext[1] > s:
        s = ext[1]
    if ext[2] > s:
        s = ext[2]
    return float(s)


def run_rounds(forward, V, F, ratio, tolerance=4, threshold=1e-3, max_rounds=100000, min_verts=0, on_round=None):
    """The PaMO.run stage-2 loop (pamo/__init__.py:147-185) around a
    per-round operator forward(points, tris, undo, n_undo, scale, thr,
    is_stuck, init) -> numpy (verts, tris, occ, map, undo)."""
    target = max(int(ratio * len(F)), min_verts)
    scale = scale_of(V)
    undo = np.zeros(0, np.int32)
    count, is_stuck, init = 0, 0, True
    hist = []
    for it in range(max_rounds):
        prev = len(F)
        out = forward(V, F, undo, len(undo), scale, threshold, is_stuck, init)
        if on_round is not None:
            on_round(it, out)
        init = False
        pv, pt, occ, mp, undo = out
        occ = occ.reshape(-1).astype(bool)
        mp = mp.reshape(-1)
        V = pv[occ]
        F = pt[pt[:, 0] >= 0]
        F = mp[F].astype(np.int32)
        hist.append(len(F))
        if len(F) <= target or len(F) <= 10:
            break
        if len(F) == prev:
            count += 1
        else:
            count = 0
            is_stuck = 0
        if count >= 2:
            is_stuck = 1
        if count == tolerance:
            break
    return V, F, np.array(hist, np.int32)

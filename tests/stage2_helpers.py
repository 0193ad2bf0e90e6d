"""Shared drivers for stage-2 parity tests."""
from __future__ import annotations

import numpy as np


def tris_mean(V, F):
    """The centring offset of PaMO.preprocess_mesh (pamo/__init__.py:60-72),
    the reference's own numpy expression on float32 points."""
    V = np.asarray(V, dtype=np.float32)
    return V[F].mean(axis=1).mean(axis=0)


def centred(V, F):
    """points - tris_mean in float32 (pamo/__init__.py:144)."""
    V = np.asarray(V, dtype=np.float32)
    return (V - tris_mean(V, F)).astype(np.float32)


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


def escape_links(left, right, n):
    """Depth-first escape link of every node of a binary tree with internal
    nodes 0..n-2 (children in left/right) and leaves n-1..2n-2: the right
    sibling of the first ancestor-or-self that is a left child, else -1."""
    import numpy as np
    N = 2 * n - 1
    parent = np.full(N, -1, dtype=np.int64)
    is_left = np.zeros(N, dtype=bool)
    for i in range(n - 1):
        parent[left[i]] = i
        parent[right[i]] = i
        is_left[left[i]] = True
    esc = np.full(N, -1, dtype=np.int64)
    # top-down: esc(left child) = right sibling; esc(right child) = esc(parent)
    order = [0]
    while order:
        x = order.pop()
        if x < n - 1:
            esc[left[x]] = right[x]
            esc[right[x]] = esc[x]
            order += [left[x], right[x]]
    return esc

"""CPU restatement of the stage-2 -> stage-3 handoff (test infrastructure
only: the tests use it as the checker; no product path imports it).

simp_cuda/pamo/__init__.py:188-194 adds the centring offset to the stage-2
vertices (float32) and builds ``trimesh.Trimesh(vertices=verts,
faces=faces)``, whose default ``process=True`` runs
``Trimesh.process`` -> ``remove_infinite_values`` and
``grouping.merge_vertices`` (trimesh 4.4.0, pinned in the reference's
env.yaml; not installed here).  The published algorithm, restated:

- the vertices become float64; each is keyed by ``round(v * 10**8)`` per
  coordinate (``tol.merge`` = 1e-8 gives 8 digits; numpy rounds half to
  even) as int64;
- the referenced vertices (every vertex when there are no faces) are
  grouped by equal keys with ``unique_rows(..., keep_order=True)``: the
  groups are ordered by their first vertex in index order;
- ``update_vertices(mask, inverse)`` keeps the first vertex of each group,
  with its own coordinates, and re-indexes the faces through ``inverse``.

Parity with trimesh itself is unpinned (trimesh is absent); the restatement
is pinned by hand cases (exact duplicates, rounding ties on both sides,
unreferenced vertices) and by an independent dictionary brute force
(``merge_vertices_brute``).
"""
from __future__ import annotations

import numpy as np


def merge_vertices(V, F):
    """(float64 vertices, int64 faces) after trimesh's process=True merge."""
    V = np.asarray(V, dtype=np.float64).reshape(-1, 3)
    F = np.asarray(F, dtype=np.int64).reshape(-1, 3)
    if len(V) == 0:
        return V, F
    if not np.isfinite(V).all():
        raise ValueError("non-finite vertex at the stage-3 handoff")
    if len(F) > 0:
        referenced = np.zeros(len(V), dtype=bool)
        referenced[F] = True
    else:
        referenced = np.ones(len(V), dtype=bool)
    keys = np.round(V * 10**8).astype(np.int64)
    sub = keys[referenced]
    # np.unique sorts stably when return_index is asked for: first occurrences
    _, first, inv = np.unique(sub, axis=0, return_index=True, return_inverse=True)
    order = first.argsort()
    u = first[order]
    i = order.argsort()[inv.reshape(-1)]
    inverse = np.zeros(len(V), dtype=np.int64)
    inverse[referenced] = i
    mask = np.nonzero(referenced)[0][u]
    return V[mask], inverse[F]


def merge_vertices_brute(V, F):
    """The same merge by a scan in index order with a dictionary of keys
    (Python's round() of a float rounds half to even)."""
    V = np.asarray(V, dtype=np.float64).reshape(-1, 3)
    F = np.asarray(F, dtype=np.int64).reshape(-1, 3)
    ref = set(int(x) for x in F.ravel()) if len(F) else set(range(len(V)))
    groups, remap, kept = {}, {}, []
    for i in range(len(V)):
        if i not in ref:
            continue
        k = tuple(int(round(float(x) * 1e8)) for x in V[i])
        if k not in groups:
            groups[k] = len(kept)
            kept.append(V[i])
        remap[i] = groups[k]
    outF = np.array([[remap[int(x)] for x in f] for f in F], dtype=np.int64).reshape(-1, 3)
    return np.array(kept, dtype=np.float64).reshape(-1, 3), outF

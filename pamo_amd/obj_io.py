"""Wavefront OBJ input and output for the example workflow of the reference
(example.py:24,39-42: `trimesh.load(path, force='mesh', process=False)` and
`trimesh.Trimesh(vertices, faces).export(path)`).

`load_obj` reads positions (`v`) and faces (`f`, position indices of
`v`, `v/vt`, `v//vn` or `v/vt/vn` corners, 1-based or negative relative
indices); polygons are triangulated as fans (0, i, i+1), as trimesh does;
every object and group goes into one mesh.  Texture coordinates and normals
are ignored, so vertices are never split at texture seams: the simplifier
needs the mesh connected through shared positions.  `save_obj` writes
`v` and `f` lines (1-based).
"""
from __future__ import annotations

import numpy as np


def load_obj(path_or_text, is_text: bool = False):
    """Returns (V float64 (n, 3), F int64 (m, 3))."""
    text = path_or_text if is_text else open(path_or_text, "r", errors="replace").read()
    verts, faces = [], []
    for line in text.splitlines():
        s = line.split("#", 1)[0].split()
        if not s:
            continue
        if s[0] == "v":
            verts.append((float(s[1]), float(s[2]), float(s[3])))
        elif s[0] == "f":
            idx = []
            for c in s[1:]:
                k = int(c.split("/", 1)[0])
                idx.append(k - 1 if k > 0 else len(verts) + k)
            for i in range(1, len(idx) - 1):
                faces.append((idx[0], idx[i], idx[i + 1]))
    V = np.array(verts, np.float64).reshape(-1, 3)
    F = np.array(faces, np.int64).reshape(-1, 3)
    if len(F) and (F.min() < 0 or F.max() >= len(V)):
        raise ValueError("OBJ face index outside the vertex list")
    return V, F


def save_obj(path, V, F):
    V = np.asarray(V, np.float64).reshape(-1, 3)
    F = np.asarray(F, np.int64).reshape(-1, 3)
    with open(path, "w") as f:
        for x, y, z in V:
            f.write(f"v {x:.9g} {y:.9g} {z:.9g}\n")
        for a, b, c in F + 1:
            f.write(f"f {a} {b} {c}\n")

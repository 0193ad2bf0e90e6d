"""Synthetic workloads for the stage-2/stage-3 path.

* ``icosphere_sh``: subdivided icosahedron with a seeded spherical-harmonic
  radial displacement (the small single-mesh case: level 5 gives 10,242
  vertices and 20,480 faces).
* ``ellipsoid_box_union``: marching-tetrahedra contour of the union of 16
  seeded rotated ellipsoids and 4 axis-aligned boxes on [-1, 1]^3 (the
  2M-face headline case at the right grid resolution).  Built by
  ``libpamo_synth.so`` (host C++); watertight and manifold by construction.
* ``offset_sheets``: a surface plus an inward offset copy with reversed
  orientation (closely spaced parallel sheets).

All generators use ``numpy.random.default_rng(seed)``; outputs are float32
positions and int32 faces.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SYNTH_LIB = None


def _synth_lib():
    global _SYNTH_LIB
    if _SYNTH_LIB is None:
        path = os.path.join(_HERE, "lib", "libpamo_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        lib.synth_mt.restype = ctypes.c_int
        lib.synth_mt.argtypes = [
            ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.POINTER(ctypes.POINTER(ctypes.c_float)), ctypes.POINTER(ctypes.c_longlong),
            ctypes.POINTER(ctypes.POINTER(ctypes.c_int)), ctypes.POINTER(ctypes.c_longlong),
        ]
        lib.synth_free.argtypes = [ctypes.c_void_p]
        _SYNTH_LIB = lib
    return _SYNTH_LIB


# ----------------------------------------------------------------- icosphere
def icosphere(level: int):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = np.array(v, dtype=np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array(f, dtype=np.int64)
    for _ in range(level):
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = V[uniq[:, 0]] + V[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(V)
        V = np.concatenate([V, mid], axis=0)
        nf = len(F)
        a = inv[:nf] + base
        b = inv[nf:2 * nf] + base
        c = inv[2 * nf:] + base
        F = np.concatenate([
            np.stack([F[:, 0], a, c], 1), np.stack([F[:, 1], b, a], 1),
            np.stack([F[:, 2], c, b], 1), np.stack([a, b, c], 1)], axis=0)
    return V, F


def _real_sph_harm(l, m, theta, phi):
    """Real orthonormal spherical harmonic Y_lm (theta polar, phi azimuth)."""
    from scipy.special import lpmv
    am = abs(m)
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    p = lpmv(am, l, np.cos(theta)) * norm
    if m > 0:
        return math.sqrt(2.0) * p * np.cos(am * phi)
    if m < 0:
        return math.sqrt(2.0) * p * np.sin(am * phi)
    return p


def icosphere_sh(level: int = 5, seed: int = 0, n_terms: int = 8, amp: float = 0.05):
    """Icosphere with radius 1 + amp * sum_i c_i Y_{l_i m_i}; l ~ U{2..6},
    m ~ U{-l..l}, c ~ N(0,1) from default_rng(seed)."""
    V, F = icosphere(level)
    rng = np.random.default_rng(seed)
    theta = np.arccos(np.clip(V[:, 2], -1.0, 1.0))
    phi = np.arctan2(V[:, 1], V[:, 0])
    r = np.ones(len(V))
    for _ in range(n_terms):
        l = int(rng.integers(2, 7))
        m = int(rng.integers(-l, l + 1))
        c = float(rng.standard_normal())
        r += amp * c * _real_sph_harm(l, m, theta, phi)
    V = V * r[:, None]
    return V.astype(np.float32), F.astype(np.int32)


# ----------------------------------------------------------------- implicit union
def _random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def union_shapes(seed: int, n_ell: int = 16, n_box: int = 4):
    """Records of 16 doubles (kind, centre, axes, rotation) for the union."""
    rng = np.random.default_rng(seed)
    recs = []
    for _ in range(n_ell):
        c = rng.uniform(-0.6, 0.6, 3)
        a = rng.uniform(0.1, 0.35, 3)
        R = _random_rotation(rng)
        recs.append(np.concatenate([[0.0], c, a, R.reshape(-1)]))
    for _ in range(n_box):
        c = rng.uniform(-0.6, 0.6, 3)
        h = rng.uniform(0.1, 0.3, 3)
        recs.append(np.concatenate([[1.0], c, h, np.eye(3).reshape(-1)]))
    return np.ascontiguousarray(np.stack(recs), dtype=np.float64)


def contour_shapes(shapes: np.ndarray, res: int, lo: float = -1.0, hi: float = 1.0, level: float = 0.0,
                   min_abs: float = 0.0):
    """Marching tetrahedra of {min over shapes = level} on a res^3 grid over
    [lo, hi]^3.  The surface is exactly the level set of the grid's
    piecewise-linear interpolant, so contours at different levels are
    disjoint.  Grid values are kept at least min_abs from the level (so that
    crossing points stay clear of grid nodes in float32)."""
    lib = _synth_lib()
    pv = ctypes.POINTER(ctypes.c_float)()
    pf = ctypes.POINTER(ctypes.c_int)()
    nv = ctypes.c_longlong()
    nf = ctypes.c_longlong()
    shapes = np.ascontiguousarray(shapes, dtype=np.float64)
    rc = lib.synth_mt(shapes.ctypes.data, len(shapes), int(res), lo, hi, float(level), float(min_abs), ctypes.byref(pv),
                      ctypes.byref(nv), ctypes.byref(pf), ctypes.byref(nf))
    if rc != 0:
        raise RuntimeError("synth_mt failed")
    V = np.ctypeslib.as_array(pv, shape=(nv.value, 3)).copy()
    F = np.ctypeslib.as_array(pf, shape=(nf.value, 3)).copy()
    lib.synth_free(ctypes.cast(pv, ctypes.c_void_p))
    lib.synth_free(ctypes.cast(pf, ctypes.c_void_p))
    return V, F


def ellipsoid_box_union(seed: int = 1, res: int = 160):
    return contour_shapes(union_shapes(seed), res)


# Grid resolution at which the seed-1 union has 1.9M-2.1M faces under
# marching tetrahedra (tuned once with ellipsoid_box_union_faces).
HEADLINE_RES = 296


def headline_mesh(seed: int = 1):
    """The 2M-face single-mesh workload (stages 2+3, ratio 0.01)."""
    return ellipsoid_box_union(seed, HEADLINE_RES)


def stress_sheets(seed: int = 500, target_faces: float = 8e6, level: float = -1e-3, min_abs: float = 1e-5):
    """Closely separated parallel sheets (intersection stress): the union's
    contour and, inside it, the contour at `level`, with reversed
    orientation, together about target_faces faces.  Being level sets of
    the same piecewise-linear function, the sheets are disjoint.  The grid
    resolution steps (res, res+1, res-1, ...) until both contours are
    watertight and manifold.  Returns (V, F, res)."""
    shapes = union_shapes(seed)
    res = max(8, int(round((0.5 * target_faces / _FACES_PER_RES2) ** 0.5)))
    for step in range(16):
        r = res + ((step + 1) // 2 if step % 2 else -(step // 2))
        Vo, Fo = contour_shapes(shapes, r, min_abs=min_abs)
        Vi, Fi = contour_shapes(shapes, r, level=level, min_abs=min_abs)
        if is_watertight_manifold(Fo, len(Vo)) and is_watertight_manifold(Fi, len(Vi)):
            V = np.concatenate([Vo, Vi], 0).astype(np.float32)
            F = np.concatenate([Fo, Fi[:, ::-1] + len(Vo)], 0).astype(np.int32)
            return V, F, r
    raise RuntimeError(f"no watertight manifold contours near grid {res} for seed {seed}")


def vertex_normals(V: np.ndarray, F: np.ndarray):
    V = V.astype(np.float64)
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    vn = np.zeros_like(V)
    for k in range(3):
        np.add.at(vn, F[:, k], n)
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-300)
    return vn


def offset_sheets(V: np.ndarray, F: np.ndarray, rel_offset: float = 1e-4):
    """Surface plus an inward copy offset by rel_offset * bbox diagonal along
    area-weighted vertex normals, with reversed orientation."""
    V64 = V.astype(np.float64)
    diag = float(np.linalg.norm(V64.max(0) - V64.min(0)))
    vn = vertex_normals(V, F)
    inner = V64 - rel_offset * diag * vn
    V2 = np.concatenate([V64, inner], 0).astype(np.float32)
    F2 = np.concatenate([F, F[:, ::-1] + len(V)], 0).astype(np.int32)
    return V2, F2


# ----------------------------------------------------------------- checks
def edge_use_counts(F: np.ndarray):
    """Directed-edge analysis: returns (n_edges, n_boundary, n_nonmanifold, n_misoriented)."""
    F = np.asarray(F, dtype=np.int64)
    d = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], 0)
    und = np.sort(d, axis=1)
    key = und[:, 0] * (int(F.max()) + 1) + und[:, 1]
    _, cnt = np.unique(key, return_counts=True)
    dkey = d[:, 0] * (int(F.max()) + 1) + d[:, 1]
    _, dcnt = np.unique(dkey, return_counts=True)
    return len(cnt), int((cnt == 1).sum()), int((cnt > 2).sum()), int((dcnt > 1).sum())


def is_watertight_manifold(F: np.ndarray, n_vertices: int | None = None):
    """Closed (every edge in exactly two faces, opposite directions), no
    repeated faces, and every vertex's faces form a single fan."""
    F = np.asarray(F, dtype=np.int64)
    if len(F) == 0:
        return False
    if np.any(F[:, 0] == F[:, 1]) or np.any(F[:, 1] == F[:, 2]) or np.any(F[:, 2] == F[:, 0]):
        return False
    ne, nb, nm, mis = edge_use_counts(F)
    if nb or nm or mis:
        return False
    # vertex fan check: around each vertex the opposite edges (b -> c) must
    # form exactly one cycle
    n = int(F.max()) + 1 if n_vertices is None else n_vertices
    a = np.concatenate([F[:, 0], F[:, 1], F[:, 2]])
    b = np.concatenate([F[:, 1], F[:, 2], F[:, 0]])
    c = np.concatenate([F[:, 2], F[:, 0], F[:, 1]])
    order = np.lexsort((b, a))
    a, b, c = a[order], b[order], c[order]
    # for each (vertex a, b) the successor is c; walk cycles per vertex
    key_ab = a * n + b
    idx_of = {int(k): i for i, k in enumerate(key_ab)} if len(a) < 200000 else None
    if idx_of is None:
        # vectorised: position of (a, c) among sorted keys
        pos = np.searchsorted(key_ab, a * n + c)
        if np.any(pos >= len(key_ab)) or np.any(key_ab[np.minimum(pos, len(key_ab) - 1)] != a * n + c):
            return False
        nxt = pos
    else:
        nxt = np.empty(len(a), dtype=np.int64)
        for i in range(len(a)):
            j = idx_of.get(int(a[i] * n + c[i]))
            if j is None:
                return False
            nxt[i] = j
    # count cycles of the permutation nxt restricted to each vertex
    seen = np.zeros(len(a), dtype=bool)
    start = np.r_[0, np.flatnonzero(np.diff(a)) + 1]
    cycles = 0
    # walk with vectorised pointer jumping: label each element by min index on its cycle
    lab = np.arange(len(a))
    cur = nxt.copy()
    for _ in range(int(np.ceil(np.log2(max(2, np.max(np.diff(np.r_[start, len(a)])))))) + 1):
        lab = np.minimum(lab, lab[cur])
        cur = cur[cur]
    cycles = int(np.sum(lab == np.arange(len(a))))
    del seen
    return cycles == len(start)


def euler_characteristic(F: np.ndarray, n_vertices: int):
    ne = edge_use_counts(F)[0]
    used = len(np.unique(np.asarray(F).reshape(-1)))
    return used - ne + len(F)


# faces per res^2 of the ellipsoid/box union under marching tetrahedra
# (2,006,528 faces at res 296 for seed 1)
_FACES_PER_RES2 = 2006528 / 296 ** 2


def sized_union(seed: int, target_faces: float):
    """The headline generator at the grid resolution whose face count is
    closest to target_faces (for this seed's shapes).  A contour whose sheets
    meet at a single grid point has a pinch vertex (not manifold); the
    resolution then steps to res+1, res-1, res+2, ... until the mesh is
    watertight and manifold."""
    res = max(8, int(round((target_faces / _FACES_PER_RES2) ** 0.5)))
    for step in range(16):
        r = res + ((step + 1) // 2 if step % 2 else -(step // 2))
        V, F = ellipsoid_box_union(seed, max(8, r))
        if is_watertight_manifold(F, len(V)):
            return V, F
    raise RuntimeError(f"no watertight manifold contour near grid {res} for seed {seed}")


def batch_workload(n: int = 96, seed0: int = 100, f_lo: float = 20e3, f_hi: float = 800e3, rng_seed: int = 3):
    """Batched workload: n meshes of the headline generator, seeds seed0..,
    face counts log-uniform in [f_lo, f_hi], ratio per mesh from
    {0.005, 0.01, 0.02, 0.05}.  Returns ([(V, F)], ratios)."""
    rng = np.random.default_rng(rng_seed)
    targets = np.exp(rng.uniform(np.log(f_lo), np.log(f_hi), n))
    ratios = rng.choice([0.005, 0.01, 0.02, 0.05], n)
    meshes = [sized_union(seed0 + i, targets[i]) for i in range(n)]
    return meshes, ratios.astype(np.float64)


def serve_stream(duration_s: float = 20.0, rate: float = 40.0, pool: int = 24, f_lo: float = 10e3,
                 f_hi: float = 2e6, seed: int = 7):
    """Served workload: Poisson arrivals at `rate` per second for duration_s;
    each job draws a mesh from a pool of `pool` distinct meshes of the
    headline generator with face counts log-uniform in [f_lo, f_hi], a ratio
    uniform in [0.002, 0.05] and stage 3 with probability 0.5.  Returns
    (pool meshes, [(arrival_s, pool index, ratio, stage3)])."""
    rng = np.random.default_rng(seed)
    targets = np.exp(rng.uniform(np.log(f_lo), np.log(f_hi), pool))
    meshes = [sized_union(700 + i, targets[i]) for i in range(pool)]
    t, jobs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration_s:
            break
        jobs.append((t, int(rng.integers(pool)), float(rng.uniform(0.002, 0.05)), bool(rng.random() < 0.5)))
    return meshes, jobs

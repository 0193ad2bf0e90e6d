"""CPU restatement of stage-1 remeshing (test infrastructure only: the
tests, smoke() and the CPU baseline use it as the checker; no product path
imports it).

Follows simp_cuda/pamo/__init__.py:60-120 (preprocess_mesh, remesh) where
the reference's own code goes, and restates the two third-party steps the
reference calls but does not ship, from their published methods:

- torchcumesh2sdf.get_sdf (pip package torchcumesh2sdf; its sign rule and
  sample placement are not in the reference): here the unsigned distance to
  the triangles at the samples within the band (closest point on a triangle,
  Ericson, Real-Time Collision Detection 5.1.5), signed by a flood fill of the
  outside from the grid border through samples farther than 0.55 / R from
  the surface.  Parity with the package is unpinned.
- pdmc.DMC (pip package pdmc, dual marching cubes after Nielson 2004) and
  the quad split: here marching cubes with one face rule for all 256
  corner-sign cases (alternating face corners: each inside corner cut off
  separately); a cell's patches are the cycles of its crossing edges joined
  face by face, oriented so that the inside lies to the right of every face
  segment seen from outside, and fanned into triangles; vertices are the
  edge crossings.  Parity with the package is unpinned; the restatement is
  pinned by closed, manifold and orientation checks on random fields and by
  a sphere's closed form.

Every function works in float32 in the operation order of the device code,
so the device results can be compared bit for bit.
"""
from __future__ import annotations

import numpy as np

EDGE_CORNER = [(0, 1), (2, 3), (4, 5), (6, 7), (0, 2), (1, 3), (4, 6), (5, 7), (0, 4), (1, 5), (2, 6), (3, 7)]
FACE_CORNER = [(0, 2, 6, 4), (1, 3, 7, 5), (0, 1, 5, 4), (2, 3, 7, 6), (0, 1, 3, 2), (4, 5, 7, 6)]


def resolution_for(target_faces: int) -> int:
    """pamo/__init__.py:136-141."""
    R = 256
    if target_faces <= 1000:
        R = 128
    if target_faces <= 50:
        R = 64
    return R


def preprocess(points: np.ndarray, faces: np.ndarray, band: float, margin: float):
    """preprocess_mesh (pamo/__init__.py:60-72), numpy float32 as written."""
    tris = points[faces].astype(np.float32)
    tris_mean = tris.mean(axis=1).mean(axis=0)
    tris = tris - tris_mean
    tris_min = tris.min(0).min(0)
    tris = tris - tris_min
    tris_max = tris.max()
    tris = (tris / tris_max + np.float32(band)) / np.float32(margin)
    return tris.astype(np.float32), tris_min, tris_max, tris_mean


FACE_NORMAL = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]


def _edge_of(a, b):
    for e, (p, q) in enumerate(EDGE_CORNER):
        if (p, q) == (a, b) or (p, q) == (b, a):
            return e
    raise ValueError


def _cpos(c):
    return np.array([c & 1, (c >> 1) & 1, c >> 2], np.float64)


def _mid(e):
    return 0.5 * (_cpos(EDGE_CORNER[e][0]) + _cpos(EDGE_CORNER[e][1]))


def case_patches(cs):
    """The patches of corner-sign case cs (bit c: corner c inside) as
    ordered edge cycles; None if the face rule gives no consistent cycles."""
    inside = [(cs >> c) & 1 for c in range(8)]
    segs = []  # (edge a, edge b, face, inside point)
    for f, q in enumerate(FACE_CORNER):
        fe = [_edge_of(q[i], q[(i + 1) % 4]) for i in range(4)]
        cross = [inside[q[i]] != inside[q[(i + 1) % 4]] for i in range(4)]
        if sum(cross) == 2:
            a, b = [fe[i] for i in range(4) if cross[i]]
            pin = np.mean([_cpos(c) for c in q if inside[c]], axis=0)
            segs.append((a, b, f, pin))
        elif sum(cross) == 4:
            for i in range(4):
                if inside[q[i]]:
                    segs.append((fe[(i + 3) % 4], fe[i], f, _cpos(q[i])))
    at = {e: [] for e in range(12)}
    for k, (a, b, f, pin) in enumerate(segs):
        at[a].append(k)
        at[b].append(k)

    def forward(k, u, v):
        a, b = _mid(u), _mid(v)
        n = np.array(FACE_NORMAL[segs[k][2]], np.float64)
        return float(np.dot(np.cross(b - a, n), segs[k][3] - 0.5 * (a + b))) > 0

    def other(k, e):
        return segs[k][1] if segs[k][0] == e else segs[k][0]

    used, patches = set(), []
    for e in range(12):
        if inside[EDGE_CORNER[e][0]] == inside[EDGE_CORNER[e][1]] or e in used:
            continue
        if len(at[e]) != 2:
            return None
        k = at[e][0]
        if not forward(k, e, other(k, e)):
            k = at[e][1]
            if not forward(k, e, other(k, e)):
                return None
        cyc, cur = [], e
        while True:
            cyc.append(cur)
            used.add(cur)
            n2 = other(k, cur)
            if not forward(k, cur, n2):
                return None
            if n2 == e:
                break
            if n2 in used or len(at[n2]) != 2:
                return None
            k = at[n2][1] if at[n2][0] == k else at[n2][0]
            cur = n2
        if len(cyc) < 3:
            return None
        patches.append(cyc)
    return patches


def case_table():
    """cyc[256, 12] (edges patch after patch in cycle order, -1 padded),
    plen[256, 4], npatch[256], ntri[256]: the layout of the device table."""
    cyc = -np.ones((256, 12), np.int8)
    plen = np.zeros((256, 4), np.int8)
    npatch = np.zeros(256, np.int8)
    ntri = np.zeros(256, np.int8)
    for cs in range(256):
        ps = case_patches(cs)
        assert ps is not None, cs
        flat = [e for p in ps for e in p]
        cyc[cs, :len(flat)] = flat
        plen[cs, :len(ps)] = [len(p) for p in ps]
        npatch[cs] = len(ps)
        ntri[cs] = sum(len(p) - 2 for p in ps)
    return cyc, plen, npatch, ntri


def _closest_dist(p, a, b, c):
    """|p - closest point of triangle abc|, float32, Ericson 5.1.5 in the
    device's operation order; p (n,3), a b c (3,)."""
    f = np.float32

    def dot(u, v):
        return (u[..., 0] * v[..., 0] + u[..., 1] * v[..., 1]) + u[..., 2] * v[..., 2]

    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = dot(ab, ap), dot(ac, ap)
    bp = p - b
    d3, d4 = dot(ab, bp), dot(ac, bp)
    vc = d1 * d4 - d3 * d2
    cp = p - c
    d5, d6 = dot(ab, cp), dot(ac, cp)
    vb = d5 * d2 - d1 * d6
    va = d3 * d6 - d5 * d4
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.empty_like(p)
        done = np.zeros(len(p), bool)

        def put(mask, val):
            m = mask & ~done
            q[m] = val[m] if val.ndim == 2 else val
            done[:] |= m

        put((d1 <= 0) & (d2 <= 0), np.broadcast_to(a, p.shape))
        put((d3 >= 0) & (d4 <= d3), np.broadcast_to(b, p.shape))
        put((vc <= 0) & (d1 >= 0) & (d3 <= 0), a + (d1 / (d1 - d3))[:, None] * ab)
        put((d6 >= 0) & (d5 <= d6), np.broadcast_to(c, p.shape))
        put((vb <= 0) & (d2 >= 0) & (d6 <= 0), a + (d2 / (d2 - d6))[:, None] * ac)
        put((va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0),
            b + ((d4 - d3) / ((d4 - d3) + (d5 - d6)))[:, None] * (c - b))
        den = f(1.0) / (va + vb + vc)
        v, w = vb * den, vc * den
        put(np.ones(len(p), bool), a + v[:, None] * ab + w[:, None] * ac)
    r = p - q
    return np.sqrt(dot(r, r)).astype(np.float32)


def udf(tris: np.ndarray, R: int, band: float) -> np.ndarray:
    """Distance to the nearest triangle at every sample ((k + 0.5) / R per
    axis), +inf beyond `band`; (R, R, R) indexed [z, y, x]."""
    f = np.float32
    k = ((np.arange(R, dtype=np.float32) + f(0.5)) * (f(1.0) / f(R))).astype(np.float32)
    Z, Y, X = np.meshgrid(k, k, k, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    best = np.full(len(P), np.inf, np.float32)
    for t in tris.astype(np.float32):
        d = _closest_dist(P, t[0], t[1], t[2])
        best = np.minimum(best, np.where(d <= f(band), d, np.float32(np.inf)))
    return best.reshape(R, R, R)


def outside(field: np.ndarray, R: int) -> np.ndarray:
    """Samples farther than 0.55 / R from the surface that a 6-connected
    path of such samples joins to the grid border."""
    from scipy import ndimage
    wall = np.float32(0.55) / np.float32(R)
    passable = field > wall
    lab, _ = ndimage.label(passable, structure=ndimage.generate_binary_structure(3, 1))
    border = np.unique(np.concatenate([lab[0].ravel(), lab[-1].ravel(), lab[:, 0].ravel(), lab[:, -1].ravel(),
                                       lab[:, :, 0].ravel(), lab[:, :, -1].ravel()]))
    border = border[border > 0]
    return np.isin(lab, border)


def marching_cubes(sdf: np.ndarray, iso: float):
    """Marching cubes of sdf ([z, y, x]) at `iso`: vertices in sample index
    units, one per crossing grid edge in the order (direction, then sample,
    x fastest), and triangles per cell (x fastest), patch and fan."""
    f = np.float32
    iso = f(iso)
    R = sdf.shape[0]
    n = R * R * R
    flat = sdf.reshape(-1)
    vid = -np.ones(3 * n, np.int64)
    verts = []
    for d in range(3):
        step = (1, R, R * R)[d]
        for s in range(n):
            q = (s % R, (s // R) % R, s // (R * R))
            if q[d] + 1 >= R:
                continue
            d0, d1 = flat[s], flat[s + step]
            if (d0 < iso) == (d1 < iso):
                continue
            t = (iso - d0) / (d1 - d0)
            p = [f(q[0]), f(q[1]), f(q[2])]
            p[d] = f(p[d] + t)
            vid[d * n + s] = len(verts)
            verts.append(p)
    cyc, plen, npatch, ntri = case_table()
    tris = []
    m = R - 1
    for z in range(m):
        for y in range(m):
            for x in range(m):
                cs = 0
                for c in range(8):
                    cs |= int(sdf[z + (c >> 2), y + ((c >> 1) & 1), x + (c & 1)] < iso) << c

                def v_of(e):
                    e = int(e)
                    c0 = EDGE_CORNER[e][0]
                    s = ((z + (c0 >> 2)) * R + (y + ((c0 >> 1) & 1))) * R + (x + (c0 & 1))
                    return int(vid[(e >> 2) * n + s])

                base = 0
                for p in range(npatch[cs]):
                    ln = plen[cs, p]
                    v0 = v_of(cyc[cs, base])
                    for j in range(1, ln - 1):
                        tris.append([v0, v_of(cyc[cs, base + j]), v_of(cyc[cs, base + j + 1])])
                    base += ln
    return np.array(verts, np.float32).reshape(-1, 3), np.array(tris, np.int64).reshape(-1, 3)


def to_frame(verts_idx: np.ndarray, R: int, band: float, margin: float, tmax, tmin) -> np.ndarray:
    """Sample index units to the centred frame (the inverse of preprocess)."""
    f = np.float32
    inv = f(1.0) / f(R)
    x = ((verts_idx.astype(np.float32) + f(0.5)) * inv * f(margin) - f(band)) * np.float32(tmax) + tmin.astype(np.float32)
    return x.astype(np.float32)


def remesh(points: np.ndarray, faces: np.ndarray, R: int):
    """The whole stage 1 on small inputs: (vertices in the centred frame,
    faces, and the intermediate arrays for comparisons)."""
    band = 3.0 / R
    margin = band * 2 + 1
    tris, tmin, tmax, tmean = preprocess(points, faces, band, margin)
    field = udf(tris.reshape(-1, 3, 3), R, np.float32(band))
    out = outside(field, R)
    sdf = np.where(out, field, -field).astype(np.float32)
    iso = np.float32(0.9 / R)
    vi, tri = marching_cubes(sdf, iso)
    return {"tris": tris, "tmin": tmin, "tmax": tmax, "tmean": tmean, "field": field, "outside": out, "sdf": sdf,
            "verts_idx": vi, "faces": tri, "verts": to_frame(vi, R, band, margin, tmax, tmin)}

"""Pins of the stage-1 oracle (oracle/stage1.py) on the CPU.

The reference calls two packages it does not ship (torchcumesh2sdf,
pdmc), so the restatement is pinned by what the mathematics fixes: every
case of the marching-cubes table forms closed, oriented patch cycles;
marching cubes of random fields gives closed, edge- and vertex-manifold,
consistently oriented surfaces (a wrong face rule or orientation fails
here); a sphere's surface has Euler characteristic 2, lies on the sphere
and faces outward; the distance field matches closed forms of a triangle;
the flood fill separates an enclosed interior and not an open sheet.
"""
import collections

import numpy as np
import pytest

from oracle import stage1 as s1


def test_case_table_patches_are_closed_oriented_cycles():
    cyc, plen, npatch, ntri = s1.case_table()  # asserts every case builds
    assert npatch[0] == 0 and npatch[255] == 0
    for c in range(8):
        for cs in (1 << c, 255 ^ (1 << c)):
            assert npatch[cs] == 1 and plen[cs, 0] == 3 and ntri[cs] == 1
    # two inside corners diagonal on the face z = 0 (corners 0 and 3): cut off separately
    assert npatch[1 | 8] == 2 and list(plen[1 | 8, :2]) == [3, 3]
    for cs in range(256):
        inside = [(cs >> c) & 1 for c in range(8)]
        crossing = sorted(e for e, (c0, c1) in enumerate(s1.EDGE_CORNER) if inside[c0] != inside[c1])
        flat = [int(e) for e in cyc[cs] if e >= 0]
        assert sorted(flat) == crossing  # each crossing edge in exactly one patch
        assert sum(plen[cs]) == len(flat) and ntri[cs] == sum(int(k) - 2 for k in plen[cs][:npatch[cs]])


def _check_closed_manifold(faces, n_verts):
    """Every directed edge once and its reverse once; every vertex's fan is
    one cycle; no degenerate triangle."""
    assert (faces[:, 0] != faces[:, 1]).all() and (faces[:, 1] != faces[:, 2]).all() and \
        (faces[:, 2] != faces[:, 0]).all()
    directed = collections.Counter()
    for a, b, c in faces:
        for u, v in ((a, b), (b, c), (c, a)):
            directed[(int(u), int(v))] += 1
    for (u, v), k in directed.items():
        assert k == 1, ("directed edge used twice", u, v)
        assert directed.get((v, u), 0) == 1, ("open or inconsistent edge", u, v)
    # vertex fans: the triangles at v chain into one cycle through their
    # edges at v (next[a] = b for the triangle (v, a, b))
    nxt = collections.defaultdict(dict)
    for a, b, c in faces:
        for v, x, y in ((a, b, c), (b, c, a), (c, a, b)):
            assert int(x) not in nxt[int(v)]
            nxt[int(v)][int(x)] = int(y)
    for v, m in nxt.items():
        start = next(iter(m))
        x, steps = start, 0
        while True:
            x = m[x]
            steps += 1
            if x == start:
                break
            assert steps <= len(m)
        assert steps == len(m), ("vertex fan is not one cycle", v)
    assert set(nxt) == set(range(n_verts))


@pytest.mark.parametrize("seed", range(6))
def test_marching_cubes_of_random_fields_is_closed_manifold_and_oriented(seed):
    rng = np.random.default_rng(seed)
    R = 9
    sdf = rng.choice(np.array([-1.0, 1.0], np.float32), size=(R, R, R), p=[0.45, 0.55])
    # the border outside, as the flood fill leaves it
    sdf[0], sdf[-1], sdf[:, 0], sdf[:, -1], sdf[:, :, 0], sdf[:, :, -1] = 1, 1, 1, 1, 1, 1
    sdf = sdf * rng.uniform(0.2, 1.0, size=sdf.shape).astype(np.float32)
    v, tri = s1.marching_cubes(sdf.astype(np.float32), 0.0)
    assert len(tri) > 100
    _check_closed_manifold(tri, len(v))


def test_marching_cubes_of_a_sphere_has_genus_zero_lies_on_it_and_faces_out():
    R = 16
    k = np.arange(R, dtype=np.float32)
    Z, Y, X = np.meshgrid(k, k, k, indexing="ij")
    c, r = np.float32(7.3), np.float32(4.6)
    sdf = (np.sqrt((X - c) ** 2 + (Y - c) ** 2 + (Z - c) ** 2) - r).astype(np.float32)
    v, tri = s1.marching_cubes(sdf, 0.0)
    _check_closed_manifold(tri, len(v))
    edges = {tuple(sorted((int(a), int(b)))) for t in tri for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))}
    assert len(v) - len(edges) + len(tri) == 2
    rad = np.linalg.norm(v - c, axis=1)
    assert np.abs(rad - r).max() < 0.1
    # outward orientation (inside = below the level): normals point away from the centre
    n = np.cross(v[tri[:, 1]] - v[tri[:, 0]], v[tri[:, 2]] - v[tri[:, 0]])
    assert (np.einsum("ij,ij->i", n, v[tri].mean(1) - c) > 0).mean() > 0.99


def test_distance_field_closed_forms():
    R = 16
    band = np.float32(3.0 / R)
    # triangle in the plane z = 0.5 (a sample plane lies at (7 + 0.5) / 16)
    t = np.array([[[0.3, 0.3, 0.5], [0.7, 0.3, 0.5], [0.3, 0.7, 0.5]]], np.float32)
    d = s1.udf(t, R, band)
    ks = (np.arange(R) + 0.5) / R
    # above the interior: |z - 0.5|
    iy, ix = 6, 6  # (0.40625, 0.40625) inside the triangle
    for iz in range(R):
        ref = abs(ks[iz] - 0.5)
        if ref <= band:
            assert d[iz, iy, ix] == pytest.approx(ref, abs=1e-6)
        else:
            assert np.isinf(d[iz, iy, ix])
    # beyond the hypotenuse x + y = 1: distance to that edge in the plane
    iz, iy, ix = 7, 8, 8  # (0.53125, 0.53125, 0.46875): past the hypotenuse, within the band
    p = np.array([ks[ix], ks[iy], ks[iz]])
    a, b = t[0, 1].astype(np.float64), t[0, 2].astype(np.float64)
    s = np.clip(np.dot(p - a, b - a) / np.dot(b - a, b - a), 0, 1)
    assert d[iz, iy, ix] == pytest.approx(np.linalg.norm(p - (a + s * (b - a))), abs=1e-6)
    # nothing farther than the band
    assert np.isinf(d[0]).all()


def _box_tris(lo, hi):
    import itertools
    V = np.array(list(itertools.product(*zip(lo, hi))), np.float32)  # corner c = (x bit 2, y bit 1, z bit 0)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    T = []
    for a, b, c, d in quads:
        T += [[V[a], V[b], V[c]], [V[a], V[c], V[d]]]
    return np.array(T, np.float32)


def test_flood_fill_separates_an_enclosed_interior_only():
    R = 20
    band = np.float32(3.0 / R)
    box = _box_tris((0.3, 0.3, 0.3), (0.7, 0.7, 0.7))
    d = s1.udf(box, R, band)
    out = s1.outside(d, R)
    c = R // 2
    assert not out[c, c, c]  # enclosed
    assert out[0, 0, 0] and out[1, c, c]
    sheet = box[:2]  # one face: an open sheet
    d2 = s1.udf(sheet, R, band)
    out2 = s1.outside(d2, R)
    assert out2[c, c, c]
    wall = np.float32(0.55) / np.float32(R)
    assert (out2 == (d2 > wall)).all()  # every passable sample is reachable


def test_whole_oracle_on_a_small_sphere_gives_a_closed_offset_surface():
    from pamo_amd import synth
    V, F = synth.icosphere_sh(2, 0)
    out = s1.remesh(V.astype(np.float32), F.astype(np.int64), 24)
    _check_closed_manifold(out["faces"], len(out["verts"]))
    # the offset surface lies about 0.9 cells outside the input surface
    R = 24
    cell = float(out["tmax"]) * (3.0 / R * 2 + 1) / R
    r_in = np.linalg.norm(V - V[F].reshape(-1, 3).mean(0), axis=1).mean()
    r_out = np.linalg.norm(out["verts"], axis=1)
    assert np.abs(r_out.mean() - (r_in + 0.9 * cell)) < 0.6 * cell

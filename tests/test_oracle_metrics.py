"""The metrics oracle (oracle/point_mesh.cpp) pinned by closed forms and by
an independent construction, and the host-side metric functions of
pamo_amd.metrics (sampling, angles).  No GPU call is made here.

The reference computes these with libigl 2.5.1 and trimesh 4.4.0, neither
installed here: the distances are pinned by the mathematics below; the
surface sampling restates trimesh's published algorithm and is checked for
its properties only (its sample positions are not compared with trimesh's
own output).
"""
import numpy as np
import pytest

from oracle import metrics as om
from pamo_amd import metrics


def _unit_cube():
    V = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float64)
    q = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    F = []
    for a, b, c, d in q:
        F += [(a, b, c), (a, c, d)]
    return V, np.array(F, np.int32)


def _grid(n=8, z=0.0):
    xs = np.linspace(0, 1, n + 1)
    V = np.array([[x, y, z] for y in xs for x in xs], np.float64)
    F = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            F += [(a, a + 1, a + n + 2), (a, a + n + 2, a + n + 1)]
    return V, np.array(F, np.int32)


T0 = (np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0]))


@pytest.mark.parametrize("p,want", [
    ((0.25, 0.25, 0.7), 0.49),     # interior: height squared
    ((-1.0, -1.0, 0.0), 2.0),      # vertex region a
    ((2.0, 0.0, 0.0), 1.0),        # vertex region b
    ((0.0, 3.0, 0.0), 4.0),        # vertex region c
    ((0.5, -1.0, 0.0), 1.0),       # edge ab
    ((1.0, 1.0, 0.0), 0.5),        # edge bc
    ((-1.0, 0.5, 3.0), 10.0),      # edge ca
    ((0.2, 0.3, 0.0), 0.0),        # on the face
])
def test_point_tri_closed_forms(p, want):
    for alt in (False, True):
        assert om.point_tri_sqdist(p, *T0, alt=alt) == pytest.approx(want, rel=1e-14, abs=1e-30)


def test_point_tri_matches_independent_construction():
    rng = np.random.default_rng(3)
    worst = 0.0
    for _ in range(20000):
        a, b, c = rng.normal(size=(3, 3))
        # points spread over every Voronoi region, some on the plane
        p = a + rng.normal(size=3) * rng.choice([0.1, 1.0, 3.0])
        if rng.random() < 0.1:
            n = np.cross(b - a, c - a)
            p = p - np.dot(p - a, n) / np.dot(n, n) * n
        d1, d2 = om.point_tri_sqdist(p, a, b, c), om.point_tri_sqdist(p, a, b, c, alt=True)
        scale = max(d2, 1e-12 * (np.abs(np.concatenate([a, b, c, p])).max() ** 2))
        worst = max(worst, abs(d1 - d2) / scale)
    assert worst < 1e-9


def test_point_mesh_brute_force_cube():
    V, F = _unit_cube()
    P = np.array([[2, .5, .5], [2, 2, 2], [.5, .5, .25], [.5, -3, .5], [.9, .9, .9]], np.float64)
    sq, idx = om.point_mesh_sqdist(P, V, F)
    assert np.allclose(sq, [1.0, 3.0, 0.0625, 9.0, 0.01], rtol=1e-14, atol=1e-15)
    # the nearest face of the inside point is on z = 0
    assert set(F[idx[2]].tolist()) <= {0, 2, 4, 6}


def test_hausdorff_closed_forms():
    V, F = _unit_cube()
    big = (V - 0.5) * 2.0 + 0.5
    assert om.hausdorff(V, F, big, F) == pytest.approx(np.sqrt(3) * 0.5, rel=1e-14)
    G, GF = _grid()
    H = G + np.array([0.0, 0.0, 0.3])
    assert om.hausdorff(G, GF, H, GF) == pytest.approx(0.3, rel=1e-14)
    assert om.hausdorff(G, GF, G, GF) == 0.0


def test_sample_surface_properties():
    V, F = _grid(4)
    V = V.copy()
    V[:, 2] = np.sin(3 * V[:, 0]) * 0.2  # uneven face areas
    s1, f1 = metrics.sample_surface(V, F, 20000, seed=42)
    s2, f2 = metrics.sample_surface(V, F, 20000, seed=42)
    assert np.array_equal(s1, s2) and np.array_equal(f1, f2)
    # every sample lies on its face
    d = [om.point_tri_sqdist(s1[i], *V[F[f1[i]]]) for i in range(0, 20000, 97)]
    assert max(d) < 1e-28
    # faces picked in proportion to their areas
    area = metrics.face_areas(V, F)
    cnt = np.bincount(f1, minlength=len(F))
    expect = 20000 * area / area.sum()
    chi2 = ((cnt - expect) ** 2 / expect).sum()
    assert chi2 < len(F) + 6 * np.sqrt(2 * len(F))


def test_face_areas_closed_form():
    V = np.array([[0, 0, 0], [3, 0, 0], [0, 4, 0], [0, 0, 2]], np.float64)
    F = np.array([[0, 1, 2], [0, 1, 3]], np.int32)
    assert np.allclose(metrics.face_areas(V, F), [6.0, 3.0], rtol=1e-15)


def test_angles_closed_forms():
    eq = np.array([[0, 0, 0], [1, 0, 0], [0.5, np.sqrt(3) / 2, 0]], np.float64)
    rt = np.array([[0, 0, 0], [4, 0, 0], [0, 3, 0]], np.float64)
    F = np.array([[0, 1, 2]], np.int32)
    assert metrics.compute_min_angle_VF(eq, F) == pytest.approx(np.pi / 3, rel=1e-12)
    assert metrics.compute_min_angle_VF(rt, F) == pytest.approx(np.arctan2(3, 4), rel=1e-12)
    ang = metrics.internal_angles(rt, F)
    assert ang.sum() == pytest.approx(np.pi, rel=1e-14)
    assert ang[0, 0] == pytest.approx(np.pi / 2, rel=1e-14)
    # degenerate (repeated corner) faces are skipped by compute_min_angle_VF
    F2 = np.array([[0, 1, 2], [0, 0, 1]], np.int32)
    assert metrics.compute_min_angle((rt, F2)) == pytest.approx(np.arctan2(3, 4), rel=1e-12)
    V = np.concatenate([eq, rt + 10, np.array([[0, 0, 5], [1, 0, 5], [0.5, 0.05, 5]])])
    FF = np.array([[0, 1, 2], [3, 4, 5], [6, 7, 8]], np.int32)
    r = metrics.compute_bad_tri_ratio(V, FF, thresholds=[1, 10, 40])
    # smallest angles: 60, 36.87, 5.71 degrees
    assert np.allclose(r, [0.0, 1 / 3, 2 / 3])

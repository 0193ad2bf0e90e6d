"""Pinning of the exact self-intersection checker (oracle/exact_check.cpp),
which the GPU tests use to state that outputs are intersection-free.

The predicates are checked against exact rational arithmetic (fractions), the
triangle tests against hand cases with known answers and against the
reference-mode predicate away from degeneracy, the mesh check against meshes
that are intersection-free by construction and against planted crossings.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import exact, oracle
from pamo_amd import synth


def _o3_frac(a, b, c, d):
    r = [[Fr(float(p[i])) - Fr(float(d[i])) for i in range(3)] for p in (a, b, c)]
    det = (r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) - r[0][1] * (r[1][0] * r[2][2] - r[1][2] * r[2][0]) +
           r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]))
    return (det > 0) - (det < 0)


def _o2_frac(a, b, c):
    det = ((Fr(float(a[0])) - Fr(float(c[0]))) * (Fr(float(b[1])) - Fr(float(c[1]))) -
           (Fr(float(a[1])) - Fr(float(c[1]))) * (Fr(float(b[0])) - Fr(float(c[0]))))
    return (det > 0) - (det < 0)


def test_orient3d_matches_rational_arithmetic():
    rng = np.random.default_rng(1)
    zeros = 0
    for k in range(1500):
        a, b, c = rng.standard_normal((3, 3))
        mode = k % 3
        if mode == 0:
            d = rng.standard_normal(3)
        elif mode == 1:  # exactly coplanar: small integers on the plane z = 2x - 3y + 1, scaled by 2^-7
            P = rng.integers(-40, 40, (4, 2)).astype(np.float64)
            pts = np.c_[P, 2 * P[:, 0] - 3 * P[:, 1] + 1] * 2.0 ** -7
            a, b, c, d = pts
        else:  # float32 near-coplanar, widely scaled coordinates (inexact differences)
            a, b, c = (rng.standard_normal((3, 3)) * np.array([1.0, 1e-9, 1e7])).astype(np.float32)
            u, v = rng.random(2)
            d = (a + np.float32(u) * (b - a) + np.float32(v) * (c - a)).astype(np.float32)
        s = exact.orient3d(a, b, c, d)
        assert s == _o3_frac(a, b, c, d)
        assert exact.orient3d(b, a, c, d) == -s  # antisymmetry
        assert exact.orient3d(b, c, a, d) == s  # even permutation
        zeros += s == 0
    assert zeros >= 400  # the coplanar branch took the exact path


def test_orient2d_matches_rational_arithmetic():
    rng = np.random.default_rng(2)
    for k in range(2000):
        if k % 2:
            P = rng.integers(-50, 50, (3, 1)).astype(np.float64)
            a, b, c = np.c_[P, 3 * P - 7] * 2.0 ** -5  # collinear
        else:
            a, b, c = (rng.standard_normal((3, 2)) * np.array([1e-8, 1e6])).astype(np.float32)
        assert exact.orient2d(a, b, c) == _o2_frac(a, b, c)


T = np.array([[0, 0, 0], [4, 0, 0], [0, 4, 0]], float)


@pytest.mark.parametrize("name,B,ib,expect", [
    ("crossing", [[1, 1, -1], [1, 1, 1], [3, -2, 0.5]], [3, 4, 5], True),
    ("above", [[1, 1, 1], [2, 1, 1], [1, 2, 2]], [3, 4, 5], False),
    ("touch at a corner", [[1, 1, 0], [1, 1, 3], [2, 2, 3]], [3, 4, 5], True),
    ("coplanar overlap", [[1, 1, 0], [5, 1, 0], [1, 5, 0]], [3, 4, 5], True),
    ("coplanar apart", [[5, 5, 0], [6, 5, 0], [5, 6, 0]], [3, 4, 5], False),
    ("coplanar edge touch", [[2, 2, 0], [5, 2, 0], [5, 5, 0]], [3, 4, 5], True),
    ("shared vertex, dihedral", [[0, 0, 0], [-1, 2, 3], [-2, -1, 3]], [0, 4, 5], False),
    ("shared vertex, opposite edge pierces", [[0, 0, 0], [1, 1, -1], [1, 1, 1]], [0, 4, 5], True),
    ("shared vertex, coplanar edge inside", [[0, 0, 0], [1, 1, 0], [-1, -1, 3]], [0, 4, 5], True),
    ("shared vertex, coplanar edge outside", [[0, 0, 0], [-1, -1, 0], [-1, -2, 3]], [0, 4, 5], False),
    ("shared vertex, collinear edge overlap", [[0, 0, 0], [2, 0, 0], [1, -1, 2]], [0, 4, 5], True),
    ("shared vertex, collinear edge away", [[0, 0, 0], [-2, 0, 0], [-1, -1, 2]], [0, 4, 5], False),
    ("shared edge, dihedral", [[4, 0, 0], [0, 0, 0], [2, -1, 1]], [1, 0, 5], False),
    ("shared edge, flat on the other side", [[4, 0, 0], [0, 0, 0], [2, -3, 0]], [1, 0, 5], False),
    ("shared edge, folded flat onto", [[4, 0, 0], [0, 0, 0], [1, 1, 0]], [1, 0, 5], True),
    ("repeated face", [[0, 0, 0], [4, 0, 0], [0, 4, 0]], [0, 1, 2], True),
])
def test_faces_intersect_hand_cases(name, B, ib, expect):
    assert exact.faces_intersect(T, [0, 1, 2], np.array(B, float), ib) == expect, name
    assert exact.faces_intersect(np.array(B, float), ib, T, [0, 1, 2]) == expect, name  # symmetric


def test_random_pairs_agree_with_reference_predicate():
    """Away from degeneracy (no shared vertices, random float32 corners), the
    exact answer equals the reference-mode predicate's (with its 1e-3 shared
    vertex rule not in play)."""
    rng = np.random.default_rng(3)
    hits = 0
    for _ in range(3000):
        A = rng.standard_normal((3, 3)).astype(np.float32)
        B = (rng.standard_normal((3, 3)) * 0.7 + rng.standard_normal(3) * 0.5).astype(np.float32)
        e = exact.faces_intersect(A, [0, 1, 2], B, [3, 4, 5])
        r = oracle.ref_guard(A, [0, 1, 2], B, [3, 4, 5], 1e-3)
        assert e == r
        hits += e
    assert 300 < hits < 2700


def test_meshes():
    V, F = synth.icosphere_sh(4, 0)
    pairs, cand, degen = exact.self_intersections(V, F)
    assert len(pairs) == 0 and degen == 0 and cand > len(F)
    # the candidate set equals the oracle LBVH's (an independent broad phase)
    assert cand == len(oracle.candidates(V, F))
    V2, F2 = synth.offset_sheets(V, F, rel_offset=1e-3)  # disjoint sheets
    assert len(exact.self_intersections(V2, F2)[0]) == 0
    # planted: one vertex pushed through to the far side of the sphere
    V3 = V.copy()
    V3[0] = -1.5 * V3[0]
    pairs, _, _ = exact.self_intersections(V3, F)
    assert len(pairs) > 0
    assert np.any(np.isin(F[pairs[:, 0]], 0) | np.isin(F[pairs[:, 1]], 0))
    # two copies, one shifted to overlap the other
    F4 = np.r_[F, F + len(V)]
    V4 = np.r_[V, V + np.float32(0.5)]
    assert len(exact.self_intersections(V4, F4)[0]) > 0
    # a degenerate face is counted, not tested
    F5 = np.r_[F, [[0, 0, 1]]]
    assert exact.self_intersections(V, F5)[2] == 1

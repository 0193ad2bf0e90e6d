"""Pins of the handoff restatement (oracle/handoff.py): trimesh's
process=True vertex merge at the stage-2 -> stage-3 handoff.

trimesh is absent, so the restatement is pinned by hand cases whose answer
follows from the published rule (keys = coordinates times 1e8 rounded half
to even; first vertex in index order kept; unreferenced vertices dropped)
and chosen so that rounding half away from zero, truncation, keeping the
last vertex of a group or keeping unreferenced vertices each fail, and by an
independent dictionary brute force on random inputs with planted
duplicates and rounding ties.
"""
import numpy as np
import pytest

from oracle import handoff as h


def _tie_case():
    a = np.float32(1.0 / 512)  # a * 1e8 = 195312.5 exactly: a rounding tie
    lo = np.nextafter(a, np.float32(-1))  # 195312.488...
    hi = np.nextafter(a, np.float32(1))  # 195312.523...
    V = np.array([[a, 0.25, 0.5],
                  [lo, 0.25, 0.5],
                  [hi, 0.25, 0.5],
                  [-a, 0.25, 0.5],
                  [0.7, 0.1, 0.3],
                  [9.0, 9.0, 9.0],  # unreferenced
                  [0.7, 0.1, 0.3],
                  [-np.nextafter(a, np.float32(-1)), 0.25, 0.5]], np.float32)
    F = np.array([[6, 1, 2], [4, 0, 7], [3, 6, 7]], np.int64)
    return V, F


def test_hand_case_ties_round_half_to_even_and_first_vertex_kept():
    V, F = _tie_case()
    Vm, Fm = h.merge_vertices(V, F)
    # keys of x: 195312, 195312, 195313, -195312, ..., -195312
    # groups in index order: {0, 1}, {2}, {3, 7}, {4, 6}; 5 dropped
    assert np.array_equal(Vm, V[[0, 2, 3, 4]].astype(np.float64))
    assert np.array_equal(Fm, np.array([[3, 0, 1], [3, 0, 2], [2, 3, 2]]))
    # half away from zero would merge 0 with 2 and keep 1 apart; truncation
    # or the floor would merge 0, 1 and 2
    keys_away = np.sign(V[:3, 0].astype(np.float64) * 1e8) * np.floor(np.abs(V[:3, 0] * 1e8) + 0.5)
    assert keys_away[0] == keys_away[2] != keys_away[1]


def test_without_faces_every_vertex_is_referenced():
    V, _ = _tie_case()
    Vm, Fm = h.merge_vertices(V, np.zeros((0, 3), np.int64))
    assert len(Fm) == 0
    assert np.array_equal(Vm, V[[0, 2, 3, 4, 5]].astype(np.float64))


def test_no_duplicates_is_the_identity_on_referenced_vertices():
    rng = np.random.default_rng(0)
    V = rng.uniform(-1, 1, (50, 3)).astype(np.float32)
    F = rng.integers(0, 50, (80, 3))
    Vm, Fm = h.merge_vertices(V, F)
    used = np.unique(F)
    assert np.array_equal(Vm, V[used].astype(np.float64))
    remap = np.full(50, -1)
    remap[used] = np.arange(len(used))
    assert np.array_equal(Fm, remap[F])


def planted(n, n_f, seed):
    """Random vertices with exact duplicates, rounding ties and their float32
    neighbours planted; random faces (a few vertices unreferenced)."""
    rng = np.random.default_rng(seed)
    V = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    k = n // 10
    src = rng.integers(0, n, k)
    dst = rng.integers(0, n, k)
    V[dst] = V[src]
    ties = rng.integers(0, n, k)
    V[ties] = (2 * rng.integers(-200, 200, (k, 3)) + 1).astype(np.float32) / np.float32(512)
    nb = rng.integers(0, n, k)
    step = np.where(rng.random((k, 3)) < 0.5, np.float32(-2), np.float32(2))
    V[nb] = np.nextafter(V[ties], step)
    F = rng.integers(0, n, (n_f, 3))
    return V, F


@pytest.mark.parametrize("seed", range(4))
def test_restatement_equals_brute_force(seed):
    V, F = planted(3000, 2500, seed)
    Vm, Fm = h.merge_vertices(V, F)
    Vb, Fb = h.merge_vertices_brute(V, F)
    assert len(Vm) < len(np.unique(F))  # something merged
    assert np.array_equal(Vm, Vb)
    assert np.array_equal(Fm, Fb)


def test_non_finite_vertex_is_an_error():
    V, F = _tie_case()
    V[4, 1] = np.nan
    with pytest.raises(ValueError):
        h.merge_vertices(V, F)

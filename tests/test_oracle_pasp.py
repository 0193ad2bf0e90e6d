"""The legacy simplifier behind PaSP on the oracle (pasp_forward /
pasp_run of oracle/pamo_oracle.cpp, restating simp_cuda/src/cusimp.cu),
pinned by a closed form and by brute force.  CPU only."""
import numpy as np
import pytest

from oracle import oracle
from pamo_amd import synth


def _octahedron(s=1.0):
    V = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32) * s
    F = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]], np.int32)
    return V, F


def _collapse_t(tau):
    q = np.float32(min(max(tau, 0.0), 10.0)) / np.float32(10.0) * np.float32(4294967296.0)
    return 0xFFFFFFFF if q >= 4294967296.0 else int(q)


@pytest.mark.parametrize("tau", [1e-3, 0.05, 0.2])
def test_octahedron_cost_closed_form(tau):
    """Unit octahedron, scale 2: every edge's midpoint lies on two of the four
    face planes around each end and 1/sqrt(3) from the other two, so the
    quadric term is (2/3 + 2/3) / 2^2 = 1/3; every edge has length sqrt(2),
    so the edge-length term, averaged and counted twice
    (cusimp.cu:374-378), is 2 * sqrt(2)/2 * tau + sqrt(2)/2 * tau; the
    skinny term vanishes (each ratio exceeds 1 at this scale)."""
    V, F = _octahedron()
    r = oracle.Round()
    oracle.pasp_forward(V, F, 2.0, tau, False, ctx=r)
    cost = r.edge_cost().astype(np.float64) / 4294967296.0 * 10.0
    want = 1.0 / 3.0 + 3.0 * np.sqrt(2.0) / 2.0 * tau
    assert len(cost) == 12
    assert np.all(np.abs(cost - want) <= 1e-6 * want)


def _brute_round(V, F, tau, cost, edges):
    """One legacy round from the oracle's edge costs, by brute force: the
    per-face minimum key over all edges touching a vertex of the face, the
    independent set, the midpoint collapses applied in edge order."""
    nV = len(V)
    key = (cost.astype(np.uint64) << np.uint64(32)) | np.arange(len(cost), dtype=np.uint64)
    vmin = np.full(nV, np.iinfo(np.uint64).max, np.uint64)
    for e, (a, b) in enumerate(edges):
        vmin[a] = min(vmin[a], key[e])
        vmin[b] = min(vmin[b], key[e])
    tri_min = vmin[F].min(axis=1)
    ring = [[] for _ in range(nV)]
    for f, t in enumerate(F):
        for x in t:
            ring[x].append(f)
    ct = _collapse_t(tau)
    P = V.copy()
    T = F.copy()
    occ = np.ones(nV, np.int32)
    for e, (u, v) in enumerate(edges):
        if int(cost[e]) > ct:
            continue
        if not all(tri_min[f] == key[e] for f in ring[u] + ring[v]):
            continue
        P[u] = (P[u] + P[v]) / np.float32(2.0)
        P[v] = 0
        occ[v] = 0
        for f in ring[u] + ring[v]:
            if u in T[f] and v in T[f]:
                T[f] = -1
            elif v in T[f]:
                T[f][T[f] == v] = u
    return P, T, occ


@pytest.mark.parametrize("seed,tau", [(0, 0.05), (1, 0.3)])
def test_round_matches_brute_force(seed, tau):
    V, F = synth.icosphere_sh(3, seed)
    V = V.astype(np.float32)
    s = float((V.max(0) - V.min(0)).max())
    r = oracle.Round()
    op, of, oo, om = oracle.pasp_forward(V, F, s, tau, False, ctx=r)
    P, T, occ = _brute_round(V, F.astype(np.int32), tau, r.edge_cost(), r.edges())
    assert occ.sum() < len(V)  # something collapsed
    assert np.array_equal(oo, occ)
    assert np.array_equal(of, T)
    assert np.array_equal(op.view(np.uint32), P.view(np.uint32))
    assert np.array_equal(om, np.concatenate([[0], np.cumsum(occ)[:-1]]))


def test_run_equals_rounds_with_compaction():
    """pasp_run = rounds (shuffle on the first) with compaction until one
    removes no face (pamo/__init__.py:224-253)."""
    V, F = synth.icosphere_sh(3, 2)
    V = V.astype(np.float32)
    s = float((V.max(0) - V.min(0)).max())
    vo, fo, rounds = oracle.pasp_run(V, F, 0.05, 1000)
    r = oracle.Round()
    v, f = V, F.astype(np.int32)
    for it in range(1000):
        n = len(f)
        p, t, occ, mp = oracle.pasp_forward(v, f, s, 0.05, it == 0, ctx=r)
        v = p[occ.astype(bool)]
        t = t[t[:, 0] >= 0]
        f = mp[t]
        if len(f) == n:
            break
    assert rounds == it + 1 and rounds > 2
    assert np.array_equal(fo, f) and np.array_equal(vo.view(np.uint32), v.view(np.uint32))

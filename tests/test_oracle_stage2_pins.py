"""More pins of the stage-2 oracle (oracle/pamo_oracle.cpp): the free edge
cost (cusimp_free.cu:446-584) against closed forms and an independent
float64 statement, the edge list rule (:306-335) against the mesh's
half-edge structure, a scripted line collapse (:234-274), and the undo
machinery (:776-878, :895-916, :1015-1040) on scripted crossings: the
5-iteration cap, the stale pts_map on the first and later calls, the stuck
prologue and the undo-candidate grid of the regular-collapse count.

Expected values come from geometry, topology and the reference's loop
structure, not from the oracle's code.
"""
import math

import numpy as np
import pytest

MAXC = 0xFFFFFFFF


def _decode(c):
    """The cost value a quantised u32 stands for (cusimp_free.cu:584)."""
    return float(c) / 4294967296.0 * 10.0


def _round(oracle_lib, V, F, thr, scale=None, ctx=None, undo=None, is_stuck=0):
    r = ctx if ctx is not None else oracle_lib.Round()
    if scale is None:
        scale = float((V.max(0) - V.min(0)).max())
    u = np.zeros(0, np.int32) if undo is None else np.asarray(undo, np.int32)
    out = r.forward(V, F, u, len(u), scale, thr, is_stuck, 0)
    return r, out


def _compact(out):
    pv, pt, occ, mp, undo = out
    occ = occ.reshape(-1).astype(bool)
    mp = mp.reshape(-1)
    F = pt[pt[:, 0] >= 0]
    return pv[occ], mp[F].astype(np.int32), mp, undo


# ----------------------------------------------------------------- A7 free edge cost
def _octahedron(r=1.0, L=1.0):
    V = np.array([[1, 0, 0], [-1, 0, 0], [0, L, 0], [0, -L, 0], [0, 0, 1], [0, 0, -1]], np.float64) * r
    F = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]], np.int32)
    return V.astype(np.float32), F


def _free_cost64(V, F, u, v, s, tau, precedence="reference", penalty=5.0, ring_average=True):
    """compute_edge_cost_kernel (cusimp_free.cu:382-585) stated in float64
    for an edge with a valid link.  The quadric term is the sum of squared
    distances of the midpoint to the planes of both rings (the closed form of
    v^T (Q_u + Q_v) v), the skinny term follows the expression as written
    (4 sqrt3 A / l01^2 + l12^2 + l20^2) or the textbook ratio
    4 sqrt3 A / (l01^2 + l12^2 + l20^2)."""
    X = V.astype(np.float64)
    ring = {w: [f for f in range(len(F)) if w in F[f]] for w in (u, v)}
    m = (X[u] + X[v]) / 2
    q = 0.0
    for w in (u, v):
        for f in ring[w]:
            a, b, c = X[F[f]]
            n = np.cross(b - a, c - a)
            n /= np.linalg.norm(n)
            q += float(np.dot(n, m - a)) ** 2
    L, ne = 0.0, 0
    for w in (u, v):
        for f in ring[w]:
            i, j, k = F[f]
            for a, b in ((i, j), (j, k), (k, i)):
                if a > b:
                    L += np.linalg.norm(X[a] - X[b])
                    ne += 1
    elen = (L / ne / s * tau if ring_average else 0.0) + np.linalg.norm(X[u] - X[v]) / s * tau
    qa, nt = 0.0, 0
    for w, other in ((u, v), (v, u)):
        for f in ring[w]:
            if other in F[f]:
                continue
            P = X[F[f]].copy()
            P[F[f] == w] = m
            A = 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
            l01, l12, l20 = (np.linalg.norm(P[0] - P[1]), np.linalg.norm(P[1] - P[2]),
                             np.linalg.norm(P[2] - P[0]))
            if precedence == "reference":
                x = 4 * math.sqrt(3) * A / l01 ** 2 + l12 ** 2 + l20 ** 2 + 1e-7
            else:
                x = 4 * math.sqrt(3) * A / (l01 ** 2 + l12 ** 2 + l20 ** 2) + 1e-7
            qa += 1.0 - min(max(x, 0.0), 1.0)
            nt += 1
    total = q / s ** 2 + elen + penalty * qa / nt * tau
    return min(max(total, 0.0), 10.0)


@pytest.mark.parametrize("r", [1.0, 0.01])
@pytest.mark.parametrize("tau", [1e-3, 0.5])
def test_free_cost_octahedron_closed_form(oracle_lib, r, tau):
    """Regular octahedron of radius r, edge (u, v) to its midpoint m:
    - quadric: each endpoint's 4 planes; m lies on the 2 planes holding the
      edge and at distance r/sqrt3 from the other 2 of each ring, so
      v^T Q v = 4 r^2 / 3 and, over s^2 = (2r)^2, 1/3;
    - ring edges all have length r sqrt2: average sqrt2 r / s * tau plus the
      edge itself sqrt2 r / s * tau = sqrt2 tau;
    - the four moved faces are far from skinny: the expression as written
      clamps to 1 and adds nothing (the textbook ratio would add
      5 (1 - 0.9575) tau).
    So every edge costs 1/3 + sqrt2 tau."""
    V, F = _octahedron(r)
    rr, _ = _round(oracle_lib, V, F, tau, scale=2.0 * r)
    C = rr.edge_cost()
    assert len(C) == 12
    want = 1.0 / 3.0 + math.sqrt(2.0) * tau
    for c in C.tolist():
        assert _decode(c) == pytest.approx(want, rel=3e-6)
    # the alternatives this rules out differ by far more than the tolerance
    E = rr.edges()
    u, v = E[0]
    alt = _free_cost64(V, F, u, v, 2.0 * r, tau, precedence="textbook")
    assert abs(alt - want) > 1e-3 * tau
    assert _free_cost64(V, F, u, v, 2.0 * r, tau) == pytest.approx(want, rel=1e-6)


def test_free_cost_skinny_term_precedence_and_penalty(oracle_lib):
    """A small (r = 1e-3) octahedron stretched 10x along y: collapsing a
    +x/+y edge leaves long thin faces whose skinny term is nonzero.  The
    oracle's cost must equal the expression as written (not the textbook
    ratio), with the x5 penalty and the ring edge-length average."""
    V, F = _octahedron(1e-3, 10.0)
    s = float((V.max(0) - V.min(0)).max())
    tau = 0.25
    rr, _ = _round(oracle_lib, V, F, tau, scale=s)
    E, C = rr.edges(), rr.edge_cost()
    checked = 0
    for (u, v), c in zip(E.tolist(), C.tolist()):
        if {u, v} != {0, 2}:
            continue
        want = _free_cost64(V, F, u, v, s, tau)
        got = _decode(c)
        assert got == pytest.approx(want, rel=3e-6)
        for kw in ({"precedence": "textbook"}, {"penalty": 1.0}, {"ring_average": False}):
            assert abs(_free_cost64(V, F, u, v, s, tau, **kw) - want) > 1e-3, kw
        checked += 1
    assert checked == 1
    # the skinny term is really active here
    assert _free_cost64(V, F, 0, 2, s, tau) - _free_cost64(V, F, 0, 2, s, tau, penalty=0.0) > 0.1


def test_free_cost_invalid_endpoints_and_stuck_round(oracle_lib):
    """Both endpoints in the invalid table -> MAX (:395-400); a stuck round
    leaves every other cost as the previous round wrote it (:402-403)."""
    V, F = _octahedron(1.0)
    r, out = _round(oracle_lib, V, F, 1e-3, scale=2.0)
    C0 = r.edge_cost().copy()
    E0 = r.edges().copy()
    # same mesh again, stuck, with vertices 0 and 2 flagged invalid
    r.forward(V, F, np.array([0, 2], np.int32), 2, 2.0, 1e-3, 1, 0)
    E1, C1 = r.edges(), r.edge_cost()
    assert np.array_equal(E1, E0)
    for e, (u, v) in enumerate(E1.tolist()):
        if {u, v} == {0, 2}:
            assert C1[e] == MAXC
        else:
            assert C1[e] == C0[e]


# ----------------------------------------------------------------- A5 edge list
def _half_edges(F):
    h = []
    for f, (i, j, k) in enumerate(F.tolist()):
        if i < 0:
            continue
        h += [(f, 0, i, j), (f, 1, j, k), (f, 2, k, i)]
    return h


@pytest.mark.parametrize("open_mesh", [False, True])
def test_edge_list_follows_half_edges(oracle_lib, open_mesh):
    """Each edge comes from the half-edge whose first vertex is the larger
    (stored smaller first), in face order then i->j, j->k, k->i
    (:306-335).  On a closed oriented mesh every undirected edge appears
    exactly once; on an open one a boundary edge whose only half-edge runs
    from the smaller vertex is dropped.  Removed rows (all -1) add none."""
    from pamo_amd import synth
    V, F = synth.icosphere(2)
    V, F = V.astype(np.float32), F.astype(np.int32)
    if open_mesh:
        F = np.delete(F, [5, 6, 50, 51, 52], axis=0)
        F[[10, 70]] = -1
    r, _ = _round(oracle_lib, V, F, 1e-3)
    E = [tuple(e) for e in r.edges().tolist()]
    H = _half_edges(F)
    directed = {(a, b) for _, _, a, b in H}
    undirected = {(min(a, b), max(a, b)) for _, _, a, b in H}
    # order: by face, then local position, over half-edges a -> b with a > b
    assert E == [(b, a) for _, _, a, b in H if a > b]
    assert len(set(E)) == len(E)
    if not open_mesh:
        assert set(E) == undirected and len(E) == 3 * len(F) // 2
    else:
        dropped = {(a, b) for (a, b) in undirected if (b, a) not in directed}  # only a -> b with a < b
        assert dropped and set(E) == undirected - dropped


# ----------------------------------------------------------------- A10 line collapse
def _bipyramid_coincident():
    ang = np.array([0, 2 * np.pi / 3, 4 * np.pi / 3])
    V = np.c_[np.cos(ang), np.sin(ang), np.zeros(3)]
    V = np.r_[V, [[0, 0, 1.0], [0, 0, -1.0]]].astype(np.float32)
    V[1] = V[0]  # vertices 0 and 1 coincide exactly
    F = np.array([[0, 1, 3], [1, 2, 3], [2, 0, 3], [1, 0, 4], [2, 1, 4], [0, 2, 4]], np.int32)
    return V, F


def test_line_collapse_scripted(oracle_lib):
    """Edge (0, 1) has coincident endpoints and fails the link condition
    (three shared neighbours -> MAX), so no regular collapse takes it; with
    a threshold that admits no regular collapse, the line-collapse pass
    (:234-274) removes vertex 1 to the origin, deletes the two faces holding
    both ends and renames 1 -> 0 in the others."""
    V, F = _bipyramid_coincident()
    r, (pv, pt, occ, mp, undo) = _round(oracle_lib, V, F, 1e-9)
    rec, nreg = r.collapsed()
    E = r.edges().tolist()
    assert nreg == 0
    assert [E[e] for e in rec.tolist()] == [[0, 1]]
    want_pt = np.array([[-1, -1, -1], [0, 2, 3], [2, 0, 3], [-1, -1, -1], [2, 0, 4], [0, 2, 4]], np.int32)
    assert np.array_equal(pt, want_pt)
    assert occ.reshape(-1).tolist() == [1, 0, 1, 1, 1]
    assert np.array_equal(pv[1], [0, 0, 0])
    keep = [0, 2, 3, 4]
    assert np.array_equal(pv[keep].view(np.uint32), V[keep].view(np.uint32))
    assert mp.reshape(-1).tolist() == [0, 1, 1, 2, 3]
    assert len(undo) == 0


def test_line_collapse_outside_the_candidate_grid_is_never_undone(oracle_lib):
    """get_undo_candidate_kernel runs over the regular-collapse count rounded
    up to blocks of 512 (:1004,1016): with no regular collapse it has no
    threads, so a line collapse whose faces cross another surface stays,
    although the guard would flag them."""
    V, F = _bipyramid_coincident()
    # a small tetrahedron piercing face (2, 0, 3) of the bipyramid
    c = (V[2] + V[0] + V[3]) / 3
    n = np.cross(V[0] - V[2], V[3] - V[2])
    n = n / np.linalg.norm(n)
    T = np.array([c + 0.3 * n, c - 0.3 * n + [0.05, 0, 0], c - 0.3 * n + [-0.03, 0.04, 0],
                  c - 0.3 * n + [-0.03, -0.04, 0.02]], np.float32)
    V2 = np.r_[V, T].astype(np.float32)
    F2 = np.r_[F, np.array([[5, 6, 7], [5, 7, 8], [5, 8, 6], [6, 8, 7]], np.int32)]
    assert oracle_lib.ref_guard(V2[[2, 0, 3]].reshape(-1), [2, 0, 3], V2[[5, 6, 7]].reshape(-1), [5, 6, 7]) or \
        oracle_lib.ref_guard(V2[[2, 0, 3]].reshape(-1), [2, 0, 3], V2[[5, 8, 6]].reshape(-1), [5, 8, 6])
    r, (pv, pt, occ, mp, undo) = _round(oracle_lib, V2, F2, 1e-9)
    rec, nreg = r.collapsed()
    assert nreg == 0 and len(rec) == 1
    assert occ.reshape(-1)[1] == 0  # kept
    assert len(r.undo_iters()) == 0 and len(undo) == 0


# ----------------------------------------------------------------- A14 undo machinery
def _interpenetrating_spheres():
    from pamo_amd import synth
    V, F = synth.icosphere(3)
    V, F = V.astype(np.float32), F.astype(np.int32)
    V2 = V + np.array([1.2, 0, 0], np.float32)
    return np.r_[V, V2], np.r_[F, F + len(V)]


def test_undo_cap_counts_the_last_batch_and_first_call_map_is_zero(oracle_lib):
    """Two unit spheres that already cross: collapses next to the crossing
    are flagged again after every undo (the original faces still cross), so
    the loop stops at its cap of 5 undo iterations (:1039).  The batch found
    after the 5th iteration is counted in the list length but neither undone
    nor written (:1037-1039).  On a context's first call pts_map has never
    been written (it is computed at the end of the call, :1042-1045), so
    every written entry maps to 0."""
    V, F = _interpenetrating_spheres()
    r, (pv, pt, occ, mp, undo) = _round(oracle_lib, V, F, 1e-3)
    it = r.undo_iters()
    assert len(it) == 5 and it.min() > 0
    assert len(undo) > 2 * it.sum()  # the uncounted-but-unwritten last batch
    assert len(undo) % 2 == 0
    assert np.all(undo == 0)


def test_undo_list_uses_the_previous_calls_map_and_stuck_prologue(oracle_lib):
    """A later call maps the undone endpoints (its own numbering) through
    the pts_map of the call before (:869-878, recomputed only at :1042); a
    stuck call starts its list with the incoming entries mapped the same way
    (:898-903)."""
    V, F = _interpenetrating_spheres()
    r = oracle_lib.Round()
    _, out1 = _round(oracle_lib, V, F, 1e-3, ctx=r)
    V1, F1, map1, undo1 = _compact(out1)
    _, out2 = _round(oracle_lib, V1, F1, 1e-3, ctx=r, undo=undo1, scale=float((V.max(0) - V.min(0)).max()))
    rec2, _ = r.collapsed()
    E2 = r.edges()
    it2 = r.undo_iters()
    assert len(it2) > 0
    written = out2[4][: 2 * int(it2.sum())].reshape(-1, 2)
    stale = {(int(map1[E2[e, 0]]), int(map1[E2[e, 1]])) for e in rec2.tolist()}
    fresh = {(int(E2[e, 0]), int(E2[e, 1])) for e in rec2.tolist()}
    assert all((int(a), int(b)) in stale for a, b in written)
    assert not all((int(a), int(b)) in fresh for a, b in written)  # the current numbering would differ
    # stuck call: the incoming list mapped through the previous call's map
    V2, F2, map2, undo2 = _compact(out2)
    _, out3 = _round(oracle_lib, V2, F2, 1e-3, ctx=r, undo=undo2, is_stuck=1,
                     scale=float((V.max(0) - V.min(0)).max()))
    assert np.array_equal(out3[4][: len(undo2)], map2[undo2])


# ----------------------------------------------------------------- opt-in optimal placement
_IU = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]


def _quadric(planes, weights=None):
    """The 10 stored entries of sum w p p^T over planes p = (n, d), n.x + d = 0."""
    M = np.zeros((4, 4))
    for k, (n, d) in enumerate(planes):
        n = np.asarray(n, float) / np.linalg.norm(n)
        p = np.r_[n, d]
        M += (1.0 if weights is None else weights[k]) * np.outer(p, p)
    return np.array([M[i, j] for i, j in _IU], np.float32)


def test_optimal_placement_closed_forms(oracle_lib):
    """Three planes meeting in a point give that point; two planes (a line)
    or none leave the system singular and keep the midpoint; a minimiser
    farther from the midpoint than the edge is long keeps the midpoint."""
    place = oracle_lib.place_optimal
    p0, p1 = np.array([0.7, 1.6, 2.9], np.float32), np.array([1.4, 2.3, 3.2], np.float32)
    mid = (p0 + p1) / np.float32(2)
    # x = 1, y = 2, z = 3
    M = _quadric([((1, 0, 0), -1.0), ((0, 1, 0), -2.0), ((0, 0, 1), -3.0)])
    assert np.allclose(place(M, p0, p1), [1.0, 2.0, 3.0], atol=1e-6)
    # three oblique planes through (1.2, 1.9, 3.0), with weights
    q = np.array([1.2, 1.9, 3.0])
    ns = [np.array([1.0, 0.3, -0.2]), np.array([-0.4, 1.0, 0.5]), np.array([0.2, -0.6, 1.0]), np.array([1, 1, 1.0])]
    planes = [(n, -float(np.dot(n / np.linalg.norm(n), q))) for n in ns]
    got = place(_quadric(planes, [1.0, 2.0, 0.5, 3.0]), p0, p1)
    assert np.allclose(got, q, atol=2e-5)
    # a line of minimisers (two planes): midpoint, bit for bit
    M2 = _quadric([((1, 0, 0), -1.0), ((0, 1, 0), -2.0)])
    assert np.array_equal(place(M2, p0, p1).view(np.uint32), mid.view(np.uint32))
    assert np.array_equal(place(np.zeros(10, np.float32), p0, p1).view(np.uint32), mid.view(np.uint32))
    # a point 10 edge lengths away: midpoint
    far = _quadric([((1, 0, 0), -10.0), ((0, 1, 0), -2.0), ((0, 0, 1), -3.0)])
    assert np.array_equal(place(far, p0, p1).view(np.uint32), mid.view(np.uint32))


def test_optimal_placement_round_moves_to_the_minimiser(oracle_lib):
    """A round with optimal placement on the stretched octahedron: every
    collapsed vertex sits where place_optimal puts it for its edge's
    quadric (from the round's own quadrics), and the default round keeps
    the midpoint."""
    from pamo_amd import synth
    V, F = synth.icosphere(2)
    V = (V * np.array([1.0, 1.7, 0.6])).astype(np.float32)
    F = F.astype(np.int32)
    s = float((V.max(0) - V.min(0)).max())
    for mode in (0, 1):
        r = oracle_lib.Round()
        r.set_placement(mode)
        pv, pt, occ, mp, undo = r.forward(V, F, np.zeros(0, np.int32), 0, s, 0.05, 0, 0)
        rec, nreg = r.collapsed()
        E = r.edges()
        Q = r.quadrics(len(V))
        moved = 0
        for e in rec[:nreg].tolist():
            u, v = E[e]
            if occ.reshape(-1)[v] != 0:
                continue  # undone
            want = oracle_lib.place_optimal(Q[u] + Q[v], V[u], V[v]) if mode else (V[u] + V[v]) / np.float32(2)
            assert np.array_equal(pv[u].view(np.uint32), np.asarray(want, np.float32).view(np.uint32))
            moved += int(mode and not np.array_equal(want, (V[u] + V[v]) / np.float32(2)))
        assert nreg > 0
        if mode:
            assert moved > 0

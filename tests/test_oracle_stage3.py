"""Pins the stage-3 CPU oracle (oracle/stage3_oracle.cpp) against what the
mathematics fixes: closed forms, finite differences, brute force, a
library eigen-solver, and the reference's own quadratic-energy test
(simp_cuda/safe_project/tests/test_dummy_energy.py, fixture written by
tests/golden/make_dense_kat.py).  CPU only."""
import os

import numpy as np
import pytest

from oracle import stage3 as s3o
from pamo_amd import stage3
from stage3_helpers import normalised, oracle_system, scene

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fd(f, x, h=1e-6):
    g = np.zeros_like(x)
    for k in range(x.size):
        e = np.zeros_like(x)
        e.flat[k] = h
        g.flat[k] = (f(x + e) - f(x - e)) / (2 * h)
    return g


# ------------------------------------------------------------ barrier (collision_energy.py:8-28)
def test_barrier_closed_form_and_derivatives():
    dh = 1e-3
    for d in [1e-5, 1e-4, 5e-4, 9.99e-4]:
        assert s3o.barrier(d, dh) == pytest.approx(-(d - dh) ** 2 * np.log(d / dh), rel=1e-14)
        assert s3o.barrier(d, dh, 1) == pytest.approx(_fd(lambda x: s3o.barrier(x[0], dh), np.array([d]), 1e-10)[0],
                                                      rel=1e-5)
        assert s3o.barrier(d, dh, 2) == pytest.approx(
            _fd(lambda x: s3o.barrier(x[0], dh, 1), np.array([d]), 1e-10)[0], rel=1e-5)
    for order in range(3):
        assert s3o.barrier(dh, dh, order) == 0.0 and s3o.barrier(2 * dh, dh, order) == 0.0
    assert s3o.barrier(0.5 * dh, dh) > 0 and s3o.barrier(0.5 * dh, dh, 1) < 0 and s3o.barrier(0.5 * dh, dh, 2) > 0


# ------------------------------------------------------------ distances (distance_kernels.py)
def _seg_dist_brute(p, a, b, n=4001):
    t = np.linspace(0, 1, n)[:, None]
    return np.min(np.linalg.norm(a + t * (b - a) - p, axis=1))


def test_point_triangle_classified_distance_is_true_distance():
    rng = np.random.default_rng(1)
    seen = set()
    for _ in range(400):
        x = rng.normal(size=(4, 3))
        t = s3o.pt_classify(x)
        seen.add(t)
        d = s3o.pair_distance(x.ravel(), True, t)
        c = s3o.closest_on_tri(x[0], x[1:].ravel())
        assert d == pytest.approx(np.linalg.norm(c - x[0]), rel=1e-9, abs=1e-12)
    assert seen == set(range(7))


def test_closest_point_on_triangle_brute_force():
    rng = np.random.default_rng(2)
    u, v = np.meshgrid(np.linspace(0, 1, 301), np.linspace(0, 1, 301))
    m = (u + v) <= 1
    u, v = u[m], v[m]
    for _ in range(30):
        x = rng.normal(size=(4, 3))
        pts = x[1] + u[:, None] * (x[2] - x[1]) + v[:, None] * (x[3] - x[1])
        brute = np.min(np.linalg.norm(pts - x[0], axis=1))
        c = s3o.closest_on_tri(x[0], x[1:].ravel())
        dc = np.linalg.norm(c - x[0])
        assert dc <= brute + 1e-12 and dc >= brute - 2e-2
        # the closest point lies in the triangle
        n = np.cross(x[2] - x[1], x[3] - x[1])
        for a, b in [(1, 2), (2, 3), (3, 1)]:
            assert np.dot(np.cross(x[b] - x[a], c - x[a]), n) >= -1e-9 * np.dot(n, n)


def test_edge_edge_classified_distance_is_true_distance():
    rng = np.random.default_rng(3)
    seen = set()
    for _ in range(300):
        x = rng.normal(size=(4, 3))
        t = s3o.ee_classify(x, 1e-3)
        seen.add(t)
        d = s3o.pair_distance(x.ravel(), False, t)
        s = np.linspace(0, 1, 801)
        P = x[0] + s[:, None] * (x[1] - x[0])
        Q = x[2] + s[:, None] * (x[3] - x[2])
        brute = np.min(np.linalg.norm(P[:, None, :] - Q[None, :, :], axis=2))
        assert d <= brute + 1e-9 and d >= brute - 5e-3
    assert {8}.issubset(seen) and len(seen) >= 5
    # nearly parallel edges fall back from EE to point-edge / point-point
    x = np.array([[0, 0, 0], [1, 0, 0], [0.2, 1e-5, 1], [0.8, 0, 1]], float)
    assert s3o.ee_classify(x, 1e-3) != 8


def test_distance_gradients_match_finite_differences():
    rng = np.random.default_rng(4)
    for is_pt, types in [(True, range(7)), (False, range(9))]:
        for t in types:
            x = rng.normal(size=12)
            g = s3o.pair_grad(x, is_pt, t)
            fd = _fd(lambda y: s3o.pair_distance(y, is_pt, t), x)
            assert np.abs(g - fd).max() < 1e-6, (is_pt, t)


# ------------------------------------------------------------ hinge (hinge_energy.py)
def test_hinge_angle_closed_form():
    # shared edge x1-x2 on the z axis, wings at azimuth 0 and phi
    for phi in [np.pi, 2.0, np.pi / 2, 4.0]:
        x = np.array([[1, 0, 0.3], [0, 0, 0], [0, 0, 1], [np.cos(phi), np.sin(phi), 0.6]], float)
        th = s3o.hinge_angle(x.ravel())
        assert abs(th) == pytest.approx(abs(np.pi - phi), abs=1e-12)
    flat = np.array([[1, 0, 0], [0, 0, 0], [0, 0, 1], [-1, 0, 0.5]], float)
    assert s3o.hinge_angle(flat.ravel()) == pytest.approx(0.0, abs=1e-12)


def test_spd_projection_properties():
    rng = np.random.default_rng(5)
    for _ in range(20):
        R = rng.normal(size=(12, 12))
        Hs = R + R.T
        P = s3o.spd_project_hinge(Hs)
        assert np.allclose(P, P.T, atol=1e-10)
        w, U = np.linalg.eigh(0.5 * (Hs[:9, :9] + Hs[:9, :9].T))
        ref9 = (U * np.maximum(w, 0)) @ U.T
        assert np.allclose(P[:9, :9], ref9, atol=1e-9)
        # translation invariance rebuild: every block row / column sums to zero
        Pb = P.reshape(4, 3, 4, 3)
        assert np.allclose(Pb.sum(axis=0), 0, atol=1e-9) and np.allclose(Pb.sum(axis=2), 0, atol=1e-9)
        assert np.linalg.eigvalsh(P).min() > -1e-9
    # identity on a translation-invariant PSD matrix
    G = rng.normal(size=(5, 12))
    T = np.kron(np.ones((4, 1)), np.eye(3)) / 2.0
    Pj = np.eye(12) - T @ T.T  # projector removing rigid translation
    Hpsd = Pj @ G.T @ G @ Pj
    assert np.allclose(s3o.spd_project_hinge(Hpsd), Hpsd, atol=1e-9)


# ------------------------------------------------------------ host preprocessing (register_mesh)
def test_voronoi_areas_closed_forms():
    eq = np.array([[0, 0, 0], [1, 0, 0], [0.5, np.sqrt(3) / 2, 0]])
    assert np.allclose(stage3.voronoi_areas(eq, np.array([[0, 1, 2]])), np.sqrt(3) / 12)
    right = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0]], float)
    # right angle at 0: circumcentre at the hypotenuse midpoint
    assert np.allclose(stage3.voronoi_areas(right, np.array([[0, 1, 2]])), [1.0, 0.5, 0.5])
    obt = np.array([[0, 0, 0], [4, 0, 0], [2, 0.5, 0]])
    assert np.allclose(stage3.voronoi_areas(obt, np.array([[0, 1, 2]])), [0.25, 0.25, 0.5])
    V, F = scene()["V"], scene()["F"]
    p = V[F].astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    assert stage3.voronoi_areas(V, F).sum() == pytest.approx(area.sum(), rel=1e-12)


def test_edges_and_hinges_closed_manifold():
    s = scene("sheets")
    F = s["F"]
    E = stage3.edges_unique(F)
    H = stage3.hinges(F)
    assert len(E) == 3 * len(F) // 2 == len(H)
    assert np.all(E[:, 0] < E[:, 1])
    key = E[:, 1].astype(np.int64) * (1 << 32) + E[:, 0]
    assert np.all(np.diff(key) > 0)  # ordered by (max, min)
    faces = {tuple(sorted(f)) for f in F.tolist()}
    for h in H[:200]:
        assert tuple(sorted((h[0], h[1], h[2]))) in faces and tuple(sorted((h[3], h[1], h[2]))) in faces
        assert h[1] < h[2]
    assert {tuple(sorted(e)) for e in H[:, 1:3].tolist()} == {tuple(e) for e in E.tolist()}


# ------------------------------------------------------------ energies of the system
def _sys(kind="plain", **kw):
    return oracle_system(scene(kind), fp64=True, **kw)


def test_energy_gradient_matches_finite_differences():
    s = scene("sheets", level=1)
    o = oracle_system(s, fp64=True)
    o.target()
    q0 = o.q()
    rng = np.random.default_rng(6)
    # small enough that every contact distance stays well above 0 (the
    # barrier's curvature grows like 1 / d^2)
    q = q0 + rng.normal(scale=5e-5, size=q0.shape)
    o.set_q(q)
    o.contacts()
    for mask in [s3o.E_M2GT, s3o.E_ELASTIC, s3o.E_HINGE, s3o.E_COLLISION]:
        o2 = oracle_system(s, fp64=True, mask=mask)
        o2.target()
        o2.set_q(q)
        g, _ = o2.diff()
        fd = _fd(lambda y: o2.energy_at(y.reshape(-1, 3), mask), q.ravel(), 1e-8).reshape(-1, 3)
        scale = max(np.abs(fd).max(), 1e-12)
        assert np.abs(-g - fd).max() < 1e-5 * scale + 1e-9, mask


def test_collision_term_is_active_on_sheets():
    o = _sys("sheets")
    o.target()
    C = o.contacts()
    assert (C[:, 4] == s3o.BLOCK_PT).sum() > 0 and (C[:, 4] == s3o.BLOCK_EE).sum() > 0
    e = o.energy()
    assert e[3] > 0 and e[5] == pytest.approx(e[:5].sum(), rel=1e-12)


def test_contacts_brute_force_semantics():
    o = _sys("sheets")
    q = o.q()
    C = o.contacts()
    s = scene("sheets")
    F = s["F"]
    thr = 2 * 1e-2 + 1e-3
    # every pair found is within the threshold, and no vertex is paired with its own face
    for row in C[:300]:
        x = q[row[:4]]
        if row[4] == s3o.BLOCK_PT:
            assert row[0] not in row[1:4]
            d = s3o.pair_distance(x.ravel(), True, s3o.pt_classify(x))
        else:
            assert len(set(row[:4].tolist())) == 4
            d = s3o.pair_distance(x.ravel(), False, s3o.ee_classify(x))
        assert d <= thr + 1e-12
    # and the count of PT pairs equals a direct count
    n_pt = 0
    for i in range(len(q)):
        for f in F:
            if i in f:
                continue
            x = q[[i, *f]]
            if s3o.pair_distance(x.ravel(), True, s3o.pt_classify(x)) <= thr:
                n_pt += 1
    assert (C[:, 4] == s3o.BLOCK_PT).sum() == n_pt


def test_elastic_hvp_is_frozen_basis_derivative():
    s = scene("plain", level=1)
    o = oracle_system(s, fp64=True, mask=s3o.E_ELASTIC)
    rng = np.random.default_rng(7)
    q = o.q() + rng.normal(scale=1e-2, size=o.q().shape)
    o.set_q(q)
    dx = rng.normal(size=q.shape)
    Hd = o.hvp(dx)
    h = 1e-6
    fd = (o.elastic_grad_basis(q + h * dx, q) - o.elastic_grad_basis(q - h * dx, q)) / (2 * h)
    assert np.abs(Hd - fd).max() < 1e-6 * max(np.abs(fd).max(), 1e-12)


def test_hinge_hessian_matches_gradient_differences():
    s = scene("plain", level=1)
    o = oracle_system(s, fp64=True, mask=s3o.E_HINGE)
    rng = np.random.default_rng(8)
    q = o.q() + rng.normal(scale=2e-2, size=o.q().shape)
    o.set_q(q)
    H = stage3.hinges(s["F"])
    # total Hessian-vector product of the raw hinge Hessians equals the gradient derivative
    dx = rng.normal(size=q.shape)
    o.diff()
    raw = o.hinge_blocks(projected=False)
    Hd = np.zeros_like(q)
    for j, hj in enumerate(H):
        v = dx[hj].ravel()
        Hd[hj] += (raw[j].reshape(12, 12) @ v).reshape(4, 3)
    o.set_q(q + 1e-6 * dx)
    gp, _ = o.diff()
    o.set_q(q - 1e-6 * dx)
    gm, _ = o.diff()
    fd = -(gp - gm) / 2e-6
    assert np.abs(Hd - fd).max() < 1e-5 * max(np.abs(fd).max(), 1e-12)


# ------------------------------------------------------------ solver
def test_dense_quadratic_kat():
    """test_dummy_energy.py: one step on E = 0.5 x^T A x + b^T x + c lands on
    solve(A, -b) (atol 1e-6).  The clamp radius is lifted: that test predates
    clamp_p_kernel, whose 1e-2 radius would hold each Newton move to 1e-2."""
    d = np.load(os.path.join(GOLDEN, "dense_kat.npz"))
    o = s3o.System(fp64=True, mask=s3o.E_DENSE, radius=1e9)
    o.set_mesh(d["x0"].astype(np.float64), np.zeros((0, 3), np.int32), np.zeros((0, 2), np.int32),
               np.zeros((0, 4), np.int32), np.ones(2))
    o.set_dense(d["A"], d["b"], d["c"])
    o.step(1)
    assert np.allclose(o.q(), d["x_gt"], atol=1e-6)
    # energy at the minimiser: c - 0.5 b^T A^-1 b
    e = o.energy()[5]
    assert e == pytest.approx(float(d["c"][0] - 0.5 * d["b"] @ np.linalg.solve(d["A"], d["b"])), rel=1e-9)


def test_step_decreases_energy_and_keeps_sheets_apart():
    s = scene("sheets")
    o = oracle_system(s, fp64=True)
    o.target()
    e0 = o.energy()[5]
    o.step(1)
    q = o.q()
    o.target()
    assert np.isfinite(q).all()
    # the two sheets never cross: every inner vertex stays inside the outer sheet
    n = len(q) // 2
    r_out = np.linalg.norm(q[:n] - q[:n].mean(0), axis=1)
    r_in = np.linalg.norm(q[n:] - q[:n].mean(0), axis=1)
    assert np.all(r_in < r_out + 1e-9)
    e1 = o.energy()[5]
    assert e1 < e0


def test_host_preparation_matches_numpy_restatement():
    """pamo_s3_prepare_host (C++, used by the device path) against the numpy
    functions of pamo_amd.stage3 on a closed mesh and on the two-sheet mesh."""
    import ctypes
    from pamo_amd import _hip
    for kind in ("plain", "sheets"):
        s = scene(kind)
        V, F = s["V"], s["F"]
        nF = len(F)
        E = np.zeros(6 * nF, np.int32)
        H = np.zeros(12 * nF, np.int32)
        vor = np.zeros(len(V), np.float32)
        ne, nh, ma = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        t = (ctypes.c_double * 3)(*s["t"])
        _hip.check(_hip.lib().pamo_s3_prepare_host(
            np.ascontiguousarray(V).ctypes.data, len(V), np.ascontiguousarray(F).ctypes.data, nF, s["scale"], t,
            E.ctypes.data, ctypes.byref(ne), H.ctypes.data, ctypes.byref(nh), vor.ctypes.data, ctypes.byref(ma)),
            "prepare")
        Vr = V.astype(np.float64) * s["scale"] + s["t"]
        assert np.array_equal(E[: 2 * ne.value].reshape(-1, 2), stage3.edges_unique(F))
        assert np.array_equal(H[: 4 * nh.value].reshape(-1, 4), stage3.hinges(F))
        assert np.allclose(vor, stage3.voronoi_areas(Vr, F).astype(np.float32), rtol=1e-6, atol=0)
        assert ma.value == pytest.approx(stage3.min_internal_angle_deg(Vr, F), abs=1e-9)

"""Stage 3 on the GPU (libpamo_hip.so, pamo_s3_*) against the CPU oracle.

Tolerances (FP32 device state, FP64 oracle unless noted):
  closest-point distances .... 2e-6 absolute (normalised frame, extent 1);
                               the points themselves may differ where two
                               input faces are equally close (ties)
  contact pairs .............. identical sets
  energies per term .......... 1e-5 relative
  gradient / diagonal / H dx . 1e-4 of the largest entry
  projected hinge blocks ..... 1e-4 of the largest entry
  PCG direction (40 its) ..... 2e-3 of the largest entry
  CCD step ................... 1e-4 absolute
  one full step .............. 2e-4 absolute against the FP32-state oracle
"""
import os

import numpy as np
import pytest
import torch

from pamo_amd import _hip, stage3
from stage3_helpers import normalised, oracle_system, scene

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def same_targets(g, o):
    """Device closest points, shared with the oracle (ties broken alike)."""
    o.set_targets(g.debug(_hip.PAMO_S3_DBG_TARGET).reshape(-1, 3))


def gpu_system(s, **cfg):
    c = stage3.Stage3Config()
    names = {1: "Mesh2GTDistanceBvhEnergyCalculator", 2: "ElasticEnergyCalculator", 4: "HingeEnergyCalculator",
             8: "CollisionBvhEnergyCalculator", 16: "DummyEnergyCalculator"}
    mask = cfg.pop("mask", None)
    if mask is not None:
        c.energy_calcs = [n for b, n in names.items() if mask & b]
    for k, v in cfg.items():
        setattr(c, k, v)
    g = stage3.Stage3System(c)
    g.set_target(torch.from_numpy(s["gt_V"]).cuda(), torch.from_numpy(s["gt_F"]).cuda(), s["scale"], s["t"])
    g.set_mesh(torch.from_numpy(s["V"]).cuda(), torch.from_numpy(s["F"]).cuda(), s["scale"], s["t"])
    return g


def _close(a, b, frac):
    scale = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() <= frac * scale


@pytest.mark.parametrize("kind", ["plain", "sheets"])
def test_registration_and_targets(kind):
    s = scene(kind)
    print(f"[stage3] targets {kind}", flush=True)
    g = gpu_system(s)
    o = oracle_system(s)
    _, v, _, _, _ = normalised(s)
    assert np.array_equal(g.get_vertices(), v)
    tg = g.debug(_hip.PAMO_S3_DBG_TARGET).reshape(-1, 3)
    to = o.target()
    dg = np.linalg.norm(tg - v, axis=1)
    do = np.linalg.norm(to - v, axis=1)
    assert np.abs(dg - do).max() < 2e-6
    same = np.abs(tg - to).max(axis=1) < 2e-6
    assert same.mean() > 0.95


def test_contacts_identical():
    s = scene("sheets")
    print("[stage3] contacts", flush=True)
    g = gpu_system(s)
    o = oracle_system(s)
    cg = g.debug(_hip.PAMO_S3_DBG_CONTACTS)
    co = o.contacts()
    assert len(co) > 0
    sg = {tuple(r) for r in cg.tolist()}
    so = {tuple(r) for r in co.tolist()}
    assert len(sg) == len(cg) and sg == so


@pytest.mark.parametrize("kind", ["plain", "sheets"])
def test_energy_terms(kind):
    s = scene(kind)
    print(f"[stage3] energy {kind}", flush=True)
    g = gpu_system(s)
    o = oracle_system(s)
    same_targets(g, o)
    eg = g.debug(_hip.PAMO_S3_DBG_ENERGY)
    eo = o.energy()
    for k in range(6):
        assert eg[k] == pytest.approx(eo[k], rel=1e-5, abs=1e-12), k
    if kind == "sheets":
        assert eo[3] > 0


@pytest.mark.parametrize("kind", ["plain", "sheets"])
def test_gradient_diagonal_hvp_hinge_blocks(kind):
    s = scene(kind)
    print(f"[stage3] diff {kind}", flush=True)
    g = gpu_system(s)
    o = oracle_system(s)
    same_targets(g, o)
    d = g.debug(_hip.PAMO_S3_DBG_DIFF).reshape(2, -1, 3)
    go, do = o.diff()
    assert _close(d[0], go, 1e-4)
    assert _close(d[1], do, 1e-4)
    hb = g.debug(_hip.PAMO_S3_DBG_HINGE).reshape(-1, 4, 4, 3, 3)
    ho = o.hinge_blocks(projected=True).reshape(-1, 4, 3, 4, 3).transpose(0, 1, 3, 2, 4)
    assert _close(hb, ho, 1e-4)
    rng = np.random.default_rng(0)
    dx = rng.normal(size=(len(go), 3))
    hg = g.debug(_hip.PAMO_S3_DBG_HVP, dx).reshape(-1, 3)
    # the oracle's product with the device's float32 dx
    ho2 = o.hvp(dx.astype(np.float32).astype(np.float64))
    assert _close(hg, ho2, 1e-4)


@pytest.mark.parametrize("kind", ["plain", "sheets"])
def test_pcg_direction_and_ccd(kind):
    s = scene(kind)
    print(f"[stage3] pcg {kind}", flush=True)
    g = gpu_system(s)
    o = oracle_system(s)
    same_targets(g, o)
    pg = g.debug(_hip.PAMO_S3_DBG_SOLVE).reshape(-1, 3)
    po = o.solve()
    assert _close(pg, po, 2e-3)
    out = g.debug(_hip.PAMO_S3_DBG_CCD, po)
    pc, st = out[:-1].reshape(-1, 3), out[-1]
    pco, sto = o.ccd(po.astype(np.float32).astype(np.float64))
    assert _close(pc, pco, 1e-4)
    assert st == pytest.approx(sto, abs=1e-4)
    assert 0.0 <= st <= 1.0


@pytest.mark.parametrize("kind", ["plain", "sheets"])
def test_cg_graph_replays_match_eager_launches(kind):
    """The PCG loop replayed from its captured HIP graph (cg_solver.py:150-156
    captures the reference's) gives the direction of the kernels launched one
    by one, bit for bit, on several replays with eager solves and host reads
    in between."""
    s = scene(kind)
    print(f"[stage3] cg graph {kind}", flush=True)
    g = gpu_system(s)
    g.set_cg_graph(False)
    p_eager = g.debug(_hip.PAMO_S3_DBG_SOLVE)  # detection, gradient, eager PCG
    assert np.isfinite(p_eager).all() and np.abs(p_eager).max() > 0
    for rep in range(3):
        g.set_cg_graph(True)
        p_graph = g.debug(_hip.PAMO_S3_DBG_RESOLVE)  # same gradient, graph replay
        assert np.array_equal(p_graph, p_eager), rep
        g.set_cg_graph(False)
        assert np.array_equal(g.debug(_hip.PAMO_S3_DBG_RESOLVE), p_eager), rep


@pytest.mark.parametrize("kind", ["plain", "sheets"])
def test_one_step_matches_oracle(kind):
    s = scene(kind)
    print(f"[stage3] step {kind}", flush=True)
    g = gpu_system(s)
    o32 = oracle_system(s, fp64=False)
    o64 = oracle_system(s, fp64=True)
    tg = g.debug(_hip.PAMO_S3_DBG_TARGET).reshape(-1, 3)
    o32.set_targets(tg)
    o64.set_targets(tg)
    g.step(1)
    o32.step(1, keep_target=True)
    o64.step(1, keep_target=True)
    qg = g.get_vertices()
    q32, q64 = o32.q(), o64.q()
    assert np.isfinite(qg).all()
    err32 = np.abs(qg - q32).max()
    err64 = np.abs(qg - q64).max()
    print(f"[stage3] step {kind}: max |gpu - oracle32| = {err32:.3e}, |gpu - oracle64| = {err64:.3e}", flush=True)
    assert err32 < 2e-4
    st = g.stats()
    assert st["newton"] == 10 and st["cg"] == 400 and st["line_search"] == 100


def test_dense_quadratic_kat_on_device():
    """simp_cuda/safe_project/tests/test_dummy_energy.py on the GPU: one step
    on the quadratic energy lands on solve(A, -b).  The clamp radius is
    lifted (that test predates clamp_p_kernel).  Tolerance: 1e-5 relative to
    |x*| (float32 state; the reference asserts atol 1e-6 in float32)."""
    print("[stage3] dense KAT", flush=True)
    d = np.load(os.path.join(GOLDEN, "dense_kat.npz"))
    c = stage3.Stage3Config()
    c.energy_calcs = ["DummyEnergyCalculator"]
    c.contact_detection_radius = 1e9
    g = stage3.Stage3System(c)
    g.set_mesh_prepared(torch.from_numpy(d["x0"]).cuda(), np.zeros((0, 3), np.int32), np.zeros((0, 2), np.int32),
                        np.zeros((0, 4), np.int32), np.ones(2, np.float32), False, 1.0, np.zeros(3))
    g.set_dense(d["A"], d["b"], d["c"])
    g.step(1)
    x = g.get_vertices()
    assert np.abs(x - d["x_gt"]).max() <= 1e-5 * np.abs(d["x_gt"]).max()


def test_process_end_to_end():
    """processing.process: five steps, mapped back to the input frame; the
    mesh moves toward the input surface and the two sheets stay apart."""
    s = scene("sheets")
    print("[stage3] process", flush=True)
    V, F = stage3.process(torch.from_numpy(s["gt_V"]).cuda(), torch.from_numpy(s["gt_F"]).cuda(),
                          torch.from_numpy(s["V"]).cuda(), torch.from_numpy(s["F"]).cuda(), 5)
    assert V.dtype == np.float64 and V.shape == s["V"].shape and np.array_equal(F, s["F"])
    assert np.isfinite(V).all()
    n = len(V) // 2
    c = V[:n].mean(0)
    assert np.all(np.linalg.norm(V[n:] - c, axis=1) < np.linalg.norm(V[:n] - c, axis=1))


def test_process_is_deterministic():
    """Contacts in a fixed order and fixed-order per-vertex sums: two runs
    on fresh systems, and a third on a reused one, give the same vertices
    bit for bit (two sheets 0.1 % apart: thousands of active contacts)."""
    s = scene("sheets", level=3)
    args = [torch.from_numpy(s[k]).cuda() for k in ("gt_V", "gt_F", "V", "F")]
    print("[stage3] process three times", flush=True)
    V1, _ = stage3.process(*args, 5)
    V2, _ = stage3.process(*args, 5)
    sy = stage3.Stage3System(stage3.Stage3Config())
    stage3.process(*args, 2, system=sy)
    V3, _ = stage3.process(*args, 5, system=sy)
    assert sy.stats()["contacts"] > 1000
    assert np.array_equal(V1.view(np.uint64), V2.view(np.uint64))
    assert np.array_equal(V1.view(np.uint64), V3.view(np.uint64))


def test_device_normalization_equals_numpy():
    """pamo_s3_normalization: numpy's sequential FP64 mean and the extents,
    bit for bit, on an off-centre cloud of 300k points."""
    rng = np.random.default_rng(3)
    V = (rng.standard_normal((300_000, 3)) * np.array([1.0, 30.0, 1e-2]) + np.array([5.0, -300.0, 2e3]))
    V = V.astype(np.float32)
    sy = stage3.Stage3System(stage3.Stage3Config())
    print("[stage3] device normalization, 300k points", flush=True)
    sc, t = stage3.normalization_transform_device(sy, torch.from_numpy(V).cuda())
    sc_ref, t_ref = stage3.normalization_transform(V.astype(np.float64))
    assert sc == sc_ref
    assert np.array_equal(np.asarray(t, np.float64), np.asarray(t_ref, np.float64))

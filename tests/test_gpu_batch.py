"""Batched and served forms (pamo_batch_run, pamo_serve_*) on the GPU.

Every job is one whole PaMO.run and must give the result of that call made
alone on a fresh object, bit for bit: stage 3 lists contacts and sums every
per-vertex term in a fixed order, and the engine normalises with numpy's
sequential FP64 mean, so its float32 vertices equal the plain form's float64
vertices rounded to float32.
"""
import numpy as np
import pytest
import torch

from pamo_amd import _hip, synth
from pamo_amd.batch import Server, simplify_batch
from pamo_amd.pamo import PaMO, simplify_device
from stage2_helpers import centred

pytestmark = pytest.mark.gpu


def _meshes():
    out = []
    for level, seed in [(4, 0), (3, 1), (4, 2), (3, 3)]:
        V, F = synth.icosphere_sh(level, seed)
        out.append((V.astype(np.float32), F.astype(np.int32)))
    V, F = synth.icosphere_sh(3, 5)
    V2, F2 = synth.offset_sheets(V, F, rel_offset=5e-3)
    out.append((V2.astype(np.float32), F2.astype(np.int32)))
    return out


def _alone(V, F, ratio, stage3, exact_guard=False):
    pm = PaMO((V, F), use_stage1=False, use_stage3=stage3, exact_guard=exact_guard)
    v, f = pm.run(torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda(), ratio, 4, 1e-3, 100000, 0)
    return np.asarray(v), np.asarray(f)


def _same(a, b, V, stage3):
    (va, fa), (vb, fb) = a, b
    assert np.array_equal(fa.astype(np.int64), fb.astype(np.int64))
    assert np.array_equal(va.astype(np.float32).view(np.uint32), vb.astype(np.float32).view(np.uint32))


def test_context_reset_equals_fresh_context():
    """A context that ran another mesh, then pamo_ctx_reset, gives the fresh
    context's result bit for bit."""
    (Va, Fa), (Vb, Fb) = _meshes()[0], _meshes()[1]
    print("case: reset vs fresh, icosphere L4 seed 0 after L3 seed 1", flush=True)
    Vca, Vcb = centred(Va, Fa), centred(Vb, Fb)
    fresh = _hip.Context(0)
    v1, f1, h1 = simplify_device(fresh, torch.from_numpy(Vca).cuda(), torch.from_numpy(Fa).cuda(), 0.05, min_verts=0)
    used = _hip.Context(0)
    simplify_device(used, torch.from_numpy(Vcb).cuda(), torch.from_numpy(Fb).cuda(), 0.05, min_verts=0)
    _hip.check(_hip.lib().pamo_ctx_reset(used.ptr), "pamo_ctx_reset")
    v2, f2, h2 = simplify_device(used, torch.from_numpy(Vca).cuda(), torch.from_numpy(Fa).cuda(), 0.05, min_verts=0)
    assert np.array_equal(h1, h2)
    assert np.array_equal(f1.cpu().numpy(), f2.cpu().numpy())
    assert np.array_equal(v1.cpu().numpy().view(np.uint32), v2.cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("n_slots", [1, 3])
def test_batch_matches_each_mesh_alone(n_slots):
    meshes = _meshes()
    ratios = [0.05, 0.1, 0.02, 0.05, 0.05]
    stage3 = [False, True, False, True, False]
    print(f"case: batch of {len(meshes)} meshes, {n_slots} slots", flush=True)
    out, secs = simplify_batch(meshes, ratios, stage3, n_slots=n_slots)
    for (V, F), r, s3, got in zip(meshes, ratios, stage3, out):
        _same(got, _alone(V, F, r, s3), V, s3)


def test_batch_and_server_exact_guard_match_plain_form():
    """The exact guard setting reaches every slot: batched and served results
    equal PaMO(..., exact_guard=True) run alone."""
    meshes = _meshes()
    ratios = [0.05, 0.1, 0.02, 0.05, 0.02]
    stage3 = [False, True, False, False, True]
    print("case: batch of 5 meshes, exact guard, 2 slots", flush=True)
    out, _ = simplify_batch(meshes, ratios, stage3, n_slots=2, exact_guard=True)
    for (V, F), r, s3, got in zip(meshes, ratios, stage3, out):
        _same(got, _alone(V, F, r, s3, exact_guard=True), V, s3)
    print("case: server, 2 jobs, exact guard, 2 slots", flush=True)
    srv = Server(n_slots=2, stage3=True, exact_guard=True)
    ids = [srv.submit(V, F, r, s3) for (V, F), r, s3 in list(zip(meshes, ratios, stage3))[3:]]
    for i, (V, F), r, s3 in zip(ids, meshes[3:], ratios[3:], stage3[3:]):
        _same(srv.result(i)[:2], _alone(V, F, r, s3, exact_guard=True), V, s3)
    srv.close()


def test_served_results_same_alone_and_in_the_mix():
    """Jobs submitted together to a 3-slot server (overlapping on the
    device) give the results of the same jobs run one at a time."""
    meshes = _meshes()
    ratios = [0.05, 0.1, 0.02, 0.05, 0.05]
    stage3 = [True, False, True, False, True]
    print("case: server, 5 jobs in a 3-slot mix vs one at a time", flush=True)
    srv = Server(n_slots=3, stage3=True)
    ids = [srv.submit(V, F, r, s3) for (V, F), r, s3 in zip(meshes, ratios, stage3)]
    mixed = [srv.result(i) for i in ids]
    srv.close()
    solo = Server(n_slots=1, stage3=True)
    alone = []
    for (V, F), r, s3 in zip(meshes, ratios, stage3):
        alone.append(solo.result(solo.submit(V, F, r, s3)))
    solo.close()
    for (V, F), s3, m, a in zip(meshes, stage3, mixed, alone):
        assert m[2]["state"] == "done" and m[2]["latency_s"] > 0
        _same(m[:2], a[:2], V, s3)


def test_many_slots_first_allocations_concurrent():
    """Twelve meshes of mixed size on six slots: every slot allocates its
    buffers while the others run (the allocation fills must be ordered on the
    slot's own streams), then reuses them for smaller meshes."""
    specs = [(5, 10), (3, 11), (4, 12), (5, 13), (3, 14), (4, 15), (5, 16), (3, 17), (4, 18), (3, 19), (4, 20),
             (5, 21)]
    meshes = []
    for level, seed in specs:
        V, F = synth.icosphere_sh(level, seed)
        meshes.append((V.astype(np.float32), F.astype(np.int32)))
    ratios = [0.05, 0.1, 0.02] * 4
    stage3 = [True, False] * 6
    print("case: batch of 12 meshes (icosphere L3-L5), 6 slots", flush=True)
    out, secs = simplify_batch(meshes, ratios, stage3, n_slots=6)
    for (V, F), r, s3, got in zip(meshes, ratios, stage3, out):
        _same(got, _alone(V, F, r, s3), V, s3)

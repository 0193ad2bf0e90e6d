"""The batched and served workloads and the failure paths, on the GPU.

- Batched stages 2+3 (mesh-of-meshes of the ellipsoid/box generator, face
  counts log-uniform, per-mesh ratios from {0.005, 0.01, 0.02, 0.05}), at a
  reduced count of 6 meshes of 20k-60k faces: each batched stage-2 output
  equals the oracle's on the reference-centred input, bit for bit.
- A Poisson request stream through pamo_serve_* (10 jobs, 40/s, face counts
  log-uniform, ratio U[0.002, 0.05], stage 3 on half by a seeded draw):
  stage-2 jobs equal the oracle bit for bit, stage-3 jobs equal PaMO.run
  alone bit for bit (float32).
- Admission under a small HBM budget: every job is larger than the budget,
  so jobs run one at a time and still give the plain results.
- Stage 3 with a tiny max_blocks: the contact radius halves until the pairs
  fit (energy.py:941-948), on the device as on the oracle.
- Stage 3 on a mesh whose smallest angle is below min_angle_thres does
  nothing (system.py:217-222,420-422).
"""
import numpy as np
import pytest
import torch

from pamo_amd import _hip, synth, stage3
from stage2_helpers import tris_mean

pytestmark = pytest.mark.gpu


def _oracle_stage2(oracle_lib, V, F, ratio):
    mean = tris_mean(V, F)
    vo, fo, _ = oracle_lib.run((V - mean).astype(np.float32), F, ratio, min_verts=0)
    return (vo + mean).astype(np.float32), fo


def _plain(V, F, ratio, s3):
    from pamo_amd.pamo import PaMO
    pm = PaMO((V, F), use_stage1=False, use_stage3=s3)
    v, f = pm.run(torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda(), ratio, 4, 1e-3, 100000, 0)
    return np.asarray(v), np.asarray(f)


def test_batch_each_mesh_equals_oracle(oracle_lib):
    from pamo_amd.batch import simplify_batch
    rng = np.random.default_rng(11)
    targets = np.exp(rng.uniform(np.log(20e3), np.log(60e3), 6))
    ratios = rng.choice([0.005, 0.01, 0.02, 0.05], 6)
    meshes = [synth.sized_union(100 + i, targets[i]) for i in range(6)]
    print(f"case: batch of 6, faces {[len(F) for _, F in meshes]}, ratios {ratios.tolist()}", flush=True)
    out, _ = simplify_batch(meshes, ratios, False, n_slots=3)
    for (V, F), (vb, fb), r in zip(meshes, out, ratios):
        vo, fo = _oracle_stage2(oracle_lib, V, F, float(r))
        assert np.array_equal(fb, fo)
        assert np.array_equal(np.asarray(vb, np.float32).view(np.uint32), vo.view(np.uint32))


def test_poisson_stream_through_the_server(oracle_lib):
    from pamo_amd.batch import Server
    import time
    rng = np.random.default_rng(7)
    pool = [synth.sized_union(700 + i, t) for i, t in enumerate(np.exp(rng.uniform(np.log(10e3), np.log(50e3), 4)))]
    jobs, t = [], 0.0
    for _ in range(10):
        t += rng.exponential(1.0 / 40.0)
        jobs.append((t, int(rng.integers(4)), float(rng.uniform(0.002, 0.05)), bool(rng.random() < 0.5)))
    assert any(s3 for *_, s3 in jobs) and not all(s3 for *_, s3 in jobs)
    print(f"case: served stream of {len(jobs)} jobs over {t:.2f}s, pool faces {[len(F) for _, F in pool]}",
          flush=True)
    dev = [(torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda()) for V, F in pool]
    srv = Server(n_slots=3, stage3=True)
    ids = []
    t0 = time.perf_counter()
    for ta, k, r, s3 in jobs:
        dt = t0 + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(dev[k][0], dev[k][1], r, s3))
    results = [srv.result(i) for i in ids]
    srv.close()
    for (ta, k, r, s3), (v, f, info) in zip(jobs, results):
        V, F = pool[k]
        assert info["state"] == "done"
        if s3:
            va, fa = _plain(V, F, r, True)
            assert np.array_equal(f.astype(np.int64), fa.astype(np.int64))
            assert np.array_equal(v.astype(np.float32).view(np.uint32), va.astype(np.float32).view(np.uint32))
        else:
            vo, fo = _oracle_stage2(oracle_lib, V, F, r)
            assert np.array_equal(f, fo)
            assert np.array_equal(v.view(np.uint32), vo.view(np.uint32))


def test_small_budget_runs_jobs_one_at_a_time():
    """An HBM budget below every job's estimate: each job is admitted only
    when no other runs (serve.hip's admission), and the results are the
    plain ones."""
    from pamo_amd.batch import Server
    import time
    meshes = [synth.icosphere_sh(4, s) for s in range(4)]
    print("case: server with a 1-byte HBM budget, 4 jobs on 4 slots", flush=True)
    srv = Server(n_slots=4, hbm_budget_bytes=1, stage3=False)
    subs, ids = [], []
    for V, F in meshes:
        subs.append(time.perf_counter())
        ids.append(srv.submit(V, F, 0.1, False))
    res = [srv.result(i) for i in ids]
    srv.close()
    spans = sorted((ts + info["queue_s"], ts + info["latency_s"]) for ts, (_, _, info) in zip(subs, res))
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert b0 >= a1 - 2e-3, (spans,)
    for (V, F), (v, f, _) in zip(meshes, res):
        va, fa = _plain(V, F, 0.1, False)
        assert np.array_equal(f, fa)


def test_contact_radius_halves_when_pairs_exceed_max_blocks():
    """Stage3Config.max_blocks below the sheets scene's pair count: the
    detection halves the radius and retries; the step then matches the
    oracle run with the same cap."""
    from stage3_helpers import oracle_system, scene
    from test_gpu_stage3 import gpu_system, same_targets
    s = scene("plain")
    # two sheets about 0.01 apart in the normalised frame: every pair lies
    # beyond d_hat and within 2 r + d_hat at the default r = 1e-2, and
    # halving r to 2.5e-3 leaves none within reach
    V = s["V"].astype(np.float64)
    s["V"] = np.concatenate([V, V * 0.98]).astype(np.float32)
    s["F"] = np.ascontiguousarray(np.concatenate([s["F"], s["F"][:, ::-1] + len(V)]), dtype=np.int32)
    g0 = gpu_system(s)
    n_pairs = len(g0.debug(_hip.PAMO_S3_DBG_CONTACTS))
    cap = n_pairs // 2
    print(f"case: sheets 0.01 apart, {n_pairs} pairs at the default radius, max_blocks {cap}", flush=True)
    assert n_pairs > 100
    g = gpu_system(s, max_blocks=cap)
    o = oracle_system(s, fp64=False, max_blocks=cap)
    same_targets(g, o)
    g.step(1)
    o.step(1, keep_target=True)
    st = g.stats()
    assert st["radius"] < 1e-2  # halved at least once
    assert st["radius"] == pytest.approx(o.stats()["radius"], rel=1e-6)
    assert np.abs(g.get_vertices() - o.q()).max() < 2e-4


def test_small_minimum_angle_does_nothing():
    """min_angle_thres above the mesh's smallest angle: register_mesh sets
    do_nothing and step() leaves the vertices as they are."""
    from stage3_helpers import scene
    from test_gpu_stage3 import gpu_system
    s = scene("plain")
    # degrees, as the reference compares them: every triangle has an angle
    # of at most 60 degrees
    g = gpu_system(s, min_angle_thres=70.0)
    print("case: do_nothing, min angle threshold 70 degrees", flush=True)
    assert g.do_nothing
    before = g.get_vertices().copy()
    g.step(1)
    assert np.array_equal(g.get_vertices(), before)
    V, F = s["V"], s["F"]
    c = stage3.Stage3Config()
    c.min_angle_thres = 70.0
    v3, f3 = stage3.process(s["gt_V"], s["gt_F"], V, F, 5, config=c)
    assert np.abs(v3 - V.astype(np.float64)).max() < 1e-6

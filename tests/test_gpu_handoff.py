"""The stage-2 -> stage-3 handoff merge (pamo_s3_merge_vertices) against the
CPU restatement of trimesh's process=True vertex merge (oracle/handoff.py):
vertices and faces equal bit for bit on the hand case, on random inputs
with planted duplicates and rounding ties, at a million vertices, and on
PaMO.run's own stage-3 input.
"""
import numpy as np
import pytest
import torch

from oracle import handoff as h
from test_oracle_handoff import _tie_case, planted

pytestmark = pytest.mark.gpu


def _system():
    from pamo_amd import stage3
    return stage3.Stage3System(stage3.Stage3Config())


def _check(system, V, F):
    v, f = system.merge_vertices(torch.from_numpy(V).cuda(), torch.from_numpy(F.astype(np.int32)).cuda())
    Vr, Fr = h.merge_vertices(V, F)
    assert np.array_equal(v.cpu().numpy().astype(np.float64), Vr)
    assert np.array_equal(f.cpu().numpy().astype(np.int64), Fr)


def test_hand_case():
    s = _system()
    V, F = _tie_case()
    print("case: handoff hand case", flush=True)
    _check(s, V, F)
    _check(s, V, np.zeros((0, 3), np.int64))


@pytest.mark.parametrize("n,n_f,seed", [(3000, 2500, 0), (40000, 30000, 1), (1 << 20, 1 << 20, 2)])
def test_planted_duplicates_and_ties(n, n_f, seed):
    s = _system()
    V, F = planted(n, n_f, seed)
    print(f"case: handoff planted n={n} faces={n_f} seed={seed}", flush=True)
    _check(s, V, F)


def test_context_reuse_with_a_smaller_input():
    s = _system()
    for n, seed in ((20000, 3), (500, 4)):
        V, F = planted(n, n, seed)
        print(f"case: handoff reuse n={n}", flush=True)
        _check(s, V, F)


def test_non_finite_vertex_is_an_error():
    s = _system()
    V, F = _tie_case()
    V[4, 1] = np.inf
    from pamo_amd._hip import PamoError
    with pytest.raises(PamoError, match="not finite"):
        s.merge_vertices(torch.from_numpy(V).cuda(), torch.from_numpy(F.astype(np.int32)).cuda())


def test_pamo_run_stage3_input_is_the_merged_stage2_result():
    from pamo_amd import synth
    from pamo_amd.pamo import PaMO
    V, F = synth.icosphere_sh(5, 2)
    print(f"case: PaMO stage 2+3 handoff, icosphere L5, {len(F)} faces", flush=True)
    pm = PaMO((V, F), use_stage1=False, use_stage3=True)
    vo, fo = pm.run(torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda(), 0.05, 4, 1e-3, 100000, 0)
    v2, f2 = pm.stage2_out
    v2 = (v2.cpu().numpy() + pm.tris_mean).astype(np.float32)
    Vr, Fr = h.merge_vertices(v2, f2.cpu().numpy())
    v3, f3 = pm.stage3_in
    assert np.array_equal(v3.cpu().numpy().astype(np.float64), Vr)
    assert np.array_equal(f3.cpu().numpy().astype(np.int64), Fr)
    assert np.array_equal(fo, Fr)
    assert len(vo) == len(Vr)

"""The drop-in path itself, PaMO(...).run(...), against the reference's own
numpy centring and the CPU oracle.

The reference centres with tris = points[triangles].cpu().numpy();
tris_mean = tris.mean(axis=1).mean(axis=0) on float32 arrays
(simp_cuda/pamo/__init__.py:60-72), runs stage 2 on points - tris_mean
(:144) and returns verts + tris_mean (:188).  The tests below run the
reference's expressions in numpy, the oracle on the same centred input, and
PaMO.run on the raw input, and compare bit for bit.
"""
import numpy as np
import pytest

from stage2_helpers import tris_mean

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("case", ["icosphere_l5", "headline_2m", "offset_cloud"])
def test_centroid_matches_numpy_float32_mean(torch_cuda, case):
    from pamo_amd import synth, _hip
    from pamo_amd.pamo import face_centroid_mean
    torch = torch_cuda
    if case == "icosphere_l5":
        V, F = synth.icosphere_sh(5, 0)
    elif case == "headline_2m":
        V, F = synth.headline_mesh(1)
    else:
        # far from the origin with uneven extents: a float32 running sum of
        # 300k values near 1000 drifts far from the exact mean, and every
        # per-face rounding shows in the result
        rng = np.random.default_rng(5)
        V = (rng.standard_normal((150001, 3)) * [3.0, 0.1, 50.0] + [1000.0, -2.0, 0.3]).astype(np.float32)
        F = rng.integers(0, len(V), (300007, 3)).astype(np.int32)
    print(f"case: centroid {case}, {len(F)} faces", flush=True)
    ref = tris_mean(V, F)
    ctx = _hip.Context(0)
    got = face_centroid_mean(ctx, torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda())
    assert got.dtype == np.float32
    assert np.array_equal(_bits(got), _bits(ref)), (got, ref, V[F].astype(np.float64).mean(axis=(0, 1)))
    if case == "offset_cloud":
        # the case is only a check if the float32 order matters
        exact = V[F].astype(np.float64).mean(axis=1).mean(axis=0)
        assert np.abs(ref - exact).max() > 1e-3


def _pamo_stage2(V, F, ratio, torch):
    from pamo_amd.pamo import PaMO
    pm = PaMO((V, F), use_stage1=False, use_stage3=False)
    vo, fo = pm.run(torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda(), ratio, min_verts=0)
    return vo, fo, pm


def test_pamo_run_icosphere_matches_oracle(torch_cuda, oracle_lib):
    """Config 1: icosphere level 5 with spherical-harmonic displacement,
    ratio 0.05, stage 2 only, through PaMO.run."""
    from pamo_amd import synth
    V, F = synth.icosphere_sh(5, 0)
    print(f"case: PaMO.run icosphere L5, {len(F)} faces, ratio 0.05", flush=True)
    vo, fo, pm = _pamo_stage2(V, F, 0.05, torch_cuda)
    mean = tris_mean(V, F)
    v_ref, f_ref, h_ref = oracle_lib.run((V - mean).astype(np.float32), F, 0.05, min_verts=0)
    assert np.array_equal(pm.face_hist, h_ref)
    assert np.array_equal(fo, f_ref)
    assert vo.dtype == np.float32
    assert np.array_equal(_bits(vo), _bits(v_ref + mean))
    assert len(fo) <= int(0.05 * len(F))


def test_pamo_run_default_min_verts_stops_after_one_round(torch_cuda, oracle_lib):
    """min_verts defaults to 1e10, so the target is never reached and the
    loop breaks after its first round (pamo/__init__.py:122,124,171)."""
    from pamo_amd import synth
    from pamo_amd.pamo import PaMO
    torch = torch_cuda
    V, F = synth.icosphere_sh(4, 1)
    pm = PaMO((V, F), use_stage1=False, use_stage3=False)
    vo, fo = pm.run(torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda(), 0.05)
    assert pm.rounds == 1
    mean = tris_mean(V, F)
    v_ref, f_ref, h_ref = oracle_lib.run((V - mean).astype(np.float32), F, 0.05)
    assert len(h_ref) == 1
    assert np.array_equal(fo, f_ref)
    assert np.array_equal(_bits(vo), _bits(v_ref + mean))


def test_pamo_run_headline_stage2_full_size_matches_oracle(torch_cuda, oracle_lib):
    """Config 2's stage 2 at full size (2,006,528 faces -> about 20k),
    through PaMO.run, bit for bit against the oracle on the reference's
    centred input."""
    import time
    from pamo_amd import synth
    V, F = synth.headline_mesh(1)
    print(f"case: PaMO.run headline stage 2, {len(F)} faces, ratio 0.01", flush=True)
    vo, fo, pm = _pamo_stage2(V, F, 0.01, torch_cuda)
    print(f"  GPU: {pm.rounds} rounds, {len(fo)} faces", flush=True)
    mean = tris_mean(V, F)
    t0 = time.perf_counter()
    v_ref, f_ref, h_ref = oracle_lib.run((V - mean).astype(np.float32), F, 0.01, min_verts=0)
    print(f"  oracle: {len(h_ref)} rounds in {time.perf_counter() - t0:.1f} s", flush=True)
    assert np.array_equal(pm.face_hist, h_ref)
    assert np.array_equal(fo, f_ref)
    assert np.array_equal(_bits(vo), _bits(v_ref + mean))
    assert len(fo) <= int(0.01 * len(F))
    assert synth.is_watertight_manifold(fo)


def test_type_errors_name_the_array(torch_cuda):
    """Float64 points and int64 faces are rejected before any device work,
    with the messages of the reference's per-round operator."""
    from pamo_amd import synth
    from pamo_amd.pamo import PaMO, CUDSP_Free
    torch = torch_cuda
    V, F = synth.icosphere_sh(2, 0)
    pm = PaMO((V, F), use_stage1=False, use_stage3=False)
    with pytest.raises(RuntimeError, match="points type"):
        pm.run(torch.from_numpy(V.astype(np.float64)).cuda(), torch.from_numpy(F).cuda(), 0.5)
    with pytest.raises(RuntimeError, match="triangles type"):
        pm.run(torch.from_numpy(V).cuda(), torch.from_numpy(F.astype(np.int64)).cuda(), 0.5)
    op = CUDSP_Free(0)
    bad = F.copy()
    bad[3, 1] = len(V)
    with pytest.raises(RuntimeError, match="face indices"):
        op.forward(torch.from_numpy(V).cuda(), torch.from_numpy(bad).cuda(), torch.zeros(0, dtype=torch.int32).cuda(),
                   0, 1.0, 1e-3, False, True)


def test_guard_walk_bound_reports_an_error(torch_cuda, oracle_lib):
    """A guard tree walk that exceeds its bound stops and fails the call with
    an error status (a corrupted tree would otherwise hold the GPU).  The
    bound is lowered to one step on a valid tree to reach that path; the
    same context then runs normally again."""
    from pamo_amd import synth, _hip
    from pamo_amd.pamo import simplify_device
    from stage2_helpers import centred
    torch = torch_cuda
    V, F = synth.icosphere_sh(4, 2)
    Vc = centred(V, F)
    ctx = _hip.Context(0)
    _hip.check(_hip.lib().pamo_ctx_set_option(ctx.ptr, _hip.PAMO_OPT_WALK_LIMIT, 1), "set_option")
    with pytest.raises(_hip.PamoError, match="exceeded its bound"):
        simplify_device(ctx, torch.from_numpy(Vc).cuda(), torch.from_numpy(F).cuda(), 0.1, min_verts=0)
    _hip.check(_hip.lib().pamo_ctx_set_option(ctx.ptr, _hip.PAMO_OPT_WALK_LIMIT, 0), "set_option")
    _hip.check(_hip.lib().pamo_ctx_reset(ctx.ptr), "reset")
    vg, fg, hg = simplify_device(ctx, torch.from_numpy(Vc).cuda(), torch.from_numpy(F).cuda(), 0.1, min_verts=0)
    vo, fo, ho = oracle_lib.run(Vc, F, 0.1, min_verts=0)
    assert np.array_equal(hg, ho)
    assert np.array_equal(fg.cpu().numpy(), fo)


def test_calls_run_on_the_callers_stream(torch_cuda, oracle_lib):
    """From a created torch stream the library call runs on that stream
    (pamo_ctx_set_stream), ordered after the work the caller queued there
    without a host wait: an input written on the stream just before the call
    is read correctly, the result equals the oracle's, and the context goes
    back to its own stream for a call from the null stream."""
    from pamo_amd import synth, _hip
    from pamo_amd.pamo import simplify_device
    torch = torch_cuda
    V, F = synth.icosphere_sh(5, 4)
    Vc = (V - tris_mean(V, F)).astype(np.float32)
    print(f"case: caller's stream, icosphere L5, {len(F)} faces", flush=True)
    vo, fo, ho = oracle_lib.run(Vc, F, 0.05, min_verts=0)
    ctx = _hip.Context(0)
    Fd = torch.from_numpy(F).cuda()
    src = torch.from_numpy(Vc).cuda()
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        Vs = torch.empty_like(src)
        Vs.copy_(src)  # queued on s; the call must see it
        Vs.mul_(1.0)
        vg, fg, hg = simplify_device(ctx, Vs, Fd, 0.05, min_verts=0)
    s.synchronize()
    assert np.array_equal(hg, ho)
    assert np.array_equal(fg.cpu().numpy(), fo)
    assert np.array_equal(_bits(vg.cpu().numpy()), _bits(vo))
    _hip.check(_hip.lib().pamo_ctx_reset(ctx.ptr), "pamo_ctx_reset")  # a fresh context's state
    v2, f2, h2 = simplify_device(ctx, src, Fd, 0.05, min_verts=0)  # null stream: the context's own
    assert np.array_equal(h2, ho) and np.array_equal(f2.cpu().numpy(), fo)

"""Stage-2 parity: the HIP path (through the C ABI) against the CPU oracle.

Integer and topology results must be bit-exact: edges, quantised costs,
collapse records, face arrays, occupancy, maps and undo lists; positions are
compared bit for bit as well (same FP32 operations, no contraction).
"""
import numpy as np
import pytest

from stage2_helpers import escape_links, centred, run_rounds

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _gpu_forward(op, torch):
    def fwd(V, F, undo, n_undo, scale, thr, is_stuck, init):
        out = op.forward(torch.from_numpy(np.ascontiguousarray(V)).cuda(),
                         torch.from_numpy(np.ascontiguousarray(F)).cuda(),
                         torch.from_numpy(np.ascontiguousarray(undo, dtype=np.int32)).cuda(),
                         n_undo, scale, thr, is_stuck, init)
        return tuple(o.cpu().numpy() for o in out)
    return fwd


def _compare_round(it, a, b):
    names = ["verts", "tris", "occ", "map", "undo"]
    for n, x, y in zip(names, a, b):
        assert x.shape == y.shape, f"round {it}: {n} shape {x.shape} vs {y.shape}"
        if x.dtype == np.float32:
            bad = np.flatnonzero(x.view(np.uint32).ravel() != y.view(np.uint32).ravel())
        else:
            bad = np.flatnonzero(x.ravel() != y.ravel())
        assert len(bad) == 0, f"round {it}: {n} differs at {bad[:8]} ({x.ravel()[bad[:4]]} vs {y.ravel()[bad[:4]]})"


def _lockstep(V, F, ratio, torch, oracle, check_debug_rounds=3, tolerance=4, options=None):
    from pamo_amd.pamo import CUDSP_Free
    from pamo_amd import _hip
    op = CUDSP_Free(0)
    for opt, val in (options or {}).items():
        _hip.check(_hip.lib().pamo_ctx_set_option(op.ctx.ptr, opt, val), "set_option")
    ref = oracle.Round()
    gpu_fwd = _gpu_forward(op, torch)
    stats = {"rounds": 0, "undo_rounds": 0}

    def fwd(V, F, undo, n_undo, scale, thr, is_stuck, init):
        a = gpu_fwd(V, F, undo, n_undo, scale, thr, is_stuck, init)
        b = ref.forward(V, F, undo, n_undo, scale, thr, is_stuck, init)
        it = stats["rounds"]
        if it < check_debug_rounds:
            e_gpu = op.debug(_hip.PAMO_DBG_EDGES).cpu().numpy().reshape(-1, 2)
            assert np.array_equal(e_gpu, ref.edges()), f"round {it}: edges"
            c_gpu = op.debug(_hip.PAMO_DBG_EDGE_COST, dtype=torch.int32).cpu().numpy().view(np.uint32)
            assert np.array_equal(c_gpu, ref.edge_cost()), f"round {it}: edge costs"
            off_o, ring_o = ref.ring(len(V))
            assert np.array_equal(op.debug(_hip.PAMO_DBG_RING_OFF).cpu().numpy(), off_o)
            assert np.array_equal(op.debug(_hip.PAMO_DBG_RING).cpu().numpy(), ring_o)
            q_gpu = op.debug(_hip.PAMO_DBG_QUADRICS, dtype=torch.float32).cpu().numpy().reshape(10, -1).T
            assert np.array_equal(q_gpu.view(np.uint32), ref.quadrics(len(V)).view(np.uint32)), f"round {it}: Q"
        rec_o, nreg = ref.collapsed()
        assert np.array_equal(op.debug(_hip.PAMO_DBG_RECORDS).cpu().numpy(), rec_o), f"round {it}: records"
        assert op.debug(_hip.PAMO_DBG_N_REGULAR) == nreg
        it_o = ref.undo_iters()
        assert op.debug(_hip.PAMO_DBG_UNDO_ITERS) == len(it_o), f"round {it}: undo iterations"
        stats["undo_rounds"] += int(len(it_o) > 0)
        _compare_round(it, a, b)
        stats["rounds"] += 1
        return b

    Vo, Fo, hist = run_rounds(fwd, V, F, ratio, tolerance=tolerance)
    return Vo, Fo, hist, stats


def test_round_parity_icosphere(torch_cuda, oracle_lib):
    from pamo_amd import synth
    V, F = synth.icosphere_sh(5, 0)
    print(f"case: icosphere L5 SH seed 0, {len(F)} faces, ratio 0.05", flush=True)
    Vo, Fo, hist, stats = _lockstep(centred(V, F), F, 0.05, torch_cuda, oracle_lib)
    assert stats["rounds"] > 10
    assert len(Fo) <= int(0.05 * len(F))
    assert synth.is_watertight_manifold(Fo)


@pytest.mark.parametrize("lanes", [1, 8])
def test_round_parity_undo_path(torch_cuda, oracle_lib, lanes):
    """Closely spaced sheets: collapses on one sheet cross the other, so the
    guard + undo loop runs; GPU and oracle must take identical decisions,
    with one lane per item and with the 8-lane cost, quadric, selection and
    undo kernels."""
    from pamo_amd import synth, _hip
    V, F = synth.icosphere_sh(4, 3)
    V2, F2 = synth.offset_sheets(V, F, rel_offset=5e-3)
    print(f"case: offset sheets L4 seed 3, {len(F2)} faces, {lanes} lanes", flush=True)
    opts = {_hip.PAMO_OPT_EDGE_LANES: lanes, _hip.PAMO_OPT_RING_LANES: lanes}
    Vo, Fo, hist, stats = _lockstep(centred(V2, F2), F2, 0.05, torch_cuda, oracle_lib, check_debug_rounds=2,
                                    options=opts)
    assert stats["undo_rounds"] > 0, "case did not exercise the undo loop"


@pytest.mark.parametrize("level", [5, 6])
def test_fused_loop_matches_oracle(torch_cuda, oracle_lib, level):
    """Level 6 (81920 faces) runs its first rounds through the tiled prefix
    scan (arrays over 65536 items), level 5 through the one-workgroup scan."""
    from pamo_amd import synth, _hip
    from pamo_amd.pamo import simplify_device
    torch = torch_cuda
    V, F = synth.icosphere_sh(level, 0)
    Vc = centred(V, F)
    print(f"case: fused loop icosphere L{level}, {len(F)} faces", flush=True)
    ctx = _hip.Context(0)
    vg, fg, hg = simplify_device(ctx, torch.from_numpy(Vc).cuda(), torch.from_numpy(F).cuda(), 0.05, min_verts=0)
    vo, fo, ho = oracle_lib.run(Vc, F, 0.05, min_verts=0)
    assert np.array_equal(hg, ho)
    assert np.array_equal(fg.cpu().numpy(), fo)
    assert np.array_equal(vg.cpu().numpy().view(np.uint32), vo.view(np.uint32))


@pytest.mark.parametrize("small_stack", [0, 1])
def test_guard_frontier_matches_oracle(torch_cuda, oracle_lib, small_stack):
    """Undo-heavy input (offset sheets) through the fused loop.  With the
    32-entry guard frontier almost every wave overflows and walks subtrees
    by escape links; both frontier sizes give the oracle's result."""
    from pamo_amd import synth, _hip
    from pamo_amd.pamo import simplify_device
    torch = torch_cuda
    V, F = synth.icosphere_sh(4, 3)
    V2, F2 = synth.offset_sheets(V, F, rel_offset=5e-3)
    Vc = centred(V2, F2)
    print(f"case: guard frontier small_stack={small_stack}, offset sheets L4 seed 3", flush=True)
    ctx = _hip.Context(0)
    _hip.check(_hip.lib().pamo_ctx_set_option(ctx.ptr, _hip.PAMO_OPT_GUARD_SMALL_STACK, small_stack), "set_option")
    vg, fg, hg = simplify_device(ctx, torch.from_numpy(Vc).cuda(), torch.from_numpy(F2).cuda(), 0.05, min_verts=0)
    octx = oracle_lib.Round()
    vo, fo, ho = oracle_lib.run(Vc, F2, 0.05, min_verts=0, ctx=octx)
    assert oracle_lib.lib().orc_guard_pairs(octx.ctx) > 0, "case evaluated no guard pairs"
    assert np.array_equal(hg, ho)
    assert np.array_equal(fg.cpu().numpy(), fo)
    assert np.array_equal(vg.cpu().numpy().view(np.uint32), vo.view(np.uint32))


@pytest.mark.parametrize("setting", [1, 2])
@pytest.mark.parametrize("case", ["stress sheets", "union 96"])
def test_exact_guard_matches_oracle(torch_cuda, oracle_lib, case, setting):
    """Exact guard mode (setting 1: filtered predicates with the exact pass
    for undecided pairs; 2: every pair through the exact pass) against the
    oracle's exact mode, bit for bit; the output has no intersecting pair
    under the exhaustive exact check."""
    from oracle import exact
    from pamo_amd import synth, _hip
    from pamo_amd.pamo import simplify_device
    torch = torch_cuda
    if case == "stress sheets":
        V, F, _ = synth.stress_sheets(500, 1e5)
        ratio, tol, iters = 0.0025, 8, 80
    else:
        V, F = synth.ellipsoid_box_union(1, 96)
        ratio, tol, iters = 0.01, 4, 100000
    Vc = centred(V, F)
    print(f"case: exact guard setting {setting}, {case}, {len(F)} faces", flush=True)
    ctx = _hip.Context(0)
    _hip.check(_hip.lib().pamo_ctx_set_option(ctx.ptr, _hip.PAMO_OPT_EXACT_GUARD, setting), "set_option")
    vg, fg, hg = simplify_device(ctx, torch.from_numpy(Vc).cuda(), torch.from_numpy(F).cuda(), ratio,
                                 tolerance=tol, iters=iters, min_verts=0)
    vo, fo, ho = oracle_lib.run(Vc, F, ratio, tolerance=tol, iters=iters, min_verts=0, pred_mode=1)
    assert np.array_equal(hg, ho)
    assert np.array_equal(fg.cpu().numpy(), fo)
    assert np.array_equal(vg.cpu().numpy().view(np.uint32), vo.view(np.uint32))
    assert len(exact.self_intersections(vo, fo)[0]) == 0


def test_fused_loop_reuses_context(torch_cuda, oracle_lib):
    """Two runs on one context (as PaMO reuses one CUDSP_Free) keep the
    reference's persistent-buffer semantics."""
    from pamo_amd import synth, _hip
    from pamo_amd.pamo import simplify_device
    torch = torch_cuda
    ctx = _hip.Context(0)
    octx = oracle_lib.Round()
    for seed, level in [(0, 4), (1, 3)]:
        V, F = synth.icosphere_sh(level, seed)
        Vc = centred(V, F)
        print(f"case: context reuse seed {seed} level {level}", flush=True)
        vg, fg, hg = simplify_device(ctx, torch.from_numpy(Vc).cuda(), torch.from_numpy(F).cuda(), 0.05, min_verts=0)
        vo, fo, ho = oracle_lib.run(Vc, F, 0.05, min_verts=0, ctx=octx)
        assert np.array_equal(hg, ho)
        assert np.array_equal(fg.cpu().numpy(), fo)


def _check_guard_tree(op, pts, tris, n_start, oracle_lib, torch):
    """Structure of the guard's LBVH after a round whose input had n_start
    faces, all live (see test_lbvh_bit_exact)."""
    from pamo_amd import _hip
    n = op.debug(_hip.PAMO_DBG_BVH_N)
    live = tris[:, 0] >= 0
    assert n == n_start  # every face live at the start of the round is a leaf
    codes = op.debug(_hip.PAMO_DBG_BVH_CODES).cpu().numpy().view(np.uint32)
    faces = op.debug(_hip.PAMO_DBG_BVH_FACES).cpu().numpy()
    left = op.debug(_hip.PAMO_DBG_BVH_LEFT).cpu().numpy()
    right = op.debug(_hip.PAMO_DBG_BVH_RIGHT).cpu().numpy()
    assert sorted(faces.tolist()) == list(range(n_start))
    keys = (codes.astype(np.uint64) << np.uint64(32)) | faces.astype(np.uint64)
    assert np.all(keys[1:] > keys[:-1])
    left_o, right_o = oracle_lib.karras(codes, faces)
    assert np.array_equal(left, left_o) and np.array_equal(right, right_o)
    boxes = op.debug(_hip.PAMO_DBG_BVH_BOXES, dtype=torch.float32).cpu().numpy().reshape(2 * n - 1, 8)
    lo, hi = boxes[:, 0:3], boxes[:, 4:7]
    # leaves: exact boxes of the final faces; empty boxes for removed faces
    lv = live[faces]
    T = pts[tris[faces[lv]]]
    assert np.array_equal(lo[n - 1:][lv], T.min(1)) and np.array_equal(hi[n - 1:][lv], T.max(1))
    assert np.all(lo[n - 1:][~lv] == np.inf) and np.all(hi[n - 1:][~lv] == -np.inf)
    # internal: exact union of children
    assert np.array_equal(lo[: n - 1], np.minimum(lo[left], lo[right]))
    assert np.array_equal(hi[: n - 1], np.maximum(hi[left], hi[right]))
    # rope links: left child (internal) / face (leaf) in lo.w, escape in hi.w
    w = boxes.view(np.int32)
    assert np.array_equal(w[: n - 1, 3], left) and np.array_equal(w[n - 1:, 3], faces)
    assert np.array_equal(w[:, 7], escape_links(left, right, n))
    return codes, faces, lv


def test_lbvh_bit_exact(torch_cuda, oracle_lib):
    """The guard's LBVH after one round.  Its leaves are the faces live at the
    start of the round (a face removed by the collapse stays a leaf with an
    empty box, since an undo may restore it); the (code, face) keys strictly
    increase; the children equal the oracle's Karras construction over the
    same keys; leaf boxes are the exact boxes of the final faces (the last
    guard refits the hierarchy the first one built); every internal box is
    exactly the union of its children's; rope links.  A round without undo
    built its tree on the final state: then the live leaves, in order, carry
    exactly the oracle's LBVH codes of the output."""
    from pamo_amd import synth, _hip
    from pamo_amd.pamo import CUDSP_Free
    torch = torch_cuda
    for seed, res in [(1, 48), (2, 40)]:
        V, F = synth.ellipsoid_box_union(seed, res)
        Vc = centred(V, F)
        op = CUDSP_Free(0)
        out = op.forward(torch.from_numpy(Vc).cuda(), torch.from_numpy(F).cuda(),
                         torch.zeros(0, dtype=torch.int32, device="cuda"), 0, float((Vc.max(0) - Vc.min(0)).max()),
                         1e-3, False, True)
        pts, tris = out[0].cpu().numpy(), out[1].cpu().numpy()
        undo_iters = op.debug(_hip.PAMO_DBG_UNDO_ITERS)
        print(f"case: LBVH after one round, seed {seed}, {len(F)} faces, {int((tris[:, 0] >= 0).sum())} live, "
              f"undo passes {undo_iters}", flush=True)
        codes, faces, lv = _check_guard_tree(op, pts, tris, len(F), oracle_lib, torch)
        if undo_iters == 0:
            codes_o, faces_o, _, _ = oracle_lib.lbvh(pts, tris)
            assert np.array_equal(codes[lv], codes_o) and np.array_equal(faces[lv], faces_o)


def test_lbvh_refit_after_undo(torch_cuda, oracle_lib):
    """Offset sheets, round by round until a round runs the undo loop: the
    tree its last guard used was refit after the undo (restored faces have
    their exact boxes again, faces still removed have empty ones), and its
    structure passes the same checks."""
    from pamo_amd import synth, _hip
    from pamo_amd.pamo import CUDSP_Free
    torch = torch_cuda
    V, F = synth.icosphere_sh(4, 3)
    V2, F2 = synth.offset_sheets(V, F, rel_offset=5e-3)
    print(f"case: LBVH refit, offset sheets L4 seed 3, {len(F2)} faces", flush=True)
    op = CUDSP_Free(0)
    fwd = _gpu_forward(op, torch)
    seen = {"n_in": 0, "checked": 0}

    def counting_fwd(Vr, Fr, *a):
        seen["n_in"] = len(Fr)
        return fwd(Vr, Fr, *a)

    def on_round(it, out):
        if seen["checked"] or op.debug(_hip.PAMO_DBG_UNDO_ITERS) == 0:
            return
        pts, tris = out[0], out[1]
        print(f"round {it}: {op.debug(_hip.PAMO_DBG_UNDO_ITERS)} undo passes, {seen['n_in']} faces in", flush=True)
        _check_guard_tree(op, pts, tris, seen["n_in"], oracle_lib, torch)
        seen["checked"] = 1

    run_rounds(counting_fwd, centred(V2, F2), F2, 0.05, on_round=on_round)
    assert seen["checked"], "no round ran the undo loop"


def test_scripted_undo_and_line_collapse_rounds_match_oracle(torch_cuda, oracle_lib):
    """The scripted cases of the oracle pins, through the C ABI: crossing
    spheres (undo cap, stale maps, a stuck call) and a line collapse outside
    the undo-candidate grid next to a piercing tetrahedron."""
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    from test_oracle_stage2_pins import _interpenetrating_spheres, _bipyramid_coincident
    from pamo_amd.pamo import CUDSP_Free
    torch = torch_cuda
    V, F = _interpenetrating_spheres()
    print(f"case: crossing spheres, {len(F)} faces, 3 scripted calls", flush=True)
    op, ref = CUDSP_Free(0), oracle_lib.Round()
    fwd = _gpu_forward(op, torch)
    scale = float((V.max(0) - V.min(0)).max())
    undo = np.zeros(0, np.int32)
    for it, stuck in enumerate((0, 0, 1)):
        a = fwd(V, F, undo, len(undo), scale, 1e-3, stuck, it == 0)
        b = ref.forward(V, F, undo, len(undo), scale, 1e-3, stuck, it == 0)
        _compare_round(it, a, b)
        assert op.debug(_hip_mod().PAMO_DBG_UNDO_ITERS) == len(ref.undo_iters())
        pv, pt, occ, mp, undo = b
        occ = occ.reshape(-1).astype(bool)
        V, F = pv[occ], mp.reshape(-1)[pt[pt[:, 0] >= 0]].astype(np.int32)
    Vb, Fb = _bipyramid_coincident()
    print("case: line collapse beside a piercing tetrahedron", flush=True)
    c = (Vb[2] + Vb[0] + Vb[3]) / 3
    n = np.cross(Vb[0] - Vb[2], Vb[3] - Vb[2])
    n = n / np.linalg.norm(n)
    T = np.array([c + 0.3 * n, c - 0.3 * n + [0.05, 0, 0], c - 0.3 * n + [-0.03, 0.04, 0],
                  c - 0.3 * n + [-0.03, -0.04, 0.02]], np.float32)
    V2 = np.r_[Vb, T].astype(np.float32)
    F2 = np.r_[Fb, np.array([[5, 6, 7], [5, 7, 8], [5, 8, 6], [6, 8, 7]], np.int32)]
    op, ref = CUDSP_Free(0), oracle_lib.Round()
    fwd = _gpu_forward(op, torch)
    s2 = float((V2.max(0) - V2.min(0)).max())
    a = fwd(V2, F2, np.zeros(0, np.int32), 0, s2, 1e-9, 0, False)
    b = ref.forward(V2, F2, np.zeros(0, np.int32), 0, s2, 1e-9, 0, False)
    _compare_round(0, a, b)
    assert a[2].reshape(-1)[1] == 0


def _hip_mod():
    from pamo_amd import _hip
    return _hip


@pytest.mark.parametrize("level", [4, 6])
def test_optimal_placement_matches_oracle(torch_cuda, oracle_lib, level):
    """Opt-in optimal placement (FP64 3x3 solve per edge in the cost and the
    collapse kernels) against the oracle's, bit for bit, through the fused
    loop; the default context on the same input keeps the midpoint."""
    from pamo_amd import synth, _hip
    from pamo_amd.pamo import simplify_device
    torch = torch_cuda
    V, F = synth.icosphere_sh(level, 5)
    Vc = centred(V, F)
    print(f"case: optimal placement, icosphere L{level}, {len(F)} faces", flush=True)
    ctx = _hip.Context(0)
    _hip.check(_hip.lib().pamo_ctx_set_option(ctx.ptr, _hip.PAMO_OPT_PLACEMENT, 1), "set_option")
    vg, fg, hg = simplify_device(ctx, torch.from_numpy(Vc).cuda(), torch.from_numpy(F).cuda(), 0.05, min_verts=0)
    vo, fo, ho = oracle_lib.run(Vc, F, 0.05, min_verts=0, placement=1)
    assert np.array_equal(hg, ho)
    assert np.array_equal(fg.cpu().numpy(), fo)
    assert np.array_equal(vg.cpu().numpy().view(np.uint32), vo.view(np.uint32))
    assert synth.is_watertight_manifold(fo)
    vm, fm, hm = oracle_lib.run(Vc, F, 0.05, min_verts=0)
    assert not (len(fm) == len(fo) and np.array_equal(vm, vo))  # the option changes the result
    # the 8-lane cost, quadric, selection and undo kernels with the solve
    # give the same result
    ctx8 = _hip.Context(0)
    _hip.check(_hip.lib().pamo_ctx_set_option(ctx8.ptr, _hip.PAMO_OPT_PLACEMENT, 1), "set_option")
    _hip.check(_hip.lib().pamo_ctx_set_option(ctx8.ptr, _hip.PAMO_OPT_EDGE_LANES, 8), "set_option")
    _hip.check(_hip.lib().pamo_ctx_set_option(ctx8.ptr, _hip.PAMO_OPT_RING_LANES, 8), "set_option")
    v8, f8, h8 = simplify_device(ctx8, torch.from_numpy(Vc).cuda(), torch.from_numpy(F).cuda(), 0.05, min_verts=0)
    assert np.array_equal(h8, ho)
    assert np.array_equal(f8.cpu().numpy(), fo)
    assert np.array_equal(v8.cpu().numpy().view(np.uint32), vo.view(np.uint32))


@pytest.mark.parametrize("case", ["icosphere_l6", "sheets_l5"])
def test_tree_carry_matches_rebuild_every_round(torch_cuda, oracle_lib, case):
    """The fused loop carries the guard tree across rounds (leaves renumbered
    at each compaction, removed faces with empty boxes) until fewer than half
    its leaves are live.  Which pairs overlap does not depend on the tree, so
    carrying, rebuilding every round and the oracle agree bit for bit."""
    from pamo_amd import synth, _hip
    from pamo_amd.pamo import simplify_device
    torch = torch_cuda
    if case == "icosphere_l6":
        V, F = synth.icosphere_sh(6, 2)
    else:
        V0, F0 = synth.icosphere_sh(5, 4)
        V, F = synth.offset_sheets(V0, F0, rel_offset=3e-3)
    Vc = centred(V, F)
    print(f"case: tree carry vs rebuild, {case}, {len(F)} faces", flush=True)
    outs = []
    for keep in (50, 90, 100):
        ctx = _hip.Context(0)
        _hip.check(_hip.lib().pamo_ctx_set_option(ctx.ptr, _hip.PAMO_OPT_TREE_KEEP, keep), "set_option")
        vg, fg, hg = simplify_device(ctx, torch.from_numpy(Vc).cuda(), torch.from_numpy(F).cuda(), 0.02,
                                     min_verts=0)
        outs.append((vg.cpu().numpy(), fg.cpu().numpy(), hg))
    vo, fo, ho = oracle_lib.run(Vc, F, 0.02, min_verts=0)
    for vg, fg, hg in outs:
        assert np.array_equal(hg, ho)
        assert np.array_equal(fg, fo)
        assert np.array_equal(vg.view(np.uint32), vo.view(np.uint32))


def test_autotuned_settings_give_the_same_result(torch_cuda, oracle_lib):
    """The autotuner's knobs (guard-tree keep percentage, guard subtree
    split, guard register budget) change the launch shape only: every
    candidate gives the oracle's result bit for bit, and the tuner returns
    one of them."""
    from pamo_amd import synth, _hip, autotune
    from pamo_amd.pamo import simplify_device
    torch = torch_cuda
    V, F = synth.icosphere_sh(6, 3)
    Vc = centred(V, F)
    Vd, Fd = torch.from_numpy(Vc).cuda(), torch.from_numpy(F).cuda()
    print(f"case: autotune knobs, icosphere L6, {len(F)} faces", flush=True)
    vo, fo, ho = oracle_lib.run(Vc, F, 0.02, min_verts=0)
    for settings in ({"guard_split_log": 0}, {"guard_split_log": 2}, {"guard_split_log": 4},
                     {"guard_split_faces": 0}, {"tree_keep": 30, "guard_split_faces": 1 << 20},
                     {"guard_waves": 4}, {"guard_waves": 4, "guard_split_log": 0},
                     {"edge_lanes": 4}, {"edge_lanes": 8, "tree_keep": 100},
                     {"ring_lanes": 4}, {"ring_lanes": 8, "edge_lanes": 4}):
        ctx = _hip.Context(0)
        autotune.apply(ctx, settings)
        vg, fg, hg = simplify_device(ctx, Vd, Fd, 0.02, min_verts=0)
        assert np.array_equal(hg, ho), settings
        assert np.array_equal(fg.cpu().numpy(), fo), settings
        assert np.array_equal(vg.cpu().numpy().view(np.uint32), vo.view(np.uint32)), settings
    res = autotune.tune_stage2(Vd, Fd, 0.02, repeats=1,
                               candidates={"tree_keep": [50, 100], "guard_split_log": [2, 3]})
    assert res["settings"]["tree_keep"] in (50, 100) and res["settings"]["guard_split_log"] in (2, 3)
    assert autotune.cached(len(F)) == res["settings"]

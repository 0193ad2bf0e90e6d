"""The Blender add-on's library (libcusimp.so, pamo_simplify on host
arrays, pamo_blender/cusimpinterface.cu:57-161) and the example workflow
(OBJ in, PaMO, OBJ out) on the GPU.

The library keeps one stage-2 context for the process, as the reference's
static CUSimp_Free; two calls in a row are compared with the oracle running
both meshes on one persistent round context, bit for bit."""
import ctypes
import os

import numpy as np
import pytest
import torch

from oracle import oracle
from pamo_amd import _hip, obj_io, synth
from pamo_amd.pamo import PaMO

pytestmark = pytest.mark.gpu


def _call(lib, V, F, threshold):
    v = np.ascontiguousarray(V, np.float32)
    t = np.ascontiguousarray(F, np.int32)
    vo, to = np.zeros_like(v), np.zeros_like(t)
    sh = np.array([len(v), len(t)], np.int32)
    lib.pamo_simplify(v.ctypes.data, t.ctypes.data, vo.ctypes.data, to.ctypes.data, sh.ctypes.data,
                      ctypes.c_float(threshold))
    return vo[:sh[0]], to[:sh[1]]


def test_blender_library_matches_oracle():
    lib = ctypes.CDLL(os.path.join(os.path.dirname(_hip.LIB_PATH), "libcusimp.so"))
    lib.pamo_simplify.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_float]
    lib.pamo_simplify.restype = None
    r = oracle.Round()
    for level, seed, thr in [(4, 7, 1e-4), (3, 8, 1e-3)]:
        V, F = synth.icosphere_sh(level, seed)
        print(f"case: blender pamo_simplify, icosphere L{level} seed {seed}, threshold {thr}", flush=True)
        vg, fg = _call(lib, V, F, thr)
        vo, fo, hist = oracle.run(V.astype(np.float32), F, 0.0, 4, thr, 1000000, 0, ctx=r)
        assert len(fo) > 10 and len(fg) < len(F)
        assert np.array_equal(fg, fo)
        assert np.array_equal(vg.view(np.uint32), vo.view(np.uint32))
    # bad input: an error message, counts set to 0, nothing read out of bounds
    V, F = synth.icosphere(1)
    bad = F.copy()
    bad[0, 0] = len(V) + 5
    vg, fg = _call(lib, V, bad, 1e-3)
    assert len(vg) == 0 and len(fg) == 0


def test_example_workflow(tmp_path):
    from pamo_amd import example
    V, F = synth.icosphere_sh(4, 9)
    src, dst = tmp_path / "in.obj", tmp_path / "out.obj"
    obj_io.save_obj(src, V, F)
    print("case: example workflow, icosphere L4 seed 9, ratio 0.1, stage 3 off", flush=True)
    example.main(["-i", str(src), "-o", str(dst), "-r", "0.1", "--disable_stage1", "--disable_stage3"])
    Vo, Fo = obj_io.load_obj(dst)
    V2, F2 = obj_io.load_obj(src)
    v, f = PaMO((V2, F2), use_stage1=False, use_stage3=False).run(
        torch.from_numpy(V2).float().cuda(), torch.from_numpy(F2).int().cuda(), 0.1, min_verts=0)
    assert np.array_equal(Fo, np.asarray(f, np.int64))
    assert np.allclose(Vo, v, rtol=0, atol=1e-6)

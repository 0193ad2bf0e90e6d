"""OBJ input/output of the example workflow (pamo_amd.obj_io) and the
Blender add-on library's exports (libcusimp.so).  No GPU call is made."""
import ctypes
import os

import numpy as np

from pamo_amd import obj_io, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_parse_corners_polygons_and_relative_indices():
    text = """# comment
o thing
v 0 0 0
v 1 0 0
v 1 1 0   # trailing comment
v 0 1 0
vt 0 0
vn 0 0 1
g side
f 1/1/1 2/1/1 3/1/1 4/1/1
f -4//1 -2//1 -1//1
s off
"""
    V, F = obj_io.load_obj(text, is_text=True)
    assert V.shape == (4, 3) and np.array_equal(V[2], [1, 1, 0])
    # quad fan (0,1,2),(0,2,3), then the relative-index triangle
    assert np.array_equal(F, [[0, 1, 2], [0, 2, 3], [0, 2, 3]])


def test_round_trip(tmp_path):
    V, F = synth.icosphere_sh(3, 0)
    p = tmp_path / "m.obj"
    obj_io.save_obj(p, V, F)
    V2, F2 = obj_io.load_obj(p)
    assert np.array_equal(F2, F.astype(np.int64))
    assert np.array_equal(V2.astype(np.float32), V.astype(np.float32))


def test_blender_library_exports_header():
    from test_cabi import _declared
    from pamo_amd import _hip
    lib = ctypes.CDLL(os.path.join(os.path.dirname(_hip.LIB_PATH), "libcusimp.so"))
    names = _declared("pamo_blender.h")
    assert names == ["pamo_simplify"]
    assert all(hasattr(lib, n) for n in names)

"""The C ABI: libpamo_hip.so loads and exports every function declared in
include/pamo_hip.h, and the Python binding declares the same set.  No GPU
call is made here."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pamo_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_header():
    from pamo_amd import _hip
    names = _declared("pamo_hip.h")
    assert len(names) >= 10
    lib = ctypes.CDLL(_hip.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in _hip.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from pamo_amd import _hip
    monkeypatch.setattr(_hip, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_hip, "_lib", None)
    import pytest
    with pytest.raises(_hip.PamoError):
        _hip.lib()


def test_mesh_types_rejected_before_device_work():
    """float64 points and int64 faces are rejected with the messages of the
    reference's per-round operator (pybind.cpp:54-63), before any device
    pointer is taken."""
    import numpy as np
    import pytest
    import torch
    from pamo_amd.pamo import _check_mesh_types, face_centroid_mean, simplify_device
    V = torch.zeros((4, 3), dtype=torch.float32)
    F = torch.tensor([[0, 1, 2], [0, 2, 3]], dtype=torch.int32)
    with pytest.raises(RuntimeError, match="points type"):
        _check_mesh_types(V.double(), F)
    with pytest.raises(RuntimeError, match="triangles type"):
        _check_mesh_types(V, F.long())
    with pytest.raises(RuntimeError, match="points type"):
        face_centroid_mean(None, torch.from_numpy(np.zeros((4, 3))), F)
    with pytest.raises(RuntimeError, match="triangles type"):
        simplify_device(None, V, F.long(), 0.5)
    with pytest.raises(RuntimeError, match="CUDA tensor"):
        _check_mesh_types(V, F)

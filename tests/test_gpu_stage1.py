"""Stage-1 remeshing on the GPU against the CPU restatement (oracle/stage1.py).

- The case table of the library equals the oracle's, case by case.
- On small inputs (an icosphere, an open sheet, two crossing boxes) every
  intermediate of the remesh equals the oracle's: the normalised triangles
  and the frame bit for bit, the distance field within 2e-7 where both are
  finite (the same samples are finite), the outside mask, the vertices in
  sample units and the triangles exactly.
- At a larger resolution the output is a closed, manifold, outward-facing
  surface about 0.9 cells outside the input.
- PaMO(..., use_stage1=True).run remeshes, simplifies to the target and
  projects with stage 3.
"""
import numpy as np
import pytest
import torch

from pamo_amd import _hip, synth
from test_oracle_stage1 import _check_closed_manifold, _box_tris

pytestmark = pytest.mark.gpu


def _remesher():
    from pamo_amd.stage1 import Remesher
    return Remesher(0)


def _dev(V, F):
    return torch.from_numpy(np.ascontiguousarray(V, np.float32)).cuda(), \
        torch.from_numpy(np.ascontiguousarray(F, np.int32)).cuda()


def _mean(V, F):
    from stage2_helpers import tris_mean
    return tris_mean(V.astype(np.float32), F)


def test_case_table_equals_oracle():
    from oracle import stage1 as o
    r = _remesher()
    t = r.debug(_hip.PAMO_S1_DBG_CASES, np.int8, 256 * 18).reshape(256, 18)
    cyc, plen, npatch, ntri = o.case_table()
    assert np.array_equal(t[:, :12], cyc)
    assert np.array_equal(t[:, 12:16], plen)
    assert np.array_equal(t[:, 16], npatch)
    assert np.array_equal(t[:, 17], ntri)


def _sheet():
    V = np.array([[0, 0, 0], [1, 0, 0.1], [1, 1, 0], [0, 1, 0.2]], np.float32)
    return V, np.array([[0, 1, 2], [0, 2, 3]], np.int32)


def _boxes():
    a = _box_tris((0.0, 0.0, 0.0), (1.0, 0.4, 0.4)).reshape(-1, 3)
    b = _box_tris((0.3, -0.3, 0.1), (0.6, 0.8, 0.3)).reshape(-1, 3)
    V = np.concatenate([a, b]).astype(np.float32)
    return V, np.arange(len(V), dtype=np.int32).reshape(-1, 3)


@pytest.mark.parametrize("case,R", [("icosphere", 24), ("sheet", 32), ("boxes", 28)])
def test_remesh_matches_oracle(case, R):
    from oracle import stage1 as o
    if case == "icosphere":
        V, F = synth.icosphere_sh(2, 3)
    elif case == "sheet":
        V, F = _sheet()
    else:
        V, F = _boxes()
    V = V.astype(np.float32)
    F = F.astype(np.int32)
    print(f"case: remesh {case}, {len(F)} faces, R {R}", flush=True)
    ref = o.remesh(V, F.astype(np.int64), R)
    r = _remesher()
    Vd, Fd = _dev(V, F)
    mean = _mean(V, F)
    assert np.array_equal(mean.view(np.uint32), ref["tmean"].astype(np.float32).view(np.uint32))
    v, f = r.remesh(Vd, Fd, mean, R)
    frame = r.debug(_hip.PAMO_S1_DBG_FRAME, np.float32, 8)
    assert np.array_equal(frame[3:6].view(np.uint32), ref["tmin"].astype(np.float32).view(np.uint32))
    assert frame[6] == np.float32(ref["tmax"])
    norm = r.debug(_hip.PAMO_S1_DBG_NORMALIZED, np.float32, 12 * len(F)).reshape(-1, 3, 4)[:, :, :3]
    assert np.array_equal(norm.view(np.uint32), ref["tris"].view(np.uint32))
    field = r.debug(_hip.PAMO_S1_DBG_FIELD, np.uint32, R ** 3).view(np.float32).reshape(R, R, R)
    fin_g, fin_o = np.isfinite(field), np.isfinite(ref["field"])
    assert np.array_equal(fin_g, fin_o)
    assert np.abs(field[fin_g] - ref["field"][fin_o]).max() <= 2e-7
    outside = r.debug(_hip.PAMO_S1_DBG_OUTSIDE, np.uint8, R ** 3).reshape(R, R, R).astype(bool)
    assert np.array_equal(outside, ref["outside"])
    vi = r.debug(_hip.PAMO_S1_DBG_VERTS, np.float32, 4 * len(v)).reshape(-1, 4)[:, :3]
    assert vi.shape == ref["verts_idx"].shape
    assert np.abs(vi - ref["verts_idx"]).max() <= 2e-5
    assert np.array_equal(f.cpu().numpy(), ref["faces"].astype(np.int32))
    assert np.abs(v.cpu().numpy() - ref["verts"]).max() <= 1e-5 * float(ref["tmax"])
    _check_closed_manifold(f.cpu().numpy(), len(v))


def test_remesh_is_closed_manifold_offset_surface_at_resolution_96():
    V, F = synth.icosphere_sh(4, 1)
    R = 96
    r = _remesher()
    Vd, Fd = _dev(V, F)
    mean = _mean(V, F)
    print(f"case: remesh icosphere L4, {len(F)} faces, R {R}", flush=True)
    v, f = r.remesh(Vd, Fd, mean, R)
    v, f = v.cpu().numpy().astype(np.float64), f.cpu().numpy()
    _check_closed_manifold(f, len(v))
    frame = r.debug(_hip.PAMO_S1_DBG_FRAME, np.float32, 8)
    band = 3.0 / R
    cell = float(frame[6]) * (band * 2 + 1) / R
    # distance of each output vertex to the input surface: about 0.9 cells
    from pamo_amd import metrics
    Vc = (V.astype(np.float64) - mean)
    sq = np.asarray(metrics.point_mesh_squared_distance(v, Vc, F)[0])
    d = np.sqrt(sq)
    assert np.abs(d - 0.9 * cell).max() < 0.5 * cell
    # outward: normals point away from the centroid of this convex input
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    assert (np.einsum("ij,ij->i", n, v[f].mean(1)) > 0).mean() > 0.999


def test_pamo_run_with_stage1():
    from pamo_amd.pamo import PaMO
    V, F = synth.icosphere_sh(5, 2)
    print(f"case: PaMO stage 1+2+3, icosphere L5, {len(F)} faces", flush=True)
    pm = PaMO((V, F), use_stage1=True, use_stage3=True)
    vo, fo = pm.run(torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda(), 0.05, 4, 1e-3, 100000, 0)
    assert pm.R == 256
    assert len(fo) <= int(0.05 * len(F))
    assert synth.is_watertight_manifold(np.asarray(fo))
    assert np.isfinite(vo).all()
    assert pm.timings["stage1_s"] > 0 and pm.timings["stage2_s"] > 0
    # the result stays close to the input surface (a displaced sphere of
    # radius about 1): distance of every output vertex to it
    from pamo_amd import metrics
    d = np.sqrt(np.asarray(metrics.point_mesh_squared_distance(np.asarray(vo, np.float64), V.astype(np.float64),
                                                               F)[0]))
    print(f"max distance to the input {d.max():.4f}, mean {d.mean():.4f}", flush=True)
    assert d.max() < 0.03

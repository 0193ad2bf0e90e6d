"""GPU surface metrics (pamo_point_mesh_sqdist, pamo_hausdorff through
pamo_amd.metrics) against the brute-force oracle, and the Hausdorff
acceptance check: on the small stage-2+3 workload (icosphere level 5 with
spherical-harmonic bumps, ratio 0.05) the output's Hausdorff distance to the
input is within 1% of the oracle's output's."""
import numpy as np
import pytest
import torch

from oracle import metrics as om
from pamo_amd import metrics, stage3, synth
from pamo_amd.pamo import PaMO
from stage3_helpers import oracle_system

pytestmark = pytest.mark.gpu


def _points_around(V, F, n, rng):
    """Points near the surface, far from it, on vertices and on faces."""
    c = V.mean(0)
    ext = np.abs(V - c).max()
    near = V[rng.integers(0, len(V), n // 4)] + rng.normal(scale=0.02 * ext, size=(n // 4, 3))
    far = c + rng.normal(scale=2 * ext, size=(n // 4, 3))
    on_v = V[rng.integers(0, len(V), n // 8)]
    on_f, _ = metrics.sample_surface(V, F, n - 2 * (n // 4) - n // 8, seed=int(rng.integers(1 << 30)))
    return np.concatenate([near, far, on_v, on_f]).astype(np.float64)


def test_point_mesh_matches_brute_force():
    V, F = synth.icosphere_sh(level=4, seed=3)
    V = V.astype(np.float64)
    P = _points_around(V, F, 4000, np.random.default_rng(5))
    print(f"[metrics] point-mesh {len(P)} points, {len(F)} faces", flush=True)
    sq, idx, C = metrics.point_mesh_squared_distance(P, V, F)
    so, io = om.point_mesh_sqdist(P, V, F)
    # the same per-face arithmetic: equal bit for bit except at exact ties
    exact = np.mean(sq == so)
    assert exact > 0.999, exact
    assert np.all(np.abs(sq - so) <= 1e-12 * so + 1e-30)
    # the reported face reaches the minimum and holds the closest point
    for i in np.flatnonzero(idx != io):
        assert om.point_tri_sqdist(P[i], *V[F[idx[i]]]) == pytest.approx(so[i], rel=1e-12, abs=1e-30)
    d = ((C - P) ** 2).sum(1)
    assert np.all(np.abs(d - sq) <= 1e-12 * sq + 1e-30)
    on = [om.point_tri_sqdist(C[i], *V[F[idx[i]]]) for i in range(0, len(P), 37)]
    assert max(on) < 1e-24


def test_hausdorff_matches_brute_force_and_closed_forms():
    A, FA = synth.icosphere_sh(level=4, seed=1)
    B, FB = synth.icosphere(3)
    print("[metrics] hausdorff bumpy sphere / icosphere", flush=True)
    hg = metrics.hausdorff(A, FA, B, FB)
    ho = om.hausdorff(A, FA, B, FB)
    assert hg == pytest.approx(ho, rel=1e-12)
    assert metrics.hausdorff(B, FB, A, FA) == hg
    # the larger of the two one-sided maxima of the point-mesh distances
    one = max(metrics.point_mesh_squared_distance(A, B, FB)[0].max(),
              metrics.point_mesh_squared_distance(B, A, FA)[0].max())
    assert hg == np.sqrt(one)
    # closed forms: a cube against the same cube doubled about its centre,
    # a plane grid against its translate
    V = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float64)
    q = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    Fc = np.array([t for a, b, c, d in q for t in ((a, b, c), (a, c, d))], np.int32)
    assert metrics.hausdorff(V, Fc, (V - 0.5) * 2 + 0.5, Fc) == pytest.approx(np.sqrt(3) / 2, rel=1e-14)
    assert metrics.hausdorff(V, Fc, V + np.array([0, 0, 0.25]), Fc) == pytest.approx(0.25, rel=1e-14)


def test_metric_inputs_checked():
    V, F = synth.icosphere(2)
    bad = F.copy()
    bad[7, 1] = len(V)
    print("[metrics] face index outside the vertices", flush=True)
    from pamo_amd import _hip
    with pytest.raises(_hip.PamoError, match="index outside"):
        metrics.hausdorff(V, bad, V, F)
    with pytest.raises(_hip.PamoError, match="index outside"):
        metrics.point_mesh_squared_distance(V, V, bad)


def test_igl_cd_hd_matches_oracle_assembly():
    A, FA = synth.icosphere_sh(level=4, seed=2)
    B, FB = synth.icosphere(3)
    print("[metrics] compute_igl_CD_HD", flush=True)
    cd, hd = metrics.compute_igl_CD_HD(A, FA, B, FB)
    ga, _ = metrics.sample_surface(A, FA, 5000, seed=42)
    gb, _ = metrics.sample_surface(B, FB, 5000, seed=42)
    cd_o = om.point_mesh_sqdist(ga, B, FB)[0].mean() + om.point_mesh_sqdist(gb, A, FA)[0].mean()
    assert cd == pytest.approx(cd_o, rel=1e-12)
    assert hd == pytest.approx(om.hausdorff(A, FA, B, FB), rel=1e-12)
    # compute_trimesh_CD_HD: samples of each mesh against the other surface
    cd2, hd2 = metrics.compute_trimesh_CD_HD((A, FA), (B, FB))
    sa, sb = om.point_mesh_sqdist(gb, A, FA)[0], om.point_mesh_sqdist(ga, B, FB)[0]
    assert cd2 == pytest.approx(sa.mean() + sb.mean(), rel=1e-12)
    assert hd2 == pytest.approx(np.sqrt(max(sa.max(), sb.max())), rel=1e-12)


def test_small_workload_hausdorff_within_one_percent_of_oracle():
    V, F = synth.icosphere_sh(level=5, seed=0)
    Vd, Fd = torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda()
    print("[metrics] icosphere level 5, ratio 0.05: stage 2 and stages 2+3", flush=True)
    v2, f2 = PaMO((V, F), use_stage1=False, use_stage3=False).run(Vd, Fd, 0.05, 4, 1e-3, 100000, 0)
    v3, f3 = PaMO((V, F), use_stage1=False, use_stage3=True).run(Vd, Fd, 0.05, 4, 1e-3, 100000, 0)
    assert np.array_equal(f3, f2.astype(np.int64))
    # the oracle's stage 3 from the same stage-2 output (stage 2 is bit-exact
    # against the oracle, tested elsewhere), five steps as PaMO.run does
    scale, t = stage3.normalization_transform(V)
    s = dict(gt_V=V.astype(np.float32), gt_F=F.astype(np.int32), V=np.asarray(v2, np.float32),
             F=np.ascontiguousarray(f2, dtype=np.int32), scale=scale, t=t)
    o = oracle_system(s)
    o.step(5)
    vo = (o.q() - t) / scale
    hd2 = metrics.hausdorff(V, F, v2, f2)
    hd3 = metrics.hausdorff(V, F, v3, f3)
    hdo = metrics.hausdorff(V, F, vo, f2)
    cd3, _ = metrics.compute_igl_CD_HD(V, F, v3, f3)
    cdo, _ = metrics.compute_igl_CD_HD(V, F, vo, f2)
    print(f"[metrics] Hausdorff: stage 2 {hd2:.6e}, stages 2+3 {hd3:.6e}, oracle stages 2+3 {hdo:.6e}; "
          f"Chamfer {cd3:.6e} vs oracle {cdo:.6e}", flush=True)
    assert abs(hd3 - hdo) <= 0.01 * hdo
    assert abs(cd3 - cdo) <= 0.01 * cdo


def test_process_eval_and_return_curve():
    """processing.process(..., eval=True, return_curve=True): the distances
    after every Newton iteration, the last one those of the result."""
    from stage3_helpers import scene
    s = scene("plain")
    print("[metrics] process return_curve", flush=True)
    args = [torch.from_numpy(s[k]).cuda() for k in ("gt_V", "gt_F", "V", "F")]
    V, F, CDs, HDs = stage3.process(*args, 2, eval=True, return_curve=True)
    assert len(CDs) == len(HDs) == 20  # 2 steps x 10 Newton iterations
    V0, F0 = stage3.process(*args, 2)
    assert np.array_equal(F, F0) and np.abs(V - V0).max() < 1e-5 * np.abs(V0).max()
    gt_n = s["gt_V"].astype(np.float64) * s["scale"] + s["t"]
    cd, hd = metrics.compute_igl_CD_HD(gt_n, s["gt_F"], V * s["scale"] + s["t"], F)
    assert cd == pytest.approx(CDs[-1], rel=1e-6) and hd == pytest.approx(HDs[-1], rel=1e-6)
    assert HDs[-1] < HDs[0] * 1.5 and np.isfinite(CDs).all()

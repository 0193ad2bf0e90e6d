"""Surface metrics of the stage-3 evaluation
(simp_cuda/safe_project/src/pamo_safe_project/metrics.py), with the distance
queries on the GPU.

- `point_mesh_squared_distance(P, V, F)` -> (sqrD, I, C): libigl 2.5.1's
  `igl.point_mesh_squared_distance`, FP64 on the device
  (`pamo_point_mesh_sqdist`).
- `hausdorff(VA, FA, VB, FB)`: `igl.hausdorff`, vertex-to-surface both ways
  (`pamo_hausdorff`).
- `sample_surface(V, F, count, seed)`: trimesh 4.4.0's area-weighted surface
  sampling (`trimesh.sample.sample_surface`), restated in numpy on the host.
  trimesh is not installed here, so the sample positions are not compared
  with trimesh's own output.
- `compute_igl_CD_HD`, `compute_trimesh_CD_HD`, `compute_min_angle`,
  `compute_min_angle_VF`, `compute_bad_tri_ratio`: the reference's functions
  (metrics.py:7-107), same arguments and results.

Inputs may be numpy arrays or CUDA tensors.  The distance entry points need
the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _hip

_ctx: dict[int, "_Metrics"] = {}


class _Metrics:
    """Owns one pamo_metrics context (tree buffers + one HIP stream)."""

    def __init__(self, device: int):
        p = _hip.c_vp()
        _hip.check(_hip.lib().pamo_metrics_create(ctypes.byref(p), int(device)), "pamo_metrics_create")
        self.ptr = p
        self.device = device

    def __del__(self):
        try:
            if self.ptr:
                _hip.lib().pamo_metrics_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


def _metrics(device: int) -> _Metrics:
    if device not in _ctx:
        _ctx[device] = _Metrics(device)
    return _ctx[device]


def _dev(a, dtype, device):
    if isinstance(a, torch.Tensor):
        t = a.to(device=device, dtype=dtype)
    else:
        t = torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(device)
    return t.reshape(-1, 3).contiguous()


def _device_of(*xs) -> torch.device:
    for x in xs:
        if isinstance(x, torch.Tensor) and x.is_cuda:
            return x.device
    return torch.device("cuda", torch.cuda.current_device())


def point_mesh_squared_distance(P, V, F, return_tensors: bool = False):
    """igl.point_mesh_squared_distance(P, V, F) -> (sqrD, I, C): squared
    distance of each point to the surface (V, F), the nearest face (ties:
    smaller index) and the closest point, in FP64."""
    dev = _device_of(P, V, F)
    m = _metrics(dev.index)
    Pd, Vd, Fd = _dev(P, torch.float64, dev), _dev(V, torch.float64, dev), _dev(F, torch.int32, dev)
    n = Pd.shape[0]
    sq = torch.empty(n, dtype=torch.float64, device=dev)
    idx = torch.empty(n, dtype=torch.int32, device=dev)
    cp = torch.empty((n, 3), dtype=torch.float64, device=dev)
    torch.cuda.current_stream(dev).synchronize()  # inputs written by torch's stream
    _hip.check(_hip.lib().pamo_point_mesh_sqdist(m.ptr, Pd.data_ptr(), n, Vd.data_ptr(), Vd.shape[0], Fd.data_ptr(),
                                                 Fd.shape[0], sq.data_ptr(), idx.data_ptr(), cp.data_ptr()),
               "pamo_point_mesh_sqdist")
    if return_tensors:
        return sq, idx, cp
    return sq.cpu().numpy(), idx.cpu().numpy().astype(np.int64), cp.cpu().numpy()


def hausdorff(VA, FA, VB, FB) -> float:
    """igl.hausdorff(VA, FA, VB, FB): max over the vertices of each mesh of
    the distance to the other mesh's surface."""
    dev = _device_of(VA, FA, VB, FB)
    m = _metrics(dev.index)
    va, fa = _dev(VA, torch.float64, dev), _dev(FA, torch.int32, dev)
    vb, fb = _dev(VB, torch.float64, dev), _dev(FB, torch.int32, dev)
    torch.cuda.current_stream(dev).synchronize()
    d = ctypes.c_double()
    _hip.check(_hip.lib().pamo_hausdorff(m.ptr, va.data_ptr(), va.shape[0], fa.data_ptr(), fa.shape[0],
                                         vb.data_ptr(), vb.shape[0], fb.data_ptr(), fb.shape[0], ctypes.byref(d)),
               "pamo_hausdorff")
    return float(d.value)


def _host(a, dtype):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=dtype)


def face_areas(V, F) -> np.ndarray:
    """trimesh's Trimesh.area_faces: half the norm of the cross product of
    the consecutive edge vectors (v1 - v0, v2 - v1), in FP64."""
    V, F = _host(V, np.float64), _host(F, np.int64)
    tri = V[F]
    vec = np.diff(tri, axis=1)
    cr = np.cross(vec[:, 0], vec[:, 1])
    return np.sqrt((cr ** 2).sum(axis=1)) / 2.0


def sample_surface(V, F, count: int, seed=None):
    """trimesh.sample.sample_surface(mesh, count, seed=seed) (trimesh 4.4.0):
    faces picked with probability proportional to area, points uniform in
    each picked triangle by the fold-back of two uniform numbers.  Returns
    (samples (count, 3) float64, face_index)."""
    V, F = _host(V, np.float64), _host(F, np.int64)
    weight_cum = np.cumsum(face_areas(V, F))
    random = np.random.default_rng(seed).random
    face_pick = random(count) * weight_cum[-1]
    face_index = np.searchsorted(weight_cum, face_pick)
    tri_origins = V[F[:, 0]]
    tri_vectors = V[F[:, 1:]].copy()
    tri_vectors -= np.tile(tri_origins, (1, 2)).reshape((-1, 2, 3))
    tri_origins = tri_origins[face_index]
    tri_vectors = tri_vectors[face_index]
    random_lengths = random((len(tri_vectors), 2, 1))
    random_test = random_lengths.sum(axis=1).reshape(-1) > 1.0
    random_lengths[random_test] -= 1.0
    random_lengths = np.abs(random_lengths)
    sample_vector = (tri_vectors * random_lengths).sum(axis=1)
    return sample_vector + tri_origins, face_index


def _mesh_vf(mesh):
    if isinstance(mesh, (tuple, list)):
        return mesh[0], mesh[1]
    return mesh.vertices, mesh.faces


def compute_trimesh_CD_HD(gt_mesh, gen_mesh, num_mesh_samples=5000, seed=42):
    """metrics.py:7-28: surface samples of each mesh against the other
    surface; (sum of the two mean squared distances, largest distance)."""
    gt_V, gt_F = _mesh_vf(gt_mesh)
    gen_V, gen_F = _mesh_vf(gen_mesh)
    gen_points, _ = sample_surface(gen_V, gen_F, num_mesh_samples, seed=seed)
    gt_points, _ = sample_surface(gt_V, gt_F, num_mesh_samples, seed=seed)
    gt_to_gen = np.sqrt(point_mesh_squared_distance(gen_points, gt_V, gt_F)[0])
    gen_to_gt = np.sqrt(point_mesh_squared_distance(gt_points, gen_V, gen_F)[0])
    chamfer = np.mean(np.square(gt_to_gen)) + np.mean(np.square(gen_to_gt))
    return chamfer, max(np.amax(gt_to_gen), np.amax(gen_to_gt))


def compute_igl_CD_HD(gt_V, gt_F, gen_V, gen_F, num_mesh_samples=5000, seed=42):
    """metrics.py:31-51: Chamfer distance from surface samples (sum of the
    two mean squared distances) and igl.hausdorff of the two meshes."""
    gt_samples, _ = sample_surface(gt_V, gt_F, num_mesh_samples, seed=seed)
    gen_samples, _ = sample_surface(gen_V, gen_F, num_mesh_samples, seed=seed)
    gt_to_gen_sq = point_mesh_squared_distance(gt_samples, gen_V, gen_F)[0]
    gen_to_gt_sq = point_mesh_squared_distance(gen_samples, gt_V, gt_F)[0]
    chamfer = np.mean(gt_to_gen_sq) + np.mean(gen_to_gt_sq)
    return chamfer, hausdorff(gt_V, gt_F, gen_V, gen_F)


def compute_min_angle_VF(V, F):
    """metrics.py:80-98: smallest interior angle (radians) over the faces
    with three distinct corners."""
    V, F = _host(V, np.float64), _host(F, np.int64)
    F = F[(F[:, 0] != F[:, 1]) & (F[:, 1] != F[:, 2]) & (F[:, 2] != F[:, 0])]
    edges = np.concatenate([F[:, None, [0, 1]], F[:, None, [1, 2]], F[:, None, [2, 0]]], axis=1)
    ev = V[edges[:, :, 0]] - V[edges[:, :, 1]]
    el = np.linalg.norm(ev, axis=-1)
    cosines = -np.sum(ev * np.roll(ev, 1, axis=1), axis=-1) / (el * np.roll(el, 1, axis=1))
    return np.min(np.arccos(np.clip(cosines, -1.0, 1.0)))


def compute_min_angle(mesh):
    """metrics.py:54-77: compute_min_angle_VF of a mesh object or (V, F)."""
    V, F = _mesh_vf(mesh)
    return compute_min_angle_VF(V, F)


def internal_angles(V, F) -> np.ndarray:
    """igl.internal_angles(V, F): the angle at each corner (radians) from
    the squared edge lengths by the law of cosines."""
    V, F = _host(V, np.float64), _host(F, np.int64)
    tri = V[F]
    # squared length of the edge opposite each corner
    l2 = np.stack([((tri[:, (k + 1) % 3] - tri[:, (k + 2) % 3]) ** 2).sum(axis=1) for k in range(3)], axis=1)
    out = np.empty_like(l2)
    for k in range(3):
        s1, s2, s3 = l2[:, k], l2[:, (k + 1) % 3], l2[:, (k + 2) % 3]
        out[:, k] = np.arccos((s3 + s2 - s1) / (2.0 * np.sqrt(s3 * s2)))
    return out


def compute_bad_tri_ratio(V, F, thresholds=(1, 5, 10)):
    """metrics.py:101-107: fraction of faces whose smallest angle (degrees)
    is below each threshold."""
    thresholds = np.array(thresholds)
    min_angles = (internal_angles(V, F) * 180 / np.pi).min(axis=1)
    bad = np.sum(min_angles[:, None] < thresholds[None, :], axis=0)
    return bad / _host(F, np.int64).shape[0]

"""Stage 3: barrier-energy safe projection on the GPU.

Host mirror of simp_cuda/safe_project/src/pamo_safe_project/ (``Stage3Config``
config.py:6-80, ``Stage3System`` system.py:20-470, ``process``
processing.py:15-108) over the pamo_s3_* C ABI of libpamo_hip.so.  The
mesh-connectivity preprocessing (unique edges, hinges, Voronoi areas, the
minimum-angle check) runs here in numpy, as the reference runs it in
trimesh / igl; all per-iteration work runs in the library.
"""
from __future__ import annotations

import ctypes
import logging

import numpy as np
import torch

from . import _hip

logger = logging.getLogger(__name__)

_ENERGY_BITS = {
    "Mesh2GTDistanceBvhEnergyCalculator": _hip.PAMO_S3_E_M2GT,
    "ElasticEnergyCalculator": _hip.PAMO_S3_E_ELASTIC,
    "HingeEnergyCalculator": _hip.PAMO_S3_E_HINGE,
    "CollisionBvhEnergyCalculator": _hip.PAMO_S3_E_COLLISION,
    "DummyEnergyCalculator": _hip.PAMO_S3_E_DENSE,
}


class Stage3Config:
    """The fields of the reference's Stage3Config (config.py:6-80) that the
    GPU path reads, with the same defaults.  ``energy_calcs`` holds class
    names; the four default terms are the reference's defaults."""

    def __init__(self):
        self.max_blocks = 1 << 25
        self.n_newton_iters = 10
        self.n_cg_iters = 40
        self.cg_precond = "jacobi"
        self.n_line_search_iters = 10
        self.detect_contact_every = "newton"
        self.energy_calcs = ["Mesh2GTDistanceBvhEnergyCalculator", "ElasticEnergyCalculator",
                             "HingeEnergyCalculator", "CollisionBvhEnergyCalculator"]
        self.system_scale = 1.0
        self.d_hat = 1e-3
        self.coll_stiffness = 1e2
        self.contact_detection_radius = 1e-2
        self.ee_classify_thres = 1e-3
        self.ccd_slackness = 0.7
        self.ccd_thickness = 1e-6
        self.ccd_max_iters = 100
        self.spd_max_iters = 3
        self.elas_young_modulus = 1e-1
        self.elas_poisson_ratio = 0.0
        self.hinge_stiffness = 1e-2
        self.mesh2gt_dist_stiffness = 1e3
        self.min_angle_thres = 0.0

    def energy_mask(self) -> int:
        m = 0
        for e in self.energy_calcs:
            name = e if isinstance(e, str) else getattr(e, "__name__", str(e))
            if name not in _ENERGY_BITS:
                raise ValueError(f"energy term {name} is not provided by pamo_amd")
            m |= _ENERGY_BITS[name]
        return m

    def to_params(self) -> _hip.S3Params:
        if self.cg_precond != "jacobi":
            raise ValueError("only the Jacobi preconditioner is provided")
        if self.detect_contact_every != "newton":
            raise ValueError("contacts are detected every Newton iteration (the reference default)")
        if self.system_scale != 1.0:
            raise ValueError("system_scale other than 1.0 is not provided")
        return _hip.S3Params(
            energy_mask=self.energy_mask(), n_newton_iters=self.n_newton_iters, n_cg_iters=self.n_cg_iters,
            n_line_search_iters=self.n_line_search_iters, max_blocks=self.max_blocks,
            ccd_max_iters=self.ccd_max_iters, d_hat=self.d_hat, coll_stiffness=self.coll_stiffness,
            contact_detection_radius=self.contact_detection_radius, ee_classify_thres=self.ee_classify_thres,
            ccd_slackness=self.ccd_slackness, ccd_thickness=self.ccd_thickness,
            elas_young_modulus=self.elas_young_modulus, elas_poisson_ratio=self.elas_poisson_ratio,
            hinge_stiffness=self.hinge_stiffness, mesh2gt_dist_stiffness=self.mesh2gt_dist_stiffness,
            min_angle_thres=float(self.min_angle_thres))


# ---------------------------------------------------------------- host preprocessing
def normalization_transform(V):
    """processing.py:15-20: centre = vertex mean, scale = 1 / max extent."""
    V = np.asarray(V, dtype=np.float64)
    center = np.mean(V, axis=0)
    extents = np.max(V, axis=0) - np.min(V, axis=0)
    scale = 1.0 / np.max(extents)
    return scale, -center * scale


def edges_unique(F):
    """trimesh ``edges_unique`` (used at system.py:179): every undirected edge
    once as (min, max), ordered by (max, min) -- the order that trimesh's
    row hash (min + max << 32) gives under np.unique."""
    F = np.asarray(F, dtype=np.int64)
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    e = np.sort(e, axis=1)
    key = e[:, 0] + (e[:, 1] << 32)
    _, first = np.unique(key, return_index=True)
    return e[first].astype(np.int32)


def hinges(F):
    """hinge_preprocess_slow_kernel (hinge_energy.py:22-57): for each face t0
    and its directed edge (a, b) with a < b, the face t1 holding (b, a) gives
    the hinge (opposite in t0, a, b, opposite in t1).  Listed in (t0, edge)
    order instead of the reference's atomic-counter order."""
    F = np.asarray(F, dtype=np.int64)
    nF = len(F)
    a = np.concatenate([F[:, 0], F[:, 1], F[:, 2]])
    b = np.concatenate([F[:, 1], F[:, 2], F[:, 0]])
    o = np.concatenate([F[:, 2], F[:, 0], F[:, 1]])
    t = np.tile(np.arange(nF), 3)
    e = np.repeat(np.arange(3), nF)
    nV = int(F.max()) + 1 if nF else 0
    key = a * nV + b
    rkey = b * nV + a
    order = np.argsort(key, kind="stable")
    sk = key[order]
    pos = np.searchsorted(sk, rkey)
    pos_c = np.minimum(pos, len(sk) - 1)
    has = (pos < len(sk)) & (sk[pos_c] == rkey) & (a < b)
    idx = np.nonzero(has)[0]
    idx = idx[np.lexsort((e[idx], t[idx]))]
    other = order[pos_c[idx]]
    return np.stack([o[idx], a[idx], b[idx], o[other]], axis=1).astype(np.int32)


def voronoi_areas(V, F):
    """Diagonal of the Voronoi mass matrix (igl.massmatrix VORONOI, used at
    system.py:226-228): Meyer et al. 2003 mixed areas -- circumcentric
    Voronoi regions for non-obtuse triangles, area/2 to the obtuse corner and
    area/4 to the others otherwise."""
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F, dtype=np.int64)
    p = V[F]
    # l[:, k] = length of the edge opposite corner k
    l = np.stack([np.linalg.norm(p[:, 1] - p[:, 2], axis=1), np.linalg.norm(p[:, 2] - p[:, 0], axis=1),
                  np.linalg.norm(p[:, 0] - p[:, 1], axis=1)], axis=1)
    l2 = l * l
    cos = np.stack([(l2[:, 1] + l2[:, 2] - l2[:, 0]) / (2 * l[:, 1] * l[:, 2]),
                    (l2[:, 2] + l2[:, 0] - l2[:, 1]) / (2 * l[:, 2] * l[:, 0]),
                    (l2[:, 0] + l2[:, 1] - l2[:, 2]) / (2 * l[:, 0] * l[:, 1])], axis=1)
    area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    bary = cos * l
    with np.errstate(invalid="ignore", divide="ignore"):
        bary = bary / bary.sum(axis=1, keepdims=True)
    part = bary * area[:, None]
    quad = np.stack([(part[:, 1] + part[:, 2]) * 0.5, (part[:, 2] + part[:, 0]) * 0.5,
                     (part[:, 0] + part[:, 1]) * 0.5], axis=1)
    for k in range(3):
        obt = cos[:, k] < 0
        quad[obt, k] = 0.5 * area[obt]
        quad[obt, (k + 1) % 3] = 0.25 * area[obt]
        quad[obt, (k + 2) % 3] = 0.25 * area[obt]
    out = np.zeros(len(V))
    np.add.at(out, F.reshape(-1), quad.reshape(-1))
    return out


def min_internal_angle_deg(V, F):
    """min of igl.internal_angles in degrees (system.py:220-224)."""
    V = np.asarray(V, dtype=np.float64)
    p = V[np.asarray(F, dtype=np.int64)]
    ang = []
    for k in range(3):
        u = p[:, (k + 1) % 3] - p[:, k]
        w = p[:, (k + 2) % 3] - p[:, k]
        lu, lw = np.linalg.norm(u, axis=1), np.linalg.norm(w, axis=1)
        with np.errstate(invalid="ignore", divide="ignore"):
            c = np.einsum("ij,ij->i", u, w) / (lu * lw)
        ang.append(np.arccos(np.clip(c, -1.0, 1.0)))
    return float(np.min(np.stack(ang)) * 180.0 / np.pi) if len(p) else 180.0


# ---------------------------------------------------------------- system
def _dev_tensor(a, dtype):
    if isinstance(a, torch.Tensor) and a.is_cuda:
        return a.to(dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


class Stage3System:
    """One pamo_s3 context: device buffers sized to the largest mesh seen,
    reused across calls like the reference's Stage3System (system.py:20-70)."""

    def __init__(self, config: Stage3Config | None = None, device=None):
        self.config = config if config is not None else Stage3Config()
        if device is None:
            device = torch.cuda.current_device()
        elif isinstance(device, str):
            device = torch.device(device).index or 0
        self.device = int(device)
        self.ptr = ctypes.c_void_p()
        self._params = self.config.to_params()
        _hip.check(_hip.lib().pamo_s3_create(ctypes.byref(self.ptr), self.device, ctypes.byref(self._params)),
                   "pamo_s3_create")
        self.n_particles = 0
        self.do_nothing = False
        self.scale = 1.0
        self.t = np.zeros(3)

    def close(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            _hip.lib().pamo_s3_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        """system.py:72-80: nothing to release; buffers are reused."""
        self.do_nothing = False

    def set_target(self, gt_V, gt_F, scale: float, t):
        """The input surface, normalised on the device in FP64."""
        v = _dev_tensor(gt_V, torch.float32)
        f = _dev_tensor(gt_F, torch.int32)
        tt = (ctypes.c_double * 3)(*[float(x) for x in t])
        torch.cuda.current_stream().synchronize()
        _hip.check(_hip.lib().pamo_s3_set_target(self.ptr, v.data_ptr(), int(v.shape[0]), f.data_ptr(),
                                                 int(f.shape[0]), float(scale), tt), "pamo_s3_set_target")

    def set_mesh(self, V, F, scale: float, t):
        """Stage3System.register_mesh (system.py:167-272) for the mesh to
        project; V is in the input frame and is normalised with (scale, t).
        The connectivity preparation (edges, hinges, Voronoi areas, minimum
        angle) runs in the library (pamo_s3_prepare_host)."""
        v = _dev_tensor(V, torch.float32)
        f = _dev_tensor(F, torch.int32)
        tt = (ctypes.c_double * 3)(*[float(x) for x in t])
        torch.cuda.current_stream().synchronize()
        _hip.check(_hip.lib().pamo_s3_set_mesh_device(self.ptr, v.data_ptr(), int(v.shape[0]), f.data_ptr(),
                                                      int(f.shape[0]), float(scale), tt), "pamo_s3_set_mesh_device")
        st = self.stats()
        self.n_particles = int(v.shape[0])
        self.n_hinges = st["n_hinges"]
        self.do_nothing = st["do_nothing"]
        self.scale, self.t = float(scale), np.asarray(t, dtype=np.float64)

    def merge_vertices(self, V, F):
        """trimesh.Trimesh(vertices=V, faces=F) with its default
        process=True, as PaMO.run builds the stage-3 input
        (pamo/__init__.py:194; trimesh 4.4.0 merge_vertices): referenced
        vertices whose coordinates times 1e8, rounded half to even, agree
        merge into the first of them; unreferenced vertices are dropped.
        Returns device (float32 vertices, int32 faces)."""
        v = _dev_tensor(V, torch.float32)
        f = _dev_tensor(F, torch.int32).clone()
        out = torch.empty_like(v)
        n = ctypes.c_int(0)
        torch.cuda.current_stream().synchronize()
        _hip.check(_hip.lib().pamo_s3_merge_vertices(self.ptr, v.data_ptr(), int(v.shape[0]), f.data_ptr(),
                                                     int(f.shape[0]), out.data_ptr(), ctypes.byref(n)),
                   "pamo_s3_merge_vertices")
        return out[: n.value], f

    def set_mesh_prepared(self, V, F, E, H, vor, do_nothing, scale: float, t):
        """set_mesh with the connectivity given (tests)."""
        v = _dev_tensor(V, torch.float32)
        f = _dev_tensor(F, torch.int32) if len(F) else None
        E = np.ascontiguousarray(E, dtype=np.int32)
        H = np.ascontiguousarray(H, dtype=np.int32)
        vor = np.ascontiguousarray(vor, dtype=np.float32)
        tt = (ctypes.c_double * 3)(*[float(x) for x in t])
        torch.cuda.current_stream().synchronize()
        _hip.check(_hip.lib().pamo_s3_set_mesh(
            self.ptr, v.data_ptr(), int(v.shape[0]), f.data_ptr() if f is not None else None, len(F), float(scale),
            tt, E.ctypes.data if len(E) else None, len(E), H.ctypes.data if len(H) else None, len(H),
            vor.ctypes.data, int(bool(do_nothing))), "pamo_s3_set_mesh")
        self.n_particles = int(v.shape[0])
        self.n_hinges = len(H)
        self.do_nothing = bool(do_nothing)
        self.scale, self.t = float(scale), np.asarray(t, dtype=np.float64)

    def register_mesh(self, V_gt, F_gt, V, F):
        """Stage3System.register_mesh with already-normalised inputs (scale 1)."""
        self.set_target(V_gt, F_gt, 1.0, np.zeros(3))
        self.set_mesh(V, F, 1.0, np.zeros(3))

    def step(self, n: int = 1, eval=False, eval_CDs=None, eval_HDs=None, gt_V=None, gt_F=None, eval_F=None):
        """Stage3System.step (system.py:411-470), n times.  With eval=True
        the Chamfer and Hausdorff distances of the current vertices (with
        faces eval_F) to (gt_V, gt_F), all normalised, are appended to
        eval_CDs / eval_HDs after every Newton iteration (system.py:466-470),
        computed by pamo_amd.metrics.compute_igl_CD_HD."""
        if not eval:
            _hip.check(_hip.lib().pamo_s3_step(self.ptr, int(n)), "pamo_s3_step")
            return
        from .metrics import compute_igl_CD_HD
        failed = []

        def on_newton(it, user):
            try:
                cd, hd = compute_igl_CD_HD(gt_V, gt_F, self.get_vertices(), eval_F)
                eval_CDs.append(cd)
                eval_HDs.append(hd)
            except Exception as e:  # raised after the step returns
                failed.append(e)

        cb = _hip.S3_NEWTON_FN(on_newton)
        for _ in range(int(n)):
            _hip.check(_hip.lib().pamo_s3_step_eval(self.ptr, cb, None), "pamo_s3_step_eval")
            if failed:
                raise failed[0]

    def get_vertices(self) -> np.ndarray:
        """Normalised positions (float32 values as float64), system.py:164-165."""
        out = np.zeros((self.n_particles, 3), dtype=np.float64)
        _hip.check(_hip.lib().pamo_s3_get_normalized(self.ptr, out.ctypes.data), "pamo_s3_get_normalized")
        return out

    def stats(self) -> dict:
        o = np.zeros(8)
        _hip.check(_hip.lib().pamo_s3_stats(self.ptr, o.ctypes.data), "pamo_s3_stats")
        return {"energy": o[0], "contacts": int(o[1]), "newton": int(o[2]), "cg": int(o[3]), "line_search": int(o[4]),
                "radius": o[5], "do_nothing": bool(o[6]), "n_hinges": int(o[7])}

    def active_contacts(self) -> int:
        """Active contacts summed over the Newton iterations since set_mesh
        (measurement)."""
        n = ctypes.c_longlong(0)
        _hip.check(_hip.lib().pamo_s3_active_contacts(self.ptr, ctypes.byref(n)), "pamo_s3_active_contacts")
        return int(n.value)

    def set_cg_graph(self, on: bool):
        """Replay the PCG loop from a captured HIP graph (default) or launch
        its kernels one by one; results do not change."""
        _hip.check(_hip.lib().pamo_s3_set_option(self.ptr, _hip.PAMO_S3_OPT_CG_GRAPH, 1 if on else 0),
                   "pamo_s3_set_option")

    # -- test access (no counterpart in the reference's API)
    def set_dense(self, A, b, c0):
        A = np.ascontiguousarray(A, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        _hip.check(_hip.lib().pamo_s3_set_dense(self.ptr, A.ctypes.data, b.ctypes.data, float(np.asarray(c0).ravel()[0]),
                                                self.n_particles), "pamo_s3_set_dense")

    def debug(self, what: int, inp=None, cap: int | None = None) -> np.ndarray:
        n3 = 3 * self.n_particles
        if cap is None:
            cap = {_hip.PAMO_S3_DBG_CONTACTS: 5 * (1 << 22), _hip.PAMO_S3_DBG_ENERGY: 6,
                   _hip.PAMO_S3_DBG_DIFF: 2 * n3, _hip.PAMO_S3_DBG_CCD: n3 + 1,
                   _hip.PAMO_S3_DBG_HINGE: 144 * max(getattr(self, "n_hinges", 0), 1)}.get(what, n3)
        out = np.zeros(max(cap, 1), dtype=np.float64)
        n = ctypes.c_longlong(0)
        if inp is not None:
            inp = np.ascontiguousarray(inp, dtype=np.float64).reshape(-1)
        _hip.check(_hip.lib().pamo_s3_debug(self.ptr, int(what), inp.ctypes.data if inp is not None else None,
                                            len(inp) if inp is not None else 0, out.ctypes.data, cap,
                                            ctypes.byref(n)), "pamo_s3_debug")
        if what == _hip.PAMO_S3_DBG_CONTACTS:
            return out[: 5 * n.value].reshape(-1, 5).astype(np.int64)
        return out[: n.value]


def normalization_transform_device(system: "Stage3System", V: torch.Tensor):
    """normalization_transform of a device float32 (n, 3) tensor, in FP64 on
    the device: the mean as numpy's sequential sum over the rows, so the
    result equals numpy's on the same values bit for bit."""
    v = V.to(torch.float32).contiguous()
    sc = ctypes.c_double()
    t = (ctypes.c_double * 3)()
    torch.cuda.current_stream().synchronize()
    _hip.check(_hip.lib().pamo_s3_normalization(system.ptr, v.data_ptr(), int(v.shape[0]), ctypes.byref(sc), t),
               "pamo_s3_normalization")
    return sc.value, np.array([t[0], t[1], t[2]])


def normalized_target(gt_V):
    """The normalised input surface of process (processing.py:57-58) from
    the caller's float64 vertices, as the reference computes it: (scale, t)
    by numpy in float64, gt_V * scale + t in float64, then float32 on the
    device.  Returns (scale, t, device tensor) for process(target=...)."""
    V = np.asarray(gt_V, dtype=np.float64)
    scale, t = normalization_transform(V)
    gn = torch.from_numpy((V * scale + t).astype(np.float32)).cuda()
    return scale, t, gn


def process(gt_V, gt_F, stage2_V, stage2_F, n_iters, system: Stage3System | None = None,
            config: Stage3Config | None = None, eval=False, return_curve=False, target=None):
    """processing.py:23-108: normalise by the input surface, n_iters steps,
    map back.  Inputs may be numpy arrays or CUDA tensors (kept on the
    device).  Returns (float64 vertices, faces) as numpy, and with
    return_curve also the Chamfer / Hausdorff distances after every Newton
    iteration (normalised frame).  eval logs the distances before the first
    step and after each step (processing.py:66-97).  target, from
    normalized_target(gt_V), supplies the normalisation and the normalised
    input surface already on the device (a caller projecting several meshes
    onto one input prepares it once); gt_V is then only used for eval."""
    if system is None:
        system = Stage3System(config if config is not None else Stage3Config())
    else:
        system.clear()
    if target is not None:
        scale, t, gn = target
        system.set_target(gn, gt_F, 1.0, np.zeros(3))
    else:
        if isinstance(gt_V, torch.Tensor) and gt_V.is_cuda:
            scale, t = normalization_transform_device(system, gt_V)
        else:
            scale, t = normalization_transform(np.asarray(gt_V))
        system.set_target(gt_V, gt_F, scale, t)
    system.set_mesh(stage2_V, stage2_F, scale, t)
    F = stage2_F.cpu().numpy() if isinstance(stage2_F, torch.Tensor) else np.asarray(stage2_F)
    if not (eval or return_curve):
        system.step(int(n_iters))
        V = (system.get_vertices() - t) / scale
        return V, F
    from .metrics import compute_igl_CD_HD
    gt_n = _host64(gt_V) * scale + t
    gt_Fn = _host64(gt_F).astype(np.int64)
    if eval:
        cd, hd = compute_igl_CD_HD(gt_n, gt_Fn, system.get_vertices(), F)
        logger.info(f"Initial CD: {cd:.3e}, HD: {hd:.3e}")
    CDs, HDs = [], []
    for i in range(int(n_iters)):
        if return_curve:
            system.step(1, eval=True, eval_CDs=CDs, eval_HDs=HDs, gt_V=gt_n, gt_F=gt_Fn, eval_F=F)
        else:
            system.step(1)
        if eval:
            cd, hd = compute_igl_CD_HD(gt_n, gt_Fn, system.get_vertices(), F)
            logger.info(f"Iter {i}: CD: {cd:.3e}, HD: {hd:.3e}")
    V = (system.get_vertices() - t) / scale
    if return_curve:
        return V, F, CDs, HDs
    return V, F


def _host64(a):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.float64)

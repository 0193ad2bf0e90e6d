"""ctypes wrapper of the stage-3 CPU oracle (oracle/stage3_oracle.cpp).

Test infrastructure: imported only by tests/ and bench.py's CPU-baseline leg,
as the checker of the HIP path.  ``fp64=True`` holds the state in double
(the error reference), ``fp64=False`` in float (tracks the GPU's rounding of
positions).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import oracle as _orc

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libstage3_oracle.so")
_lib = None

c_int, c_double, c_ll, c_vp = ctypes.c_int, ctypes.c_double, ctypes.c_longlong, ctypes.c_void_p

E_M2GT, E_ELASTIC, E_HINGE, E_COLLISION, E_DENSE = 1, 2, 4, 8, 16
E_DEFAULT = 15
BLOCK_PT, BLOCK_EE = 3, 4


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            _orc.build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "s3o_create": (c_vp, [c_int, c_vp]),
            "s3o_destroy": (None, [c_vp]),
            "s3o_set_target": (None, [c_vp, c_vp, c_int, c_vp, c_int]),
            "s3o_set_mesh": (None, [c_vp, c_vp, c_int, c_vp, c_int, c_vp, c_int, c_vp, c_int, c_vp, c_int]),
            "s3o_set_dense": (None, [c_vp, c_vp, c_vp, c_double]),
            "s3o_set_q": (None, [c_vp, c_vp]),
            "s3o_get_q": (None, [c_vp, c_vp]),
            "s3o_target": (None, [c_vp, c_vp]),
            "s3o_contacts": (c_ll, [c_vp, c_vp, c_ll]),
            "s3o_energy": (None, [c_vp, c_vp]),
            "s3o_diff": (None, [c_vp, c_vp, c_vp]),
            "s3o_hvp": (None, [c_vp, c_vp, c_vp]),
            "s3o_solve": (None, [c_vp, c_vp]),
            "s3o_ccd": (None, [c_vp, c_vp, c_vp, c_vp]),
            "s3o_hinge_blocks": (None, [c_vp, c_vp, c_int]),
            "s3o_step": (None, [c_vp, c_int, c_int]),
            "s3o_set_targets": (None, [c_vp, c_vp]),
            "s3o_stats": (None, [c_vp, c_vp]),
            "s3o_pt_classify": (c_int, [c_vp]),
            "s3o_ee_classify": (c_int, [c_vp, c_double]),
            "s3o_pair_distance": (c_double, [c_vp, c_int, c_int]),
            "s3o_pair_grad": (None, [c_vp, c_int, c_int, c_vp]),
            "s3o_hinge_angle": (c_double, [c_vp]),
            "s3o_closest_on_tri": (None, [c_vp, c_vp, c_vp]),
            "s3o_barrier": (c_double, [c_double, c_double, c_int]),
            "s3o_spd_project_hinge": (None, [c_vp, c_vp]),
            "s3o_elastic_grad_basis": (None, [c_vp, c_vp, c_vp, c_vp]),
            "s3o_energy_at": (c_double, [c_vp, c_vp, c_int]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


DEFAULTS = dict(mask=E_DEFAULT, n_newton=10, n_cg=40, n_ls=10, max_blocks=1 << 25, ccd_max_iters=100, d_hat=1e-3,
                kappa=1e2, radius=1e-2, ee_thres=1e-3, ccd_slack=0.7, ccd_thick=1e-6, young=1e-1, poisson=0.0,
                k_hinge=1e-2, k_m2gt=1e3)
_ORDER = ["mask", "n_newton", "n_cg", "n_ls", "max_blocks", "ccd_max_iters", "d_hat", "kappa", "radius", "ee_thres",
          "ccd_slack", "ccd_thick", "young", "poisson", "k_hinge", "k_m2gt"]


class System:
    """One oracle system; arrays are normalised-frame float64."""

    def __init__(self, fp64: bool = True, **params):
        p = dict(DEFAULTS)
        p.update(params)
        self._p = _d([p[k] for k in _ORDER])
        self.h = lib().s3o_create(int(fp64), self._p.ctypes.data)
        self.nV = 0

    def __del__(self):
        try:
            lib().s3o_destroy(self.h)
        except Exception:
            pass

    def set_target(self, V, F):
        V, F = _d(V), _i(F)
        lib().s3o_set_target(self.h, V.ctypes.data, len(V), F.ctypes.data, len(F))

    def set_mesh(self, V, F, E, H, vor, do_nothing=False):
        V, F, E, H, vor = _d(V), _i(F).reshape(-1, 3), _i(E).reshape(-1, 2), _i(H).reshape(-1, 4), _d(vor)
        self.nV, self.nH = len(V), len(H)
        lib().s3o_set_mesh(self.h, V.ctypes.data, len(V), F.ctypes.data, len(F), E.ctypes.data, len(E),
                           H.ctypes.data, len(H), vor.ctypes.data, int(do_nothing))

    def set_dense(self, A, b, c):
        A, b = _d(A), _d(b)
        lib().s3o_set_dense(self.h, A.ctypes.data, b.ctypes.data, float(np.asarray(c).ravel()[0]))

    def set_q(self, q):
        q = _d(q)
        lib().s3o_set_q(self.h, q.ctypes.data)

    def q(self):
        o = np.zeros((self.nV, 3))
        lib().s3o_get_q(self.h, o.ctypes.data)
        return o

    def target(self):
        o = np.zeros((self.nV, 3))
        lib().s3o_target(self.h, o.ctypes.data)
        return o

    def contacts(self, cap=1 << 22):
        o = np.zeros((cap, 5), dtype=np.int32)
        n = lib().s3o_contacts(self.h, o.ctypes.data, cap)
        return o[: min(n, cap)].astype(np.int64)

    def energy(self):
        o = np.zeros(6)
        lib().s3o_energy(self.h, o.ctypes.data)
        return o

    def diff(self):
        g, d = np.zeros((self.nV, 3)), np.zeros((self.nV, 3))
        lib().s3o_diff(self.h, g.ctypes.data, d.ctypes.data)
        return g, d

    def hvp(self, dx):
        dx = _d(dx)
        o = np.zeros((self.nV, 3))
        lib().s3o_hvp(self.h, dx.ctypes.data, o.ctypes.data)
        return o

    def solve(self):
        o = np.zeros((self.nV, 3))
        lib().s3o_solve(self.h, o.ctypes.data)
        return o

    def ccd(self, p):
        p = _d(p)
        o = np.zeros((self.nV, 3))
        st = ctypes.c_double(0)
        lib().s3o_ccd(self.h, p.ctypes.data, o.ctypes.data, ctypes.byref(st))
        return o, st.value

    def hinge_blocks(self, projected=True):
        o = np.zeros((self.nH, 144))
        lib().s3o_hinge_blocks(self.h, o.ctypes.data, int(projected))
        return o

    def step(self, n=1, keep_target=False):
        lib().s3o_step(self.h, int(n), int(keep_target))

    def set_targets(self, t):
        t = _d(t)
        lib().s3o_set_targets(self.h, t.ctypes.data)

    def stats(self):
        o = np.zeros(4)
        lib().s3o_stats(self.h, o.ctypes.data)
        return {"energy": o[0], "contacts": int(o[1]), "newton": int(o[2]), "radius": o[3]}

    def elastic_grad_basis(self, x, basis_x):
        x, b = _d(x), _d(basis_x)
        o = np.zeros((self.nV, 3))
        lib().s3o_elastic_grad_basis(self.h, x.ctypes.data, b.ctypes.data, o.ctypes.data)
        return o

    def energy_at(self, x, mask):
        x = _d(x)
        return lib().s3o_energy_at(self.h, x.ctypes.data, int(mask))


# single functions
def pt_classify(x12):
    return lib().s3o_pt_classify(_d(x12).ctypes.data)


def ee_classify(x12, thres=1e-3):
    return lib().s3o_ee_classify(_d(x12).ctypes.data, float(thres))


def pair_distance(x12, is_pt, t):
    return lib().s3o_pair_distance(_d(x12).ctypes.data, int(is_pt), int(t))


def pair_grad(x12, is_pt, t):
    o = np.zeros(12)
    lib().s3o_pair_grad(_d(x12).ctypes.data, int(is_pt), int(t), o.ctypes.data)
    return o


def hinge_angle(x12):
    return lib().s3o_hinge_angle(_d(x12).ctypes.data)


def closest_on_tri(p, tri):
    o = np.zeros(3)
    lib().s3o_closest_on_tri(_d(p).ctypes.data, _d(tri).ctypes.data, o.ctypes.data)
    return o


def barrier(d, dh, order=0):
    return lib().s3o_barrier(float(d), float(dh), int(order))


def spd_project_hinge(H):
    H = _d(H).reshape(144)
    o = np.zeros(144)
    lib().s3o_spd_project_hinge(H.ctypes.data, o.ctypes.data)
    return o.reshape(12, 12)

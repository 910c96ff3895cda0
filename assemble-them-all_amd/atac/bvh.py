"""BVH contact variant: a triangle-mesh body with exact signed distance, the
reference's BodyBVHObj (simulation/redmax/Body/BodyBVHObj.cpp) and
ForceGeneralBVHContact (simulation/redmax/Force/ForceGeneralBVHContact.cpp).

The mesh is prepared once (BVH, face / edge / vertex pseudonormals; atac_mesh_create);
distance queries (BodyBVHObj::distance_normal_parallel) and force evaluations
(detection + wrench, p6 column, Km / Dm) run on the GPU through the C ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, lib
from .batch import make_instance


def _vertices(V):
    v = np.asarray(V, dtype=np.float64)
    if v.ndim != 2 or 3 not in v.shape:
        raise ValueError("vertices must be 3 x n or n x 3")
    if v.shape[0] == 3 and v.shape[1] != 3:
        v = v.T
    return np.ascontiguousarray(v)


def _faces(F):
    f = np.asarray(F)
    if f.ndim != 2 or 3 not in f.shape:
        raise ValueError("faces must be 3 x m or m x 3")
    if f.shape[0] == 3 and f.shape[1] != 3:
        f = f.T
    return np.ascontiguousarray(f, dtype=np.int64)


class BVHMesh:
    """A triangle mesh in its body frame, resident on the GPU with its BVH."""

    def __init__(self, V, F):
        self.V = _vertices(V)
        self.F = _faces(F)
        h = C.c_void_p()
        check(lib().atac_mesh_create(self.V.ctypes.data_as(C.c_void_p), len(self.V), self.F.ctypes.data_as(C.c_void_p),
                                     len(self.F), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def distance_normal(self, xw, E2=None, col_th=0.0):
        """BodyBVHObj::distance_normal_parallel for world points xw (n x 3) with the
        body at E2 (4x4, identity by default): (d n, normal n x 3, face n); d =
        signed distance + col_th inside the vertices' bounding box, 10 outside."""
        x = np.ascontiguousarray(np.asarray(xw, dtype=np.float64).reshape(-1, 3))
        E = np.ascontiguousarray(np.eye(4) if E2 is None else np.asarray(E2, dtype=np.float64).reshape(4, 4))
        n = len(x)
        d = np.empty(n)
        nrm = np.empty((n, 3))
        face = np.empty(n, dtype=np.int32)
        check(lib().atac_mesh_distance_normal_host(self._h, L.dptr(E), float(col_th), x.ctypes.data_as(C.c_void_p), n,
                                                   d.ctypes.data_as(C.c_void_p), nrm.ctypes.data_as(C.c_void_p),
                                                   face.ctypes.data_as(C.c_void_p)))
        return d, nrm, face

    def force(self, xi, E1, phi1, E2, phi2, col_th, kn, kt, mu, damping, scale=1.0, contact_scale=1.0,
              deriv=False, p6=False):
        """ForceGeneralBVHContact evaluation for the contact body's points xi (n x 3,
        body frame): dict with fm (12), n_active, and with deriv Km, Dm (12 x 12),
        with p6 the contact-scale column dfm_dp6 (12)."""
        x = np.ascontiguousarray(np.asarray(xi, dtype=np.float64).reshape(-1, 3))
        inst = make_instance(0, 0, 0, E1, phi1, E2, phi2, col_th, kn, kt, mu, damping, scale, contact_scale)
        flags = (L.EVAL_DERIV if deriv else 0) | (L.EVAL_P6 if p6 else 0)
        fm = np.zeros(12)
        Km = np.zeros(144) if deriv else None
        Dm = np.zeros(144) if deriv else None
        p6v = np.zeros(12) if p6 else None
        na = C.c_int64(0)
        vp = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
        check(lib().atac_bvh_force_host(self._h, C.byref(inst), x.ctypes.data_as(C.c_void_p), len(x), flags, vp(fm),
                                        vp(Km), vp(Dm), vp(p6v), C.byref(na)))
        out = {"fm": fm, "n_active": int(na.value)}
        if deriv:
            out["Km"] = Km.reshape(12, 12)
            out["Dm"] = Dm.reshape(12, 12)
        if p6:
            out["dfm_dp6"] = p6v
        return out

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().atac_mesh_destroy(self._h)
                self._h = None
        except Exception:
            pass

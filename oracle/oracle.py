"""ctypes front end of the FP64 CPU oracle (oracle/atac_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker. The product path never imports it.

Every function here forwards to the C restatement; the reference routine each one
follows is cited in atac_oracle.cpp (BodySDFObj.cpp:123-377,
CollisionDetection.cpp:124-147, ForceGeneralSDFContact.cpp:147-447 and the p6 lines)
and levelset_oracle.cpp (sdfgen makelevelset3.cpp:6-230, BodySDFObj.cpp:38-77,
Mesh.cpp:54-95).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

MODE_REFERENCE = 0  # FD stencils, eps = 1e-5 (BodySDFObj.cpp:176-205)
MODE_ANALYTIC = 1   # closed-form trilinear partials

FLAG_DERIV = 1      # Km / Dm
FLAG_P6 = 2         # contact-scale column of dfm_dp


class OracleGrid(C.Structure):
    _fields_ = [
        ("values", C.POINTER(C.c_float)),
        ("n", C.c_int32 * 3),
        ("min_box", C.c_double * 3),
        ("max_box", C.c_double * 3),
        ("dxyz", C.c_double * 3),
    ]


class OracleInstance(C.Structure):
    _fields_ = [
        ("grid", C.c_int32),
        ("mode", C.c_int32),
        ("points", C.POINTER(C.c_double)),
        ("n_points", C.c_int64),
        ("E1", C.c_double * 16),
        ("phi1", C.c_double * 6),
        ("E2", C.c_double * 16),
        ("phi2", C.c_double * 6),
        ("col_th", C.c_double),
        ("kn", C.c_double),
        ("kt", C.c_double),
        ("mu", C.c_double),
        ("damping", C.c_double),
        ("scale", C.c_double),
        ("contact_scale", C.c_double),
    ]


def build() -> None:
    """Compile liboracle.so from oracle/atac_oracle.cpp (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.oracle_query_sdf.argtypes = [C.POINTER(OracleGrid), dp, C.c_int64, C.c_int, dp]
        L.oracle_query_dsdf.argtypes = [C.POINTER(OracleGrid), dp, C.c_int64, C.c_int, dp, dp]
        L.oracle_distance.argtypes = [C.POINTER(OracleGrid), dp, C.c_double, dp, C.c_int64, dp]
        L.oracle_collision.argtypes = [C.POINTER(OracleGrid), dp, dp, C.c_double, C.c_int, dp, dp, C.c_int64, dp, dp]
        L.oracle_eval_instance.argtypes = [C.POINTER(OracleGrid), C.POINTER(OracleInstance), C.c_uint, dp, dp, dp, dp,
                                           C.POINTER(C.c_uint8), C.POINTER(C.c_int64), dp]
        L.oracle_eval_batch.argtypes = [C.POINTER(OracleGrid), C.POINTER(OracleInstance), C.c_int64, C.c_uint, C.c_int,
                                        dp, dp, dp, dp, C.POINTER(C.c_int64), dp]
        L.oracle_force_fixed_contacts.argtypes = [C.POINTER(OracleGrid), dp, C.c_int64, dp, dp, dp, dp, C.c_double, C.c_int,
                                                  C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                                  C.c_uint, dp, dp, dp, dp]
        L.oracle_se3_exp.argtypes = [dp, dp]
        vp = C.c_void_p
        fp = C.POINTER(C.c_float)
        L.oracle_make_level_set3.argtypes = [vp, C.c_int64, vp, C.c_int64, fp, C.c_float, C.c_float, C.c_float,
                                             C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]
        L.oracle_sdf_mesh_grid.argtypes = [dp, C.c_int64, C.c_double, C.c_int, fp, fp, dp, C.POINTER(C.c_int)]
        L.oracle_mesh_normals.argtypes = [vp, C.c_int64, vp, C.c_int64, vp, vp, vp, vp, C.POINTER(C.c_int64)]
        L.oracle_mesh_distance_normal.argtypes = [vp, C.c_int64, vp, C.c_int64, dp, C.c_double, vp, C.c_int64, vp, vp, vp]
        L.oracle_bvh_eval.argtypes = [vp, C.c_int64, vp, C.c_int64, C.POINTER(OracleInstance), C.c_uint, dp, dp, dp, dp,
                                      C.POINTER(C.c_uint8), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class GridRef:
    """Keeps the numpy buffer of a grid alive next to its ctypes descriptor."""

    def __init__(self, values: np.ndarray, min_box, max_box, dxyz):
        v = np.ascontiguousarray(values, dtype=np.float32)
        assert v.ndim == 3, "values must be (nk, nj, ni)"
        self.values = v
        nk, nj, ni = v.shape
        self.c = OracleGrid()
        self.c.values = v.ctypes.data_as(C.POINTER(C.c_float))
        self.c.n[:] = (ni, nj, nk)
        self.c.min_box[:] = [float(x) for x in min_box]
        self.c.max_box[:] = [float(x) for x in max_box]
        self.c.dxyz[:] = [float(x) for x in dxyz]

    @classmethod
    def from_grid(cls, g):
        return cls(g.values, g.min_box, g.max_box, g.dxyz)


def query_sdf(grid: GridRef, x, outside_accurate=False) -> np.ndarray:
    x = _f64(x).reshape(-1, 3)
    out = np.empty(len(x))
    lib().oracle_query_sdf(C.byref(grid.c), _dp(x), len(x), int(outside_accurate), _dp(out))
    return out


def query_dsdf(grid: GridRef, x, mode=MODE_REFERENCE):
    x = _f64(x).reshape(-1, 3)
    g = np.empty((len(x), 3))
    h = np.empty((len(x), 3, 3))
    lib().oracle_query_dsdf(C.byref(grid.c), _dp(x), len(x), mode, _dp(g), _dp(h))
    return g, h


def distance(grid: GridRef, E2, col_th, xw) -> np.ndarray:
    xw = _f64(xw).reshape(-1, 3)
    out = np.empty(len(xw))
    lib().oracle_distance(C.byref(grid.c), _dp(_f64(E2)), float(col_th), _dp(xw), len(xw), _dp(out))
    return out


COLLISION_DERIV_LAYOUT = [
    ("dd_dxw", (3,)), ("dd_dq2", (6,)), ("dn_dxw", (3, 3)), ("dn_dq2", (3, 6)),
    ("dddot_dxw", (3,)), ("dddot_dxwdot", (3,)), ("dddot_dq2", (6,)), ("dddot_dphi2", (6,)),
    ("dtdot_dxw", (3, 3)), ("dtdot_dxwdot", (3, 3)), ("dtdot_dq2", (3, 6)), ("dtdot_dphi2", (3, 6)),
    ("dxi2_dxw", (3, 3)), ("dxi2_dq2", (3, 6)),
]


def collision(grid: GridRef, E2, phi2, col_th, xw, xw_dot, mode=MODE_REFERENCE, derivs=False):
    """BodySDFObj::collision for n points. Returns dict of arrays."""
    xw = _f64(xw).reshape(-1, 3)
    xd = _f64(xw_dot).reshape(-1, 3)
    n = len(xw)
    vals = np.empty((n, 11))
    der = np.empty((n, 135)) if derivs else None
    lib().oracle_collision(C.byref(grid.c), _dp(_f64(E2)), _dp(_f64(phi2)), float(col_th), mode, _dp(xw), _dp(xd), n,
                           _dp(vals), _dp(der) if derivs else None)
    out = {"d": vals[:, 0], "n": vals[:, 1:4], "ddot": vals[:, 4], "tdot": vals[:, 5:8], "xi2": vals[:, 8:11]}
    if derivs:
        off = 0
        for name, shape in COLLISION_DERIV_LAYOUT:
            cnt = int(np.prod(shape))
            out[name] = der[:, off:off + cnt].reshape((n,) + shape)
            off += cnt
    return out


def make_instance(grid_index, points, E1, phi1, E2, phi2, col_th, kn, kt, mu, damping, scale=1.0, contact_scale=1.0,
                  mode=MODE_REFERENCE):
    inst = OracleInstance()
    pts = _f64(points).reshape(-1, 3)
    inst.grid = grid_index
    inst.mode = mode
    inst.points = _dp(pts)
    inst.n_points = len(pts)
    inst.E1[:] = list(_f64(E1).ravel())
    inst.phi1[:] = list(_f64(phi1).ravel())
    inst.E2[:] = list(_f64(E2).ravel())
    inst.phi2[:] = list(_f64(phi2).ravel())
    inst.col_th, inst.kn, inst.kt, inst.mu, inst.damping = col_th, kn, kt, mu, damping
    inst.scale, inst.contact_scale = scale, contact_scale
    inst._keep = pts  # keep the point buffer alive
    return inst


def eval_instance(grids, inst: OracleInstance, flags=0):
    """One force object: detection + force (+ Km, Dm, p6 column). Returns dict."""
    garr = (OracleGrid * len(grids))(*[g.c for g in grids])
    fm = np.zeros(12)
    Km = np.zeros(144)
    Dm = np.zeros(144)
    p6 = np.zeros(12)
    act = np.zeros(inst.n_points, dtype=np.uint8)
    na = C.c_int64(0)
    scale = np.zeros(3)
    lib().oracle_eval_instance(garr, C.byref(inst), flags, _dp(fm), _dp(Km), _dp(Dm), _dp(p6),
                               act.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(na), _dp(scale))
    # scale: sum over contacts of ||fm_c||_2, ||Km_c||_F, ||Dm_c||_F (per-contact contributions)
    return {"fm": fm, "Km": Km.reshape(12, 12), "Dm": Dm.reshape(12, 12), "dfm_dp6": p6, "active": act.astype(bool),
            "n_active": int(na.value), "scale_fm": scale[0], "scale_Km": scale[1], "scale_Dm": scale[2]}


def eval_batch(grids, instances, flags=0, threads=0):
    garr = (OracleGrid * len(grids))(*[g.c for g in grids])
    n = len(instances)
    iarr = (OracleInstance * n)(*instances)
    fm = np.zeros((n, 12))
    Km = np.zeros((n, 144)) if flags & FLAG_DERIV else None
    Dm = np.zeros((n, 144)) if flags & FLAG_DERIV else None
    p6 = np.zeros((n, 12)) if flags & FLAG_P6 else None
    na = np.zeros(n, dtype=np.int64)
    scale = np.zeros((n, 3))
    lib().oracle_eval_batch(garr, iarr, n, flags, threads, _dp(fm), _dp(Km) if Km is not None else None,
                            _dp(Dm) if Dm is not None else None, _dp(p6) if p6 is not None else None,
                            na.ctypes.data_as(C.POINTER(C.c_int64)), _dp(scale))
    # scale_*: per instance, the sum over its contacts of ||fm_c||_2, ||Km_c||_F, ||Dm_c||_F
    out = {"fm": fm, "n_active": na, "scale_fm": scale[:, 0], "scale_Km": scale[:, 1], "scale_Dm": scale[:, 2]}
    if Km is not None:
        out["Km"] = Km.reshape(n, 12, 12)
        out["Dm"] = Dm.reshape(n, 12, 12)
    if p6 is not None:
        out["dfm_dp6"] = p6
    return out


def force_fixed_contacts(grid: GridRef, xi, E1, phi1, E2, phi2, col_th, kn, kt, mu, damping, scale=1.0,
                         contact_scale=1.0, mode=MODE_REFERENCE, derivs=False):
    """Force over a fixed contact list (no detection), as the reference FD self-test uses."""
    xi = _f64(xi).reshape(-1, 3)
    fm = np.zeros(12)
    Km = np.zeros(144)
    Dm = np.zeros(144)
    p6 = np.zeros(12)
    lib().oracle_force_fixed_contacts(C.byref(grid.c), _dp(xi), len(xi), _dp(_f64(E1)), _dp(_f64(phi1)), _dp(_f64(E2)),
                                      _dp(_f64(phi2)), float(col_th), mode, kn, kt, mu, damping, scale, contact_scale,
                                      1 if derivs else 0, _dp(fm), _dp(Km), _dp(Dm), _dp(p6))
    return {"fm": fm, "Km": Km.reshape(12, 12), "Dm": Dm.reshape(12, 12), "dfm_dp6": p6}


def se3_exp(q) -> np.ndarray:
    E = np.zeros(16)
    lib().oracle_se3_exp(_dp(_f64(q)), _dp(E))
    return E.reshape(4, 4)


def print_error_metric(a, b):
    """print_error's metric (Utils.h:362-381): (error, error_rel) with
    error = ||a-b||, error_rel = error / max(min(||a||, ||b||), 1e-4)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    err = float(np.linalg.norm(a - b))
    rel = err / max(min(float(np.linalg.norm(a)), float(np.linalg.norm(b))), 1e-4)
    return err, rel


def make_level_set3(x, tri, origin, d, n, exact_band=1, with_closest=False):
    """sdfgen::make_level_set3 restated: x (nv x 3 FP32 vertices), tri (nf x 3),
    origin and spacing d (FP32), n = (ni, nj, nk). Returns phi (nk, nj, ni) FP32
    (and the closest triangle per node)."""
    xv = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    tv = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1, 3)
    ni, nj, nk = (int(a) for a in n)
    phi = np.empty((nk, nj, ni), dtype=np.float32)
    cl = np.empty((nk, nj, ni), dtype=np.int32)
    o = (C.c_float * 3)(*[float(np.float32(a)) for a in origin])
    rc = lib().oracle_make_level_set3(xv.ctypes.data_as(C.c_void_p), len(xv), tv.ctypes.data_as(C.c_void_p), len(tv), o,
                                      float(np.float32(d[0])), float(np.float32(d[1])), float(np.float32(d[2])),
                                      ni, nj, nk, int(exact_band), phi.ctypes.data_as(C.c_void_p),
                                      cl.ctypes.data_as(C.c_void_p))
    if rc:
        raise ValueError("oracle_make_level_set3: bad arguments")
    return (phi, cl) if with_closest else phi


def sdf_mesh_grid(V, dx, min_size):
    """precompute_SDF's grid placement: V (nv x 3 FP64). Returns (min_box FP32,
    max_box FP32, spacing FP64, (ni, nj, nk))."""
    v = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
    mn = (C.c_float * 3)()
    mx = (C.c_float * 3)()
    d = np.zeros(3)
    n = (C.c_int * 3)()
    if lib().oracle_sdf_mesh_grid(_dp(v), len(v), float(dx), int(min_size), mn, mx, _dp(d), n):
        raise ValueError("oracle_sdf_mesh_grid: bad arguments")
    return np.array(mn[:], dtype=np.float32), np.array(mx[:], dtype=np.float32), d, tuple(int(a) for a in n)


# ---------------------------------------------------------------- BVH contact variant
def _mesh_args(V, F):
    v = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(F, dtype=np.int64).reshape(-1, 3)
    return v, f


def mesh_normals(V, F):
    """(face normals nf x 3, edge index of the edge opposite each corner nf x 3,
    edge normals ne x 3, angle-weighted vertex normals nv x 3)."""
    v, f = _mesh_args(V, F)
    FN = np.zeros((len(f), 3))
    FE = np.zeros((len(f), 3), dtype=np.int64)
    VN = np.zeros((len(v), 3))
    ne = C.c_int64(0)
    EN = np.zeros((3 * len(f), 3))
    lib().oracle_mesh_normals(v.ctypes.data_as(C.c_void_p), len(v), f.ctypes.data_as(C.c_void_p), len(f),
                              FN.ctypes.data_as(C.c_void_p), FE.ctypes.data_as(C.c_void_p),
                              EN.ctypes.data_as(C.c_void_p), VN.ctypes.data_as(C.c_void_p), C.byref(ne))
    return FN, FE, EN[:ne.value], VN


def mesh_distance_normal(V, F, xw, E2=None, col_th=0.0):
    """BodyBVHObj::distance_normal_parallel by brute force: (d, normal, face)."""
    v, f = _mesh_args(V, F)
    x = np.ascontiguousarray(xw, dtype=np.float64).reshape(-1, 3)
    E = _f64(np.eye(4) if E2 is None else E2).reshape(4, 4)
    n = len(x)
    d = np.zeros(n)
    nrm = np.zeros((n, 3))
    face = np.zeros(n, dtype=np.int64)
    lib().oracle_mesh_distance_normal(v.ctypes.data_as(C.c_void_p), len(v), f.ctypes.data_as(C.c_void_p), len(f), _dp(E),
                                      float(col_th), x.ctypes.data_as(C.c_void_p), n, d.ctypes.data_as(C.c_void_p),
                                      nrm.ctypes.data_as(C.c_void_p), face.ctypes.data_as(C.c_void_p))
    return d, nrm, face


def bvh_eval(V, F, inst: OracleInstance, flags=0):
    """One ForceGeneralBVHContact evaluation (flags: bit0 Km/Dm, bit1 p6 column)."""
    v, f = _mesh_args(V, F)
    fm = np.zeros(12)
    Km = np.zeros(144)
    Dm = np.zeros(144)
    p6 = np.zeros(12)
    act = np.zeros(inst.n_points, dtype=np.uint8)
    na = C.c_int64(0)
    lib().oracle_bvh_eval(v.ctypes.data_as(C.c_void_p), len(v), f.ctypes.data_as(C.c_void_p), len(f), C.byref(inst), flags,
                          _dp(fm), _dp(Km), _dp(Dm), _dp(p6), act.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(na))
    return {"fm": fm, "Km": Km.reshape(12, 12), "Dm": Dm.reshape(12, 12), "dfm_dp6": p6, "active": act.astype(bool),
            "n_active": int(na.value)}


"""Mesh -> SDF grid and distance queries: the reference's SDFMesh
(simulation/redmax/Mesh.cpp:41-198; python_interface.cpp:283-292 binds
SDFMesh(V, F, dx, load_path, save_path), .distance(X), .min_distance(X)) and the grid
construction of BodySDFObj::precompute_SDF (Body/BodySDFObj.cpp:38-77).

The grid is built on the GPU by atac_sdf_from_mesh (the level-set construction of
sdfgen::make_level_set3, node for node the same values) on the grid
atac_sdf_mesh_grid places around the mesh; queries run through atac_sdf_distance.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L
from ._lib import check, lib


def mesh_grid(V, dx, min_size=20):
    """(min_box FP32, max_box FP32, spacing FP64, node counts) that precompute_SDF
    places around the vertices V (3 x nv or nv x 3)."""
    v = _vertices(V)
    mn = (C.c_float * 3)()
    mx = (C.c_float * 3)()
    d = np.zeros(3)
    n = (C.c_int32 * 3)()
    check(lib().atac_sdf_mesh_grid(v.ctypes.data_as(C.c_void_p), len(v), float(dx), int(min_size), mn, mx, L.dptr(d), n))
    return (np.array(mn[:], dtype=np.float32), np.array(mx[:], dtype=np.float32), d, tuple(int(a) for a in n))


def level_set(V, F, origin, d, n, exact_band=1):
    """Signed distance grid (FP32, shape (nk, nj, ni)) of the closed triangle mesh
    (V: vertices, F: triangles, either 3 x m or m x 3), grid node (0,0,0) at
    `origin`, FP32 spacing d and node counts n = (ni, nj, nk)."""
    v = _vertices(V).astype(np.float32)
    f = _faces(F)
    o = (C.c_float * 3)(*[float(np.float32(a)) for a in origin])
    ni, nj, nk = (int(a) for a in n)
    phi = np.empty((nk, nj, ni), dtype=np.float32)
    check(lib().atac_sdf_from_mesh(v.ctypes.data_as(C.c_void_p), len(v), f.ctypes.data_as(C.c_void_p), len(f), o,
                                   float(np.float32(d[0])), float(np.float32(d[1])), float(np.float32(d[2])), ni, nj, nk,
                                   int(exact_band), phi.ctypes.data_as(C.c_void_p)))
    return phi


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
    if f.size and f.min() < 0:
        raise ValueError("negative vertex index")
    return np.ascontiguousarray(f, dtype=np.uint32)


class SDFMesh:
    """SDF of a triangle mesh with trilinear distance queries (Mesh.cpp:41-198).

    V: 3 x nv vertices (Eigen's column layout) or nv x 3; F: triangles likewise.
    The grid is loaded from `load_path` when that file exists (the .sdf text format
    of SDFMesh::load_SDF), else built on the GPU (spacing min(dx, bbox / 20) per
    axis, one cell of padding) and written to `save_path` when given."""

    MIN_SIZE = 20   # Mesh.cpp:75

    def __init__(self, V, F, dx, load_path="", save_path=""):
        self.V = _vertices(V)
        self.F = _faces(F)
        loaded = bool(load_path) and self.load_SDF(load_path)
        if not loaded:
            self.precompute_SDF(dx)
            if save_path:
                self.save_SDF(save_path)
        self._dev = None

    def precompute_SDF(self, dx):
        mn, mx, d, n = mesh_grid(self.V, dx, self.MIN_SIZE)
        self.values = level_set(self.V, self.F, mn, d.astype(np.float32), n)
        self.min_box = mn.astype(np.float64)
        self.max_box = mx.astype(np.float64)
        self.dxyz = d
        self._dev = None

    def load_SDF(self, path):
        """The .sdf text format: ni nj nk / min_box / max_box / dx dy dz / values (x fastest)."""
        if not os.path.exists(path):
            return False
        with open(path) as f:
            tok = f.read().split()
        ni, nj, nk = (int(t) for t in tok[:3])
        nums = np.array([float(t) for t in tok[3:]], dtype=np.float64)
        self.min_box, self.max_box, self.dxyz = nums[0:3].copy(), nums[3:6].copy(), nums[6:9].copy()
        vals = nums[9:9 + ni * nj * nk]
        if len(vals) != ni * nj * nk:
            raise ValueError(f"{path}: expected {ni * nj * nk} values, found {len(vals)}")
        self.values = vals.astype(np.float32).reshape(nk, nj, ni)
        self._dev = None
        return True

    def save_SDF(self, path):
        """Written as SDFMesh::save_SDF does (Mesh.cpp:123-136): std::ostream's default
        6-significant-digit format, so a saved grid reads back rounded to that."""
        nk, nj, ni = self.values.shape
        g = lambda a: format(float(a), ".6g")
        with open(path, "w") as f:
            f.write(f"{ni} {nj} {nk}\n")
            f.write(" ".join(g(a) for a in self.min_box) + "\n")
            f.write(" ".join(g(a) for a in self.max_box) + "\n")
            f.write(" ".join(g(a) for a in self.dxyz) + "\n")
            f.write("\n".join(g(v) for v in self.values.reshape(-1)) + "\n")

    def clear_saved_SDF(self, path):
        if os.path.exists(path):
            os.remove(path)

    def _grid(self):
        if self._dev is None:
            from .batch import DeviceGrid
            self._dev = DeviceGrid(self.values, self.min_box, self.max_box, self.dxyz)
        return self._dev

    def distance(self, X):
        """query_SDF at each point (3 x n or n x 3): the trilinear value inside the
        box, FLT_MAX outside (Mesh.cpp:143-198)."""
        x = _vertices(X) if np.asarray(X).ndim == 2 else np.asarray(X, dtype=np.float64).reshape(1, 3)
        out = np.empty(len(x))
        eye = np.eye(4)
        check(lib().atac_sdf_distance_host(self._grid().handle, L.dptr(eye), 0.0, x.ctypes.data_as(C.c_void_p), len(x),
                                           out.ctypes.data_as(C.c_void_p)))
        return out

    def min_distance(self, X):
        return float(self.distance(X).min())

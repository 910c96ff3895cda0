"""TEST ONLY: ctypes front end of libhostcheck.so, the contact library's device
arithmetic compiled for the host (see hostcheck.hip)."""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libhostcheck.so")
_REPO = os.path.dirname(os.path.dirname(_HERE))


def build():
    cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I" + os.path.join(_REPO, "assemble-them-all_amd", "csrc"), "-I" + os.path.join(_REPO, "include"),
           os.path.join(_HERE, "hostcheck.hip"), "-o", _LIB]
    r = subprocess.run(cmd, stderr=subprocess.PIPE, text=True)
    if r.returncode != 0:
        errs = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln]
        raise RuntimeError("building libhostcheck.so failed:\n" + "\n".join(errs[-40:]))


_lib = None


def _stale():
    if not os.path.exists(_LIB):
        return True
    csrc = os.path.join(_REPO, "assemble-them-all_amd", "csrc")
    srcs = [os.path.join(_HERE, "hostcheck.hip"), os.path.join(_REPO, "include", "atac.h")]
    srcs += [os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith(".h")]
    return max(os.path.getmtime(f) for f in srcs) > os.path.getmtime(_LIB)


def lib():
    global _lib
    if _lib is None:
        if _stale():
            build()
        _lib = C.CDLL(_LIB)
    return _lib


def _grid_args(g):
    vals = np.ascontiguousarray(g.values, dtype=np.float32)
    nk, nj, ni = vals.shape
    n = np.array([ni, nj, nk], dtype=np.int32)
    return vals, n, np.ascontiguousarray(g.min_box, dtype=np.float64), np.ascontiguousarray(g.max_box, dtype=np.float64), \
        np.ascontiguousarray(g.dxyz, dtype=np.float64)


def _vp(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def eval_instance(grid, pts_f32, inst, deriv=False, precise=False, g_fm=None):
    """Host run of the device math for one force object. inst: atac._lib.Instance."""
    vals, n, mn, mx, d = _grid_args(grid)
    p = np.ascontiguousarray(pts_f32, dtype=np.float32).reshape(-1, 3)
    px, py, pz = [np.ascontiguousarray(p[:, k]) for k in range(3)]
    npts = len(p)
    fm = np.zeros(12); p6 = np.zeros(12)
    Km = np.zeros(144) if deriv else None
    Dm = np.zeros(144) if deriv else None
    na = C.c_int64(0)
    act = np.zeros(npts, dtype=np.uint8)
    slow = np.zeros(npts, dtype=np.uint8)
    g = np.ascontiguousarray(g_fm, dtype=np.float64) if g_fm is not None else None
    vjp = np.zeros(26) if g is not None else None
    flags = 8 if precise else 0
    lib().hc_eval(_vp(vals), _vp(n), _vp(mn), _vp(mx), _vp(d), _vp(px), _vp(py), _vp(pz), C.c_int64(npts), C.byref(inst),
                  C.c_uint32(flags), _vp(fm), _vp(p6), _vp(Km), _vp(Dm), C.byref(na), _vp(act), _vp(slow), _vp(g), _vp(vjp))
    out = {"fm": fm, "dfm_dp6": p6, "n_active": int(na.value), "active": act.astype(bool), "slow": slow.astype(bool)}
    if deriv:
        out["Km"] = Km.reshape(12, 12)
        out["Dm"] = Dm.reshape(12, 12)
    if vjp is not None:
        out["vjp"] = vjp
    return out


def collision(grid, E2, phi2, col_th, xw, xw_dot, precise=False, derivs=True):
    """Host run of collision_kernel: values (n, 11) and derivative values (n, 135)."""
    vals, n, mn, mx, d = _grid_args(grid)
    x = np.ascontiguousarray(xw, dtype=np.float64).reshape(-1, 3)
    xd = np.ascontiguousarray(xw_dot, dtype=np.float64).reshape(-1, 3)
    E = np.ascontiguousarray(E2, dtype=np.float64)
    ph = np.ascontiguousarray(phi2, dtype=np.float64)
    out = np.zeros((len(x), 11))
    der = np.zeros((len(x), 135)) if derivs else None
    lib().hc_collision(_vp(vals), _vp(n), _vp(mn), _vp(mx), _vp(d), _vp(E), _vp(ph), C.c_double(col_th),
                       C.c_int(1 if precise else 0), _vp(x), _vp(xd), C.c_int64(len(x)), _vp(out), _vp(der))
    return out, der


def detect_values(grid, pts_f32, inst):
    vals, n, mn, mx, d = _grid_args(grid)
    p = np.ascontiguousarray(pts_f32, dtype=np.float32).reshape(-1, 3)
    px, py, pz = [np.ascontiguousarray(p[:, k]) for k in range(3)]
    out = np.zeros(len(p), dtype=np.float32)
    lib().hc_detect(_vp(vals), _vp(n), _vp(mn), _vp(mx), _vp(d), _vp(px), _vp(py), _vp(pz), C.c_int64(len(p)),
                    C.byref(inst), _vp(out))
    return out


def detect_decisions(grid, pts_f32, inst):
    """(bounded decision, full-interpolation decision, settled-by-block-bound, the
    4-point batched decision, its slow-path flag, that flag recomputed point by point
    from fast_cell_test) per point."""
    vals, n, mn, mx, d = _grid_args(grid)
    p = np.ascontiguousarray(pts_f32, dtype=np.float32).reshape(-1, 3)
    px, py, pz = [np.ascontiguousarray(p[:, k]) for k in range(3)]
    dec = np.zeros(len(p), dtype=np.uint8)
    full = np.zeros(len(p), dtype=np.uint8)
    settled = np.zeros(len(p), dtype=np.uint8)
    batched = np.zeros(len(p), dtype=np.uint8)
    slow = np.zeros(len(p), dtype=np.uint8)
    slow_ref = np.zeros(len(p), dtype=np.uint8)
    lib().hc_detect_decisions(_vp(vals), _vp(n), _vp(mn), _vp(mx), _vp(d), _vp(px), _vp(py), _vp(pz),
                              C.c_int64(len(p)), C.byref(inst), _vp(dec), _vp(full), _vp(settled), _vp(batched),
                              _vp(slow), _vp(slow_ref))
    return (dec.astype(bool), full.astype(bool), settled.astype(bool), batched.astype(bool), slow.astype(bool),
            slow_ref.astype(bool))


def group_cull(grid, pts_f32, inst, gsize=256):
    """Per group of gsize consecutive points: (group_live, a point of it reads a
    non-inactive table entry, the summed-area count equals the brute-force count)."""
    vals, n, mn, mx, d = _grid_args(grid)
    p = np.ascontiguousarray(pts_f32, dtype=np.float32).reshape(-1, 3)
    px, py, pz = [np.ascontiguousarray(p[:, k]) for k in range(3)]
    ng = (len(p) + gsize - 1) // gsize
    live = np.zeros(ng, dtype=np.uint8)
    needed = np.zeros(ng, dtype=np.uint8)
    exact = np.zeros(ng, dtype=np.uint8)
    lib().hc_group_cull(_vp(vals), _vp(n), _vp(mn), _vp(mx), _vp(d), _vp(px), _vp(py), _vp(pz), C.c_int64(len(p)),
                        C.byref(inst), C.c_int(gsize), _vp(live), _vp(needed), _vp(exact))
    return live.astype(bool), needed.astype(bool), exact.astype(bool)

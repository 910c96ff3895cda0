"""ctypes binding of libatac_contact.so (the C ABI declared in include/atac.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libatac_contact.so")

ATAC_OK = 0
ATAC_ERR_INVALID = -1
ATAC_ERR_HIP = -2
ATAC_ERR_NOMEM = -3
ATAC_ERR_STATE = -4

EVAL_FORCE = 0
EVAL_DERIV = 1
EVAL_P6 = 2
EVAL_CHECKPOINT = 4
EVAL_PRECISE = 8
EVAL_ALL_ACTIVE = 16
EVAL_FUSED_ADJOINT = 32

BVH_NOALLOC_MAX_POINTS = 1048576   # ATAC_BVH_NOALLOC_MAX_POINTS

EXPORTED = [
    "atac_last_error", "atac_version", "atac_device_init",
    "atac_grid_create", "atac_grid_destroy",
    "atac_points_create", "atac_points_destroy", "atac_points_permutation",
    "atac_batch_create", "atac_batch_destroy", "atac_batch_set_instances", "atac_batch_retune", "atac_batch_info",
    "atac_batch_eval", "atac_batch_vjp", "atac_batch_eval_vjp", "atac_batch_graph_capture", "atac_batch_graph_launch",
    "atac_batch_active_ids", "atac_batch_touched_values", "atac_eval_plain", "atac_eval_plain_vjp", "atac_batch_timing",
    "atac_batch_workspace",
    "atac_batch_live_points",
    "atac_sdf_distance", "atac_sdf_collision",
    "atac_batch_eval_host", "atac_batch_eval_instances_host", "atac_batch_vjp_host", "atac_sdf_distance_host", "atac_sdf_collision_host",
    "atac_sdf_from_mesh", "atac_sdf_mesh_grid",
    "atac_mesh_create", "atac_mesh_destroy", "atac_mesh_distance_normal", "atac_mesh_distance_normal_host",
    "atac_bvh_force_host", "atac_bvh_force_noalloc",
]


class AtacError(RuntimeError):
    """A non-zero status from the contact library (the message is atac_last_error())."""

    def __init__(self, code, msg):
        super().__init__(f"[atac {code}] {msg}")
        self.code = code


class Instance(C.Structure):
    _fields_ = [
        ("grid", C.c_int32), ("points", C.c_int32), ("segment", C.c_int32), ("reserved", C.c_int32),
        ("E1", C.c_double * 16), ("phi1", C.c_double * 6), ("E2", C.c_double * 16), ("phi2", C.c_double * 6),
        ("col_th", C.c_double), ("kn", C.c_double), ("kt", C.c_double), ("mu", C.c_double), ("damping", C.c_double),
        ("scale", C.c_double), ("contact_scale", C.c_double),
    ]


class Tuning(C.Structure):
    _fields_ = [("tile_points", C.c_int32), ("detect_block", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("order", C.c_int32), ("points_per_lane", C.c_int32), ("max_active", C.c_int32),
                ("group_cull", C.c_int32)]


class Outputs(C.Structure):
    _fields_ = [("fm", C.c_void_p), ("Km", C.c_void_p), ("Dm", C.c_void_p), ("dfm_dp6", C.c_void_p),
                ("n_active", C.c_void_p)]


class Gradients(C.Structure):
    _fields_ = [("g_q1", C.c_void_p), ("g_q2", C.c_void_p), ("g_phi1", C.c_void_p), ("g_phi2", C.c_void_p),
                ("g_kn", C.c_void_p), ("g_kt", C.c_void_p)]


_lib = None


def lib():
    """Load the library; raises when it has not been built (there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AtacError(ATAC_ERR_STATE, f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                                        "(make -C assemble-them-all_amd)")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dp = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)
    u32 = C.c_uint32
    L.atac_last_error.restype = C.c_char_p
    L.atac_last_error.argtypes = []
    L.atac_version.argtypes = []
    L.atac_device_init.argtypes = [i32, C.POINTER(i32)]
    L.atac_grid_create.argtypes = [vp, i32, i32, i32, dp, dp, dp, C.POINTER(vp)]
    L.atac_grid_destroy.argtypes = [vp]
    L.atac_points_create.argtypes = [vp, i64, vp, i32, C.c_double, C.POINTER(vp)]
    L.atac_points_destroy.argtypes = [vp]
    L.atac_points_permutation.argtypes = [vp, i32, vp, i64]
    L.atac_batch_create.argtypes = [vp, i32, vp, i32, vp, i32, vp, C.POINTER(vp)]
    L.atac_batch_destroy.argtypes = [vp]
    L.atac_batch_set_instances.argtypes = [vp, vp, i32, vp]
    L.atac_batch_retune.argtypes = [vp, vp]
    L.atac_batch_info.argtypes = [vp, C.POINTER(i64), C.POINTER(i32), C.POINTER(i32), vp]
    L.atac_batch_eval.argtypes = [vp, u32, vp, vp]
    L.atac_batch_vjp.argtypes = [vp, vp, vp, vp]
    L.atac_batch_eval_vjp.argtypes = [vp, u32, vp, vp, vp, vp]
    L.atac_batch_graph_capture.argtypes = [vp, u32, vp, vp, vp]
    L.atac_batch_graph_launch.argtypes = [vp, vp]
    L.atac_batch_active_ids.argtypes = [vp, i32, vp, i64, C.POINTER(i64)]
    L.atac_batch_touched_values.argtypes = [vp, C.POINTER(i64)]
    L.atac_eval_plain.argtypes = [vp, u32, vp, vp]
    L.atac_eval_plain_vjp.argtypes = [vp, u32, vp, vp, vp, vp]
    L.atac_batch_timing.argtypes = [vp, i32, vp, vp]
    L.atac_batch_workspace.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i32)]
    L.atac_batch_live_points.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    L.atac_sdf_distance.argtypes = [vp, dp, C.c_double, vp, i64, vp, vp]
    L.atac_sdf_collision.argtypes = [vp, dp, dp, C.c_double, u32, vp, vp, i64, vp, vp, vp]
    L.atac_batch_eval_host.argtypes = [vp, u32, vp, vp, vp, vp, vp]
    L.atac_batch_eval_instances_host.argtypes = [vp, vp, i32, u32, vp, vp, vp, vp, vp]
    L.atac_batch_vjp_host.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    L.atac_sdf_distance_host.argtypes = [vp, dp, C.c_double, vp, i64, vp]
    L.atac_sdf_collision_host.argtypes = [vp, dp, dp, C.c_double, u32, vp, vp, i64, vp, vp]
    fp = C.POINTER(C.c_float)
    L.atac_sdf_from_mesh.argtypes = [vp, i64, vp, i64, fp, C.c_float, C.c_float, C.c_float, i32, i32, i32, i32, vp]
    L.atac_sdf_mesh_grid.argtypes = [vp, i64, C.c_double, i32, fp, fp, dp, C.POINTER(i32)]
    L.atac_mesh_create.argtypes = [vp, i64, vp, i64, C.POINTER(vp)]
    L.atac_mesh_destroy.argtypes = [vp]
    L.atac_mesh_distance_normal.argtypes = [vp, dp, C.c_double, vp, i64, vp, vp, vp, vp]
    L.atac_mesh_distance_normal_host.argtypes = [vp, dp, C.c_double, vp, i64, vp, vp, vp]
    L.atac_bvh_force_host.argtypes = [vp, C.POINTER(Instance), vp, i64, u32, vp, vp, vp, vp, C.POINTER(i64)]
    L.atac_bvh_force_noalloc.argtypes = [vp, C.POINTER(Instance), vp, i64, u32, vp, vp, vp]
    for name in EXPORTED:
        if name not in ("atac_last_error",):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc):
    if rc != ATAC_OK:
        msg = lib().atac_last_error()
        raise AtacError(rc, msg.decode() if msg else "")
    return rc


def dptr(a):
    """C double* of a contiguous numpy array."""
    return a.ctypes.data_as(C.POINTER(C.c_double))
